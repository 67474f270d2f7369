"""Decoder of Huffman-coded hierarchical JPEG files with sequential DCT frames, for the golden tests.

rust-compress's own decoder wrote the "recorded" decodes of its hierarchical DCT fixtures
(jpeg/examples/make_hier.rs:86-96, checked by jpeg/tests/hierarchical.rs:260-291). Those pictures run through the
same f32 dequantisation + IDCT, fancy upsampling and YCbCr -> RGB conversion as `reconstruct`
(jpeg/src/decode.rs:801-903 against :1312-1343, :1769-1829 and :1680-1697), so reproducing them sample for sample pins
the oracle's `reconstruct` to the reference's arithmetic. This module does the frame bookkeeping; the arithmetic of
each stage is the oracle's:
  * entropy decoding (T.81 Annex F.2.2; J.2.3.1: a differential frame's DC is decoded without prediction) and the
    markers DHP, EXP, SOF0/SOF5, DQT, DHT, SOS (T.81 Annex B, J.1);
  * per frame, oracle.idct_plane, unclamped (decode.rs:836-844): level shift 128 for the first frame, 0 for a
    differential one; the plane is cropped to the component's size (decode.rs:852-856);
  * a differential frame is added to its reference, expanded by `expand` (jpeg/src/hierarchical.rs:47-68, restated
    below) as the preceding EXP marker asks (decode.rs:857-873);
  * at the end every plane is clamped to 0..255, brought to full size by oracle.upsample and converted by
    oracle.ycc_to_rgb (finish_hier, decode.rs:882-903).
Test helper only: arithmetic coding, progressive and lossless frames are not handled.
"""
import numpy as np

from oracle import oracle as O
from jpeg_scan import ZIGZAG, _Bits, _decode_sym, _extend, _huff_table

I32_LO, I32_HI = -(2 ** 31) // 4, (2 ** 31 - 1) // 4   # i32::MIN / 4, i32::MAX / 4 (Rust division truncates)


def expand(p, horizontal, vertical):
    """jpeg/src/hierarchical.rs:47-68: upsample an int plane by two, bilinear with one truncating division (toward
    zero, as Rust's integer division), the last row and column replicated."""
    h, w = p.shape
    nw, nh = (2 * w if horizontal else w), (2 * h if vertical else h)
    out = np.zeros((nh, nw), np.int64)

    def at(x, y):
        return int(p[min(y, h - 1), min(x, w - 1)])

    def tdiv(a, b):
        q = abs(a) // b
        return q if a >= 0 else -q

    for oy in range(nh):
        y, odd_y = (oy // 2, oy % 2 == 1) if vertical else (oy, False)
        for ox in range(nw):
            x, odd_x = (ox // 2, ox % 2 == 1) if horizontal else (ox, False)
            if not odd_x and not odd_y:
                v = at(x, y)
            elif odd_x and not odd_y:
                v = tdiv(at(x, y) + at(x + 1, y), 2)
            elif odd_y and not odd_x:
                v = tdiv(at(x, y) + at(x, y + 1), 2)
            else:
                v = tdiv(at(x, y) + at(x + 1, y) + at(x, y + 1) + at(x + 1, y + 1), 4)
            out[oy, ox] = v
    return out


def _scan(data, pos, frame, sel, dc_t, ac_t):
    """One sequential Huffman scan (interleaved or not) into the frame's components; returns the position after the
    entropy-coded data."""
    comps = frame["comps"]
    by_id = {c["id"]: c for c in comps}
    if len(sel) == 1:
        c = by_id[sel[0][0]]
        cw = -(-frame["width"] * c["h"] // frame["hmax"])
        ch = -(-frame["height"] * c["v"] // frame["vmax"])
        order = [(c, by, bx) for by in range(-(-ch // 8)) for bx in range(-(-cw // 8))]
    else:
        order = []
        for m_y in range(frame["mcus_y"]):
            for m_x in range(frame["mcus_x"]):
                for cid, _, _ in sel:
                    c = by_id[cid]
                    for v in range(c["v"]):
                        for hh in range(c["h"]):
                            order.append((c, m_y * c["v"] + v, m_x * c["h"] + hh))
    tabs = {cid: (dc_t[td], ac_t[ta]) for cid, td, ta in sel}
    pred = {cid: 0 for cid, _, _ in sel}
    br = _Bits(data, pos)
    for c, by, bx in order:
        dct, act = tabs[c["id"]]
        if frame["differential"]:
            pred[c["id"]] = 0
        t = _decode_sym(br, dct)
        pred[c["id"]] += _extend(br.bits(t), t)
        blk = np.zeros(64, np.int64)
        blk[0] = pred[c["id"]]
        k = 1
        while k < 64:
            rs = _decode_sym(br, act)
            r, s = rs >> 4, rs & 15
            if s == 0:
                if r == 15:
                    k += 16
                    continue
                break
            k += r
            blk[k] = _extend(br.bits(s), s)
            k += 1
        off = (by * c["blocks_w"] + bx) * 64
        c["coefs"][off: off + 64] = blk.astype(np.int16)
    return br.p


def _reconstruct_into(frame, hier):
    """decode.rs:810-879 for a DCT frame."""
    center = 0 if frame["differential"] else 128
    for c in frame["comps"]:
        bh = len(c["coefs"]) // 64 // c["blocks_w"]
        plane = O.idct_plane(c["coefs"], c["blocks_w"], bh, c["qtable"], center, I32_LO, I32_HI).astype(np.int64)
        cw = -(-frame["width"] * c["h"] // frame["hmax"])
        ch = -(-frame["height"] * c["v"] // frame["vmax"])
        plane = plane[:ch, :cw]
        idx = next(i for i, d in enumerate(hier["comps"]) if d["id"] == c["id"])
        if frame["differential"]:
            up = expand(hier["planes"][idx], *frame["expand"])
            assert up.shape[0] >= ch and up.shape[1] >= cw and up.shape[0] - ch <= 1 and up.shape[1] - cw <= 1
            plane = up[:ch, :cw] + plane
        hier["planes"][idx] = plane


def decode(data):
    """Decode a hierarchical JPEG of SOF0 + SOF5 Huffman frames to Gray8 (h, w) or RGB8 (h, w, 3)."""
    assert data[0] == 0xFF and data[1] == 0xD8
    pos = 2
    qt, dc_t, ac_t = {}, {}, {}
    hier, frame, exp = None, None, (False, False)
    while True:
        assert data[pos] == 0xFF, hex(pos)
        marker = data[pos + 1]
        if marker == 0xD9:
            break
        length = (data[pos + 2] << 8) | data[pos + 3]
        body = data[pos + 4: pos + 2 + length]
        pos += 2 + length
        if marker == 0xDE:                      # DHP (B.3.2)
            assert body[0] == 8, "8-bit samples only"
            comps = [dict(id=body[6 + 3 * i], h=body[7 + 3 * i] >> 4, v=body[7 + 3 * i] & 15) for i in range(body[5])]
            hier = dict(width=(body[3] << 8) | body[4], height=(body[1] << 8) | body[2], comps=comps,
                        planes=[None] * len(comps))
        elif marker == 0xDF:                    # EXP (B.3.3)
            exp = (body[0] >> 4 == 1, body[0] & 15 == 1)
        elif marker == 0xDB:
            b = body
            while b:
                pq, tq = b[0] >> 4, b[0] & 15
                n = 128 if pq else 64
                raw = [(b[1 + 2 * k] << 8 | b[2 + 2 * k]) if pq else b[1 + k] for k in range(64)]
                t = np.zeros(64, np.int64)
                for k in range(64):
                    t[ZIGZAG[k]] = raw[k]
                assert t.max() <= 255, "16-bit tables are not needed by these fixtures"
                qt[tq] = t.astype(np.uint8)
                b = b[1 + n:]
        elif marker == 0xC4:
            b = body
            while b:
                tc, th = b[0] >> 4, b[0] & 15
                counts = list(b[1:17])
                n = sum(counts)
                (dc_t if tc == 0 else ac_t)[th] = _huff_table(counts, list(b[17:17 + n]))
                b = b[17 + n:]
        elif marker in (0xC0, 0xC1, 0xC5):
            if frame is not None:
                _reconstruct_into(frame, hier)
            assert hier is not None, "not a hierarchical file"
            h, w = (body[1] << 8) | body[2], (body[3] << 8) | body[4]
            comps = [dict(id=body[6 + 3 * i], h=body[7 + 3 * i] >> 4, v=body[7 + 3 * i] & 15, tq=body[8 + 3 * i])
                     for i in range(body[5])]
            hmax, vmax = max(c["h"] for c in comps), max(c["v"] for c in comps)
            mx, my = -(-w // (8 * hmax)), -(-h // (8 * vmax))
            for c in comps:
                c["blocks_w"] = mx * c["h"]
                c["coefs"] = np.zeros(mx * c["h"] * my * c["v"] * 64, np.int16)
                c["qtable"] = None
            frame = dict(width=w, height=h, comps=comps, hmax=hmax, vmax=vmax, mcus_x=mx, mcus_y=my,
                         differential=marker == 0xC5, expand=exp)
            exp = (False, False)
        elif marker == 0xDA:
            sel = [(body[1 + 2 * i], body[2 + 2 * i] >> 4, body[2 + 2 * i] & 15) for i in range(body[0])]
            for cid, _, _ in sel:
                c = next(c for c in frame["comps"] if c["id"] == cid)
                if c["qtable"] is None:         # the table in force when the component's first scan began
                    c["qtable"] = qt[c["tq"]].copy()
            pos = _scan(data, pos, frame, sel, dc_t, ac_t)
        elif 0xC0 <= marker <= 0xCF and marker not in (0xC4, 0xC8, 0xCC):
            raise ValueError(f"frame type {marker:#x} is not handled")
    _reconstruct_into(frame, hier)
    w, h = hier["width"], hier["height"]
    hmax = max(c["h"] for c in hier["comps"])
    vmax = max(c["v"] for c in hier["comps"])
    full = []
    for c, plane in zip(hier["comps"], hier["planes"]):
        cw, ch = -(-w * c["h"] // hmax), -(-h * c["v"] // vmax)
        assert plane.shape[0] >= ch and plane.shape[1] >= cw
        samples = np.clip(plane, 0, 255).astype(np.uint8)
        full.append(O.upsample(samples, cw, ch, hmax // c["h"], vmax // c["v"], w, h))
    if len(full) == 1:
        return full[0]
    assert len(full) == 3
    return O.ycc_to_rgb(*full, rgb_direct=bytes(c["id"] for c in hier["comps"]) == b"RGB")

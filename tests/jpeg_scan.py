"""Minimal baseline (SOF0, Huffman, sequential) JPEG reader for the golden tests.

It recovers the quantised coefficients and tables of the first frame of a JPEG stream, so that the reference's own
fixture files can be compared with build_coefs output coefficient for coefficient. Test helper only; it follows
ITU-T T.81 (Annex F.2.2 decoding, Annex B marker syntax) and handles what rust-compress's baseline writer emits
(jpeg/src/encode.rs:568-600): one interleaved scan, no restart markers.
"""
import numpy as np

ZIGZAG = [0, 1, 8, 16, 9, 2, 3, 10, 17, 24, 32, 25, 18, 11, 4, 5, 12, 19, 26, 33, 40, 48, 41, 34, 27, 20, 13, 6, 7,
          14, 21, 28, 35, 42, 49, 56, 57, 50, 43, 36, 29, 22, 15, 23, 30, 37, 44, 51, 58, 59, 52, 45, 38, 31, 39, 46,
          53, 60, 61, 54, 47, 55, 62, 63]


class _Bits:
    def __init__(self, data, pos):
        self.d = data
        self.p = pos
        self.acc = 0
        self.n = 0

    def bit(self):
        if self.n == 0:
            b = self.d[self.p]
            self.p += 1
            if b == 0xFF:
                nxt = self.d[self.p]
                if nxt == 0x00:
                    self.p += 1
                else:
                    raise ValueError("marker inside entropy data")
            self.acc = b
            self.n = 8
        self.n -= 1
        return (self.acc >> self.n) & 1

    def bits(self, k):
        v = 0
        for _ in range(k):
            v = (v << 1) | self.bit()
        return v


def _huff_table(counts, vals):
    """Canonical code -> symbol map (T.81 Annex C)."""
    table = {}
    code = 0
    k = 0
    for length in range(1, 17):
        for _ in range(counts[length - 1]):
            table[(length, code)] = vals[k]
            k += 1
            code += 1
        code <<= 1
    return table


def _decode_sym(br, table):
    code = 0
    for length in range(1, 17):
        code = (code << 1) | br.bit()
        s = table.get((length, code))
        if s is not None:
            return s
    raise ValueError("bad Huffman code")


def _extend(v, t):
    return v - (1 << t) + 1 if t and v < (1 << (t - 1)) else v


def first_frame(data):
    """Parse segments up to and including the first scan.

    Returns dict(width, height, comps=[dict(id,h,v,tq,blocks_w,blocks_h,coefs)], qtables={slot: natural uint8[64]},
    dqt_raw={slot: zigzag bytes})."""
    assert data[0] == 0xFF and data[1] == 0xD8
    pos = 2
    qt, dqt_raw, dc_t, ac_t = {}, {}, {}, {}
    frame = None
    while pos < len(data):
        assert data[pos] == 0xFF, hex(pos)
        marker = data[pos + 1]
        length = (data[pos + 2] << 8) | data[pos + 3]
        body = data[pos + 4: pos + 2 + length]
        pos += 2 + length
        if marker == 0xDB:
            b = body
            while b:
                pq, tq = b[0] >> 4, b[0] & 15
                assert pq == 0
                raw = bytes(b[1:65])
                dqt_raw[tq] = raw
                t = np.zeros(64, np.uint8)
                for k in range(64):
                    t[ZIGZAG[k]] = raw[k]
                qt[tq] = t
                b = b[65:]
        elif marker in (0xC0, 0xC1):
            h = (body[1] << 8) | body[2]
            w = (body[3] << 8) | body[4]
            comps = []
            for i in range(body[5]):
                cid, hv, tq = body[6 + 3 * i: 9 + 3 * i]
                comps.append(dict(id=cid, h=hv >> 4, v=hv & 15, tq=tq))
            frame = dict(width=w, height=h, comps=comps)
        elif marker == 0xC4:
            b = body
            while b:
                tc, th = b[0] >> 4, b[0] & 15
                counts = list(b[1:17])
                n = sum(counts)
                vals = list(b[17:17 + n])
                (dc_t if tc == 0 else ac_t)[th] = _huff_table(counts, vals)
                b = b[17 + n:]
        elif marker == 0xDA:
            ns = body[0]
            sel = []
            for i in range(ns):
                cid, tt = body[1 + 2 * i], body[2 + 2 * i]
                sel.append((cid, tt >> 4, tt & 15))
            _decode_scan(data, pos, frame, sel, dc_t, ac_t)
            frame["qtables"] = qt
            frame["dqt_raw"] = dqt_raw
            return frame
    raise ValueError("no scan")


def _decode_scan(data, pos, frame, sel, dc_t, ac_t):
    comps = frame["comps"]
    hmax = max(c["h"] for c in comps)
    vmax = max(c["v"] for c in comps)
    w, h = frame["width"], frame["height"]
    by_id = {c["id"]: c for c in comps}
    if len(sel) == 1:
        c = by_id[sel[0][0]]
        cw = -(-w * c["h"] // hmax)
        ch = -(-h * c["v"] // vmax)
        bw, bh = -(-cw // 8), -(-ch // 8)
        c["blocks_w"], c["blocks_h"] = bw, bh
        c["coefs"] = np.zeros(bw * bh * 64, np.int16)
        order = [(c, by, bx) for by in range(bh) for bx in range(bw)]
    else:
        mx = -(-w // (8 * hmax))
        my = -(-h // (8 * vmax))
        for c in comps:
            c["blocks_w"], c["blocks_h"] = mx * c["h"], my * c["v"]
            c["coefs"] = np.zeros(c["blocks_w"] * c["blocks_h"] * 64, np.int16)
        order = []
        for m_y in range(my):
            for m_x in range(mx):
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
        blk = np.zeros(64, np.int32)
        t = _decode_sym(br, dct)
        diff = _extend(br.bits(t), t)
        pred[c["id"]] += diff
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
        c["coefs"][off: off + 64] = blk

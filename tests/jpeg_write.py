"""Minimal baseline JPEG writer for the golden tests: serialises a CoefImage with the Annex K example Huffman
tables (ITU-T T.81 K.3) so that an independent decoder (libjpeg through Pillow) can decode oracle or GPU
coefficients. Test helper only; it is not the reference's optimised-table writer."""
import struct

import numpy as np

from jpeg_scan import ZIGZAG

DC_BITS = [[0, 1, 5, 1, 1, 1, 1, 1, 1, 0, 0, 0, 0, 0, 0, 0], [0, 3, 1, 1, 1, 1, 1, 1, 1, 1, 1, 0, 0, 0, 0, 0]]
DC_VALS = [list(range(12)), list(range(12))]
AC_BITS = [[0, 2, 1, 3, 3, 2, 4, 3, 5, 5, 4, 4, 0, 0, 1, 0x7D], [0, 2, 1, 2, 4, 4, 3, 4, 7, 5, 4, 4, 0, 1, 2, 0x77]]
AC_VALS = [
    bytes.fromhex(
        "01020300041105122131410613516107227114328191a1082342b1c11552d1f02433627282090a161718191a25262728292a3435363738"
        "393a434445464748494a535455565758595a636465666768696a737475767778797a838485868788898a92939495969798999aa2a3a4a5"
        "a6a7a8a9aab2b3b4b5b6b7b8b9bac2c3c4c5c6c7c8c9cad2d3d4d5d6d7d8d9dae1e2e3e4e5e6e7e8e9eaf1f2f3f4f5f6f7f8f9fa"),
    bytes.fromhex(
        "000102031104052131061241510761711322328108144291a1b1c109233352f0156272d10a162434e125f11718191a262728292a353637"
        "38393a434445464748494a535455565758595a636465666768696a737475767778797a82838485868788898a92939495969798999aa2a3"
        "a4a5a6a7a8a9aab2b3b4b5b6b7b8b9bac2c3c4c5c6c7c8c9cad2d3d4d5d6d7d8d9dae2e3e4e5e6e7e8e9eaf2f3f4f5f6f7f8f9fa"),
]


def _codes(bits, vals):
    out = {}
    code = 0
    k = 0
    for length in range(1, 17):
        for _ in range(bits[length - 1]):
            out[vals[k]] = (code, length)
            k += 1
            code += 1
        code <<= 1
    return out


class _W:
    def __init__(self):
        self.out = bytearray()
        self.acc = 0
        self.n = 0

    def put(self, v, k):
        for i in range(k - 1, -1, -1):
            self.acc = (self.acc << 1) | ((v >> i) & 1)
            self.n += 1
            if self.n == 8:
                self.out.append(self.acc)
                if self.acc == 0xFF:
                    self.out.append(0)
                self.acc = 0
                self.n = 0

    def finish(self):
        if self.n:
            self.put((1 << (8 - self.n)) - 1, 8 - self.n)


def _cat(v):
    a = abs(int(v))
    return a.bit_length()


def _bits_of(v, s):
    v = int(v)
    return v if v >= 0 else v + (1 << s) - 1


def _seg(marker, body):
    return bytes([0xFF, marker]) + struct.pack(">H", len(body) + 2) + bytes(body)


def baseline(ci):
    comps = ci.comps
    gray = len(comps) == 1
    hmax = max(c.h for c in comps)
    vmax = max(c.v for c in comps)
    out = bytearray(b"\xff\xd8")
    dqt = bytearray()
    for slot, q in enumerate(ci.qtables):
        if q is not None:
            dqt.append(slot)
            dqt.extend(int(q[n]) for n in ZIGZAG)
    out += _seg(0xDB, dqt)
    sof = bytearray([8]) + struct.pack(">HH", ci.height, ci.width) + bytes([len(comps)])
    for c in comps:
        sof += bytes([c.id, (c.h << 4) | c.v, c.tq])
    out += _seg(0xC0, sof)
    for t in range(1 if gray else 2):
        out += _seg(0xC4, bytes([t]) + bytes(DC_BITS[t]) + bytes(DC_VALS[t]))
        out += _seg(0xC4, bytes([0x10 | t]) + bytes(AC_BITS[t]) + AC_VALS[t])
    sos = bytearray([len(comps)])
    for i, c in enumerate(comps):
        t = min(i, 1)
        sos += bytes([c.id, (t << 4) | t])
    sos += bytes([0, 63, 0])
    out += _seg(0xDA, sos)
    codes = [(_codes(DC_BITS[t], DC_VALS[t]), _codes(AC_BITS[t], list(AC_VALS[t]))) for t in range(2)]
    w = _W()
    pred = [0] * len(comps)
    mx = -(-ci.width // (8 * hmax))
    my = -(-ci.height // (8 * vmax))

    def block(ci_idx, coefs):
        dc_c, ac_c = codes[min(ci_idx, 1)]
        diff = int(coefs[0]) - pred[ci_idx]
        pred[ci_idx] = int(coefs[0])
        s = _cat(diff)
        w.put(*dc_c[s])
        if s:
            w.put(_bits_of(diff, s), s)
        run = 0
        for k in range(1, 64):
            v = int(coefs[k])
            if v == 0:
                run += 1
                continue
            while run > 15:
                w.put(*ac_c[0xF0])
                run -= 16
            s = _cat(v)
            w.put(*ac_c[(run << 4) | s])
            w.put(_bits_of(v, s), s)
            run = 0
        if run:
            w.put(*ac_c[0x00])

    if gray:
        c = comps[0]
        bw = -(-ci.width // 8)
        bh = -(-ci.height // 8)
        for by in range(bh):
            for bx in range(bw):
                off = (by * c.blocks_w + bx) * 64
                block(0, c.coefs[off:off + 64])
    else:
        for m_y in range(my):
            for m_x in range(mx):
                for i, c in enumerate(comps):
                    for v in range(c.v):
                        for hh in range(c.h):
                            off = ((m_y * c.v + v) * c.blocks_w + m_x * c.h + hh) * 64
                            block(i, c.coefs[off:off + 64])
    w.finish()
    out += w.out
    out += b"\xff\xd9"
    return bytes(out)

"""ctypes view of the CPU oracle (oracle/liboracle.so).

TEST INFRASTRUCTURE: this is the checker that the HIP path is compared with. Only tests/, the smoke check in
__graft_entry__.py and the cpu_baseline leg of bench.py import it; the product package never does.

The functions mirror the reference's Rust API on numpy arrays:
  build_coefs(img, quality, subsampling)          jpeg/src/encode.rs:426 (trellis off)
  reconstruct(ci)                                 jpeg/src/decode.rs:334
  ssimulacra2(a, b)                               core/src/ssimulacra2.rs:471
  find_quality(img, target, max_quality, subs)    jpeg/src/target.rs:29 with the CLI's wiring (cli/src/pipeline.rs:207)
Images are numpy uint8 arrays of shape (h, w, 3) for RGB or (h, w) for Gray.
"""
import ctypes
import os
import subprocess
from dataclasses import dataclass

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None

AUTO, S444, S422, S420 = 0, 1, 2, 3


def build():
    """Compile liboracle.so from oracle.c (gcc, -ffp-contract=off), every time it is called."""
    subprocess.run(["make", "-s", "-B", "-C", _HERE], check=True)


def lib():
    """The oracle library; RCG_ORACLE_LIB names another build of oracle.c (the sanitizer test loads its own)."""
    global _LIB
    if _LIB is None:
        path = os.environ.get("RCG_ORACLE_LIB") or os.path.join(_HERE, "liboracle.so")
        if not os.path.exists(path):
            build()
        L = ctypes.CDLL(path)
        P = ctypes.c_void_p
        I = ctypes.c_int
        D = ctypes.c_double
        L.orc_build_coefs.argtypes = [P, I, I, I, I, I, P, P, P, P]
        L.orc_build_coefs_tables.argtypes = [P, I, I, I, I, I, P, P, P, P, P]
        L.orc_scale_base.argtypes = [P, I, P]
        L.orc_reconstruct.argtypes = [I, I, I, P, P, P, P, P, P, I, I, P]
        L.orc_idct_plane.argtypes = [P, I, I, P, ctypes.c_int32, ctypes.c_int32, ctypes.c_int32, P]
        L.orc_idct_plane.restype = None
        L.orc_upsample.argtypes = [P, ctypes.c_size_t, I, I, I, I, I, I, P]
        L.orc_upsample.restype = None
        L.orc_ycc_to_rgb.argtypes = [P, P, P, ctypes.c_size_t, I, P]
        L.orc_ycc_to_rgb.restype = None
        L.orc_cmyk_to_rgb.argtypes = [P, P, P, P, ctypes.c_size_t, I, P]
        L.orc_cmyk_to_rgb.restype = None
        L.orc_ssimulacra2.argtypes = [P, P, I, I, I, P, P, P]
        L.orc_find_quality.argtypes = [P, I, I, I, D, I, I, P, P, P]
        L.orc_bisect_table.argtypes = [P, D, I, P, P]
        L.orc_scale_table.argtypes = [I, P]
        L.orc_fdct_block.argtypes = [P]
        L.orc_idct_block.argtypes = [P]
        L.orc_round_f32.argtypes = [ctypes.c_float]
        L.orc_round_f32.restype = ctypes.c_int32
        L.orc_build_scaled.argtypes = [P, I, I, I, I, I, P, P, P, P, P, P]
        L.orc_build_lengths.argtypes = [P, I, I, P]
        L.orc_build_lengths.restype = None
        L.orc_rdo_quantize_ac.argtypes = [P, ctypes.c_float, P, P, P]
        L.orc_rdo_quantize_ac.restype = None
        L.orc_trellis.argtypes = [I, I, I, P, P, P, P, P, P, P, P, P]
        L.orc_requantize.argtypes = [I, I, I, P, P, P, P, P, P, P, I, I, I, P, P, P, P]
        L.orc_ssim2_pool.argtypes = [P, I]
        L.orc_ssim2_pool.restype = D
        _LIB = L
    return _LIB


def _ptr(a):
    return a.ctypes.data_as(ctypes.c_void_p)


def sampling(channels, quality, subsampling):
    """(hmax, vmax) as build_coefs picks them (jpeg/src/encode.rs:428-439)."""
    if channels == 1:
        return 1, 1
    if subsampling == S420:
        return 2, 2
    if subsampling == S422:
        return 2, 1
    if subsampling == S444:
        return 1, 1
    return (2, 2) if quality < 90 else (1, 1)


@dataclass
class CoefComponent:
    """jpeg/src/encode.rs:56-69"""
    id: int
    h: int
    v: int
    tq: int
    blocks_w: int
    coefs: np.ndarray  # int16, 64 zigzag coefficients per block, raster block order


@dataclass
class CoefImage:
    """jpeg/src/encode.rs:71-83 (metadata omitted: not on the pixel path)."""
    width: int
    height: int
    comps: list
    qtables: list  # natural-order uint8[64] per slot, or None
    adobe_transform: int = -1  # the Adobe APP14 transform, -1 without the marker (encode.rs:82)


def _channels(img):
    return 1 if img.ndim == 2 else img.shape[2]


def build_coefs(img, quality, subsampling=S420, tables=None):
    """tables: optional (luma, chroma) natural-order uint8[64] used instead of the quality's tables."""
    img = np.ascontiguousarray(img, dtype=np.uint8)
    h, w = img.shape[:2]
    ch = _channels(img)
    hmax, vmax = sampling(ch, quality, subsampling)
    pw = -(-w // (8 * hmax)) * 8 * hmax
    ph = -(-h // (8 * vmax)) * 8 * vmax
    ybw, ybh = pw // 8, ph // 8
    cbw, cbh = pw // hmax // 8, ph // vmax // 8
    y = np.zeros(ybw * ybh * 64, np.int16)
    cb = np.zeros(cbw * cbh * 64 if ch == 3 else 64, np.int16)
    cr = np.zeros_like(cb)
    qt = np.zeros(128, np.uint8)
    if tables is None:
        rc = lib().orc_build_coefs(_ptr(img), w, h, ch, quality, subsampling, _ptr(y), _ptr(cb), _ptr(cr), _ptr(qt))
    else:
        qt[:64] = tables[0]
        qt[64:] = tables[1] if tables[1] is not None else tables[0]
        rc = lib().orc_build_coefs_tables(_ptr(img), w, h, ch, quality, subsampling, _ptr(qt[:64]), _ptr(qt[64:]),
                                          _ptr(y), _ptr(cb), _ptr(cr))
    if rc != 0:
        raise ValueError(f"orc_build_coefs failed: {rc}")
    if ch == 1:
        comps = [CoefComponent(1, 1, 1, 0, ybw, y)]
        return CoefImage(w, h, comps, [qt[:64].copy(), None, None, None])
    comps = [CoefComponent(1, hmax, vmax, 0, ybw, y), CoefComponent(2, 1, 1, 1, cbw, cb),
             CoefComponent(3, 1, 1, 1, cbw, cr)]
    return CoefImage(w, h, comps, [qt[:64].copy(), qt[64:].copy(), None, None])


def build_scaled(img, quality, subsampling=S420):
    """build_coefs plus the trellis hand-off values: (CoefImage, [float32 array per component]) where each array
    holds 64 values a[k] = dct[f]*div[f] per block (jpeg/src/encode.rs:206-211), in the coefficients' layout."""
    img = np.ascontiguousarray(img, dtype=np.uint8)
    h, w = img.shape[:2]
    ch = _channels(img)
    hmax, vmax = sampling(ch, quality, subsampling)
    pw = -(-w // (8 * hmax)) * 8 * hmax
    ph = -(-h // (8 * vmax)) * 8 * vmax
    ybw, ybh = pw // 8, ph // 8
    cbw, cbh = pw // hmax // 8, ph // vmax // 8
    y = np.zeros(ybw * ybh * 64, np.int16)
    cb = np.zeros(cbw * cbh * 64 if ch == 3 else 64, np.int16)
    cr = np.zeros_like(cb)
    sy = np.zeros(len(y), np.float32)
    scb = np.zeros(len(cb), np.float32)
    scr = np.zeros(len(cr), np.float32)
    rc = lib().orc_build_scaled(_ptr(img), w, h, ch, quality, subsampling, _ptr(y), _ptr(cb), _ptr(cr), _ptr(sy),
                                _ptr(scb), _ptr(scr))
    if rc != 0:
        raise ValueError(f"orc_build_scaled failed: {rc}")
    qy = scale_table(quality)
    if ch == 1:
        return CoefImage(w, h, [CoefComponent(1, 1, 1, 0, ybw, y)], [qy, None, None, None]), [sy]
    comps = [CoefComponent(1, hmax, vmax, 0, ybw, y), CoefComponent(2, 1, 1, 1, cbw, cb),
             CoefComponent(3, 1, 1, 1, cbw, cr)]
    return CoefImage(w, h, comps, [qy, scale_table(quality), None, None]), [sy, scb, scr]


def build_lengths(freqs, max_len):
    """core/src/util/huffman.rs:10-88"""
    f = np.ascontiguousarray(freqs, dtype=np.uint32)
    out = np.zeros(len(f), np.uint8)
    lib().orc_build_lengths(_ptr(f), len(f), max_len, _ptr(out))
    return out


def rdo_quantize_ac(table, kappa, sizes, a):
    """jpeg/src/rdo.rs: Rdo::new(table, kappa, sizes).quantize_ac(a, out); returns out (positions 1..63 set)."""
    t = np.ascontiguousarray(table, dtype=np.uint8)
    sz = np.ascontiguousarray(sizes, dtype=np.uint8)
    av = np.ascontiguousarray(a, dtype=np.float32)
    out = np.zeros(64, np.int16)
    lib().orc_rdo_quantize_ac(_ptr(t), ctypes.c_float(kappa), _ptr(sz), _ptr(av), _ptr(out))
    return out


def trellis(ci, scaled):
    """The trellis pass of build_coefs (jpeg/src/encode.rs:527-537) over the plain coefficients of `ci` and their
    hand-off values; returns the trellis coefficients per component."""
    n = len(ci.comps)
    ch = np.array([c.h for c in ci.comps], np.int32)
    cv = np.array([c.v for c in ci.comps], np.int32)
    bw = np.array([c.blocks_w for c in ci.comps], np.int32)
    bh = np.array([len(c.coefs) // 64 // c.blocks_w for c in ci.comps], np.int32)
    plain = [np.ascontiguousarray(c.coefs, dtype=np.int16) for c in ci.comps]
    sc = [np.ascontiguousarray(x, dtype=np.float32) for x in scaled]
    outs = [np.zeros_like(p) for p in plain]
    pp = (ctypes.c_void_p * n)(*[x.ctypes.data for x in plain])
    sp = (ctypes.c_void_p * n)(*[x.ctypes.data for x in sc])
    op = (ctypes.c_void_p * n)(*[x.ctypes.data for x in outs])
    qy = np.ascontiguousarray(ci.qtables[0], dtype=np.uint8)
    qc = np.ascontiguousarray(ci.qtables[1] if ci.qtables[1] is not None else ci.qtables[0], dtype=np.uint8)
    lib().orc_trellis(ci.width, ci.height, n, _ptr(ch), _ptr(cv), _ptr(bw), _ptr(bh), ctypes.cast(pp, ctypes.c_void_p),
                      ctypes.cast(sp, ctypes.c_void_p), _ptr(qy), _ptr(qc), ctypes.cast(op, ctypes.c_void_p))
    return outs


def requantize(ci, quality, trellis=False, adobe=None):
    """jpeg/src/encode.rs:107-171: CoefImage::requantize(quality, trellis); returns a new CoefImage that keeps the
    Adobe transform. adobe: the source's Adobe transform (-1: none); by default the one `ci` carries."""
    if adobe is None:
        adobe = getattr(ci, "adobe_transform", -1)
        adobe = -1 if adobe is None else int(adobe)
    n = len(ci.comps)
    i32 = lambda v: np.array(v, np.int32)
    tq = i32([c.tq for c in ci.comps])
    ch, cv = i32([c.h for c in ci.comps]), i32([c.v for c in ci.comps])
    bw = i32([c.blocks_w for c in ci.comps])
    bh = i32([len(c.coefs) // 64 // c.blocks_w for c in ci.comps])
    src = [np.ascontiguousarray(c.coefs, dtype=np.int16) for c in ci.comps]
    outs = [np.zeros_like(x) for x in src]
    qts = [None if q is None else np.ascontiguousarray(q, dtype=np.uint8) for q in list(ci.qtables) + [None] * 4][:4]
    sp = (ctypes.c_void_p * n)(*[x.ctypes.data for x in src])
    op = (ctypes.c_void_p * n)(*[x.ctypes.data for x in outs])
    qp = (ctypes.c_void_p * 4)(*[None if q is None else q.ctypes.data for q in qts])
    tabs = np.zeros((4, 64), np.uint8)
    present = np.zeros(4, np.int32)
    new_tq = np.zeros(n, np.int32)
    lib().orc_requantize(ci.width, ci.height, n, _ptr(tq), _ptr(ch), _ptr(cv), _ptr(bw), _ptr(bh),
                         ctypes.cast(sp, ctypes.c_void_p), ctypes.cast(qp, ctypes.c_void_p), adobe, quality,
                         1 if trellis else 0, ctypes.cast(op, ctypes.c_void_p), _ptr(tabs), _ptr(present),
                         _ptr(new_tq))
    comps = [CoefComponent(c.id, c.h, c.v, int(t), c.blocks_w, o) for c, t, o in zip(ci.comps, new_tq, outs)]
    return CoefImage(ci.width, ci.height, comps, [tabs[s].copy() if present[s] else None for s in range(4)], adobe)


def reconstruct(ci):
    """jpeg/src/decode.rs:334: Gray8 (h, w) for one component, RGB8 (h, w, 3) for three or four. The colour rule
    follows finish (decode.rs:1423-1445): components stored as RGB when the Adobe transform is 0, or when there is no
    Adobe marker and the ids are "RGB"; four components go through cmyk_to_rgb (decode.rs:1454)."""
    n = len(ci.comps)
    ch = np.array([c.h for c in ci.comps], np.int32)
    cv = np.array([c.v for c in ci.comps], np.int32)
    bw = np.array([c.blocks_w for c in ci.comps], np.int32)
    bh = np.array([len(c.coefs) // 64 // c.blocks_w for c in ci.comps], np.int32)
    coefs = [np.ascontiguousarray(c.coefs, dtype=np.int16) for c in ci.comps]
    qts = [np.ascontiguousarray(ci.qtables[c.tq], dtype=np.uint8) for c in ci.comps]
    cp = (ctypes.c_void_p * n)(*[c.ctypes.data for c in coefs])
    qp = (ctypes.c_void_p * n)(*[q.ctypes.data for q in qts])
    ids = bytes(c.id for c in ci.comps)
    adobe = getattr(ci, "adobe_transform", -1)
    adobe = -1 if adobe is None else int(adobe)
    rgb_direct = 1 if adobe == 0 or (adobe < 0 and ids == b"RGB") else 0
    out = np.zeros((ci.height, ci.width) if n == 1 else (ci.height, ci.width, 3), np.uint8)
    rc = lib().orc_reconstruct(ci.width, ci.height, n, _ptr(ch), _ptr(cv), _ptr(bw), _ptr(bh),
                               ctypes.cast(cp, ctypes.c_void_p), ctypes.cast(qp, ctypes.c_void_p), rgb_direct, adobe,
                               _ptr(out))
    if rc != 0:
        raise ValueError(f"orc_reconstruct failed: {rc}")
    return out


def idct_plane(coefs, blocks_w, blocks_h, qtable, center, lo, hi):
    """jpeg/src/decode.rs:1312-1343 idct_planes::<i32>(center, lo, max) for one component: int32 samples
    round_f32(idct(dequantised block)) + center clamped to [lo, hi], shape (blocks_h*8, blocks_w*8)."""
    c = np.ascontiguousarray(coefs, dtype=np.int16)
    q = np.ascontiguousarray(qtable, dtype=np.uint8)
    out = np.zeros((blocks_h * 8, blocks_w * 8), np.int32)
    lib().orc_idct_plane(_ptr(c), blocks_w, blocks_h, _ptr(q), center, lo, hi, _ptr(out))
    return out


def upsample(plane, cw, ch, fx, fy, w, h):
    """jpeg/src/decode.rs:1769-1829: a u8 plane (rows of any stride >= cw) brought to w x h."""
    p = np.ascontiguousarray(plane, dtype=np.uint8)
    out = np.zeros((h, w), np.uint8)
    lib().orc_upsample(_ptr(p), p.shape[1], cw, ch, fx, fy, w, h, _ptr(out))
    return out


def ycc_to_rgb(y, cb, cr, rgb_direct=False):
    """jpeg/src/decode.rs:1427-1445 (and samples_to_image, :1680-1697, for 8 bits): three (h, w) u8 planes to RGB8."""
    planes = [np.ascontiguousarray(p, dtype=np.uint8) for p in (y, cb, cr)]
    out = np.zeros(planes[0].shape + (3,), np.uint8)
    lib().orc_ycc_to_rgb(_ptr(planes[0]), _ptr(planes[1]), _ptr(planes[2]), planes[0].size, 1 if rgb_direct else 0,
                         _ptr(out))
    return out


def cmyk_to_rgb(c, m, y, k, adobe_transform=-1):
    """jpeg/src/decode.rs:1454-1478: four (h, w) u8 planes to RGB8."""
    planes = [np.ascontiguousarray(p, dtype=np.uint8) for p in (c, m, y, k)]
    out = np.zeros(planes[0].shape + (3,), np.uint8)
    lib().orc_cmyk_to_rgb(*[_ptr(p) for p in planes], planes[0].size, adobe_transform, _ptr(out))
    return out


def ssimulacra2(a, b, with_stats=False):
    a = np.ascontiguousarray(a, dtype=np.uint8)
    b = np.ascontiguousarray(b, dtype=np.uint8)
    if a.shape != b.shape:
        raise ValueError("images differ in size")
    h, w = a.shape[:2]
    score = ctypes.c_double()
    stats = np.zeros((6, 3, 6), np.float64)
    ns = ctypes.c_int()
    rc = lib().orc_ssimulacra2(_ptr(a), _ptr(b), w, h, _channels(a), ctypes.byref(score), _ptr(stats),
                               ctypes.byref(ns))
    if rc != 0:
        raise ValueError("SSIMULACRA 2 needs at least 8x8 pixels and a finite score")
    if with_stats:
        return score.value, stats[: ns.value].copy()
    return score.value


def ssim2_pool(stats):
    stats = np.ascontiguousarray(stats, dtype=np.float64)
    return lib().orc_ssim2_pool(_ptr(stats), stats.shape[0])


def find_quality(img, target, max_quality=100, subsampling=S420):
    img = np.ascontiguousarray(img, dtype=np.uint8)
    h, w = img.shape[:2]
    q = ctypes.c_int()
    s = ctypes.c_double()
    steps = ctypes.c_int()
    rc = lib().orc_find_quality(_ptr(img), w, h, _channels(img), float(target), max_quality, subsampling,
                                ctypes.byref(q), ctypes.byref(s), ctypes.byref(steps))
    if rc != 0:
        raise ValueError("search needs an image of at least 8x8 pixels")
    return q.value, s.value, steps.value


def bisect_table(scores, target, max_quality):
    scores = np.ascontiguousarray(scores, dtype=np.float64)
    q = ctypes.c_int()
    s = ctypes.c_double()
    lib().orc_bisect_table(_ptr(scores), float(target), max_quality, ctypes.byref(q), ctypes.byref(s))
    return q.value, s.value


def scale_table(quality, base=None):
    out = np.zeros(64, np.uint8)
    if base is None:
        lib().orc_scale_table(quality, _ptr(out))
    else:
        base = np.ascontiguousarray(base, dtype=np.uint8)
        lib().orc_scale_base(_ptr(base), quality, _ptr(out))
    return out


def fdct(block):
    b = np.ascontiguousarray(block, dtype=np.float32).copy()
    lib().orc_fdct_block(_ptr(b))
    return b


def idct(block):
    b = np.ascontiguousarray(block, dtype=np.float32).copy()
    lib().orc_idct_block(_ptr(b))
    return b


def round_f32(x):
    return lib().orc_round_f32(float(x))

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
    """Compile liboracle.so from oracle.c (gcc, -ffp-contract=off)."""
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "liboracle.so")
        if not os.path.exists(path):
            build()
        L = ctypes.CDLL(path)
        P = ctypes.c_void_p
        I = ctypes.c_int
        D = ctypes.c_double
        L.orc_build_coefs.argtypes = [P, I, I, I, I, I, P, P, P, P]
        L.orc_build_coefs_tables.argtypes = [P, I, I, I, I, I, P, P, P, P, P]
        L.orc_scale_base.argtypes = [P, I, P]
        L.orc_reconstruct.argtypes = [I, I, I, P, P, P, P, P, P, I, P]
        L.orc_ssimulacra2.argtypes = [P, P, I, I, I, P, P, P]
        L.orc_find_quality.argtypes = [P, I, I, I, D, I, I, P, P, P]
        L.orc_bisect_table.argtypes = [P, D, I, P, P]
        L.orc_scale_table.argtypes = [I, P]
        L.orc_fdct_block.argtypes = [P]
        L.orc_idct_block.argtypes = [P]
        L.orc_round_f32.argtypes = [ctypes.c_float]
        L.orc_round_f32.restype = ctypes.c_int32
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


def reconstruct(ci):
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
    rgb_direct = 1 if ids == b"RGB" else 0
    out = np.zeros((ci.height, ci.width) if n == 1 else (ci.height, ci.width, n), np.uint8)
    rc = lib().orc_reconstruct(ci.width, ci.height, n, _ptr(ch), _ptr(cv), _ptr(bw), _ptr(bh),
                               ctypes.cast(cp, ctypes.c_void_p), ctypes.cast(qp, ctypes.c_void_p), rgb_direct, _ptr(out))
    if rc != 0:
        raise ValueError(f"orc_reconstruct failed: {rc}")
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

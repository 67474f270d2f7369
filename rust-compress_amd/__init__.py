"""rust-compress_amd: the JPEG target-quality inner loop of rust-compress on MI355X (gfx950), through librcgpu.so.

This module is the Python view of the C ABI in include/rcgpu.h. It mirrors the reference's Rust API for the path:

  build_coefs(img, quality, subsampling)   rc_jpeg::encode::build_coefs      (jpeg/src/encode.rs:426), trellis off
  reconstruct(ci)                          rc_jpeg::decode::reconstruct      (jpeg/src/decode.rs:334)
  ssimulacra2(reference, distorted)        rc_core::ssimulacra2::ssimulacra2 (core/src/ssimulacra2.rs:471)
  find_quality(img, target, ...)           rc_jpeg::target::find_quality_with as the CLI wires it for
                                           --target-ssimulacra2 (cli/src/pipeline.rs:207-247)

Images are numpy uint8 arrays, (h, w, 3) for RGB8 or (h, w) for Gray8. Errors raise RcgError with the status code
and message of the library (Invalid/Unsupported map rc_core::error::Error). There is no CPU fallback: without the
built library every call raises.
"""
import ctypes
import os
from dataclasses import dataclass, field

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "librcgpu.so")

OK, INVALID, UNSUPPORTED, DEVICE, OOM = 0, 1, 2, 3, 4
AUTO, S444, S422, S420 = 0, 1, 2, 3
FP32, FP64 = 0, 1
SEARCH_PLAIN, SEARCH_TRELLIS, SEARCH_REQUANT = 0, 1, 2

_STATUS = {INVALID: "invalid", UNSUPPORTED: "unsupported", DEVICE: "device", OOM: "out of memory"}


class RcgError(RuntimeError):
    def __init__(self, status, message):
        super().__init__(f"{_STATUS.get(status, status)}: {message}")
        self.status = status


class Layout(ctypes.Structure):
    _fields_ = [("width", ctypes.c_uint32), ("height", ctypes.c_uint32), ("ncomp", ctypes.c_uint32),
                ("hmax", ctypes.c_uint32), ("vmax", ctypes.c_uint32), ("blocks_w", ctypes.c_uint32 * 3),
                ("blocks_h", ctypes.c_uint32 * 3)]


class _CoefComponent(ctypes.Structure):
    _fields_ = [("id", ctypes.c_uint8), ("h", ctypes.c_uint8), ("v", ctypes.c_uint8), ("tq", ctypes.c_uint8),
                ("blocks_w", ctypes.c_uint32), ("blocks_h", ctypes.c_uint32), ("coefs", ctypes.c_void_p)]


class _CoefImage(ctypes.Structure):
    _fields_ = [("width", ctypes.c_uint32), ("height", ctypes.c_uint32), ("ncomp", ctypes.c_uint32),
                ("comps", _CoefComponent * 4), ("qtables", ctypes.c_void_p * 4), ("adobe_transform", ctypes.c_int)]


class _Found(ctypes.Structure):
    _fields_ = [("quality", ctypes.c_int), ("score", ctypes.c_double), ("steps", ctypes.c_int)]


MAX_CAND = 16


class _Requantized(ctypes.Structure):
    _fields_ = [("coefs", ctypes.c_void_p * 4), ("qtables", (ctypes.c_uint8 * 64) * 4), ("present", ctypes.c_int * 4),
                ("tq", ctypes.c_uint8 * 4)]


class _JpegInfo(ctypes.Structure):
    _fields_ = [("width", ctypes.c_uint32), ("height", ctypes.c_uint32), ("ncomp", ctypes.c_uint32),
                ("id", ctypes.c_uint8 * 4), ("h", ctypes.c_uint8 * 4), ("v", ctypes.c_uint8 * 4),
                ("tq", ctypes.c_uint8 * 4), ("blocks_w", ctypes.c_uint32 * 4), ("blocks_h", ctypes.c_uint32 * 4),
                ("qtables", (ctypes.c_uint8 * 64) * 4), ("present", ctypes.c_int * 4), ("adobe_transform", ctypes.c_int),
                ("source_quality", ctypes.c_int), ("coded_rgb", ctypes.c_int), ("progressive", ctypes.c_int),
                ("arithmetic", ctypes.c_int), ("lossless", ctypes.c_int), ("hierarchical", ctypes.c_int),
                ("precision", ctypes.c_int), ("orientation", ctypes.c_int), ("table16", ctypes.c_int)]


class BatchJob(ctypes.Structure):
    """rcg_batch_job: one image and its candidate (quality, subsampling) list."""
    _fields_ = [("pixels", ctypes.c_void_p), ("width", ctypes.c_uint32), ("height", ctypes.c_uint32),
                ("channels", ctypes.c_uint32), ("ncand", ctypes.c_uint32), ("quality", ctypes.c_int * MAX_CAND),
                ("subsampling", ctypes.c_int * MAX_CAND), ("target", ctypes.c_double),
                ("winner", ctypes.c_void_p * 3)]


class BatchResult(ctypes.Structure):
    _fields_ = [("scores", ctypes.c_double * MAX_CAND), ("settled", ctypes.c_int * MAX_CAND),
                ("chosen_index", ctypes.c_int), ("chosen_quality", ctypes.c_int), ("sweep_index", ctypes.c_int),
                ("sweep_quality", ctypes.c_int), ("status", ctypes.c_int), ("winner_written", ctypes.c_int)]


class Tune(ctypes.Structure):
    """rcg_tune: tile policy of the SSIMULACRA2 scale kernel (rows per tile by scale height)."""
    _fields_ = [("wide_nt", ctypes.c_int), ("narrow_w", ctypes.c_uint32), ("nclass", ctypes.c_int),
                ("min_h", ctypes.c_uint32 * 8), ("th", ctypes.c_int * 8)]

    def as_dict(self):
        return {"wide_nt": self.wide_nt, "narrow_w": self.narrow_w,
                "classes": [[int(self.min_h[i]), int(self.th[i])] for i in range(self.nclass)]}


class ServerOpts(ctypes.Structure):
    _fields_ = [("hbm_budget", ctypes.c_size_t), ("max_active", ctypes.c_int), ("speculate", ctypes.c_int),
                ("tune", ctypes.c_void_p)]


class Request(ctypes.Structure):
    _fields_ = [("id", ctypes.c_uint64), ("pixels", ctypes.c_void_p), ("width", ctypes.c_uint32),
                ("height", ctypes.c_uint32), ("channels", ctypes.c_uint32), ("target", ctypes.c_double),
                ("max_quality", ctypes.c_int), ("subsampling", ctypes.c_int), ("coefs", ctypes.c_void_p * 3)]


class Completion(ctypes.Structure):
    _fields_ = [("id", ctypes.c_uint64), ("status", ctypes.c_int), ("quality", ctypes.c_int),
                ("score", ctypes.c_double), ("steps", ctypes.c_int), ("ticks", ctypes.c_int),
                ("submit_s", ctypes.c_double), ("admit_s", ctypes.c_double), ("done_s", ctypes.c_double),
                ("coefs_written", ctypes.c_int)]


class ServerStats(ctypes.Structure):
    _fields_ = [("ticks", ctypes.c_uint64), ("requests_done", ctypes.c_uint64), ("mean_batch", ctypes.c_double),
                ("peak_footprint", ctypes.c_size_t), ("busy_s", ctypes.c_double)]


_LIB = None


def lib():
    """Load librcgpu.so (built by __graft_entry__.build() or rust-compress_amd/build.py)."""
    global _LIB
    if _LIB is None:
        if not os.path.exists(LIB_PATH):
            raise RcgError(DEVICE, f"{LIB_PATH} is not built; run __graft_entry__.build()")
        L = ctypes.CDLL(LIB_PATH)
        P, I, U, D = ctypes.c_void_p, ctypes.c_int, ctypes.c_uint32, ctypes.c_double
        L.rcg_create.argtypes = [I, ctypes.POINTER(P)]
        L.rcg_destroy.argtypes = [P]
        L.rcg_destroy.restype = None
        L.rcg_last_error.restype = ctypes.c_char_p
        L.rcg_version.restype = ctypes.c_char_p
        L.rcg_coef_layout.argtypes = [U, U, U, I, I, ctypes.POINTER(Layout)]
        L.rcg_quant_tables.argtypes = [I, P]
        L.rcg_build_coefs.argtypes = [P, P, U, U, U, I, I, P, P]
        L.rcg_build_coefs_dev.argtypes = [P, P, U, U, U, I, I, P]
        L.rcg_reconstruct.argtypes = [P, ctypes.POINTER(_CoefImage), P]
        L.rcg_reconstruct_dev.argtypes = [P, ctypes.POINTER(_CoefImage), P]
        L.rcg_reconstruct_dev_noalloc.argtypes = [P, ctypes.POINTER(_CoefImage), P]
        L.rcg_ssimulacra2.argtypes = [P, P, P, U, U, U, I, ctypes.POINTER(D), P, ctypes.POINTER(I)]
        L.rcg_ssimulacra2_dev.argtypes = [P, P, P, U, U, U, I, ctypes.POINTER(D), P, ctypes.POINTER(I)]
        L.rcg_find_quality.argtypes = [P, P, U, U, U, D, I, I, ctypes.POINTER(_Found), P, P]
        L.rcg_find_quality_dev.argtypes = [P, P, U, U, U, D, I, I, ctypes.POINTER(_Found)]
        L.rcg_synth_rgb8.argtypes = [ctypes.c_uint64, U, U, P]
        L.rcg_score_batch.argtypes = [P, ctypes.POINTER(BatchJob), I, ctypes.POINTER(BatchResult)]
        L.rcg_score_batch_dev.argtypes = [P, ctypes.POINTER(BatchJob), I, ctypes.POINTER(BatchResult)]
        L.rcg_server_create.argtypes = [I, ctypes.POINTER(ServerOpts), ctypes.POINTER(P)]
        L.rcg_server_submit.argtypes = [P, ctypes.POINTER(Request)]
        L.rcg_server_poll.argtypes = [P, ctypes.POINTER(Completion), I, I, ctypes.POINTER(I)]
        L.rcg_server_get_stats.argtypes = [P, ctypes.POINTER(ServerStats)]
        L.rcg_server_destroy.argtypes = [P]
        L.rcg_server_destroy.restype = None
        L.rcg_device_alloc.argtypes = [P, ctypes.c_size_t, ctypes.POINTER(P)]
        L.rcg_device_free.argtypes = [P, P]
        L.rcg_memcpy_h2d.argtypes = [P, P, P, ctypes.c_size_t]
        L.rcg_memcpy_d2h.argtypes = [P, P, P, ctypes.c_size_t]
        L.rcg_synchronize.argtypes = [P]
        L.rcg_device_copy.argtypes = [P, P, P, ctypes.c_size_t]
        L.rcg_batch_plan_create.argtypes = [P, ctypes.POINTER(BatchJob), I, ctypes.POINTER(P)]
        L.rcg_batch_plan_run.argtypes = [P, ctypes.POINTER(BatchResult)]
        L.rcg_batch_plan_destroy.argtypes = [P]
        L.rcg_batch_plan_destroy.restype = None
        L.rcg_get_tune.argtypes = [P, ctypes.POINTER(Tune)]
        L.rcg_set_tune.argtypes = [P, ctypes.POINTER(Tune)]
        L.rcg_autotune.argtypes = [P, I, ctypes.POINTER(Tune)]
        L.rcg_write_jpeg.argtypes = [ctypes.POINTER(_CoefImage), I, ctypes.POINTER(ctypes.POINTER(ctypes.c_uint8)),
                                     ctypes.POINTER(ctypes.c_size_t)]
        L.rcg_free.argtypes = [P]
        L.rcg_free.restype = None
        L.rcg_huffman_lengths.argtypes = [P, ctypes.c_size_t, I, P]
        L.rcg_build_scaled.argtypes = [P, P, U, U, U, I, I, P, P, P]
        L.rcg_trellis.argtypes = [ctypes.POINTER(_CoefImage), P, P]
        L.rcg_rdo_quantize_ac.argtypes = [P, ctypes.c_float, P, P, P]
        L.rcg_requantize.argtypes = [ctypes.POINTER(_CoefImage), I, I, ctypes.POINTER(_Requantized)]
        L.rcg_requant_find_quality.argtypes = [P, ctypes.POINTER(_CoefImage), D, I, I, ctypes.POINTER(_Found),
                                               ctypes.POINTER(_Requantized)]
        L.rcg_find_quality_trellis.argtypes = [P, P, U, U, U, D, I, I, ctypes.POINTER(_Found), P, P]
        L.rcg_search_bytes.argtypes = [P, U, U, U, I, I, ctypes.POINTER(ctypes.c_size_t)]
        L.rcg_jpeg_parse.argtypes = [P, ctypes.c_size_t, ctypes.POINTER(_JpegInfo)]
        L.rcg_jpeg_read_coefs.argtypes = [P, ctypes.c_size_t, P]
        L.rcg_profile.argtypes = [P, I]
        L.rcg_profile_get.argtypes = [P, ctypes.c_char_p, ctypes.POINTER(D), ctypes.POINTER(I)]
        _LIB = L
    return _LIB


def _check(rc):
    if rc != OK:
        raise RcgError(rc, lib().rcg_last_error().decode(errors="replace"))


def _ptr(a):
    return a.ctypes.data_as(ctypes.c_void_p)


def _channels(img):
    return 1 if img.ndim == 2 else img.shape[2]


def synth_rgb8(seed, width, height):
    """Seeded synthetic RGB8 image (sinusoids + rectangles + Gaussian noise), shape (height, width, 3)."""
    out = np.empty((height, width, 3), np.uint8)
    _check(lib().rcg_synth_rgb8(seed, width, height, _ptr(out)))
    return out


def quant_tables(quality):
    out = np.zeros(128, np.uint8)
    _check(lib().rcg_quant_tables(quality, _ptr(out)))
    return out[:64].copy(), out[64:].copy()


def coef_layout(width, height, channels, quality, subsampling):
    L = Layout()
    _check(lib().rcg_coef_layout(width, height, channels, quality, subsampling, ctypes.byref(L)))
    return L


@dataclass
class CoefComponent:
    """rc_jpeg::encode::CoefComponent (jpeg/src/encode.rs:56-69)."""
    id: int
    h: int
    v: int
    tq: int
    blocks_w: int
    coefs: np.ndarray  # int16, 64 zigzag coefficients per block, raster block order


@dataclass
class CoefImage:
    """rc_jpeg::encode::CoefImage (jpeg/src/encode.rs:71-83) without metadata."""
    width: int
    height: int
    comps: list
    qtables: list  # natural-order uint8[64] per slot, or None
    adobe_transform: int = -1


@dataclass
class Found:
    """rc_jpeg::target::Found (jpeg/src/target.rs:11-16)."""
    quality: int
    score: float
    coefs: CoefImage = None
    steps: int = 0


class Context:
    """One device context: a HIP stream and a device workspace on `device`."""

    def __init__(self, device=0):
        self._h = ctypes.c_void_p()
        _check(lib().rcg_create(device, ctypes.byref(self._h)))

    def close(self):
        if self._h:
            lib().rcg_destroy(self._h)
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    @property
    def handle(self):
        return self._h

    # ------------------------------------------------------------------ host-buffer drop-in API

    def build_coefs(self, img, quality, subsampling=S420):
        img = np.ascontiguousarray(img, dtype=np.uint8)
        h, w = img.shape[:2]
        ch = _channels(img)
        L = coef_layout(w, h, ch, quality, subsampling)
        bufs = [np.empty(L.blocks_w[c] * L.blocks_h[c] * 64, np.int16) for c in range(L.ncomp)]
        ptrs = (ctypes.c_void_p * 3)(*([b.ctypes.data for b in bufs] + [None] * (3 - len(bufs))))
        qt = np.zeros(128, np.uint8)
        _check(lib().rcg_build_coefs(self._h, _ptr(img), w, h, ch, quality, subsampling,
                                     ctypes.cast(ptrs, ctypes.c_void_p), _ptr(qt)))
        if ch == 1:
            return CoefImage(w, h, [CoefComponent(1, 1, 1, 0, L.blocks_w[0], bufs[0])],
                             [qt[:64].copy(), None, None, None])
        comps = [CoefComponent(1, L.hmax, L.vmax, 0, L.blocks_w[0], bufs[0]),
                 CoefComponent(2, 1, 1, 1, L.blocks_w[1], bufs[1]),
                 CoefComponent(3, 1, 1, 1, L.blocks_w[2], bufs[2])]
        return CoefImage(w, h, comps, [qt[:64].copy(), qt[64:].copy(), None, None])

    def build_scaled(self, img, quality, subsampling=S420):
        """build_coefs plus the trellis hand-off values (rcg_build_scaled): (CoefImage, [float32 per component])."""
        img = np.ascontiguousarray(img, dtype=np.uint8)
        h, w = img.shape[:2]
        ch = _channels(img)
        L = coef_layout(w, h, ch, quality, subsampling)
        bufs = [np.empty(L.blocks_w[c] * L.blocks_h[c] * 64, np.int16) for c in range(L.ncomp)]
        scs = [np.empty(L.blocks_w[c] * L.blocks_h[c] * 64, np.float32) for c in range(L.ncomp)]
        ptrs = (ctypes.c_void_p * 3)(*([b.ctypes.data for b in bufs] + [None] * (3 - len(bufs))))
        sptrs = (ctypes.c_void_p * 3)(*([b.ctypes.data for b in scs] + [None] * (3 - len(scs))))
        qt = np.zeros(128, np.uint8)
        _check(lib().rcg_build_scaled(self._h, _ptr(img), w, h, ch, quality, subsampling,
                                      ctypes.cast(ptrs, ctypes.c_void_p), ctypes.cast(sptrs, ctypes.c_void_p),
                                      _ptr(qt)))
        if ch == 1:
            ci = CoefImage(w, h, [CoefComponent(1, 1, 1, 0, L.blocks_w[0], bufs[0])], [qt[:64].copy(), None, None, None])
        else:
            ci = CoefImage(w, h, [CoefComponent(1, L.hmax, L.vmax, 0, L.blocks_w[0], bufs[0]),
                                  CoefComponent(2, 1, 1, 1, L.blocks_w[1], bufs[1]),
                                  CoefComponent(3, 1, 1, 1, L.blocks_w[2], bufs[2])],
                           [qt[:64].copy(), qt[64:].copy(), None, None])
        return ci, scs

    def requant_find_quality(self, ci, target, max_quality=100, trellis=False):
        """The JPEG -> JPEG target search (rcg_requant_find_quality): Found with the winner's CoefImage."""
        keep = []
        s = _coef_image_struct(ci, keep, device=False)
        r, outs = _requantized(ci)
        f = _Found()
        _check(lib().rcg_requant_find_quality(self._h, ctypes.byref(s), target, max_quality, 1 if trellis else 0,
                                              ctypes.byref(f), ctypes.byref(r)))
        return Found(f.quality, f.score, _from_requantized(ci, r, outs), f.steps)

    def search_bytes(self, width, height, channels, subsampling=AUTO, mode=SEARCH_PLAIN):
        """Device bytes one search call takes from this context (rcg_search_bytes), for HBM admission."""
        n = ctypes.c_size_t()
        _check(lib().rcg_search_bytes(self._h, width, height, channels, subsampling, mode, ctypes.byref(n)))
        return n.value

    def reconstruct(self, ci):
        keep = []
        c_ci = _coef_image_struct(ci, keep, device=False)
        n = len(ci.comps)
        out = np.empty((ci.height, ci.width) if n == 1 else (ci.height, ci.width, 3), np.uint8)
        _check(lib().rcg_reconstruct(self._h, ctypes.byref(c_ci), _ptr(out)))
        return out

    def ssimulacra2(self, reference, distorted, precision=FP32, with_stats=False):
        a = np.ascontiguousarray(reference, dtype=np.uint8)
        b = np.ascontiguousarray(distorted, dtype=np.uint8)
        if a.shape != b.shape:
            raise RcgError(INVALID, "images differ in size")
        h, w = a.shape[:2]
        score = ctypes.c_double()
        stats = np.zeros((6, 3, 6), np.float64)
        ns = ctypes.c_int()
        _check(lib().rcg_ssimulacra2(self._h, _ptr(a), _ptr(b), w, h, _channels(a), precision,
                                     ctypes.byref(score), _ptr(stats), ctypes.byref(ns)))
        if with_stats:
            return score.value, stats[: ns.value].copy()
        return score.value

    def find_quality(self, img, target, max_quality=100, subsampling=S420, trellis=False):
        """The target search (rcg_find_quality; rcg_find_quality_trellis with trellis, the CLI default)."""
        img = np.ascontiguousarray(img, dtype=np.uint8)
        h, w = img.shape[:2]
        ch = _channels(img)
        f = _Found()
        # The chosen quality's layout is only known afterwards; size for every layout the search can meet (under
        # Auto both 4:4:4 and 4:2:0, whose 16-pixel MCUs can give luma an extra block row or column).
        Ls = [coef_layout(w, h, ch, q, subsampling) for q in ((100, 1) if subsampling == AUTO else (100,))]
        bufs = [np.zeros(max(L.blocks_w[c] * L.blocks_h[c] for L in Ls) * 64, np.int16) for c in range(Ls[0].ncomp)]
        ptrs = (ctypes.c_void_p * 3)(*([b.ctypes.data for b in bufs] + [None] * (3 - len(bufs))))
        qt = np.zeros(128, np.uint8)
        fn = lib().rcg_find_quality_trellis if trellis else lib().rcg_find_quality
        _check(fn(self._h, _ptr(img), w, h, ch, float(target), max_quality, subsampling, ctypes.byref(f),
                  ctypes.cast(ptrs, ctypes.c_void_p), _ptr(qt)))
        Lq = coef_layout(w, h, ch, f.quality, subsampling)
        if ch == 1:
            comps = [CoefComponent(1, 1, 1, 0, Lq.blocks_w[0], bufs[0][: Lq.blocks_w[0] * Lq.blocks_h[0] * 64])]
            qts = [qt[:64].copy(), None, None, None]
        else:
            comps = [CoefComponent(1 + c, Lq.hmax if c == 0 else 1, Lq.vmax if c == 0 else 1, min(c, 1),
                                   Lq.blocks_w[c], bufs[c][: Lq.blocks_w[c] * Lq.blocks_h[c] * 64]) for c in range(3)]
            qts = [qt[:64].copy(), qt[64:].copy(), None, None]
        return Found(f.quality, f.score, CoefImage(w, h, comps, qts), f.steps)

    # ------------------------------------------------------------------ device buffers

    def upload(self, array):
        """Copy a numpy array into a new device buffer; returns a DeviceBuffer."""
        a = np.ascontiguousarray(array)
        buf = DeviceBuffer(self, a.nbytes)
        _check(lib().rcg_memcpy_h2d(self._h, buf.ptr, a.ctypes.data_as(ctypes.c_void_p), a.nbytes))
        return buf

    def synchronize(self):
        _check(lib().rcg_synchronize(self._h))

    # ------------------------------------------------------------------ batched scoring

    def _batch_jobs(self, images, candidates, targets, device_ptrs, winners=False):
        """The rcg_batch_job array; with winners, also per job the buffers for the winner's coefficients (numpy
        arrays for host images, DeviceBuffers for device ones), sized for every candidate's grid."""
        n = len(images) if device_ptrs is None else len(device_ptrs)
        jobs = (BatchJob * n)()
        keep = []
        wbufs = []
        per_image = isinstance(candidates[0], (list, tuple)) and isinstance(candidates[0][0], (list, tuple))
        for i in range(n):
            cand = candidates[i] if per_image else candidates
            if device_ptrs is None:
                img = np.ascontiguousarray(images[i], dtype=np.uint8)
                keep.append(img)
                jobs[i].pixels = img.ctypes.data
                jobs[i].height, jobs[i].width = img.shape[:2]
                jobs[i].channels = _channels(img)
            else:
                ptr, w, h, ch = device_ptrs[i]
                jobs[i].pixels = ptr
                jobs[i].width, jobs[i].height, jobs[i].channels = w, h, ch
            jobs[i].ncand = len(cand)
            for k, (q, sub) in enumerate(cand):
                jobs[i].quality[k] = q
                jobs[i].subsampling[k] = sub
            jobs[i].target = float(targets[i] if isinstance(targets, (list, tuple, np.ndarray)) else targets)
            if winners:
                w, h, ch = jobs[i].width, jobs[i].height, jobs[i].channels
                sizes = _max_grid(w, h, ch, [q for q, _ in cand], [s for _, s in cand])
                bufs = []
                for c, nc in enumerate(sizes):
                    if device_ptrs is None:
                        b = np.zeros(nc, np.int16)
                        jobs[i].winner[c] = b.ctypes.data
                    else:
                        b = DeviceBuffer(self, max(nc, 1) * 2)
                        jobs[i].winner[c] = b.ptr
                    bufs.append(b)
                wbufs.append(bufs)
        return (jobs, keep, wbufs) if winners else (jobs, keep)

    def batch_plan(self, device_ptrs, candidates, targets, winners=False):
        """Capture the batched scoring of device-resident images as a hipGraph (see BatchPlan)."""
        if winners:
            jobs, _, wbufs = self._batch_jobs(None, candidates, targets, device_ptrs, winners=True)
            return BatchPlan(self, jobs, wbufs)
        jobs, _ = self._batch_jobs(None, candidates, targets, device_ptrs)
        return BatchPlan(self, jobs)

    def score_batch(self, images, candidates, targets, device_ptrs=None, winners=False):
        """Score every image against its candidate list [(quality, subsampling), ...] in one launch sequence.

        images: list of numpy arrays (host) or, with device_ptrs, a list of (ptr, width, height, channels).
        candidates: one list for every image, or a list of lists. Returns a list of BatchResult (scores, settled
        flags, chosen_* = the reference's bisection over the sorted list, sweep_* = exhaustive choice); with
        winners, (results, [CoefImage of the chosen candidate, or None]) -- Found.coefs of jpeg/src/target.rs:11-16."""
        out = self._batch_jobs(images, candidates, targets, device_ptrs, winners)
        jobs, keep = out[0], out[1]
        n = len(jobs)
        res = (BatchResult * n)()
        fn = lib().rcg_score_batch if device_ptrs is None else lib().rcg_score_batch_dev
        _check(fn(self._h, jobs, n, res))
        del keep
        if not winners:
            return list(res)
        return list(res), [_winner_image(jobs[i], res[i], out[2][i]) for i in range(n)]

    # ------------------------------------------------------------------ tile policy

    def get_tune(self):
        t = Tune()
        _check(lib().rcg_get_tune(self._h, ctypes.byref(t)))
        return t

    def set_tune(self, tune):
        _check(lib().rcg_set_tune(self._h, ctypes.byref(tune)))

    def autotune(self, jobs=8):
        """Measure and install the tile policy for launches of `jobs` same-size pairs (rcg_autotune); jobs=0 tunes
        for throughput, for the launches of a large batch."""
        t = Tune()
        _check(lib().rcg_autotune(self._h, jobs, ctypes.byref(t)))
        return t

    # ------------------------------------------------------------------ profiling

    def profile(self, enable):
        _check(lib().rcg_profile(self._h, 1 if enable else 0))

    def profile_get(self, kernel):
        ms = ctypes.c_double()
        n = ctypes.c_int()
        _check(lib().rcg_profile_get(self._h, kernel.encode(), ctypes.byref(ms), ctypes.byref(n)))
        return ms.value, n.value

    # ------------------------------------------------------------------ device-resident API (raw device pointers)

    def build_coefs_dev(self, d_pixels, width, height, channels, quality, subsampling, d_coefs):
        ptrs = (ctypes.c_void_p * 3)(*(list(d_coefs) + [None] * (3 - len(d_coefs))))
        _check(lib().rcg_build_coefs_dev(self._h, d_pixels, width, height, channels, quality, subsampling,
                                         ctypes.cast(ptrs, ctypes.c_void_p)))

    def reconstruct_dev(self, width, height, comps, qtables, d_pixels, adobe_transform=-1):
        """rcg_reconstruct_dev: comps are (id, h, v, tq, blocks_w, blocks_h, device coefficient pointer) tuples,
        qtables host natural-order tables by slot (or None); d_pixels receives Gray8 or RGB8 on the device."""
        s = _CoefImage()
        s.width, s.height, s.ncomp = width, height, len(comps)
        keep = []
        for i, (cid, h, v, tq, bw, bh, ptr) in enumerate(comps):
            c = s.comps[i]
            c.id, c.h, c.v, c.tq, c.blocks_w, c.blocks_h, c.coefs = cid, h, v, tq, bw, bh, ptr
        for slot, q in enumerate(qtables):
            if q is not None:
                qa = np.ascontiguousarray(q, dtype=np.uint8)
                keep.append(qa)
                s.qtables[slot] = qa.ctypes.data
        s.adobe_transform = adobe_transform
        _check(lib().rcg_reconstruct_dev(self._h, ctypes.byref(s), d_pixels))

    def ssimulacra2_dev(self, d_ref, d_dist, width, height, channels, precision=FP32):
        score = ctypes.c_double()
        ns = ctypes.c_int()
        _check(lib().rcg_ssimulacra2_dev(self._h, d_ref, d_dist, width, height, channels, precision,
                                         ctypes.byref(score), None, ctypes.byref(ns)))
        return score.value

    def find_quality_dev(self, d_pixels, width, height, channels, target, max_quality=100, subsampling=S420):
        f = _Found()
        _check(lib().rcg_find_quality_dev(self._h, d_pixels, width, height, channels, float(target), max_quality,
                                          subsampling, ctypes.byref(f)))
        return Found(f.quality, f.score, None, f.steps)


def _coef_image_struct(ci, keep, device):
    s = _CoefImage()
    s.width = ci.width
    s.height = ci.height
    s.ncomp = len(ci.comps)
    for i, c in enumerate(ci.comps):
        s.comps[i].id = c.id
        s.comps[i].h = c.h
        s.comps[i].v = c.v
        s.comps[i].tq = c.tq
        s.comps[i].blocks_w = c.blocks_w
        if device:
            s.comps[i].blocks_h = c.blocks_h
            s.comps[i].coefs = c.coefs
        else:
            co = np.ascontiguousarray(c.coefs, dtype=np.int16)
            keep.append(co)
            s.comps[i].blocks_h = len(co) // 64 // c.blocks_w
            s.comps[i].coefs = co.ctypes.data
    for slot, q in enumerate(ci.qtables):
        if q is not None:
            qa = np.ascontiguousarray(q, dtype=np.uint8)
            keep.append(qa)
            s.qtables[slot] = qa.ctypes.data
    s.adobe_transform = getattr(ci, "adobe_transform", -1)
    return s


class DeviceBuffer:
    """A device allocation on the context's GPU (rcg_device_alloc); .ptr is the raw device address."""

    def __init__(self, ctx, nbytes):
        self._ctx = ctx
        self.nbytes = nbytes
        p = ctypes.c_void_p()
        _check(lib().rcg_device_alloc(ctx.handle, nbytes, ctypes.byref(p)))
        self.ptr = p.value

    def write(self, array):
        """Copy a numpy array into this buffer, from its start (rcg_memcpy_h2d); the address stays the same."""
        a = np.ascontiguousarray(array)
        assert a.nbytes <= self.nbytes
        _check(lib().rcg_memcpy_h2d(self._ctx.handle, self.ptr, a.ctypes.data_as(ctypes.c_void_p), a.nbytes))
        return self

    def download(self, out):
        assert out.nbytes <= self.nbytes and out.flags["C_CONTIGUOUS"]
        _check(lib().rcg_memcpy_d2h(self._ctx.handle, out.ctypes.data_as(ctypes.c_void_p), self.ptr, out.nbytes))
        return out

    def free(self):
        if self.ptr:
            _check(lib().rcg_device_free(self._ctx.handle, self.ptr))
            self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def _max_grid(w, h, ch, qualities, subsamplings):
    """Coefficient count of each component's largest grid over the candidates (rcg_coef_layout)."""
    sizes = [0, 0, 0]
    for q, s in zip(qualities, subsamplings):
        L = coef_layout(w, h, ch, q, s)
        for c in range(L.ncomp):
            sizes[c] = max(sizes[c], L.blocks_w[c] * L.blocks_h[c] * 64)
    return sizes[:1] if ch == 1 else sizes


def _coef_image_at(w, h, ch, quality, subsampling, arrays):
    """A CoefImage over winner buffers: build_coefs's layout and tables at (quality, subsampling)."""
    L = coef_layout(w, h, ch, quality, subsampling)
    qy, qc = quant_tables(quality)
    comps = []
    for c in range(L.ncomp):
        n = L.blocks_w[c] * L.blocks_h[c] * 64
        hh, vv = (L.hmax, L.vmax) if c == 0 else (1, 1)
        comps.append(CoefComponent(c + 1, hh, vv, min(c, 1), L.blocks_w[c], np.array(arrays[c][:n], np.int16)))
    return CoefImage(w, h, comps, [qy, qc if L.ncomp == 3 else None, None, None])


def _winner_image(job, res, bufs):
    if res.status != OK or not res.winner_written:
        return None
    arrays = []
    for b in bufs:
        if isinstance(b, DeviceBuffer):
            arrays.append(b.download(np.empty(b.nbytes // 2, np.int16)))
        else:
            arrays.append(b)
    k = res.chosen_index
    return _coef_image_at(job.width, job.height, job.channels, job.quality[k], job.subsampling[k], arrays)


class BatchPlan:
    """rcg_batch_plan: the batched launch sequence captured once as a hipGraph, replayed by run()."""

    def __init__(self, ctx, jobs, winner_buffers=None):
        self._ctx = ctx
        self._n = len(jobs)
        self._jobs = jobs
        self._wbufs = winner_buffers
        self._h = ctypes.c_void_p()
        _check(lib().rcg_batch_plan_create(ctx.handle, jobs, self._n, ctypes.byref(self._h)))

    def run(self):
        res = (BatchResult * self._n)()
        _check(lib().rcg_batch_plan_run(self._h, res))
        return list(res)

    def winners(self, results):
        """The chosen candidates' coefficients of the last run (plans made with winners=True)."""
        return [_winner_image(self._jobs[i], results[i], self._wbufs[i]) for i in range(self._n)]

    def close(self):
        if self._h:
            lib().rcg_batch_plan_destroy(self._h)
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Server:
    """Continuous-batching target-quality service on one GPU (rcg_server_*)."""

    def __init__(self, device=0, hbm_budget=0, max_active=0, speculate=1, tune=None):
        self._h = ctypes.c_void_p()
        self._bufs = {}
        o = ServerOpts(hbm_budget, max_active, speculate, ctypes.addressof(tune) if tune is not None else None)
        _check(lib().rcg_server_create(device, ctypes.byref(o), ctypes.byref(self._h)))

    def submit(self, req_id, img, target, max_quality=100, subsampling=S420, coefs=False):
        """Queue a search. coefs: also return the chosen quality's coefficients (winner(completion))."""
        img = np.ascontiguousarray(img, dtype=np.uint8)
        h, w = img.shape[:2]
        ch = _channels(img)
        r = Request(req_id, img.ctypes.data, w, h, ch, float(target), max_quality, subsampling)
        if coefs:
            qs = [100, 1] if subsampling == AUTO else [100]
            bufs = [np.zeros(n, np.int16) for n in _max_grid(w, h, ch, qs, [subsampling] * len(qs))]
            for c, b in enumerate(bufs):
                r.coefs[c] = b.ctypes.data
            self._bufs[req_id] = (bufs, w, h, ch, subsampling)
        try:
            _check(lib().rcg_server_submit(self._h, ctypes.byref(r)))
        except RcgError:
            self._bufs.pop(req_id, None)
            raise

    def winner(self, completion):
        """The CoefImage of a completed request submitted with coefs=True (None if it failed); call once."""
        bufs = self._bufs.pop(completion.id, None)
        if bufs is None or completion.status != OK or not completion.coefs_written:
            return None
        arrays, w, h, ch, sub = bufs
        return _coef_image_at(w, h, ch, completion.quality, sub, arrays)

    def poll(self, max_items=64, timeout_ms=100):
        buf = (Completion * max_items)()
        n = ctypes.c_int()
        _check(lib().rcg_server_poll(self._h, buf, max_items, timeout_ms, ctypes.byref(n)))
        return [buf[i] for i in range(n.value)]

    def stats(self):
        st = ServerStats()
        _check(lib().rcg_server_get_stats(self._h, ctypes.byref(st)))
        return st

    def close(self):
        if self._h:
            lib().rcg_server_destroy(self._h)
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


_DEFAULT = None


def default_context():
    global _DEFAULT
    if _DEFAULT is None:
        _DEFAULT = Context(0)
    return _DEFAULT


def build_coefs(img, quality, subsampling=S420):
    return default_context().build_coefs(img, quality, subsampling)


WRITE_BASELINE, WRITE_PROGRESSIVE, WRITE_BEST = 0, 1, 2


def write_jpeg(ci, mode=WRITE_BASELINE):
    """Entropy-code a host CoefImage into JPEG bytes on the CPU (no device): WRITE_BASELINE is
    rc_jpeg::encode::write_jpeg, WRITE_PROGRESSIVE rc_jpeg::progressive::write_progressive, WRITE_BEST
    rc_jpeg::encode::write_best."""
    keep = []
    s = _coef_image_struct(ci, keep, device=False)
    out = ctypes.POINTER(ctypes.c_uint8)()
    n = ctypes.c_size_t()
    _check(lib().rcg_write_jpeg(ctypes.byref(s), mode, ctypes.byref(out), ctypes.byref(n)))
    try:
        return ctypes.string_at(out, n.value)
    finally:
        lib().rcg_free(out)


def trellis(ci, scaled):
    """The host trellis pass (rcg_trellis) over a plain CoefImage and its hand-off values: a CoefImage with the
    trellis coefficients (jpeg/src/encode.rs:527-537)."""
    keep = []
    s = _coef_image_struct(ci, keep, device=False)
    sc = [np.ascontiguousarray(x, dtype=np.float32) for x in scaled]
    outs = [np.zeros(len(c.coefs), np.int16) for c in ci.comps]
    sp = (ctypes.c_void_p * 4)(*([x.ctypes.data for x in sc] + [None] * (4 - len(sc))))
    op = (ctypes.c_void_p * 4)(*([x.ctypes.data for x in outs] + [None] * (4 - len(outs))))
    _check(lib().rcg_trellis(ctypes.byref(s), ctypes.cast(sp, ctypes.c_void_p), ctypes.cast(op, ctypes.c_void_p)))
    comps = [CoefComponent(c.id, c.h, c.v, c.tq, c.blocks_w, o) for c, o in zip(ci.comps, outs)]
    return CoefImage(ci.width, ci.height, comps, list(ci.qtables), getattr(ci, "adobe_transform", -1))


def _requantized(ci):
    r = _Requantized()
    outs = [np.zeros(len(c.coefs), np.int16) for c in ci.comps]
    for i, o in enumerate(outs):
        r.coefs[i] = o.ctypes.data
    return r, outs


def _from_requantized(ci, r, outs):
    comps = [CoefComponent(c.id, c.h, c.v, int(r.tq[i]), c.blocks_w, o) for i, (c, o) in enumerate(zip(ci.comps, outs))]
    tabs = [np.frombuffer(bytes(r.qtables[s]), np.uint8).copy() if r.present[s] else None for s in range(4)]
    return CoefImage(ci.width, ci.height, comps, tabs, getattr(ci, "adobe_transform", -1))


def requantize(ci, quality, trellis=False):
    """rc_jpeg::encode::CoefImage::requantize(quality, trellis) on the host (rcg_requantize): a new CoefImage."""
    keep = []
    s = _coef_image_struct(ci, keep, device=False)
    r, outs = _requantized(ci)
    _check(lib().rcg_requantize(ctypes.byref(s), quality, 1 if trellis else 0, ctypes.byref(r)))
    return _from_requantized(ci, r, outs)


@dataclass
class JpegInfo:
    """rc_jpeg::decode::Info (jpeg/src/decode.rs:311-331) with the layout, tables and EXIF orientation."""
    width: int
    height: int
    sampling: list           # (h, v) per component
    ids: list
    tq: list
    blocks: list             # (blocks_w, blocks_h) per component, MCU-padded
    qtables: list            # natural-order uint8[64] per slot, or None
    adobe_transform: int
    source_quality: int
    coded_rgb: bool
    cmyk: bool
    progressive: bool
    arithmetic: bool
    lossless: bool
    hierarchical: bool
    precision: int
    orientation: int
    table16: bool


def jpeg_info(data):
    """The frame facts of a JPEG file (rcg_jpeg_parse); host only."""
    buf = np.frombuffer(bytes(data), np.uint8)
    i = _JpegInfo()
    _check(lib().rcg_jpeg_parse(_ptr(buf), len(buf), ctypes.byref(i)))
    n = i.ncomp
    return JpegInfo(i.width, i.height, [(i.h[c], i.v[c]) for c in range(n)], [i.id[c] for c in range(n)],
                    [i.tq[c] for c in range(n)], [(i.blocks_w[c], i.blocks_h[c]) for c in range(n)],
                    [np.frombuffer(bytes(i.qtables[s]), np.uint8).copy() if i.present[s] else None for s in range(4)],
                    i.adobe_transform, i.source_quality, bool(i.coded_rgb), n == 4, bool(i.progressive),
                    bool(i.arithmetic), bool(i.lossless), bool(i.hierarchical), i.precision, i.orientation,
                    bool(i.table16))


def read_jpeg(data, info=None):
    """The quantised coefficients of a Huffman-coded 8-bit DCT JPEG (rcg_jpeg_read_coefs) as a CoefImage, the
    equivalent of rc_jpeg::decode::parse(data)?.into_coefs() (jpeg/src/decode.rs:529-569); host only."""
    info = info or jpeg_info(data)
    buf = np.frombuffer(bytes(data), np.uint8)
    outs = [np.zeros(bw * bh * 64, np.int16) for bw, bh in info.blocks]
    ptrs = (ctypes.c_void_p * 4)(*([o.ctypes.data for o in outs] + [None] * (4 - len(outs))))
    _check(lib().rcg_jpeg_read_coefs(_ptr(buf), len(buf), ctypes.cast(ptrs, ctypes.c_void_p)))
    comps = [CoefComponent(info.ids[c], info.sampling[c][0], info.sampling[c][1], info.tq[c], info.blocks[c][0],
                           outs[c]) for c in range(len(outs))]
    return CoefImage(info.width, info.height, comps, list(info.qtables), info.adobe_transform)


def rdo_quantize_ac(table, kappa, sizes, a):
    """rc_jpeg::rdo::Rdo::new(table, kappa, sizes).quantize_ac(a, out) (rcg_rdo_quantize_ac); returns out."""
    t = np.ascontiguousarray(table, dtype=np.uint8)
    sz = np.ascontiguousarray(sizes, dtype=np.uint8)
    av = np.ascontiguousarray(a, dtype=np.float32)
    out = np.zeros(64, np.int16)
    _check(lib().rcg_rdo_quantize_ac(t.ctypes.data, ctypes.c_float(kappa), sz.ctypes.data, av.ctypes.data,
                                     out.ctypes.data))
    return out


def huffman_lengths(freqs, max_len):
    """rc_core::util::huffman::build_lengths: length-limited Huffman code lengths of the symbol frequencies."""
    f = np.ascontiguousarray(freqs, dtype=np.uint32)
    out = np.zeros(len(f), dtype=np.uint8)
    _check(lib().rcg_huffman_lengths(f.ctypes.data, len(f), max_len, out.ctypes.data))
    return out


def reconstruct(ci):
    return default_context().reconstruct(ci)


def ssimulacra2(reference, distorted, precision=FP32):
    return default_context().ssimulacra2(reference, distorted, precision)


def find_quality(img, target, max_quality=100, subsampling=S420):
    return default_context().find_quality(img, target, max_quality, subsampling)

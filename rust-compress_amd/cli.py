"""Command-line driver of the JPEG target-quality path on one GPU: ``rc --target-ssimulacra2 SCORE -f jpeg FILES``.

Run as ``python -m rust-compress_amd [OPTIONS] INPUT...``. It mirrors the reference CLI (cli/src/cli.rs:56-293,
cli/src/main.rs:18-97, cli/src/pipeline.rs, cli/src/report.rs) for the part this library builds: JPEG output from
PNG or JPEG sources, with a quality (-q) or a SSIMULACRA2 target, trellis on by default. Arguments, their
validation (exit status 2), the routing of JPEG sources (DCT-domain requantisation or a pixel round trip,
pipeline.rs:285-420), the never-grow rule (pipeline.rs:745-752), the per-file report lines and the summary are the
reference's. Other output formats, the other targets and metadata (ICC, EXIF, XMP are not written) are reported per
file as unsupported.

Scheduling replaces the reference's rayon pool gated by a host-memory Budget (cli/src/budget.rs:11-50,
main.rs:57-75; pipeline.rs:881):
* pixel sources searched without trellis go to one continuous-batching server (rcg_server_*) that admits requests
  against the HBM budget and scores candidates of many files in shared launches;
* the trellis search (the default), JPEG -> JPEG requantisation searches and fixed-quality encodes run on a small
  pool of device contexts; each job first reserves its device footprint (rcg_search_bytes) against the HBM budget
  and waits while running jobs hold too much; a job larger than the whole budget runs, alone. Host threads keep
  several jobs in flight, so one file's host trellis overlaps another's device work.
The reference pyramid of each file's source is built once per search and stays resident while its candidates are
scored.
"""
import argparse
import io
import os
import queue
import sys
import tempfile
import threading
import time
from concurrent.futures import ThreadPoolExecutor

import numpy as np

from . import (AUTO, INVALID, OOM, S420, S422, S444, SEARCH_PLAIN, SEARCH_REQUANT, SEARCH_TRELLIS, UNSUPPORTED,
               WRITE_BEST, Context, RcgError, Server, jpeg_info, read_jpeg, requantize, trellis as trellis_pass,
               write_jpeg)

PNG_SIGNATURE = b"\x89PNG\r\n\x1a\n"
JPEG_EXT = {"jpg", "jpeg", "jpe", "jfif"}
READ_EXT = JPEG_EXT | {"png"}
ALL_EXT = READ_EXT | {"webp", "avif", "gif"}
MAX_SIDE = 65535


class FileError(Exception):
    """A per-file failure, already phrased as the reference reports it (rc_core::error::Error's Display)."""


def _fail(kind, msg):
    return FileError(f"{kind}: {msg}")


def _from_rcg(e):
    kind = {INVALID: "invalid data", UNSUPPORTED: "unsupported", OOM: "out of memory"}.get(e.status, "device error")
    text = str(e).split(": ", 1)[-1]
    return FileError(f"{kind}: {text}")


# ----------------------------------------------------------------------------------------------- arguments

def _number(lo, hi, what, example, upper_name=None):
    def parse(s):
        try:
            v = float(s)
        except ValueError:
            raise argparse.ArgumentTypeError(f"expected a number such as {example}")
        if not (lo < v <= hi):
            raise argparse.ArgumentTypeError(f"{what} target must be in ({lo:g}, {upper_name or f'{hi:g}'}]")
        return v
    return parse


def _int_range(lo, hi):
    def parse(s):
        try:
            v = int(s)
        except ValueError:
            raise argparse.ArgumentTypeError(f"expected an integer in {lo}..={hi}")
        if not lo <= v <= hi:
            raise argparse.ArgumentTypeError(f"{v} is not in {lo}..={hi}")
        return v
    return parse


def _hex_color(s):
    """cli/src/cli.rs:245-250."""
    t = s[1:] if s.startswith("#") else s
    if len(t) != 6 or not all(c in "0123456789abcdefABCDEF" for c in t):
        raise argparse.ArgumentTypeError("expected 6 hex digits, e.g. ffffff")
    v = int(t, 16)
    return (v >> 16 & 255, v >> 8 & 255, v & 255)


def _suffix(s):
    if not s or "/" in s or "\\" in s:
        raise argparse.ArgumentTypeError("expected a non-empty suffix without path separators")
    return s


def parser():
    p = argparse.ArgumentParser(prog="rc", description="JPEG target-quality compression on the GPU (gfx950).")
    p.add_argument("inputs", nargs="+", metavar="INPUT", help="image files and/or directories")
    p.add_argument("-r", "--recursive", action="store_true", help="descend into subdirectories")
    p.add_argument("--min-suffix", default="min", type=_suffix, help="suffix of default outputs (<name>.<suffix>.<ext>)")
    p.add_argument("-o", "--output", help="output file (single input) or directory")
    p.add_argument("-q", "--quality", type=_int_range(1, 100), default=80,
                   help="JPEG quality ceiling (1-100): sources at or below it are transcoded, better ones requantised")
    p.add_argument("-f", "--format", choices=["keep", "png", "jpeg", "webp", "avif", "gif"], default="keep")
    p.add_argument("--subsampling", choices=["auto", "444", "422", "420"], default="auto")
    p.add_argument("--lossy", action="store_true", help="allow re-encoding a hierarchical JPEG")
    g = p.add_mutually_exclusive_group()
    g.add_argument("--target-ssim", type=_number(0.0, 1.0, "SSIM", "0.97"), metavar="SSIM")
    g.add_argument("--target-psnr", type=_number(0.0, 99.0, "PSNR", "40"), metavar="DB")
    g.add_argument("--target-ms-ssim", type=_number(0.0, 1.0, "SSIM", "0.97"), metavar="SSIM")
    g.add_argument("--target-ssimulacra2", type=_number(0.0, 100.0, "SSIMULACRA 2", "80"), metavar="SCORE",
                   help="lowest quality that still scores at least this SSIMULACRA 2 (0-100]")
    g.add_argument("--target-size", type=_int_range(1, 2 ** 64 - 1), metavar="BYTES")
    p.add_argument("--no-trellis", action="store_true", help="disable trellis quantisation")
    p.add_argument("--background", type=_hex_color, default=(255, 255, 255), help="hex RRGGBB for flattening alpha")
    p.add_argument("--force", action="store_true", help="overwrite existing outputs")
    p.add_argument("--dry-run", action="store_true", help="report without writing")
    p.add_argument("--quiet", action="store_true", help="only report failures and the summary")
    p.add_argument("-j", "--jobs", type=int, default=None, help="files in flight at once (default: CPUs, at most 16)")
    p.add_argument("--max-memory", type=_int_range(1, 4194304), default=65536, metavar="MIB",
                   help="HBM budget of concurrently running jobs, MiB (default 65536)")
    p.add_argument("--contexts", type=_int_range(1, 16), default=4, help="device contexts (streams) in the pool")
    p.add_argument("--device", type=int, default=0, help="HIP device index")
    return p


# ----------------------------------------------------------------------------------------------- reports

def human(n):
    """cli/src/report.rs:44-52."""
    units = ["B", "KB", "MB", "GB"]
    v, u = float(n), 0
    while v >= 1024.0 and u < len(units) - 1:
        v /= 1024.0
        u += 1
    return f"{n} B" if u == 0 else f"{v:.1f} {units[u]}"


def percent(a, b):
    return "n/a" if a == 0 else f"{(b - a) * 100.0 / a:+.1f}%"


def sanitize(s):
    return "".join(f"\\x{ord(c):02x}" if ord(c) < 32 or 127 <= ord(c) < 160 else c for c in s)


class Report:
    """cli/src/report.rs:16-97."""

    def __init__(self, job):
        self.input, self.output = job
        self.in_size = self.out_size = 0
        self.status = "no_gain"
        self.note = None
        self.error = None

    def line(self, dry_run):
        name = self.input
        if self.status == "failed":
            return sanitize(f"{name}  FAILED: {self.error}")
        if self.status == "no_gain":
            return sanitize(f"{name}  {human(self.in_size)}  (no gain, kept original size)"
                            + (f"  [{self.note}]" if self.note else ""))
        tail = (f"  [{self.note}]" if self.note else "") + ("  [dry run]" if dry_run else "")
        if self.status == "larger":
            tail += "  [larger than the original: format change or resize]"
        return sanitize(f"{name}  {human(self.in_size)} -> {human(self.out_size)}  "
                        f"({percent(self.in_size, self.out_size)}){tail}")


def summary_line(reports, problems, secs):
    """cli/src/report.rs:99-134."""
    n = {k: 0 for k in ("compressed", "larger", "no_gain", "failed")}
    bin_ = bout = 0
    for r in reports:
        n[r.status] += 1
        if r.status != "failed":
            bin_ += r.in_size
            bout += r.out_size
    failed = n["failed"] + len(problems)
    larger = f", {n['larger']} larger" if n["larger"] else ""
    return (f"{len(reports)} file(s): {n['compressed']} compressed{larger}, {n['no_gain']} no gain, {failed} failed; "
            f"{human(bin_)} -> {human(bout)} ({percent(bin_, bout)}) in {secs:.2f}s"), failed


# ----------------------------------------------------------------------------------------------- job collection

def _ext(path):
    return os.path.splitext(path)[1][1:].lower()


def _is_default_output(path, suffix):
    stem = os.path.splitext(os.path.basename(path))[0]
    return stem.endswith("." + suffix)


def _walk(d, recursive, suffix, out):
    for name in sorted(os.listdir(d)):
        p = os.path.join(d, name)
        if os.path.isdir(p):
            if recursive:
                _walk(p, recursive, suffix, out)
        elif os.path.isfile(p) and _ext(p) in ALL_EXT and not _is_default_output(p, suffix):
            out.append(p)


def _out_ext(args, path):
    e = os.path.splitext(path)[1][1:]
    if args.format == "keep" or (args.format == "jpeg" and e.lower() in JPEG_EXT) or (
            args.format == "png" and e.lower() == "png"):
        return e
    return {"jpeg": "jpg"}.get(args.format, args.format)


def collect(args):
    """(input, output) pairs plus problems with inputs (cli/src/collect.rs:76-117,213-300)."""
    problems, jobs = [], []
    n_files = sum(1 for p in args.inputs if not os.path.isdir(p))
    any_dir = any(os.path.isdir(p) for p in args.inputs)
    out = args.output
    dir_mode = out is not None and (os.path.isdir(out) or n_files != 1 or any_dir or out.endswith(("/", "\\")))
    for inp in args.inputs:
        if os.path.isdir(inp):
            files = []
            try:
                _walk(inp, args.recursive, args.min_suffix, files)
            except OSError as e:
                problems.append(f"{inp}: {e}")
                continue
            pairs = [(f, os.path.relpath(f, inp)) for f in files]
        elif os.path.isfile(inp):
            pairs = [(inp, os.path.basename(inp))]
        else:
            problems.append(f"{inp}: no such file or directory" if not os.path.exists(inp) else
                            f"{inp}: not a regular file (pipes and devices are not supported)")
            continue
        for f, rel in pairs:
            ext = _out_ext(args, f)
            if out is None:
                stem = os.path.splitext(os.path.basename(f))[0]
                dst = os.path.join(os.path.dirname(f), f"{stem}.{args.min_suffix}.{ext}")
            elif dir_mode:
                dst = os.path.join(out, os.path.splitext(rel)[0] + (f".{ext}" if ext else ""))
            else:
                dst = out
            jobs.append((f, dst))
    return jobs, problems


# ----------------------------------------------------------------------------------------------- pixels

def _blend(c, a, bg):
    """core/src/image.rs:186-190: (c*a + bg*(255-a) + 127) / 255."""
    a = a.astype(np.uint32)
    return ((c.astype(np.uint32) * a + np.uint32(bg) * (255 - a) + 127) // 255).astype(np.uint8)


def decode_png(data, bg):
    """A PNG source as Gray8 or RGB8, alpha flattened onto bg (Image::flatten, core/src/image.rs:145-178)."""
    from PIL import Image  # the PNG container is decoded by Pillow; its samples are lossless
    if len(data) < 33 or data[12:16] != b"IHDR":
        raise _fail("invalid data", "truncated PNG")
    depth, ctype = data[24], data[25]
    if depth == 16 and ctype != 0:
        raise _fail("unsupported", "16-bit colour PNG sources are not part of this build")
    try:
        im = Image.open(io.BytesIO(data))
        im.load()
    except Exception as e:  # noqa: BLE001 - any decoder complaint is a per-file error
        raise _fail("invalid data", f"PNG: {e}")
    if im.mode in ("I;16", "I;16B", "I"):
        v = np.asarray(im).astype(np.uint32)
        return ((v * 255 + 32767) // 65535).astype(np.uint8)
    if im.mode == "P":
        im = im.convert("RGBA" if "transparency" in im.info else "RGB")
    elif im.mode == "1":
        im = im.convert("L")
    elif im.mode in ("L", "RGB") and "transparency" in im.info:
        im = im.convert("LA" if im.mode == "L" else "RGBA")
    a = np.asarray(im)
    if im.mode == "L" or im.mode == "RGB":
        return np.ascontiguousarray(a)
    if im.mode == "LA":
        return _blend(a[:, :, 0], a[:, :, 1], bg[0])
    if im.mode == "RGBA":
        return np.stack([_blend(a[:, :, c], a[:, :, 3], bg[c]) for c in range(3)], axis=2)
    raise _fail("unsupported", f"PNG colour mode {im.mode}")


def upright(img, orientation):
    """Image::oriented (core/src/image.rs:194-219): EXIF orientation 2..8 baked into the pixels."""
    if orientation == 2:
        return img[:, ::-1]
    if orientation == 3:
        return img[::-1, ::-1]
    if orientation == 4:
        return img[::-1, :]
    if orientation in (5, 6, 7, 8):
        t = np.swapaxes(img, 0, 1)
        return {5: t, 6: t[:, ::-1], 7: t[::-1, ::-1], 8: t[::-1, :]}[orientation]
    return img


# ----------------------------------------------------------------------------------------------- device side

class HbmBudget:
    """cli/src/budget.rs:11-50 over device memory: acquire(bytes) waits while running jobs hold too much; nothing is
    reserved beyond the whole budget and nothing waits when nothing else runs."""

    def __init__(self, limit):
        self.limit = int(limit)
        self.used = 0
        self.cv = threading.Condition()
        self.peak = 0

    def acquire(self, nbytes):
        nbytes = min(int(nbytes), self.limit)
        with self.cv:
            while self.used > 0 and self.used + nbytes > self.limit:
                self.cv.wait()
            self.used += nbytes
            self.peak = max(self.peak, self.used)
        return _Permit(self, nbytes)


class _Permit:
    def __init__(self, budget, nbytes):
        self.budget, self.nbytes = budget, nbytes

    def __enter__(self):
        return self

    def __exit__(self, *a):
        with self.budget.cv:
            self.budget.used -= self.nbytes
            self.budget.cv.notify_all()


class Gpu:
    """The device side of one CLI run: a pool of contexts, the HBM budget, and the serving loop (made on first use)."""

    def __init__(self, device, contexts, budget_bytes):
        self.device = device
        self.budget = HbmBudget(budget_bytes)
        self.pool = queue.Queue()
        self._made = 0
        self._ncontexts = contexts
        self._lock = threading.Lock()
        self._server = None
        self._waiters = {}
        self._next_id = 0
        self._poller = None
        self._poll_error = None
        self._stop = False

    def _context(self):
        try:
            return self.pool.get_nowait()
        except queue.Empty:
            pass
        with self._lock:
            if self._made < self._ncontexts:
                self._made += 1
                try:
                    return Context(self.device)
                except Exception:
                    self._made -= 1
                    raise
        return self.pool.get()

    def run(self, need, fn):
        """fn(ctx) on a pooled context once `need` device bytes fit the budget."""
        with self.budget.acquire(need):
            ctx = self._context()
            try:
                return fn(ctx)
            finally:
                self.pool.put(ctx)

    def footprint(self, w, h, ch, subsampling, mode):
        ctx = self._context()
        try:
            return ctx.search_bytes(w, h, ch, subsampling, mode)
        finally:
            self.pool.put(ctx)

    # The continuous-batching server for searches without trellis.
    def search(self, img, target, subsampling):
        with self._lock:
            if self._server is None:
                self._server = Server(self.device, hbm_budget=self.budget.limit)
                self._poller = threading.Thread(target=self._poll, daemon=True)
                self._poller.start()
            if self._poll_error is not None:
                raise self._poll_error
            self._next_id += 1
            rid = self._next_id
            ev = threading.Event()
            self._waiters[rid] = [ev, None]
        try:
            self._server.submit(rid, img, target, 100, subsampling, coefs=True)
        except RcgError:
            with self._lock:
                self._waiters.pop(rid, None)
            raise
        ev.wait()
        with self._lock:
            return self._waiters.pop(rid)[1]

    def _poll(self):
        try:
            while not self._stop:
                for c in self._server.poll(64, 50):
                    with self._lock:
                        w = self._waiters.get(c.id)
                    if w is not None:
                        w[1] = (c.status, c.quality, c.score, self._server.winner(c))
                        w[0].set()
        except RcgError as e:
            # A failing poll ends the serving loop: every waiting search, and every later one, fails with it.
            with self._lock:
                self._poll_error = e
                for w in self._waiters.values():
                    if w[1] is None:
                        w[1] = (e.status, 0, 0.0, None)
                    w[0].set()

    def close(self):
        self._stop = True
        if self._poller is not None:
            self._poller.join()
        if self._server is not None:
            self._server.close()
        while not self.pool.empty():
            self.pool.get_nowait().close()


# ----------------------------------------------------------------------------------------------- one file

SUBSAMPLING = {"auto": AUTO, "444": S444, "422": S422, "420": S420}


def _coefs_at(ctx, img, quality, subsampling, trellis):
    """build_coefs(img, EncodeOptions{quality, subsampling, trellis}) (jpeg/src/encode.rs:426-540) on the device,
    with the trellis pass on the host."""
    if trellis and quality < 100:
        ci, scaled = ctx.build_scaled(img, quality, subsampling)
        return trellis_pass(ci, scaled)
    return ctx.build_coefs(img, quality, subsampling)


class Pipeline:
    def __init__(self, args, gpu):
        self.args = args
        self.gpu = gpu
        self.trellis = not args.no_trellis
        self.subsampling = SUBSAMPLING[args.subsampling]

    def encode_pixels(self, img, quality):
        """encode_jpeg_pixels (cli/src/pipeline.rs:224-247): (CoefImage, note)."""
        h, w = img.shape[:2]
        ch = 1 if img.ndim == 2 else 3
        target = self.args.target_ssimulacra2
        if target is None:
            if w > MAX_SIDE or h > MAX_SIDE:
                raise _fail("unsupported", "JPEG dimensions are limited to 65535 pixels")
            need = self.gpu.footprint(w, h, ch, self.subsampling, SEARCH_TRELLIS)
            return self.gpu.run(need, lambda ctx: _coefs_at(ctx, img, quality, self.subsampling, self.trellis)), None
        if w < 8 or h < 8:
            raise _fail("unsupported", "--target-ssimulacra2 needs an image of at least 8x8 pixels")
        if not self.trellis:
            # The serving loop returns the chosen quality's coefficients with the completion (Found.coefs).
            st, q, score, ci = self.gpu.search(img, target, self.subsampling)
            if st != 0 or ci is None:
                raise _from_rcg(RcgError(st or 3, "search failed"))
            return ci, f"q{q}, ssimulacra2 {score:.4f}"
        need = self.gpu.footprint(w, h, ch, self.subsampling, SEARCH_TRELLIS)
        found = self.gpu.run(need, lambda ctx: ctx.find_quality(img, target, 100, self.subsampling, trellis=True))
        return found.coefs, f"q{found.quality}, ssimulacra2 {found.score:.4f}"

    def jpeg_to_jpeg(self, data):
        """cli/src/pipeline.rs:285-420: (bytes or None to keep the original, note, forced)."""
        a = self.args
        info = jpeg_info(data)
        three = len(info.sampling) == 3
        is444 = three and all(s == (1, 1) for s in info.sampling)
        is422 = three and info.sampling[0] == (2, 1) and all(s == (1, 1) for s in info.sampling[1:])
        is420 = three and info.sampling[0] == (2, 2) and all(s == (1, 1) for s in info.sampling[1:])
        effective_q = min(a.quality, info.source_quality)
        resample = three and {"auto": is444 and effective_q < 90 and a.target_ssim is None,
                              "420": not is420, "422": not is422, "444": not is444}[a.subsampling]
        if info.lossless and not a.lossy and info.orientation == 1:
            return None, "lossless JPEG left unchanged (lossless recoding is not part of this build)", False
        if info.hierarchical and not a.lossy and info.orientation == 1:
            return None, "hierarchical JPEG left unchanged (use --lossy to re-encode)", False
        if info.arithmetic:
            raise _fail("unsupported", "arithmetic-coded JPEG sources are not part of this build")
        if info.orientation == 1 and not resample and not info.coded_rgb and info.precision == 8 \
                and not info.lossless and not info.hierarchical:
            ci = read_jpeg(data, info)
            if info.table16:
                raise _fail("unsupported", "16-bit quantization table")
            target = a.target_ssimulacra2
            if target is not None:
                if info.width < 8 or info.height < 8:
                    raise _fail("unsupported", "--target-ssimulacra2 needs an image of at least 8x8 pixels")
                need = self.gpu.footprint(info.width, info.height, len(ci.comps), AUTO, SEARCH_REQUANT)
                found = self.gpu.run(need, lambda ctx: ctx.requant_find_quality(ci, target, 100, self.trellis))
                return write_jpeg(found.coefs, WRITE_BEST), f"q{found.quality}, ssimulacra2 {found.score:.4f}", False
            before = [None if t is None else t.copy() for t in ci.qtables]
            out = requantize(ci, a.quality, self.trellis)
            note = None
            if info.cmyk:
                same = all((x is None and y is None) or (x is not None and y is not None and np.array_equal(x, y))
                           for x, y in zip(before, out.qtables))
                note = "CMYK, lossless" if same else "CMYK, requantized"
            return write_jpeg(out, WRITE_BEST), note, False
        # The pixel route: decode (on the device), bake in the orientation, encode.
        if info.lossless or info.hierarchical or info.precision != 8:
            raise _fail("unsupported", "decoding this JPEG frame type is not part of this build")
        ci = read_jpeg(data, info)
        w, h = info.width, info.height
        need = self.gpu.footprint(w, h, 1 if len(ci.comps) == 1 else 3, AUTO, SEARCH_PLAIN)
        img = self.gpu.run(need, lambda ctx: ctx.reconstruct(ci))
        img = np.ascontiguousarray(upright(img, info.orientation))
        coefs, note = self.encode_pixels(img, effective_q)
        return write_jpeg(coefs, WRITE_BEST), note, False

    def process(self, job):
        """cli/src/pipeline.rs:680-752,842-889 for JPEG output."""
        a = self.args
        r = Report(job)
        try:
            with open(r.input, "rb") as f:
                data = f.read()
        except OSError as e:
            r.status, r.error = "failed", f"i/o error: {e.strerror or e}"
            return r
        r.in_size = len(data)
        try:
            if data.startswith(PNG_SIGNATURE):
                in_fmt = "png"
            elif data[:2] == b"\xff\xd8":
                in_fmt = "jpeg"
            elif data[:4] == b"RIFF" and data[8:12] == b"WEBP":
                in_fmt = "webp"
            elif data[:6] in (b"GIF87a", b"GIF89a"):
                in_fmt = "gif"
            elif data[4:8] == b"ftyp":
                in_fmt = "avif"
            else:
                ext = _ext(r.input)
                in_fmt = "jpeg" if ext in JPEG_EXT else (ext if ext in ALL_EXT else None)
            if in_fmt is None:
                raise _fail("unsupported", "unrecognized image format (supported: PNG, JPEG, WebP, AVIF, GIF)")
            out_fmt = in_fmt if a.format == "keep" else a.format
            if out_fmt != "jpeg":
                raise _fail("unsupported", "this build writes JPEG only (use -f jpeg)")
            if a.target_psnr is not None or a.target_ms_ssim is not None or a.target_size is not None:
                raise _fail("unsupported", "--target-psnr/--target-ms-ssim/--target-size need lossy WebP, AVIF or "
                                           "--lossy PNG")
            if a.target_ssim is not None:
                raise _fail("unsupported", "--target-ssim is not part of this build (use --target-ssimulacra2)")
            if not a.force and not a.dry_run and os.path.exists(r.output):
                raise _fail("unsupported", f"{r.output} already exists (use --force)")
            if in_fmt not in ("png", "jpeg"):
                raise _fail("unsupported", f"{in_fmt.upper()} sources are not part of this build")
            if in_fmt == "jpeg":
                encoded, note, forced = self.jpeg_to_jpeg(data)
            else:
                img = decode_png(data, a.background)
                coefs, note = self.encode_pixels(img, a.quality)
                encoded, forced = write_jpeg(coefs, WRITE_BEST), False
            r.note = note
            if encoded is None or (in_fmt == out_fmt and not forced and len(encoded) >= len(data)):
                # Never make the file bigger: keep the original bytes (copy them when writing elsewhere).
                if not a.dry_run:
                    _write(r.output, data)
                r.status, r.out_size = "no_gain", len(data)
                return r
            if not a.dry_run:
                _write(r.output, encoded)
            r.out_size = len(encoded)
            r.status = "larger" if len(encoded) > len(data) else "compressed"
        except FileError as e:
            r.status, r.error = "failed", str(e)
        except RcgError as e:
            r.status, r.error = "failed", str(_from_rcg(e))
        except OSError as e:
            r.status, r.error = "failed", f"i/o error: {e.strerror or e}"
        except (MemoryError, ValueError) as e:
            # One file's failure stays that file's (main.rs:57-75 reports per file and carries on).
            r.status, r.error = "failed", f"internal error: {type(e).__name__}: {e}"
        return r


def _write(path, data):
    """Write through a temporary file in the same directory and rename it into place."""
    d = os.path.dirname(os.path.abspath(path))
    os.makedirs(d, exist_ok=True)
    fd, tmp = tempfile.mkstemp(dir=d, prefix=".rc-", suffix=".tmp")
    try:
        with os.fdopen(fd, "wb") as f:
            f.write(data)
        os.replace(tmp, path)
    except BaseException:
        if os.path.exists(tmp):
            os.unlink(tmp)
        raise


# ----------------------------------------------------------------------------------------------- main

def main(argv=None):
    args = parser().parse_args(argv)
    if args.jobs is not None and args.jobs < 1:
        print("error: --jobs must be at least 1", file=sys.stderr)
        return 2
    start = time.time()
    jobs, problems = collect(args)
    for p in problems:
        print(f"error: {sanitize(p)}", file=sys.stderr)
    if not jobs:
        if not problems:
            print("error: no supported image files found (PNG, JPEG, WebP, AVIF, GIF; use -r to descend into "
                  "subdirectories)", file=sys.stderr)
        return 1
    nthreads = args.jobs or min(16, os.cpu_count() or 1)
    gpu = Gpu(args.device, args.contexts, args.max_memory << 20)
    pipe = Pipeline(args, gpu)
    out_lock = threading.Lock()

    def one(job):
        r = pipe.process(job)
        with out_lock:
            if r.status == "failed":
                print(r.line(args.dry_run), file=sys.stderr, flush=True)
            elif not args.quiet:
                print(r.line(args.dry_run), flush=True)
        return r

    try:
        with ThreadPoolExecutor(max_workers=nthreads) as ex:
            reports = list(ex.map(one, jobs))
    finally:
        gpu.close()
    line, failed = summary_line(reports, problems, time.time() - start)
    print(line, flush=True)
    return 1 if failed else 0


if __name__ == "__main__":
    sys.exit(main())

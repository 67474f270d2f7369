"""The allocation-free entry point rcg_reconstruct_dev_noalloc, through the C ABI, against rcg_reconstruct_dev (the
entry point it is a variant of) and the oracle: bit-exact on two calls in a row, the second with a smaller image
(its planes overwrite part of the first call's in the shared scratch array), on gray, 4:2:0, 4:4:4 and YCCK images,
and the refusal of an image whose sample planes exceed the array. The allocation-free entry point always takes the
two-kernel form (IDCT into the scratch planes, then the colour stage), where rcg_reconstruct_dev takes the fused
kernel for the gray, 4:2:0 and 4:4:4 images: those comparisons check the two forms against each other too."""
import ctypes
import os

import numpy as np
import pytest

from oracle import oracle as O

pytestmark = pytest.mark.gpu
GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


class _DevComp:
    def __init__(self, c, ptr):
        self.id, self.h, self.v, self.tq, self.blocks_w = c.id, c.h, c.v, c.tq, c.blocks_w
        self.blocks_h = len(c.coefs) // 64 // c.blocks_w
        self.coefs = ptr


class _DevImage:
    def __init__(self, ci, comps):
        self.width, self.height, self.qtables = ci.width, ci.height, ci.qtables
        self.adobe_transform = getattr(ci, "adobe_transform", -1)
        self.comps = comps


def _device_image(ctx, rc, ci, keep):
    comps = []
    for c in ci.comps:
        buf = ctx.upload(np.ascontiguousarray(c.coefs, dtype=np.int16))
        keep.append(buf)
        comps.append(_DevComp(c, buf.ptr))
    return rc._coef_image_struct(_DevImage(ci, comps), keep, device=True)


def _reconstruct(ctx, rc, fn, ci):
    keep = []
    s = _device_image(ctx, rc, ci, keep)
    ch = 1 if len(ci.comps) == 1 else 3
    out = rc.DeviceBuffer(ctx, ci.width * ci.height * ch)
    keep.append(out)
    rc._check(fn(ctx.handle, ctypes.byref(s), out.ptr))
    host = np.empty((ci.height, ci.width) if ch == 1 else (ci.height, ci.width, 3), np.uint8)
    return out.download(host)


def _images(rc):
    big = O.build_coefs(rc.synth_rgb8(5, 1531, 1027), 80, O.S420)        # 6.3 MB of planes
    small = O.build_coefs(rc.synth_rgb8(6, 333, 197), 90, O.S444)        # 0.2 MB
    gray = O.build_coefs(rc.synth_rgb8(7, 250, 99)[:, :, 1].copy(), 60, O.S444)
    with open(os.path.join(GOLD, "ycck.jpg"), "rb") as f:
        ycck = rc.read_jpeg(f.read())
    return big, small, gray, ycck


def test_no_alloc_reconstruct_equals_reconstruct_dev_on_calls_in_a_row(ctx, rc):
    L = rc.lib()
    for i, ci in enumerate(_images(rc)):
        print(f"no-alloc call {i}: {ci.width}x{ci.height}, {len(ci.comps)} components", flush=True)
        got = _reconstruct(ctx, rc, L.rcg_reconstruct_dev_noalloc, ci)
        want = _reconstruct(ctx, rc, L.rcg_reconstruct_dev, ci)
        np.testing.assert_array_equal(got, want)
        np.testing.assert_array_equal(got, O.reconstruct(ci))
    # And once more at the first size after the smaller ones, then the smaller one again.
    big, small, _, _ = _images(rc)
    for ci in (big, small):
        np.testing.assert_array_equal(_reconstruct(ctx, rc, L.rcg_reconstruct_dev_noalloc, ci), O.reconstruct(ci))


def test_no_alloc_refuses_planes_beyond_its_scratch(ctx, rc):
    # 2048 x 1024 blocks of one component: 128 MiB of planes against the 64 MiB array. The check runs on the host
    # before anything is enqueued (the CPU test of the ABI shows it refuses without a context too).
    buf = rc.DeviceBuffer(ctx, 1 << 16)
    s = rc._CoefImage()
    s.width, s.height, s.ncomp = 16384, 8192, 1
    c = s.comps[0]
    c.id, c.h, c.v, c.tq, c.blocks_w, c.blocks_h, c.coefs = 1, 1, 1, 0, 2048, 1024, buf.ptr
    q = np.ones(64, np.uint8)
    s.qtables[0] = q.ctypes.data
    s.adobe_transform = -1
    out = rc.DeviceBuffer(ctx, 1 << 16)
    assert rc.lib().rcg_reconstruct_dev_noalloc(ctx.handle, ctypes.byref(s), out.ptr) == rc.UNSUPPORTED
    assert b"allocation-free scratch" in rc.lib().rcg_last_error()

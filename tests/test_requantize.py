"""DCT-domain requantisation on the host (rcg_requantize) against what pins it.

* cli/tests/cmyk.rs:38-54: at q100 nothing is requantised to a finer table than the source had, so every
  coefficient and every table is unchanged (the four-component CMYK/YCCK files and the 1/3-component ones).
* Idempotence: requantising to q twice is requantising once (the new tables are never finer than the old).
* The product equals the oracle's restatement (jpeg/src/encode.rs:107-171) bit for bit, plain and with the trellis,
  including the table-slot assignment.
* A requantised image written by the entropy writer decodes (libjpeg) to the oracle's reconstruction of it within
  an independent decoder's IDCT rounding.
No device is used.
"""
import importlib
import io
import os

import numpy as np
import pytest
from PIL import Image

from oracle import oracle as O
import jpeg_scan

rc = importlib.import_module("rust-compress_amd")
GOLD = os.path.join(os.path.dirname(__file__), "golden")


def _frame(name):
    with open(os.path.join(GOLD, name), "rb") as f:
        fr = jpeg_scan.first_frame(f.read())
    comps = [rc.CoefComponent(c["id"], c["h"], c["v"], c["tq"], c["blocks_w"], np.asarray(c["coefs"], np.int16))
             for c in fr["comps"]]
    return rc.CoefImage(fr["width"], fr["height"], comps, [fr["qtables"].get(i) for i in range(4)])


def _synth(seed, w, h, q, subs, gray=False):
    img = rc.synth_rgb8(seed, w, h)
    if gray:
        img = np.ascontiguousarray(img[:, :, 0])
    ci = O.build_coefs(img, q, subs)
    return rc.CoefImage(ci.width, ci.height,
                        [rc.CoefComponent(c.id, c.h, c.v, c.tq, c.blocks_w, c.coefs) for c in ci.comps],
                        list(ci.qtables) + [None] * (4 - len(ci.qtables)))


FILES = ["cmyk.jpg", "ycck.jpg", "ycck_444.jpg", "rgb8_420.jpg", "rgb8_444.jpg", "gray8.jpg"]


@pytest.mark.parametrize("name", FILES)
@pytest.mark.parametrize("trellis", [False, True])
def test_q100_changes_no_coefficient_or_table(name, trellis):
    ci = _frame(name)
    out = rc.requantize(ci, 100, trellis)
    for a, b in zip(ci.comps, out.comps):
        np.testing.assert_array_equal(a.coefs, b.coefs)
    used = sorted({c.tq for c in ci.comps})
    for s in used:
        np.testing.assert_array_equal(out.qtables[out.comps[[c.tq for c in ci.comps].index(s)].tq], ci.qtables[s])


@pytest.mark.parametrize("name", FILES)
@pytest.mark.parametrize("q", [90, 60, 25])
def test_requantize_is_idempotent(name, q):
    once = rc.requantize(_frame(name), q)
    twice = rc.requantize(once, q)
    for a, b in zip(once.comps, twice.comps):
        np.testing.assert_array_equal(a.coefs, b.coefs)
        assert a.tq == b.tq


CASES = [(name, q, t) for name in FILES for q in (95, 70, 40, 10) for t in (False, True)]


@pytest.mark.parametrize("name,q,trellis", CASES)
def test_product_equals_the_oracle(name, q, trellis):
    ci = _frame(name)
    got = rc.requantize(ci, q, trellis)
    want = O.requantize(ci, q, trellis)
    for g, w in zip(got.comps, want.comps):
        np.testing.assert_array_equal(g.coefs, w.coefs)
        assert g.tq == w.tq
    for s in range(4):
        assert (got.qtables[s] is None) == (want.qtables[s] is None)
        if got.qtables[s] is not None:
            np.testing.assert_array_equal(got.qtables[s], want.qtables[s])


@pytest.mark.parametrize("w,h,q0,q,subs,gray", [(97, 61, 92, 50, O.S420, False), (64, 40, 85, 70, O.S444, False),
                                               (50, 31, 95, 30, O.S422, False), (45, 33, 90, 60, O.S444, True)])
def test_requantized_file_decodes_to_the_oracle_reconstruction(w, h, q0, q, subs, gray):
    src = _synth(300 + w, w, h, q0, subs, gray)
    for trellis in (False, True):
        r = rc.requantize(src, q, trellis)
        want = O.requantize(src, q, trellis)
        for a, b in zip(r.comps, want.comps):
            np.testing.assert_array_equal(a.coefs, b.coefs)
        dec = np.asarray(Image.open(io.BytesIO(rc.write_jpeg(r)))).astype(np.int32)
        d = np.abs(dec - O.reconstruct(want).astype(np.int32))
        assert d.max() <= 3, (w, h, trellis, d.max())


def test_requantize_to_a_coarser_table_changes_coefficients_by_the_step_ratio():
    # A q60 source requantised to q30: every table entry grows, and a coefficient c becomes round(c * old/new)
    # (checked where old/new is exactly 1/2, zigzag position 4 of these tables; odd c are ties, rounded to even).
    src = _synth(3, 96, 64, 60, O.S444)
    out = rc.requantize(src, 30)
    old = src.qtables[0][jpeg_scan.ZIGZAG]
    new = out.qtables[out.comps[0].tq][jpeg_scan.ZIGZAG]
    assert np.all(new >= old) and np.any(new > old)
    co = src.comps[0].coefs.reshape(-1, 64)
    ro = out.comps[0].coefs.reshape(-1, 64)
    half = np.where(2 * old.astype(int) == new.astype(int))[0]
    assert len(half) > 0
    for k in half:
        c = co[:, k].astype(np.int64)
        odd = c % 2 != 0
        want = np.where(odd, np.where((c // 2) % 2 == 0, c // 2, c // 2 + 1), c // 2)  # c/2 rounded half to even
        np.testing.assert_array_equal(ro[:, k], want)

"""Pins the CPU oracle (oracle/oracle.c) to the reference's own golden data and to closed forms.

Every oracle function is checked here by something other than a retyped formula:
  * ssimulacra2: the 54 py-ssimulacra2 pairs of fixtures/ssim2/expected.csv (cli/tests/ssimulacra2.rs:7-38) at the
    reference's 1e-4 tolerance, identity = exactly 100, and the invariants of core/src/ssimulacra2.rs:541-633;
  * build_coefs: the first frame of three hierarchical fixtures, which the reference wrote as
    write_jpeg(build_coefs(picture, q85, trellis off)) (jpeg/src/hierarchical.rs:209-225), coefficient for
    coefficient, plus the DQT bytes; DCT agreement with a naive f64 DCT (jpeg/src/dct.rs:167-202);
  * reconstruct: DC-only closed form, a naive f64 inverse DCT, and libjpeg (through Pillow) decoding the
    reference's baseline fixtures and files written from oracle coefficients (4:4:4, 4:2:2, 4:2:0, odd sizes);
  * find_quality: bisection against exhaustive search on monotone score tables, and jpeg/tests/jpeg.rs:97-119.
"""
import csv
import io
import math
import os

import numpy as np
import pytest
from PIL import Image

from oracle import oracle as O
import hier_inputs
import jpeg_scan
import jpeg_write

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def _png(name):
    return np.asarray(Image.open(os.path.join(GOLD, name)).convert("RGB"))


def _pairs():
    with open(os.path.join(GOLD, "ssim2", "expected.csv")) as f:
        return [r for r in csv.reader(f) if r]


# ---------------------------------------------------------------------------------------------- SSIMULACRA2

def test_ssimulacra2_matches_every_golden_pair():
    rows = _pairs()
    assert len(rows) >= 50
    worst = 0.0
    for name, a, b, want in rows:
        got = O.ssimulacra2(_png("ssim2/" + a), _png("ssim2/" + b))
        assert math.isfinite(got)
        worst = max(worst, abs(got - float(want)))
        assert abs(got - float(want)) <= 1e-4, (name, got, want)
    # The reference reports its own worst case as about 7e-6 (cli/tests/ssimulacra2.rs:7).
    assert worst < 2e-5


def test_ssimulacra2_identity_is_exactly_100():
    img = _png("ssim2/identical_96x72.ref.png")
    assert O.ssimulacra2(img, img) == 100.0


def _noise_img(w, h, seed):
    rng = np.random.default_rng(seed)
    return rng.integers(0, 256, (h, w, 3), dtype=np.uint8)


def test_ssimulacra2_monotone_in_noise_and_repeatable():
    base = _png("ssim2/size_129x97.ref.png").astype(np.int32)
    rng = np.random.default_rng(3)
    noise = rng.standard_normal(base.shape)
    scores = []
    for amp in (2, 6, 18, 50):
        d = np.clip(base + np.round(noise * amp), 0, 255).astype(np.uint8)
        scores.append(O.ssimulacra2(base.astype(np.uint8), d))
    assert all(scores[i] > scores[i + 1] for i in range(3)), scores
    d = np.clip(base + np.round(noise * 6), 0, 255).astype(np.uint8)
    assert O.ssimulacra2(base.astype(np.uint8), d) == O.ssimulacra2(base.astype(np.uint8), d)


def test_ssimulacra2_gray_equals_replicated_rgb():
    rng = np.random.default_rng(5)
    a = rng.integers(0, 256, (40, 33), dtype=np.uint8)
    b = np.clip(a.astype(np.int32) + rng.integers(-9, 10, a.shape), 0, 255).astype(np.uint8)
    g = O.ssimulacra2(a, b)
    c = O.ssimulacra2(np.repeat(a[:, :, None], 3, 2), np.repeat(b[:, :, None], 3, 2))
    assert g == c


def test_ssimulacra2_rejects_small_and_mismatched():
    a = np.zeros((7, 30, 3), np.uint8)
    with pytest.raises(ValueError):
        O.ssimulacra2(a, a)
    with pytest.raises(ValueError):
        O.ssimulacra2(np.zeros((8, 8, 3), np.uint8), np.zeros((8, 9, 3), np.uint8))


def test_ssimulacra2_scale_count_follows_the_short_side():
    # A scale is added only while both sides are >= 8 (sides halve rounding up): 8x8 -> 1 scale, 15x15 -> 2,
    # 31x65 -> 3 (31,16,8), 256x144 -> 5 (144 ... 9, then 5), 320x256 -> 6.
    for (w, h, n) in [(8, 8, 1), (15, 15, 2), (16, 16, 2), (31, 65, 3), (64, 64, 4), (256, 144, 5), (320, 256, 6)]:
        a = _noise_img(w, h, 1)
        b = _noise_img(w, h, 2)
        _, stats = O.ssimulacra2(a, b, with_stats=True)
        assert stats.shape[0] == n, (w, h)


# ---------------------------------------------------------------------------------------------- DCT KATs

def _naive_dct(b):
    out = np.zeros((8, 8))
    for u in range(8):
        for v in range(8):
            s = 0.0
            for y in range(8):
                for x in range(8):
                    s += b[y, x] * math.cos((2 * x + 1) * u * math.pi / 16) * math.cos((2 * y + 1) * v * math.pi / 16)
            a = lambda k: 1 / math.sqrt(2) if k == 0 else 1.0
            out[u, v] = 0.25 * a(u) * a(v) * s
    return out


AAN = np.array([1.0, 1.3870398, 1.306563, 1.1758756, 1.0, 0.78569495, 0.5411961, 0.27589938], np.float32)


def _lcg_block(seed):
    s = seed
    b = np.zeros((8, 8), np.float32)
    for y in range(8):
        for x in range(8):
            s = (s * 1664525 + 1013904223) & 0xFFFFFFFF
            b[y, x] = float((s >> 24) - 128)
    return b


def test_round_f32_known_answers():
    assert [O.round_f32(v) for v in (2.4, 2.6, -2.6, -0.4, 0.5, 1.5, 2.5, -0.5, -1.5)] == [2, 3, -3, 0, 0, 2, 2, 0, -2]


def test_fdct_matches_naive_dct():
    for seed in range(20):
        b = _lcg_block(seed)
        want = _naive_dct(b.astype(np.float64))
        got = O.fdct(b)
        for u in range(8):
            for v in range(8):
                g = got[u, v] / (8.0 * AAN[u] * AAN[v])
                assert abs(g - want[u, v]) < 1e-2


def test_idct_inverts_naive_dct():
    for seed in range(20):
        b = _lcg_block(seed + 100)
        c = _naive_dct(b.astype(np.float64))
        inp = np.zeros((8, 8), np.float32)
        for v in range(8):
            for u in range(8):
                inp[v, u] = c[u, v] * AAN[u] * AAN[v]
        out = O.idct(inp)
        assert np.abs(out - b).max() < 1e-2


def test_flat_block_closed_forms():
    # fdct of a constant c is 64c at DC and 0 elsewhere; the quantised DC is round(8c / q00).
    for q in (30, 75, 95):
        qt = O.scale_table(q)
        for level in (0, 17, 128, 200, 255):
            img = np.full((8, 8), level, np.uint8)
            ci = O.build_coefs(img, q)
            co = ci.comps[0].coefs
            c = level - 128
            assert abs(int(co[0]) - 8 * c / int(qt[0])) <= 0.5 + 1e-6
            assert np.all(co[1:] == 0)
            # DC-only inverse: every pixel is clamp(round(dc*q00/8) + 128).
            rec = O.reconstruct(ci)
            want = min(255, max(0, O.round_f32(np.float32(co[0] * int(qt[0])) / np.float32(8)) + 128))
            assert np.all(rec == want)


def test_scale_table_known_values():
    # IJG scaling of the base table: q50 is the base itself, q100 all ones; q85 is what the hierarchical
    # fixtures carry in their DQT (checked against the file bytes below).
    assert O.scale_table(100).tolist() == [1] * 64
    base = O.scale_table(50)
    assert base[:8].tolist() == [16, 16, 16, 18, 25, 37, 56, 85]
    assert O.scale_table(1).max() == 255


# ---------------------------------------------------------------------------------------------- build_coefs goldens

def _hier_frame(name):
    with open(os.path.join(GOLD, "hier", name), "rb") as f:
        return jpeg_scan.first_frame(f.read())


# The hierarchical fixtures were written when the encoder still quantised toward the Annex K tables (their DQT
# holds Annex K scaled to q85, not PERCEPTUAL_Q). Every pixel->coefficient step is independent of the table's
# values, so the frames are compared with build_coefs run on the tables the file carries, and the tables
# themselves pin the IJG scaling formula (jpeg/src/quality.rs:6-14) applied to the Annex K bases.
STD_LUMA_Q = [16, 11, 10, 16, 24, 40, 51, 61, 12, 12, 14, 19, 26, 58, 60, 55, 14, 13, 16, 24, 40, 57, 69, 56, 14, 17,
              22, 29, 51, 87, 80, 62, 18, 22, 37, 56, 68, 109, 103, 77, 24, 35, 55, 64, 81, 104, 113, 92, 49, 64, 78,
              87, 103, 121, 120, 101, 72, 92, 95, 98, 112, 100, 103, 99]
STD_CHROMA_Q = [17, 18, 24, 47] + [99] * 4 + [18, 21, 26, 66] + [99] * 4 + [24, 26, 56] + [99] * 5 + [47, 66] + \
    [99] * 38


def _compare_frame(frame, img, subsampling):
    tabs = frame["qtables"]
    np.testing.assert_array_equal(tabs[0], O.scale_table(85, STD_LUMA_Q))
    if 1 in tabs:
        np.testing.assert_array_equal(tabs[1], O.scale_table(85, STD_CHROMA_Q))
    ci = O.build_coefs(img, 85, subsampling, tables=(tabs[0], tabs.get(1)))
    assert (frame["width"], frame["height"]) == (ci.width, ci.height)
    assert len(frame["comps"]) == len(ci.comps)
    for fc, cc in zip(frame["comps"], ci.comps):
        assert (fc["h"], fc["v"], fc["blocks_w"]) == (cc.h, cc.v, cc.blocks_w)
        np.testing.assert_array_equal(fc["coefs"], cc.coefs)


def test_build_coefs_matches_hierarchical_colour_frame():
    # hier_dct_huffman.jpg: picture(64,48,Rgb) shrunk twice to 16x12, q85, 4:2:0 (pads 12 -> 16 rows).
    img = hier_inputs.shrink(hier_inputs.shrink(hier_inputs.picture(64, 48, 3)))
    assert img.shape == (12, 16, 3)
    _compare_frame(_hier_frame("hier_dct_huffman.jpg"), img, O.S420)


def test_build_coefs_matches_hierarchical_gray_frame():
    img = hier_inputs.shrink(hier_inputs.shrink(hier_inputs.picture(61, 47, 1)))
    assert img.shape == (12, 16)
    _compare_frame(_hier_frame("hier_gray_dct_then_lossless.jpg"), img, O.S444)


def test_build_coefs_matches_hierarchical_refine_frame_with_edge_padding():
    # Shape::Refine keeps 61x47: the encoder replicates the last column and row out to 64x48.
    img = hier_inputs.picture(61, 47, 1)
    fr = _hier_frame("hier_gray_refine_dct.jpg")
    _compare_frame(fr, img, O.S444)
    assert fr["dqt_raw"][0] == bytes(O.scale_table(85, STD_LUMA_Q)[jpeg_scan.ZIGZAG])


# ---------------------------------------------------------------------------------------------- reconstruct

def _libjpeg_decode(data):
    return np.asarray(Image.open(io.BytesIO(data)))


def _coefimage_from_frame(fr):
    comps = [O.CoefComponent(c["id"], c["h"], c["v"], c["tq"], c["blocks_w"], c["coefs"]) for c in fr["comps"]]
    qts = [fr["qtables"].get(i) for i in range(4)]
    return O.CoefImage(fr["width"], fr["height"], comps, qts)


@pytest.mark.parametrize("name", ["rgb8_420.jpg", "rgb8_444.jpg", "gray8.jpg"])
def test_reconstruct_agrees_with_libjpeg_on_reference_files(name):
    # The reference checks reconstruct == its own decoder on these files (jpeg/tests/jpeg.rs:121-130). Here the
    # independent decoder is libjpeg (integer IDCT, same fancy upsampling and colour constants), so samples may
    # differ by IDCT rounding only.
    with open(os.path.join(GOLD, name), "rb") as f:
        data = f.read()
    ours = O.reconstruct(_coefimage_from_frame(jpeg_scan.first_frame(data))).astype(np.int32)
    theirs = _libjpeg_decode(data).astype(np.int32)
    d = np.abs(ours - theirs)
    assert d.max() <= 3 and d.mean() < 0.35, (d.max(), d.mean())


def _adobe_transform(data):
    """The transform byte of a file's Adobe APP14 segment (T.872 6.5.3), or -1 without one."""
    pos = 2
    while pos + 4 <= len(data) and data[pos] == 0xFF and data[pos + 1] != 0xDA:
        length = data[pos + 2] << 8 | data[pos + 3]
        body = data[pos + 4: pos + 2 + length]
        if data[pos + 1] == 0xEE and body[:5] == b"Adobe" and len(body) >= 12:
            return body[11]
        pos += 2 + length
    return -1


@pytest.mark.parametrize("name,expected", [("cmyk.jpg", "cmyk_expected.png"), ("ycck.jpg", "ycck_expected.png"),
                                           ("ycck_444.jpg", "ycck_444_expected.png")])
def test_four_component_reconstruct_matches_the_reference_fixtures(name, expected):
    # cli/tests/cmyk.rs:14-27,58-72: the expected PNGs are Pillow's (libjpeg's) RGB conversion of the files, and the
    # reference's decode must be within a mean of 2 levels of them. cmyk_to_rgb (decode.rs:1454-1478) with the
    # file's Adobe transform (0: inverted CMYK, 2: YCCK).
    with open(os.path.join(GOLD, name), "rb") as f:
        data = f.read()
    ci = _coefimage_from_frame(jpeg_scan.first_frame(data))
    ci.adobe_transform = _adobe_transform(data)
    assert ci.adobe_transform in (0, 2) and len(ci.comps) == 4
    ours = O.reconstruct(ci).astype(np.int32)
    want = np.asarray(Image.open(os.path.join(GOLD, expected)).convert("RGB")).astype(np.int32)
    d = np.abs(ours - want)
    assert d.mean() < 2.0 and d.max() <= 4, (d.mean(), d.max())
    # A plausible slip -- the ink amounts read without the inversion an Adobe marker means -- is far off.
    ci.adobe_transform = -1
    assert np.abs(O.reconstruct(ci).astype(np.int32) - want).mean() > 20


@pytest.mark.parametrize("w,h", [(67, 45), (203, 77), (33, 31), (17, 9)])
@pytest.mark.parametrize("subs", [O.S444, O.S422, O.S420])
def test_reconstruct_agrees_with_libjpeg_on_oracle_coefficients(w, h, subs):
    img = _png("rgb8.png") if (w, h) == (67, 45) else _smooth(w, h)
    ci = O.build_coefs(img, 88, subs)
    data = jpeg_write.baseline(ci)
    ours = O.reconstruct(ci).astype(np.int32)
    theirs = _libjpeg_decode(data).astype(np.int32)
    d = np.abs(ours - theirs)
    assert d.max() <= 3 and d.mean() < 0.35, (d.max(), d.mean())


def _smooth(w, h):
    # smooth_image of jpeg/tests/jpeg.rs:13-24, plus a little texture.
    y, x = np.mgrid[0:h, 0:w]
    out = np.stack([(x * 3 + y * 2 + c * 40) % 256 for c in range(3)], -1)
    return out.astype(np.uint8)


def _psnr(a, b):
    mse = np.mean((a.astype(np.float64) - b.astype(np.float64)) ** 2)
    return 99.0 if mse == 0 else 10 * math.log10(255.0 * 255.0 / mse)


@pytest.mark.parametrize("q,floor", [(30, 19.0), (75, 25.0), (95, 34.5)])
def test_round_trip_psnr_floors(q, floor):
    # jpeg/tests/jpeg.rs:26-59: 4:4:4 and gray fidelity floors on the rgb8/gray8 fixtures.
    rgb = _png("rgb8.png")
    gray = np.asarray(Image.open(os.path.join(GOLD, "gray8.png")).convert("L"))
    assert _psnr(O.reconstruct(O.build_coefs(rgb, q, O.S444)), rgb) > floor
    assert _psnr(O.reconstruct(O.build_coefs(gray, q, O.S444)), gray) > floor


def test_s422_fidelity():
    # jpeg/tests/jpeg.rs:132-152
    w, h = 203, 77
    y, x = np.mgrid[0:h, 0:w]
    img = np.stack([(x * 5 // 4) % 256, (y * 3) % 256, (128 + x // 4 + y) % 256], -1).astype(np.uint8)
    ci = O.build_coefs(img, 92, O.S422)
    assert (ci.comps[0].h, ci.comps[0].v) == (2, 1)
    assert _psnr(O.reconstruct(ci), img) > 30.0


# ---------------------------------------------------------------------------------------------- search

def test_bisection_equals_exhaustive_search_on_monotone_tables():
    rng = np.random.default_rng(11)
    for _ in range(300):
        steps = np.sort(rng.uniform(-20, 100, 100))
        target = float(rng.uniform(-30, 110))
        q, s = O.bisect_table(steps, target, 100)
        passing = [i + 1 for i in range(100) if steps[i] >= target]
        want = passing[0] if passing else 100
        assert q == want
        assert s == steps[q - 1]


def test_bisection_on_non_monotone_table_returns_a_passing_quality():
    # target.rs:18-22: a non-monotone stretch can only raise the answer, never return a failing one
    # (unless even max_quality fails).
    rng = np.random.default_rng(12)
    for _ in range(300):
        steps = np.sort(rng.uniform(0, 100, 100)) + rng.normal(0, 3, 100)
        target = float(rng.uniform(0, 100))
        q, s = O.bisect_table(steps, target, 100)
        assert s >= target or q == 100


def test_search_meets_target_and_lower_quality_misses():
    img = _png("ssim2/size_129x97.ref.png")
    last = 0
    for target in (40.0, 60.0, 75.0, 85.0):
        q, s, steps = O.find_quality(img, target, 100, O.S420)
        assert s >= target or q == 100
        assert q >= last
        last = q
        assert steps <= 9
        if q > 1:
            ci = O.build_coefs(img, q - 1, O.S420)
            assert O.ssimulacra2(img, O.reconstruct(ci)) < target + 0.5

"""The host JPEG reader (rcg_jpeg_parse, rcg_jpeg_read_coefs) against what pins it.

* The reference's own pair rgb8_prog.jpg / rgb8_opt.jpg: a progressive file with successive approximation and the
  baseline file it was transcoded from hold the same coefficients (jpeg/tests/jpeg_decode.rs:6,77 decode both).
* Baseline fixtures (RGB 4:2:0/4:4:4, gray, CMYK, YCCK): the coefficients equal those of the test suite's own
  baseline Huffman reader, which the writer tests pin against libjpeg.
* libjpeg (through Pillow): progressive, restart-interval and optimised-table encodes of one image hold exactly the
  coefficients of its plain baseline encode, for every subsampling and for gray.
* The entropy writer's baseline and progressive files read back to the coefficients written.
* The quality estimate (jpeg/src/quality.rs:17-22): IJG tables scaled to quality q estimate q; 12-bit and lossless
  files report 100 (jpeg/tests/arith.rs:168).
* EXIF orientation (decode.rs:572-620) and the reference's error behaviour: a flood of repeated scans is refused,
  a few repeats are fine (jpeg_decode.rs:105-134); frames the reader does not decode are Unsupported.
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


def _read(name):
    with open(os.path.join(GOLD, name), "rb") as f:
        return f.read()


def _same(a, b):
    return len(a.comps) == len(b.comps) and all(np.array_equal(x.coefs, y.coefs) for x, y in zip(a.comps, b.comps))


BASELINE = ["rgb8.jpg", "rgb8_420.jpg", "rgb8_444.jpg", "gray8.jpg", "rgb8_opt.jpg", "cmyk.jpg", "ycck.jpg",
            "ycck_444.jpg"]


@pytest.mark.parametrize("name", BASELINE)
def test_baseline_fixtures_match_the_independent_reader(name):
    data = _read(name)
    info = rc.jpeg_info(data)
    ci = rc.read_jpeg(data, info)
    fr = jpeg_scan.first_frame(data)
    assert (info.width, info.height) == (fr["width"], fr["height"])
    for got, want in zip(ci.comps, fr["comps"]):
        assert (got.id, got.h, got.v, got.tq, got.blocks_w) == (want["id"], want["h"], want["v"], want["tq"],
                                                                 want["blocks_w"])
        np.testing.assert_array_equal(got.coefs, np.asarray(want["coefs"], np.int16))
    for s, t in fr["qtables"].items():
        np.testing.assert_array_equal(ci.qtables[s], t)
    assert not info.progressive and not info.coded_rgb and info.precision == 8


def test_the_progressive_fixture_holds_its_baseline_twins_coefficients():
    prog = _read("rgb8_prog.jpg")
    info = rc.jpeg_info(prog)
    assert info.progressive
    assert _same(rc.read_jpeg(prog), rc.read_jpeg(_read("rgb8_opt.jpg")))
    # And it decodes (oracle reconstruction) to what libjpeg shows, within the IDCT rounding of another decoder.
    d = np.abs(O.reconstruct(rc.read_jpeg(prog)).astype(int) - np.asarray(Image.open(io.BytesIO(prog))).astype(int))
    assert d.max() <= 3


def test_four_component_facts():
    assert rc.jpeg_info(_read("cmyk.jpg")).adobe_transform == 0
    y = rc.jpeg_info(_read("ycck.jpg"))
    assert y.cmyk and y.adobe_transform == 2 and y.sampling == [(2, 2), (1, 1), (1, 1), (2, 2)]


def _pil(img, **kw):
    b = io.BytesIO()
    Image.fromarray(img).save(b, "JPEG", **kw)
    return b.getvalue()


VARIANTS = [dict(progressive=True), dict(progressive=True, restart_marker_blocks=3), dict(restart_marker_rows=1),
            dict(optimize=True, restart_marker_blocks=7), dict(progressive=True, optimize=True)]


@pytest.mark.parametrize("sub", [0, 1, 2, "gray"])
def test_libjpeg_variants_hold_the_baseline_coefficients(sub):
    img = rc.synth_rgb8(5, 203, 77)
    kw = {}
    if sub == "gray":
        img = np.ascontiguousarray(img[:, :, 1])
    else:
        kw["subsampling"] = sub
    base = rc.read_jpeg(_pil(img, quality=83, **kw))
    for v in VARIANTS:
        data = _pil(img, quality=83, **kw, **v)
        if "restart_marker_blocks" in v or "restart_marker_rows" in v:
            assert any(data[i] == 0xFF and 0xD0 <= data[i + 1] <= 0xD7 for i in range(len(data) - 1))
        assert rc.jpeg_info(data).progressive == bool(v.get("progressive"))
        assert _same(rc.read_jpeg(data), base), v


@pytest.mark.parametrize("w,h,ch,subs", [(67, 45, 3, O.S420), (33, 17, 3, O.S444), (50, 31, 3, O.S422),
                                         (23, 9, 1, O.S444), (129, 70, 3, O.S420)])
def test_writer_output_reads_back(w, h, ch, subs):
    rng = np.random.default_rng(w)
    img = rng.integers(0, 256, size=(h, w, ch) if ch == 3 else (h, w), dtype=np.uint8)
    ci = O.build_coefs(img, 80, subs)
    ours = rc.CoefImage(ci.width, ci.height,
                        [rc.CoefComponent(c.id, c.h, c.v, c.tq, c.blocks_w, np.asarray(c.coefs, np.int16))
                         for c in ci.comps], list(ci.qtables) + [None] * (4 - len(ci.qtables)))
    for mode in (rc.WRITE_BASELINE, rc.WRITE_PROGRESSIVE):
        back = rc.read_jpeg(rc.write_jpeg(ours, mode))
        for s, t in enumerate(ours.qtables):
            if t is not None:
                np.testing.assert_array_equal(back.qtables[s], t)
        if mode == rc.WRITE_BASELINE:
            assert _same(back, ours)
            continue
        # Progressive AC scans hold one component each and code only the blocks that cover its samples (T.81
        # A.2.2); the MCU-padding blocks keep their DC (from the interleaved DC scan) and no AC.
        hmax = max(c.h for c in ours.comps)
        vmax = max(c.v for c in ours.comps)
        for a, b in zip(ours.comps, back.comps):
            bw = a.blocks_w
            bh = len(a.coefs) // 64 // bw
            rw = -(-(-(-w * a.h // hmax)) // 8)
            rh = -(-(-(-h * a.v // vmax)) // 8)
            A = a.coefs.reshape(bh, bw, 64)
            B = b.coefs.reshape(bh, bw, 64)
            np.testing.assert_array_equal(B[:rh, :rw], A[:rh, :rw])
            np.testing.assert_array_equal(B[..., 0], A[..., 0])
            pad = np.ones((bh, bw), bool)
            pad[:rh, :rw] = False
            assert not B[pad][:, 1:].any()


@pytest.mark.parametrize("q", [50, 60, 75, 83, 90, 95])
def test_quality_estimate_of_ijg_tables(q):
    img = rc.synth_rgb8(1, 64, 64)
    assert rc.jpeg_info(_pil(img, quality=q)).source_quality == q


def test_quality_estimate_of_frames_without_meaningful_tables():
    assert rc.jpeg_info(_read("jpeg12_444.jpg")).source_quality == 100
    assert rc.jpeg_info(_read("lossless_gray_p8_pred1.jpg")).source_quality == 100


@pytest.mark.parametrize("orientation", [1, 2, 3, 4, 5, 6, 7, 8])
def test_exif_orientation(orientation):
    ex = Image.Exif()
    ex[0x0112] = orientation
    img = rc.synth_rgb8(2, 24, 16)
    assert rc.jpeg_info(_pil(img, quality=90, exif=ex.tobytes())).orientation == orientation


def test_no_exif_means_orientation_1():
    assert rc.jpeg_info(_read("rgb8.jpg")).orientation == 1


@pytest.mark.parametrize("name,flag", [("arith_420_seq.jpg", "arithmetic"), ("jpeg12_444.jpg", None),
                                       ("lossless_gray_p8_pred1.jpg", "lossless"),
                                       ("hier_dct_progressive.jpg", "hierarchical")])
def test_frames_the_reader_does_not_decode_are_described_and_unsupported(name, flag):
    data = _read(name)
    info = rc.jpeg_info(data)
    if flag:
        assert getattr(info, flag)
    else:
        assert info.precision == 12
    with pytest.raises(rc.RcgError) as e:
        rc.read_jpeg(data, info)
    assert e.value.status == rc.UNSUPPORTED


def _with_repeated_scan(jpeg, copies):
    """The third scan of a progressive file (a first AC pass, which may legally be decoded again) repeated."""
    soss = [i for i in range(len(jpeg) - 1) if jpeg[i] == 0xFF and jpeg[i + 1] == 0xDA]
    sos = soss[2]
    end = sos + 2 + (jpeg[sos + 2] << 8 | jpeg[sos + 3])
    while not (jpeg[end] == 0xFF and jpeg[end + 1] != 0 and not 0xD0 <= jpeg[end + 1] <= 0xD7):
        end += 1
    return jpeg[:end] + jpeg[sos:end] * copies + jpeg[end:]


def test_a_flood_of_repeated_scans_is_refused_but_a_few_repeats_are_fine():
    prog = _read("rgb8_prog.jpg")
    assert _same(rc.read_jpeg(_with_repeated_scan(prog, 5)), rc.read_jpeg(prog))
    with pytest.raises(rc.RcgError) as e:
        rc.jpeg_info(_with_repeated_scan(prog, 1500))
    assert "too many scans" in str(e.value)


def test_malformed_files_are_invalid():
    good = _read("rgb8.jpg")
    for bad in (b"", b"\x89PNG\r\n\x1a\n", good[:2], good[:300], b"\xff\xd8\xff\xd9"):
        with pytest.raises(rc.RcgError) as e:
            rc.jpeg_info(bad)
        assert e.value.status == rc.INVALID, bad[:8]


def test_truncated_entropy_data_decodes_with_zero_bits():
    # The bit reader feeds zeros past the end of the data (decode.rs:31-45); a file cut inside its last scan and
    # closed with EOI still reads, and its first blocks are the intact file's.
    good = _read("rgb8.jpg")
    cut = good[: len(good) - 200] + b"\xff\xd9"
    a, b = rc.read_jpeg(good), rc.read_jpeg(cut)
    np.testing.assert_array_equal(a.comps[0].coefs[:64 * 4], b.comps[0].coefs[:64 * 4])


@pytest.mark.parametrize("name", ["rgb8.jpg", "rgb8_prog.jpg", "ycck.jpg", "arith_420_seq.jpg"])
def test_hostile_inputs_fail_cleanly(name):
    # Truncations, byte flips and random runs over real files (cli/tests/hostile.rs feeds the reference the same
    # kind): every call returns OK, Invalid or Unsupported, and a file that parses reads into buffers of its size.
    rng = np.random.default_rng(len(name))
    src = np.frombuffer(_read(name), np.uint8)
    for it in range(150):
        d = src.copy()
        kind = it % 3
        if kind == 0:
            d = d[: int(rng.integers(0, len(d) + 1))]
        elif kind == 1:
            idx = rng.integers(0, len(d), size=int(rng.integers(1, 9)))
            d[idx] ^= rng.integers(1, 256, size=len(idx)).astype(np.uint8)
        else:
            a = int(rng.integers(0, len(d)))
            n = int(rng.integers(1, 64))
            d[a:a + n] = rng.integers(0, 256, size=len(d[a:a + n])).astype(np.uint8)
        data = d.tobytes()
        try:
            info = rc.jpeg_info(data)
        except rc.RcgError as e:
            assert e.status in (rc.INVALID, rc.UNSUPPORTED)
            continue
        if sum(bw * bh for bw, bh in info.blocks) > 1 << 22:
            continue
        try:
            ci = rc.read_jpeg(data, info)
        except rc.RcgError as e:
            assert e.status in (rc.INVALID, rc.UNSUPPORTED)
            continue
        assert all(len(c.coefs) == bw * bh * 64 for c, (bw, bh) in zip(ci.comps, info.blocks))

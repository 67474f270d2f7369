"""The command-line driver (python -m rust-compress_amd) against the reference CLI's own tests.

* cli/tests/ssimulacra2.rs:40-123: a JPEG source gets a quality and the report names the metric; a higher target
  never gives a smaller file; conflicting or out-of-range targets exit with status 2; an image under 8x8 is a clean
  per-file error naming "8x8".
* cli/tests/cmyk.rs:38-80: at -q 100 CMYK/YCCK files are transcoded losslessly with their Adobe flag; at -q 30 they
  are requantised to smaller files.
* The chosen quality equals the oracle's bisection with the reference's make (build_coefs with the trellis, or
  without it through the serving queue), and the written file holds exactly those coefficients.
* Host parts without a device: HBM admission (cli/src/budget.rs:128-170), alpha flattening (core/src/image.rs:
  160-190), orientation (image.rs:194-219), output naming (cli/src/collect.rs:85-117), report lines.
"""
import importlib
import io
import os
import subprocess
import sys
import threading
import time

import numpy as np
import pytest
from PIL import Image

from oracle import oracle as O

rc = importlib.import_module("rust-compress_amd")
cli = importlib.import_module("rust-compress_amd.cli")
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLD = os.path.join(ROOT, "tests", "golden")


def _put(d, name, fixture):
    p = os.path.join(d, name)
    with open(os.path.join(GOLD, fixture), "rb") as f, open(p, "wb") as g:
        g.write(f.read())
    return p


def _run(*args):
    env = dict(os.environ, PYTHONPATH=ROOT + os.pathsep + os.environ.get("PYTHONPATH", ""))
    return subprocess.run([sys.executable, "-m", "rust-compress_amd", *map(str, args)], capture_output=True, text=True,
                          env=env, timeout=600)


# ------------------------------------------------------------------------------------------------- no device

def test_conflicting_or_out_of_range_targets_are_rejected(tmp_path):
    pic = _put(tmp_path, "pic.png", "rgb8.png")
    assert _run(pic, "--target-ssim", "0.9", "--target-ssimulacra2", "70").returncode == 2
    assert _run(pic, "--target-ssimulacra2", "70", "--target-size", "5000").returncode == 2
    for bad in ("0", "101", "abc", "nan", "-5"):
        assert _run(pic, "--target-ssimulacra2", bad).returncode == 2, bad
    assert _run(pic, "-q", "0").returncode == 2
    assert _run(pic, "--background", "12345").returncode == 2


def test_an_image_under_8x8_is_a_clean_per_file_error(tmp_path):
    tiny = os.path.join(tmp_path, "tiny.png")
    Image.new("RGB", (4, 4), (200, 30, 30)).save(tiny)
    r = _run(tiny, "-f", "jpeg", "--target-ssimulacra2", "70", "--force")
    assert r.returncode == 1
    assert "8x8" in r.stderr + r.stdout
    assert "Traceback" not in r.stderr


def test_no_inputs_found_is_an_error(tmp_path):
    r = _run(os.path.join(tmp_path, "missing.png"))
    assert r.returncode == 1 and "no such file" in r.stderr


def test_other_formats_and_targets_are_reported_per_file(tmp_path, capsys):
    pic = _put(tmp_path, "pic.png", "rgb8.png")
    assert cli.main([pic, "--dry-run"]) == 1
    assert "JPEG only" in capsys.readouterr().err
    assert cli.main([pic, "-f", "jpeg", "--target-psnr", "40", "--dry-run"]) == 1
    assert "--target-psnr" in capsys.readouterr().err


@pytest.mark.parametrize("name", ["cmyk.jpg", "ycck.jpg", "ycck_444.jpg"])
def test_cmyk_at_full_quality_is_transcoded_losslessly_with_its_colour_flag(tmp_path, name):
    src = _put(tmp_path, "in.jpg", name)
    out = os.path.join(tmp_path, "out.jpg")
    assert cli.main([src, "-o", out, "-q", "100", "--quiet"]) == 0
    a = rc.read_jpeg(open(src, "rb").read())
    b_bytes = open(out, "rb").read()
    assert len(b_bytes) <= os.path.getsize(src)
    b = rc.read_jpeg(b_bytes)
    assert len(b.comps) == 4 and a.adobe_transform == b.adobe_transform
    for x, y in zip(a.comps, b.comps):
        np.testing.assert_array_equal(x.coefs, y.coefs)
    for s in range(4):
        assert (a.qtables[s] is None) == (b.qtables[s] is None)
        if a.qtables[s] is not None:
            np.testing.assert_array_equal(a.qtables[s], b.qtables[s])


@pytest.mark.parametrize("name", ["cmyk.jpg", "ycck.jpg", "ycck_444.jpg"])
@pytest.mark.parametrize("extra", [[], ["--no-trellis"]])
def test_cmyk_and_ycck_are_requantized_to_the_requested_quality(tmp_path, name, extra):
    src = _put(tmp_path, "in.jpg", name)
    out = os.path.join(tmp_path, "out.jpg")
    assert cli.main([src, "-o", out, "-q", "30", "--quiet"] + extra) == 0
    data = open(out, "rb").read()
    assert len(data) < os.path.getsize(src)
    # The file holds requantize(source, 30, trellis) exactly (the host product is pinned to the oracle).
    want = O.requantize(rc.read_jpeg(open(src, "rb").read()), 30, not extra)
    got = rc.read_jpeg(data)
    for x, y in zip(got.comps, want.comps):
        np.testing.assert_array_equal(x.coefs, y.coefs)


@pytest.mark.parametrize("name", ["rgb8_420.jpg", "rgb8_444.jpg", "gray8.jpg", "rgb8_prog.jpg"])
def test_a_jpeg_source_is_requantised_to_the_quality_and_never_grows(tmp_path, capsys, name):
    # Without a target a JPEG stays in the DCT domain (pipeline.rs:388-404): requantize(source, -q 80, trellis),
    # written if smaller, else the original bytes are copied (pipeline.rs:745-749). 4:4:4 sources below q90 would
    # be resampled under Auto only with a target search; here they stay too.
    src = _put(tmp_path, "in.jpg", name)
    out = os.path.join(tmp_path, "out.jpg")
    assert cli.main([src, "-o", out, "--subsampling", "444" if name == "rgb8_444.jpg" else "auto"]) == 0
    data = open(out, "rb").read()
    assert len(data) <= os.path.getsize(src)
    text = capsys.readouterr().out
    assert "in.jpg" in text and "1 file(s)" in text
    if "(no gain, kept original size)" in text:
        assert data == open(src, "rb").read()
        return
    want = O.requantize(rc.read_jpeg(open(src, "rb").read()), 80, True)
    for x, y in zip(rc.read_jpeg(data).comps, want.comps):
        np.testing.assert_array_equal(x.coefs, y.coefs)


def test_existing_output_needs_force(tmp_path, capsys):
    src = _put(tmp_path, "in.jpg", "rgb8_420.jpg")
    out = _put(tmp_path, "out.jpg", "gray8.jpg")
    assert cli.main([src, "-o", out]) == 1
    assert "already exists (use --force)" in capsys.readouterr().err
    assert cli.main([src, "-o", out, "--force"]) == 0


def test_hierarchical_and_lossless_sources_are_left_unchanged(tmp_path, capsys):
    for name, note in [("hier_dct_progressive.jpg", "hierarchical JPEG left unchanged"),
                       ("lossless_gray_p8_pred1.jpg", "lossless JPEG left unchanged")]:
        src = _put(tmp_path, name, name)
        out = os.path.join(tmp_path, "o_" + name)
        assert cli.main([src, "-o", out]) == 0
        assert open(out, "rb").read() == open(src, "rb").read()
        assert note in capsys.readouterr().out


def test_output_naming(tmp_path):
    a = _put(tmp_path, "a.png", "rgb8.png")
    b = _put(tmp_path, "b.JPG", "rgb8.jpg")
    sub = os.path.join(tmp_path, "sub")
    os.makedirs(sub)
    _put(sub, "c.jpg", "rgb8.jpg")
    _put(tmp_path, "old.min.jpg", "rgb8.jpg")
    args = cli.parser().parse_args([str(tmp_path), "-f", "jpeg"])
    jobs, problems = cli.collect(args)
    assert not problems
    assert sorted(jobs) == sorted([(a, os.path.join(tmp_path, "a.min.jpg")), (b, os.path.join(tmp_path, "b.min.JPG"))])
    args = cli.parser().parse_args([str(tmp_path), "-r", "-f", "jpeg", "-o", os.path.join(tmp_path, "out")])
    jobs, _ = cli.collect(args)
    assert (os.path.join(sub, "c.jpg"), os.path.join(tmp_path, "out", "sub", "c.jpg")) in jobs
    args = cli.parser().parse_args([a, "-f", "jpeg", "-o", os.path.join(tmp_path, "x.jpg")])
    assert cli.collect(args)[0] == [(a, os.path.join(tmp_path, "x.jpg"))]


def test_report_lines():
    assert cli.human(999) == "999 B" and cli.human(2048) == "2.0 KB" and cli.human(3 << 20) == "3.0 MB"
    assert cli.percent(200, 150) == "-25.0%" and cli.percent(0, 5) == "n/a"
    r = cli.Report(("a\x1b.png", "o.jpg"))
    r.status, r.in_size, r.out_size, r.note = "compressed", 2000, 1000, "q42, ssimulacra2 70.1234"
    assert r.line(False) == "a\\x1b.png  2.0 KB -> 1000 B  (-50.0%)  [q42, ssimulacra2 70.1234]"


def test_a_second_big_job_waits_until_the_first_is_done():
    budget = cli.HbmBudget(100)
    started = threading.Event()
    first = budget.acquire(60)
    first.__enter__()

    def second():
        with budget.acquire(60):
            started.set()

    t = threading.Thread(target=second)
    t.start()
    time.sleep(0.15)
    assert not started.is_set()
    first.__exit__(None, None, None)
    t.join(5)
    assert started.is_set()


def test_small_jobs_run_together_and_an_oversized_one_runs_alone():
    budget = cli.HbmBudget(100)
    with budget.acquire(30), budget.acquire(30), budget.acquire(30):
        assert budget.used == 90
    with budget.acquire(1000):
        assert budget.used == 100  # clamped to the whole budget: nothing else fits beside it
    assert budget.used == 0


def test_alpha_is_flattened_onto_the_background():
    rng = np.random.default_rng(3)
    rgba = rng.integers(0, 256, size=(9, 11, 4), dtype=np.uint8)
    buf = io.BytesIO()
    Image.fromarray(rgba, "RGBA").save(buf, "PNG")
    got = cli.decode_png(buf.getvalue(), (10, 200, 255))
    c = rgba[:, :, :3].astype(np.int64)
    a = rgba[:, :, 3:].astype(np.int64)
    bg = np.array([10, 200, 255])
    np.testing.assert_array_equal(got, (c * a + bg * (255 - a) + 127) // 255)
    # Fully opaque and fully transparent pixels are exact.
    assert np.array_equal(cli._blend(np.array([77], np.uint8), np.array([255], np.uint8), 3), [77])
    assert np.array_equal(cli._blend(np.array([77], np.uint8), np.array([0], np.uint8), 3), [3])


def test_sixteen_bit_gray_is_scaled_to_eight_bits():
    v = np.array([[0, 1, 128, 129, 32767, 32896, 65534, 65535]], np.uint16)
    buf = io.BytesIO()
    Image.fromarray(v).save(buf, "PNG")
    np.testing.assert_array_equal(cli.decode_png(buf.getvalue(), (255, 255, 255)),
                                  (v.astype(np.int64) * 255 + 32767) // 65535)


@pytest.mark.parametrize("orientation", range(1, 9))
def test_upright_matches_the_exif_mapping(orientation):
    # core/src/image.rs:204-212: source (sx, sy) lands at (dx, dy).
    h, w = 3, 5
    img = np.arange(h * w).reshape(h, w)
    got = cli.upright(img, orientation)
    dw, dh = (h, w) if orientation >= 5 else (w, h)
    assert got.shape == (dh, dw)
    for sy in range(h):
        for sx in range(w):
            dx, dy = {1: (sx, sy), 2: (w - 1 - sx, sy), 3: (w - 1 - sx, h - 1 - sy), 4: (sx, h - 1 - sy),
                      5: (sy, sx), 6: (h - 1 - sy, sx), 7: (h - 1 - sy, w - 1 - sx), 8: (sy, w - 1 - sx)}[orientation]
            assert got[dy, dx] == img[sy, sx]


# ------------------------------------------------------------------------------------------------- on the GPU

def _oracle_search(img, target, make):
    """jpeg/src/target.rs:37-55 with the oracle's reconstruct and FP64 SSIMULACRA2."""
    def score(ci):
        return O.ssimulacra2(img, O.reconstruct(ci))
    lo, hi = 1, 100
    best_q, best_s = hi, score(make(hi))
    if best_s < target:
        return best_q, best_s
    while lo < hi:
        mid = lo + (hi - lo) // 2
        s = score(make(mid))
        if s >= target:
            hi, best_q, best_s = mid, mid, s
        else:
            lo = mid + 1
    return best_q, best_s


def _trellis_make(img, subs):
    def make(q):
        ci, scaled = O.build_scaled(img, q, subs)
        if q >= 100:
            return ci
        t = O.trellis(ci, scaled)
        for c, coefs in zip(ci.comps, t):
            c.coefs = np.asarray(coefs, np.int16)
        return ci
    return make


def _note_quality(text):
    import re
    m = re.search(r"\[q(\d+), ssimulacra2 ([0-9.]+)\]", text)
    assert m, text
    return int(m.group(1)), float(m.group(2))


@pytest.mark.gpu
def test_jpeg_from_a_jpeg_source_picks_a_quality_and_names_the_metric(tmp_path):
    src = _put(tmp_path, "photo.jpg", "rgb8_444.jpg")
    out = os.path.join(tmp_path, "t.jpg")
    r = _run(src, "--target-ssimulacra2", "60", "-o", out)
    assert r.returncode == 0, r.stderr
    assert "ssimulacra2" in r.stdout
    assert Image.open(out).size == (67, 45)


@pytest.mark.gpu
@pytest.mark.parametrize("extra", [[], ["--no-trellis"]])
def test_jpeg_from_pixels_and_a_higher_target_never_gives_a_smaller_file(tmp_path, extra):
    src = _put(tmp_path, "pic.png", "rgb8.png")
    sizes = {}
    for t in ("30", "85"):
        out = os.path.join(tmp_path, f"o{t}.jpg")
        assert cli.main([src, "-f", "jpeg", "--target-ssimulacra2", t, "-o", out, "--quiet"] + extra) == 0
        sizes[t] = os.path.getsize(out)
    assert sizes["30"] <= sizes["85"], sizes


@pytest.mark.gpu
@pytest.mark.parametrize("trellis", [True, False])
@pytest.mark.parametrize("target", [55.0, 80.0])
def test_cli_choice_and_file_equal_the_oracle_search(tmp_path, capsys, trellis, target):
    src = _put(tmp_path, "pic.png", "rgb8.png")
    img = np.asarray(Image.open(src).convert("RGB"))
    out = os.path.join(tmp_path, "o.jpg")
    args = [src, "-f", "jpeg", "--target-ssimulacra2", str(target), "-o", out] + ([] if trellis else ["--no-trellis"])
    assert cli.main(args) == 0
    q, s = _note_quality(capsys.readouterr().out)
    make = _trellis_make(img, O.AUTO) if trellis else (lambda qq: O.build_coefs(img, qq, O.AUTO))
    want_q, want_s = _oracle_search(img, target, make)
    assert q == want_q and abs(s - want_s) < 0.02
    got = rc.read_jpeg(open(out, "rb").read())
    for x, y in zip(got.comps, make(want_q).comps):
        np.testing.assert_array_equal(x.coefs, np.asarray(y.coefs, np.int16))


@pytest.mark.gpu
@pytest.mark.parametrize("trellis", [True, False])
def test_jpeg_to_jpeg_target_search_equals_the_oracle(tmp_path, capsys, trellis):
    # rgb8_420.jpg is 4:2:0: it stays in the DCT domain (pipeline.rs:369-386).
    src = _put(tmp_path, "in.jpg", "rgb8_420.jpg")
    out = os.path.join(tmp_path, "o.jpg")
    assert cli.main([src, "--target-ssimulacra2", "50", "-o", out, "--force"] +
                    ([] if trellis else ["--no-trellis"])) == 0
    text = capsys.readouterr().out
    ci = rc.read_jpeg(open(src, "rb").read())
    ref = O.reconstruct(ci)
    want_q, _ = _oracle_search(ref, 50.0, lambda qq: O.requantize(ci, qq, trellis))
    if "(no gain, kept original size)" in text:
        assert open(out, "rb").read() == open(src, "rb").read()
    else:
        q, _ = _note_quality(text)
        assert q == want_q
        got = rc.read_jpeg(open(out, "rb").read())
        for x, y in zip(got.comps, O.requantize(ci, want_q, trellis).comps):
            np.testing.assert_array_equal(x.coefs, y.coefs)


@pytest.mark.gpu
@pytest.mark.parametrize("extra", [[], ["--no-trellis"]])
def test_many_files_in_one_run_get_what_each_gets_alone(tmp_path, capsys, extra):
    d = os.path.join(tmp_path, "in")
    os.makedirs(d)
    for i, (w, h) in enumerate([(67, 45), (128, 96), (40, 300), (250, 17), (96, 96), (9, 8)]):
        Image.fromarray(rc.synth_rgb8(70 + i, w, h)).save(os.path.join(d, f"f{i}.png"))
    out = os.path.join(tmp_path, "out")
    assert cli.main([d, "-f", "jpeg", "--target-ssimulacra2", "70", "-o", out, "-j", "6", "--contexts", "3"] + extra) == 0
    together = capsys.readouterr().out
    for i in range(6):
        src = os.path.join(d, f"f{i}.png")
        line = next(l for l in together.splitlines() if l.startswith(src))
        alone_out = os.path.join(tmp_path, f"alone{i}.jpg")
        assert cli.main([src, "-f", "jpeg", "--target-ssimulacra2", "70", "-o", alone_out] + extra) == 0
        alone = capsys.readouterr().out
        assert _note_quality(line)[0] == _note_quality(alone)[0]
        assert open(os.path.join(out, f"f{i}.jpg"), "rb").read() == open(alone_out, "rb").read()


@pytest.mark.gpu
def test_an_exif_orientation_is_baked_into_the_pixels(tmp_path, capsys):
    # pipeline.rs:366-420: a JPEG with an orientation to apply leaves the DCT domain; the decoded pixels are turned
    # upright (core/src/image.rs:194-219) and searched as pixels, so a 6 (90 degrees clockwise) swaps the sides.
    img = rc.synth_rgb8(31, 96, 64)
    ex = Image.Exif()
    ex[0x0112] = 6
    src = os.path.join(tmp_path, "rot.jpg")
    Image.fromarray(img).save(src, "JPEG", quality=98, exif=ex.tobytes())
    out = os.path.join(tmp_path, "o.jpg")
    assert cli.main([src, "--target-ssimulacra2", "60", "-o", out, "--no-trellis"]) == 0
    text = capsys.readouterr().out
    assert "(no gain" not in text  # a q98 source searched down to 60 shrinks, so the re-encode is written
    q, _ = _note_quality(text)
    assert Image.open(out).size == (64, 96)
    upright = np.ascontiguousarray(cli.upright(O.reconstruct(rc.read_jpeg(open(src, "rb").read())), 6))
    want_q, _ = _oracle_search(upright, 60.0, lambda qq: O.build_coefs(upright, qq, O.AUTO))
    assert q == want_q


def test_a_failing_serving_loop_fails_its_searches_instead_of_hanging(monkeypatch):
    class FailingServer:
        def __init__(self, *a, **k):
            pass

        def submit(self, *a, **k):
            pass

        def poll(self, *a, **k):
            time.sleep(0.05)
            raise rc.RcgError(rc.DEVICE, "poll failed")

        def close(self):
            pass

    monkeypatch.setattr(cli, "Server", FailingServer)
    gpu = cli.Gpu(0, 1, 1 << 30)
    results = []
    t = threading.Thread(target=lambda: results.append(gpu.search(np.zeros((8, 8, 3), np.uint8), 70.0, rc.AUTO)))
    t.start()
    t.join(10)
    assert not t.is_alive() and results and results[0][0] == rc.DEVICE
    with pytest.raises(rc.RcgError):
        gpu.search(np.zeros((8, 8, 3), np.uint8), 70.0, rc.AUTO)
    gpu.close()

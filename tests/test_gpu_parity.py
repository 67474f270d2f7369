"""GPU parity: the HIP path (through the C ABI) against the CPU oracle on the same inputs.

Integer and f32-DCT work (coefficients, reconstructed pixels) must be bit-exact. SSIMULACRA2 in FP32 must be within
0.02 of the oracle's f64 score (the precision bound this path states); the FP64 kernel within 1e-6. The search must
choose the oracle's quality."""
import csv
import os

import numpy as np
import pytest
from PIL import Image

from oracle import oracle as O
import jpeg_scan

pytestmark = pytest.mark.gpu
GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
TOL32 = 0.02
TOL64 = 1e-6


def _assert_coefs_equal(gpu, ref):
    assert len(gpu.comps) == len(ref.comps)
    for a, b in zip(gpu.comps, ref.comps):
        assert (a.h, a.v, a.blocks_w) == (b.h, b.v, b.blocks_w)
        assert a.coefs.shape == b.coefs.shape
        if not np.array_equal(a.coefs, b.coefs):
            bad = np.nonzero(a.coefs != b.coefs)[0]
            raise AssertionError(f"component {a.id}: {len(bad)} coefficients differ, first at {bad[:8]}")
    for s in (0, 1):
        if ref.qtables[s] is not None:
            np.testing.assert_array_equal(gpu.qtables[s], ref.qtables[s])


def _smooth(w, h, ch):
    y, x = np.mgrid[0:h, 0:w]
    out = np.stack([(x * 3 + y * 2 + c * 40) % 256 for c in range(ch)], -1).astype(np.uint8)
    return out[:, :, 0] if ch == 1 else out


# jpeg/tests/jpeg.rs:65-77 sizes, both colour types, all three subsamplings.
@pytest.mark.parametrize("w,h", [(1, 1), (7, 9), (8, 8), (9, 9), (17, 1), (1, 17), (33, 31), (250, 3), (67, 45),
                                 (513, 129)])
@pytest.mark.parametrize("ch", [1, 3])
@pytest.mark.parametrize("subs", [O.S444, O.S422, O.S420])
def test_build_coefs_and_reconstruct_odd_and_tiny(ctx, w, h, ch, subs):
    img = _smooth(w, h, ch)
    if ch == 3:
        img = np.ascontiguousarray(img)
    gpu = ctx.build_coefs(img, 90, subs)
    ref = O.build_coefs(img, 90, subs)
    _assert_coefs_equal(gpu, ref)
    np.testing.assert_array_equal(ctx.reconstruct(gpu), O.reconstruct(ref))


@pytest.mark.parametrize("q", [1, 30, 75, 95, 100])
def test_build_coefs_c1_512(ctx, rc, q):
    img = rc.synth_rgb8(1, 512, 512)
    gpu = ctx.build_coefs(img, q, rc.S420)
    ref = O.build_coefs(img, q, O.S420)
    _assert_coefs_equal(gpu, ref)
    np.testing.assert_array_equal(ctx.reconstruct(gpu), O.reconstruct(ref))


def test_build_coefs_and_reconstruct_4096_all_c2_cases(ctx, rc):
    img = rc.synth_rgb8(2, 4096, 4096)
    for subs in (rc.S420, rc.S444):
        for q in (50, 75, 90):
            gpu = ctx.build_coefs(img, q, subs)
            ref = O.build_coefs(img, q, subs)
            _assert_coefs_equal(gpu, ref)
            np.testing.assert_array_equal(ctx.reconstruct(gpu), O.reconstruct(ref))


def test_build_coefs_odd_3000x2001_auto(ctx, rc):
    img = rc.synth_rgb8(3, 3000, 2001)
    for q in (60, 95):
        gpu = ctx.build_coefs(img, q, rc.AUTO)
        ref = O.build_coefs(img, q, O.AUTO)
        _assert_coefs_equal(gpu, ref)
        np.testing.assert_array_equal(ctx.reconstruct(gpu), O.reconstruct(ref))


@pytest.mark.parametrize("name", ["rgb8_420.jpg", "rgb8_444.jpg", "gray8.jpg"])
def test_reconstruct_reference_files(ctx, rc, name):
    with open(os.path.join(GOLD, name), "rb") as f:
        fr = jpeg_scan.first_frame(f.read())
    comps = [rc.CoefComponent(c["id"], c["h"], c["v"], c["tq"], c["blocks_w"], c["coefs"]) for c in fr["comps"]]
    ci = rc.CoefImage(fr["width"], fr["height"], comps, [fr["qtables"].get(i) for i in range(4)])
    oc = O.CoefImage(fr["width"], fr["height"],
                     [O.CoefComponent(c.id, c.h, c.v, c.tq, c.blocks_w, c.coefs) for c in comps], ci.qtables)
    np.testing.assert_array_equal(ctx.reconstruct(ci), O.reconstruct(oc))


def _random_coef_image(rng, w, h, factors, ids, adobe, lib):
    """A coefficient image with random coefficients (DC and AC beyond what 8-bit samples need, so clamping is
    exercised) on build_coefs' MCU-padded block grids (jpeg/src/encode.rs:440-443), for any sampling factors."""
    hmax = max(f[0] for f in factors)
    vmax = max(f[1] for f in factors)
    mx, my = -(-w // (8 * hmax)), -(-h // (8 * vmax))
    comps, qt = [], [rng.integers(1, 64, 64).astype(np.uint8), rng.integers(1, 64, 64).astype(np.uint8)]
    for i, (ch, cv) in enumerate(factors):
        bw, bh = mx * ch, my * cv
        co = rng.integers(-120, 121, bw * bh * 64).astype(np.int16)
        co[::64] = rng.integers(-1100, 1100, bw * bh)
        comps.append(lib.CoefComponent(ids[i], ch, cv, 0 if i == 0 else 1, bw, co))
    return lib.CoefImage(w, h, comps, [qt[0], qt[1], None, None], adobe)


# The fused reconstruct kernel (Gray, 4:4:4, 4:2:2, 4:4:0, 4:2:0 with the luma at full resolution, YCbCr- and
# RGB-coded) and the two-kernel form it falls back to (mixed chroma factors, subsampled luma), at sizes that end
# inside a 256x64 tile, span several and straddle one chroma ring block, against the oracle (decode.rs:1390-1451).
@pytest.mark.parametrize("w,h", [(1, 1), (13, 7), (257, 65), (300, 131), (513, 200), (1030, 70)])
def test_reconstruct_every_layout_and_colour_coding(ctx, rc, w, h):
    rng = np.random.default_rng(w * 1000 + h)
    cases = [([(1, 1)], b"\x01", -1), ([(2, 2)], b"\x01", -1)]  # a gray file may declare 2x2 sampling
    for fy in ((1, 1), (2, 1), (1, 2), (2, 2)):
        cases += [([fy, (1, 1), (1, 1)], b"\x01\x02\x03", -1), ([fy, (1, 1), (1, 1)], b"RGB", -1),
                  ([fy, (1, 1), (1, 1)], b"\x01\x02\x03", 0)]
    cases += [([(2, 2), (1, 1), (2, 1)], b"\x01\x02\x03", -1), ([(1, 1), (2, 2), (2, 2)], b"\x01\x02\x03", -1)]
    for factors, ids, adobe in cases:
        dev = _random_coef_image(rng, w, h, factors, ids, adobe, rc)
        ora = O.CoefImage(w, h, [O.CoefComponent(c.id, c.h, c.v, c.tq, c.blocks_w, c.coefs) for c in dev.comps],
                          dev.qtables, adobe)
        got = ctx.reconstruct(dev)
        want = O.reconstruct(ora)
        assert np.array_equal(got, want), (factors, ids, adobe, np.argwhere(got != want)[:4])


def test_reconstruct_errors(ctx, rc):
    img = _smooth(32, 32, 3)
    ci = ctx.build_coefs(img, 80, rc.S420)
    bad = rc.CoefImage(ci.width, ci.height, ci.comps, [ci.qtables[0], None, None, None])
    with pytest.raises(rc.RcgError) as e:
        ctx.reconstruct(bad)
    assert e.value.status == rc.INVALID
    odd = rc.CoefImage(ci.width, ci.height, [ci.comps[0], rc.CoefComponent(2, 3, 1, 1, ci.comps[1].blocks_w,
                                                                           ci.comps[1].coefs), ci.comps[2]],
                       ci.qtables)
    with pytest.raises(rc.RcgError) as e:
        ctx.reconstruct(odd)
    assert e.value.status == rc.UNSUPPORTED


# ------------------------------------------------------------------------------------------------ SSIMULACRA2

def _png(name):
    return np.asarray(Image.open(os.path.join(GOLD, name)).convert("RGB"))


def test_ssimulacra2_golden_pairs_fp32_and_fp64(ctx, rc):
    rows = [r for r in csv.reader(open(os.path.join(GOLD, "ssim2", "expected.csv"))) if r]
    worst32 = worst64 = 0.0
    for name, a, b, want in rows:
        A, B = _png("ssim2/" + a), _png("ssim2/" + b)
        oracle = O.ssimulacra2(A, B)
        s32 = ctx.ssimulacra2(A, B, rc.FP32)
        s64 = ctx.ssimulacra2(A, B, rc.FP64)
        worst32 = max(worst32, abs(s32 - oracle))
        worst64 = max(worst64, abs(s64 - oracle))
        assert abs(s32 - oracle) <= TOL32, (name, s32, oracle)
        assert abs(s64 - oracle) <= TOL64, (name, s64, oracle)
        assert abs(s64 - float(want)) <= 1e-4, (name, s64, want)
    print(f"golden pairs: worst |fp32 - oracle| = {worst32:.3e}, worst |fp64 - oracle| = {worst64:.3e}")


def test_ssimulacra2_identity_is_100(ctx, rc):
    img = _png("ssim2/identical_96x72.ref.png")
    assert ctx.ssimulacra2(img, img, rc.FP64) == 100.0
    assert abs(ctx.ssimulacra2(img, img, rc.FP32) - 100.0) < 1e-9


def test_ssimulacra2_gray_input(ctx, rc):
    rng = np.random.default_rng(5)
    a = rng.integers(0, 256, (40, 33), dtype=np.uint8)
    b = np.clip(a.astype(np.int32) + rng.integers(-9, 10, a.shape), 0, 255).astype(np.uint8)
    assert abs(ctx.ssimulacra2(a, b, rc.FP64) - O.ssimulacra2(a, b)) <= TOL64
    assert abs(ctx.ssimulacra2(a, b, rc.FP32) - O.ssimulacra2(a, b)) <= TOL32


def test_ssimulacra2_errors(ctx, rc):
    with pytest.raises(rc.RcgError) as e:
        ctx.ssimulacra2(np.zeros((7, 30, 3), np.uint8), np.zeros((7, 30, 3), np.uint8))
    assert e.value.status == rc.INVALID
    with pytest.raises(rc.RcgError):
        ctx.ssimulacra2(np.zeros((8, 8, 3), np.uint8), np.zeros((8, 9, 3), np.uint8))


@pytest.mark.parametrize("w,h", [(2048, 2048), (3000, 2001)])
def test_ssimulacra2_precision_config(ctx, rc, w, h):
    # Distorted = reference decode at q60 4:2:0; per-scale and final |error| against the f64 oracle.
    img = rc.synth_rgb8(11, w, h)
    dist = O.reconstruct(O.build_coefs(img, 60, O.S420))
    want, wstats = O.ssimulacra2(img, dist, with_stats=True)
    got, gstats = ctx.ssimulacra2(img, dist, rc.FP32, with_stats=True)
    got64, g64 = ctx.ssimulacra2(img, dist, rc.FP64, with_stats=True)
    assert gstats.shape == wstats.shape
    per_scale = np.abs(gstats - wstats).max(axis=(1, 2))
    print(f"{w}x{h}: score {want:.6f}, fp32 |d| {abs(got - want):.3e}, fp64 |d| {abs(got64 - want):.3e}, "
          f"per-scale max |d stat| {per_scale}")
    assert abs(got - want) <= TOL32
    assert abs(got64 - want) <= TOL64
    assert np.abs(g64 - wstats).max() <= 1e-9


# ------------------------------------------------------------------------------------------------ search

@pytest.mark.parametrize("target", [60.0, 70.0, 80.0])
def test_find_quality_matches_oracle(ctx, rc, target):
    for seed, (w, h) in ((21, (256, 256)), (22, (301, 199))):
        img = rc.synth_rgb8(seed, w, h)
        found = ctx.find_quality(img, target, 100, rc.S420)
        q, s, steps = O.find_quality(img, target, 100, O.S420)
        assert found.quality == q, (seed, target, found.quality, q)
        assert abs(found.score - s) <= TOL32
        assert found.steps == steps
        _assert_coefs_equal(found.coefs, O.build_coefs(img, q, O.S420))


@pytest.mark.parametrize("w,h", [(9, 8), (8, 9), (24, 8), (8, 40), (17, 9), (67, 45), (1009, 1009)])
@pytest.mark.parametrize("target", [30.0, 97.0])
def test_auto_search_where_420_pads_more_than_444(ctx, rc, w, h, target):
    # Under Auto a search meets 4:4:4 (q >= 90) and 4:2:0 (q < 90); at these sizes the 4:2:0 luma grid, padded to
    # 16-pixel MCUs, has more blocks than the 4:4:4 one. Every buffer of the search must hold both.
    img = rc.synth_rgb8(w * 7 + h, w, h)
    for trellis in (False, True):
        found = ctx.find_quality(img, target, 100, rc.AUTO, trellis=trellis)
        if trellis:
            want, scores, make = _oracle_trellis_search(img, target, O.AUTO)
            s = scores[want]
        else:
            want, s, _ = O.find_quality(img, target, 100, O.AUTO)
            make = lambda q: O.build_coefs(img, q, O.AUTO)
        assert found.quality == want, (w, h, trellis, found.quality, want)
        assert abs(found.score - s) <= TOL32
        _assert_coefs_equal(found.coefs, make(want))


def test_find_quality_rejects_small(ctx, rc):
    with pytest.raises(rc.RcgError) as e:
        ctx.find_quality(np.zeros((7, 100, 3), np.uint8), 70.0)
    assert e.value.status == rc.UNSUPPORTED


CANDS = [(q, 3) for q in (40, 50, 60, 70, 80, 85, 90, 95)]


def test_score_batch_mixed_sizes_matches_oracle(ctx, rc):
    rng = np.random.default_rng(9)
    sizes = [(int(rng.integers(64, 700)), int(rng.integers(64, 700))) for _ in range(6)] + [(8, 8), (65, 9)]
    imgs = [rc.synth_rgb8(100 + i, w, h) for i, (w, h) in enumerate(sizes)]
    res = ctx.score_batch(imgs, CANDS, 70.0)
    for img, r in zip(imgs, res):
        assert r.status == 0
        oscores = []
        for k, (q, sub) in enumerate(CANDS):
            o = O.ssimulacra2(img, O.reconstruct(O.build_coefs(img, q, sub)))
            oscores.append(o)
            tol = TOL64 if r.settled[k] else TOL32
            assert abs(r.scores[k] - o) <= tol, (img.shape, q, r.scores[k], o)
        # The reference's search semantics over the candidate list, on the oracle's scores.
        qs = [q for q, _ in CANDS]
        order = np.argsort(qs, kind="stable")
        lo, hi = 0, len(qs) - 1
        best = hi
        if oscores[order[hi]] >= 70.0:
            while lo < hi:
                mid = lo + (hi - lo) // 2
                if oscores[order[mid]] >= 70.0:
                    hi = mid
                    best = mid
                else:
                    lo = mid + 1
        assert r.chosen_quality == qs[order[best]]


def test_score_batch_mixed_layouts_and_gray_match_oracle(ctx, rc):
    # Per-image candidate lists over every subsampling mode (Auto included), RGB8 and Gray8 sources of odd sizes in
    # one batch: every score within the bound of the f64 oracle's, the choice the reference's bisection over the
    # oracle's scores of the list.
    gray = np.ascontiguousarray(rc.synth_rgb8(42, 201, 97)[:, :, 1])
    imgs = [rc.synth_rgb8(41, 333, 250), gray, rc.synth_rgb8(43, 97, 64)]
    cands = [[(40, rc.S444), (60, rc.S422), (80, rc.S420), (92, rc.AUTO)],
             [(30, rc.S420), (55, rc.S444), (75, rc.AUTO), (95, rc.S422)],
             [(50, rc.AUTO), (70, rc.S422), (85, rc.S444), (90, rc.AUTO)]]
    res = ctx.score_batch(imgs, cands, 70.0)
    for img, cl, r in zip(imgs, cands, res):
        assert r.status == 0
        osc = [O.ssimulacra2(img, O.reconstruct(O.build_coefs(img, q, sub))) for q, sub in cl]
        for k, (q, sub) in enumerate(cl):
            assert abs(r.scores[k] - osc[k]) <= (TOL64 if r.settled[k] else TOL32), (img.shape, q, sub)
        lo, hi = 0, len(cl) - 1
        best = hi
        if osc[hi] >= 70.0:
            while lo < hi:
                mid = lo + (hi - lo) // 2
                if osc[mid] >= 70.0:
                    hi = mid
                    best = mid
                else:
                    lo = mid + 1
        assert r.chosen_quality == cl[best][0], (img.shape, [round(x, 3) for x in osc])


def test_score_batch_extreme_aspect_ratios_match_oracle(ctx, rc):
    # The longest side JPEG allows (encode.rs:417) and thin strips in one batch, through the column tiling's main and
    # tail tiles: scores within the bound of the oracle's, one or two scales each.
    sizes = [(65535, 9), (4097, 17), (9, 3001), (1050, 8), (8, 8)]
    imgs = [rc.synth_rgb8(80 + i, w, h) for i, (w, h) in enumerate(sizes)]
    cands = [(45, rc.S420), (75, rc.S444), (92, rc.AUTO)]
    res = ctx.score_batch(imgs, cands, 70.0)
    for img, r in zip(imgs, res):
        assert r.status == 0, img.shape
        for k, (q, sub) in enumerate(cands):
            want = O.ssimulacra2(img, O.reconstruct(O.build_coefs(img, q, sub)))
            assert abs(r.scores[k] - want) <= (TOL64 if r.settled[k] else TOL32), (img.shape, q, r.scores[k], want)


def test_batch_scores_do_not_depend_on_the_rest_of_the_batch(ctx, rc):
    # A launch over many images tiles the reference caches taller than one image alone does; the cache, and so every
    # score, must not depend on that tiling: each image's scores in a 24-image batch equal its scores alone, bit for bit.
    rng = np.random.default_rng(21)
    sizes = [(int(rng.integers(200, 1300)), int(rng.integers(64, 1100))) for _ in range(22)] + [(3000, 40), (9, 700)]
    imgs = [rc.synth_rgb8(300 + i, w, h) for i, (w, h) in enumerate(sizes)]
    together = ctx.score_batch(imgs, CANDS, 70.0)
    for img, r in zip(imgs, together):
        alone = ctx.score_batch([img], CANDS, 70.0)[0]
        assert r.status == 0 and alone.status == 0
        assert list(r.scores) == list(alone.scores), img.shape
        assert r.chosen_quality == alone.chosen_quality


def test_column_tilings_agree_across_paths_and_with_the_oracle(ctx, rc):
    # Scale-0 column tilings of the FP32 kernel: 1000x600 four 256-thread tiles and a 64-thread tail column,
    # 1050x600 four 256-thread tiles and a 128-thread tail column, 348x400 three 128-thread tiles, 52x200 one
    # 64-thread tile, 1400x380 six 256-thread tiles (the later scales take other forms). Each image's scores in a
    # batch with the others equal its scores alone bit for bit, lie within the FP32 bound of the oracle, and the
    # pair entry point on the oracle's decode gives the batch's score bit for bit.
    sizes = [(1000, 600), (1050, 600), (348, 400), (52, 200), (1400, 380)]
    imgs = [rc.synth_rgb8(70 + i, w, h) for i, (w, h) in enumerate(sizes)]
    cands = [(50, rc.S420), (85, rc.S420)]
    together = ctx.score_batch(imgs, cands, 70.0)
    for img, r in zip(imgs, together):
        alone = ctx.score_batch([img], cands, 70.0)[0]
        assert r.status == 0 and alone.status == 0
        assert list(r.scores[:2]) == list(alone.scores[:2]), img.shape
        for k, (q, sub) in enumerate(cands):
            dist = O.reconstruct(O.build_coefs(img, q, sub))
            want = O.ssimulacra2(img, dist)
            pair = ctx.ssimulacra2(img, dist, rc.FP32)
            assert abs(pair - want) <= TOL32, (img.shape, q, pair, want)
            if not r.settled[k]:
                assert r.scores[k] == pair, (img.shape, q, r.scores[k], pair)


def test_large_call_scores_equal_single_candidate_calls(ctx, rc):
    # A call of 2^25 candidate-pixels or more uploads its descriptors in two parts and stages the candidate launches
    # while the encode and reference launches run; a call below that (one candidate of the same image) uploads them
    # at once. The scores agree bit for bit.
    img = rc.synth_rgb8(31, 4096, 2048)
    cands = [(q, s) for s in (rc.S420, rc.S444) for q in (50, 75, 90)]
    together = ctx.score_batch([img], cands, 70.0)[0]
    assert together.status == 0
    for k, cand in enumerate(cands):
        alone = ctx.score_batch([img], [cand], 70.0)[0]
        assert alone.status == 0
        assert alone.scores[0] == together.scores[k], cand


def test_score_batch_reports_per_job_errors(ctx, rc):
    good = rc.synth_rgb8(1, 64, 64)
    res = ctx.score_batch([good, np.zeros((7, 64, 3), np.uint8)], CANDS, 70.0)
    assert res[0].status == 0 and res[0].chosen_quality > 0
    assert res[1].status == rc.UNSUPPORTED


def test_tile_policy_roundtrip_and_validation(ctx, rc):
    t = ctx.get_tune()
    assert t.nclass >= 1 and t.min_h[t.nclass - 1] == 0
    bad = rc.Tune.from_buffer_copy(t)
    bad.th[0] = 7  # odd tile heights break the 2x2 pyramid pairing
    with pytest.raises(rc.RcgError):
        ctx.set_tune(bad)
    ctx.set_tune(t)


@pytest.mark.parametrize("jobs", [4, 0])
def test_autotuned_policy_keeps_parity(rc, jobs):
    # jobs = 0 is the throughput mode (launches of many rounds, as a large batch runs).
    c = rc.Context(0)
    t = c.autotune(jobs)
    assert 1 <= t.nclass <= 7
    grid = (2, 4, 6, 8, 12, 16, 24, 32, 40, 48, 56, 64, 72, 80, 88, 96, 104, 112, 128, 144, 160, 176, 208, 256, 320,
            384, 512)
    assert all(t.th[i] in grid for i in range(t.nclass))
    img = rc.synth_rgb8(12, 700, 530)
    dist = O.reconstruct(O.build_coefs(img, 60, O.S420))
    assert abs(c.ssimulacra2(img, dist, rc.FP32) - O.ssimulacra2(img, dist)) <= TOL32
    imgs = [rc.synth_rgb8(40 + i, w, h) for i, (w, h) in enumerate([(1500, 1100), (300, 900), (64, 80)])]
    for img, r in zip(imgs, c.score_batch(imgs, CANDS, 70.0)):
        for k, (q, sub) in enumerate(CANDS[:3]):
            o = O.ssimulacra2(img, O.reconstruct(O.build_coefs(img, q, sub)))
            assert abs(r.scores[k] - o) <= (TOL64 if r.settled[k] else TOL32), (img.shape, q)
    c.close()


def test_graph_plan_replays_match_eager(ctx, rc):
    # The captured batch is replayed several times with eager work and synchronisations in between; every replay
    # must give exactly the eager scores and choices.
    imgs = [rc.synth_rgb8(700 + i, w, h) for i, (w, h) in enumerate([(300, 200), (97, 131), (640, 480), (64, 64)])]
    dev = [ctx.upload(im) for im in imgs]
    ptrs = [(d.ptr, im.shape[1], im.shape[0], 3) for d, im in zip(dev, imgs)]
    eager = ctx.score_batch(None, CANDS, 70.0, device_ptrs=ptrs)
    plan = ctx.batch_plan(ptrs, CANDS, 70.0)
    for rep in range(3):
        got = plan.run()
        for g, e in zip(got, eager):
            assert list(g.scores[:len(CANDS)]) == list(e.scores[:len(CANDS)]), rep
            assert g.chosen_quality == e.chosen_quality and g.sweep_quality == e.sweep_quality
        # eager work on the same context between replays
        again = ctx.score_batch(None, CANDS, 70.0, device_ptrs=ptrs)
        assert list(again[0].scores[:len(CANDS)]) == list(eager[0].scores[:len(CANDS)])
        ctx.synchronize()
    plan.close()
    for d in dev:
        d.free()


def test_graph_plan_sees_new_pixels_at_the_same_addresses(ctx, rc):
    # A plan captures device addresses, not pixel values: after new pixels are written into the same buffers, a
    # replay must give exactly what a fresh eager batch gives on the new pixels, and the oracle's scores.
    shapes = [(300, 200), (97, 131), (640, 480)]
    first = [rc.synth_rgb8(720 + i, w, h) for i, (w, h) in enumerate(shapes)]
    dev = [ctx.upload(im) for im in first]
    ptrs = [(d.ptr, im.shape[1], im.shape[0], 3) for d, im in zip(dev, first)]
    plan = ctx.batch_plan(ptrs, CANDS, 70.0)
    before = plan.run()
    for seed in (730, 740):
        new = [rc.synth_rgb8(seed + i, w, h) for i, (w, h) in enumerate(shapes)]
        for d, im in zip(dev, new):
            d.write(im)
        ctx.synchronize()
        got = plan.run()
        eager = ctx.score_batch(None, CANDS, 70.0, device_ptrs=ptrs)
        for g, e, b in zip(got, eager, before):
            assert list(g.scores[:len(CANDS)]) == list(e.scores[:len(CANDS)]), seed
            assert g.chosen_quality == e.chosen_quality and g.sweep_quality == e.sweep_quality
            assert list(g.scores[:len(CANDS)]) != list(b.scores[:len(CANDS)])
        for g, im in zip(got, new):
            for k, (q, sub) in enumerate(CANDS):
                o = O.ssimulacra2(im, O.reconstruct(O.build_coefs(im, q, sub)))
                assert abs(g.scores[k] - o) <= (TOL64 if g.settled[k] else TOL32), (im.shape, q, sub)
    plan.close()
    for d in dev:
        d.free()


def test_device_entry_points_match_host_ones(ctx, rc):
    img = rc.synth_rgb8(31, 333, 250)
    d = ctx.upload(img)
    L = rc.coef_layout(333, 250, 3, 80, rc.S420)
    outs = [rc.DeviceBuffer(ctx, L.blocks_w[c] * L.blocks_h[c] * 64 * 2) for c in range(3)]
    ctx.build_coefs_dev(d.ptr, 333, 250, 3, 80, rc.S420, [o.ptr for o in outs])
    host = ctx.build_coefs(img, 80, rc.S420)
    for o, comp in zip(outs, host.comps):
        got = o.download(np.empty(len(comp.coefs), np.int16))
        np.testing.assert_array_equal(got, comp.coefs)
    dist = O.reconstruct(O.build_coefs(img, 50, O.S420))
    dd = ctx.upload(dist)
    assert ctx.ssimulacra2_dev(d.ptr, dd.ptr, 333, 250, 3) == ctx.ssimulacra2(img, dist)
    f1 = ctx.find_quality_dev(d.ptr, 333, 250, 3, 70.0)
    f2 = ctx.find_quality(img, 70.0)
    assert (f1.quality, f1.score, f1.steps) == (f2.quality, f2.score, f2.steps)


@pytest.mark.parametrize("w,h,ch", [(333, 250, 3), (129, 37, 1), (1030, 17, 3)])
def test_unaligned_device_sources(ctx, rc, w, h, ch):
    # Source pixels at every byte offset of a 16-byte chunk: the encode stage stages rows as aligned 16-byte chunks,
    # and a chunk that reaches outside the source buffer is read byte by byte.
    import ctypes
    img = rc.synth_rgb8(5, w, h)
    if ch == 1:
        img = np.ascontiguousarray(img[:, :, 1])
    host = ctx.build_coefs(img, 80, rc.S420)
    want = ctx.find_quality(img, 70.0)
    buf = rc.DeviceBuffer(ctx, img.nbytes + 64)
    junk = np.full(img.nbytes + 64, 0xA5, np.uint8)
    L = rc.coef_layout(w, h, ch, 80, rc.S420)
    outs = [rc.DeviceBuffer(ctx, L.blocks_w[c] * L.blocks_h[c] * 64 * 2) for c in range(len(host.comps))]
    for off in (1, 3, 7, 8, 13, 15, 16):
        for start, a in ((0, junk), (off, img)):
            a = np.ascontiguousarray(a)
            rc._check(rc.lib().rcg_memcpy_h2d(ctx.handle, buf.ptr + start, a.ctypes.data_as(ctypes.c_void_p),
                                              a.nbytes))
        ctx.build_coefs_dev(buf.ptr + off, w, h, ch, 80, rc.S420, [o.ptr for o in outs])
        for o, comp in zip(outs, host.comps):
            got = o.download(np.empty(len(comp.coefs), np.int16))
            np.testing.assert_array_equal(got, comp.coefs, err_msg=f"offset {off}")
        f = ctx.find_quality_dev(buf.ptr + off, w, h, ch, 70.0)
        assert (f.quality, f.score, f.steps) == (want.quality, want.score, want.steps), off


def test_written_files_from_device_coefficients_are_byte_identical(ctx, rc):
    # The stage after the search: the device's coefficients, entropy-coded on the host, give the same file as the
    # oracle's coefficients (the coefficients are bit-exact), for baseline and progressive output; and the file
    # decodes (libjpeg) to the device reconstruction within an independent decoder's IDCT rounding.
    import io
    for (w, h, q, subs) in [(333, 250, 80, O.S420), (97, 61, 92, O.S444), (640, 480, 55, O.S422)]:
        img = rc.synth_rgb8(77 + w, w, h)
        dev = ctx.build_coefs(img, q, subs)
        ora = O.build_coefs(img, q, subs)
        as_rc = rc.CoefImage(ora.width, ora.height,
                             [rc.CoefComponent(c.id, c.h, c.v, c.tq, c.blocks_w, c.coefs) for c in ora.comps],
                             list(ora.qtables) + [None] * (4 - len(ora.qtables)))
        for mode in (rc.WRITE_BASELINE, rc.WRITE_PROGRESSIVE, rc.WRITE_BEST):
            assert rc.write_jpeg(dev, mode) == rc.write_jpeg(as_rc, mode), (w, h, mode)
        dec = np.asarray(Image.open(io.BytesIO(rc.write_jpeg(dev, rc.WRITE_BASELINE)))).astype(np.int32)
        d = np.abs(dec - ctx.reconstruct(dev).astype(np.int32))
        assert d.max() <= 3, (w, h, d.max())


@pytest.mark.parametrize("w,h,q,subs", [(333, 250, 80, O.S420), (97, 61, 92, O.S444), (640, 480, 55, O.S422),
                                        (64, 48, 30, O.S420)])
def test_trellis_hand_off_is_bit_exact(ctx, rc, w, h, q, subs):
    # The device's hand-off values (a[k] = dct[f]*div[f], unrounded) equal the oracle's bit for bit, so the host
    # trellis over them gives the reference's trellis coefficients.
    img = rc.synth_rgb8(500 + w, w, h)
    dev, dsc = ctx.build_scaled(img, q, subs)
    ora, osc = O.build_scaled(img, q, subs)
    for a, b, x, y in zip(dev.comps, ora.comps, dsc, osc):
        np.testing.assert_array_equal(a.coefs, b.coefs)
        np.testing.assert_array_equal(x.view(np.uint32), y.view(np.uint32))
    got = rc.trellis(dev, dsc)
    for g, want in zip(got.comps, O.trellis(ora, osc)):
        np.testing.assert_array_equal(g.coefs, want)


def _oracle_trellis_search(img, target, subs):
    # cli/src/pipeline.rs:238 with the reference's default trellis: make(q) = build_coefs(img, q, trellis on)
    # (encode.rs:527-537), metric = f64 SSIMULACRA2, bisection of jpeg/src/target.rs:37-55.
    cache = {}

    def make(q):
        ci, scaled = O.build_scaled(img, q, subs)
        if q < 100:
            for c, t in zip(ci.comps, O.trellis(ci, scaled)):
                c.coefs = np.asarray(t, np.int16)
        return ci

    def score(q):
        if q not in cache:
            cache[q] = O.ssimulacra2(img, O.reconstruct(make(q)))
        return cache[q]

    lo, hi = 1, 100
    best = hi
    if score(hi) >= target:
        while lo < hi:
            mid = lo + (hi - lo) // 2
            if score(mid) >= target:
                hi = mid
                best = mid
            else:
                lo = mid + 1
    return best, cache, make


@pytest.mark.parametrize("w,h,subs,gray,target", [(256, 192, O.S420, False, 70.0), (301, 199, O.AUTO, False, 85.0),
                                                  (97, 61, O.S444, False, 55.0), (120, 90, O.S420, True, 75.0),
                                                  (200, 120, O.S422, False, 92.0)])
def test_trellis_search_chooses_the_oracles_quality(ctx, rc, w, h, subs, gray, target):
    img = rc.synth_rgb8(700 + w, w, h)
    if gray:
        img = np.ascontiguousarray(img[:, :, 0])
    found = ctx.find_quality(img, target, 100, subs, trellis=True)
    want, scores, make = _oracle_trellis_search(img, target, subs)
    assert found.quality == want, (found.quality, want, sorted(scores.items()))
    assert abs(found.score - scores[want]) <= TOL32
    assert found.steps == len(scores)
    for a, b in zip(found.coefs.comps, make(want).comps):
        np.testing.assert_array_equal(a.coefs, b.coefs)


def test_trellis_search_reuses_the_context(ctx, rc):
    # Two searches in a row on one context (the arena and reference state are reset per call) agree with fresh ones.
    a = rc.synth_rgb8(1, 160, 120)
    b = rc.synth_rgb8(2, 96, 64)
    fa = ctx.find_quality(a, 70.0, 100, rc.S420, trellis=True)
    fb = ctx.find_quality(b, 70.0, 100, rc.S420, trellis=True)
    fa2 = ctx.find_quality(a, 70.0, 100, rc.S420, trellis=True)
    assert (fa.quality, fa.score) == (fa2.quality, fa2.score)
    want_b, _, _ = _oracle_trellis_search(b, 70.0, O.S420)
    assert fb.quality == want_b


def test_jpeg_reader_feeds_the_requant_search(ctx, rc):
    # A file read by the product's reader searches like the oracle's view of the same coefficients.
    path = os.path.join(os.path.dirname(__file__), "golden", "rgb8_prog.jpg")
    data = open(path, "rb").read()
    ci = rc.read_jpeg(data)
    found = ctx.requant_find_quality(ci, 60.0, 100, True)
    want, _ = _oracle_requant_search(ci, 60.0, True)
    assert found.quality == want
    np.testing.assert_array_equal(ctx.reconstruct(ci), O.reconstruct(ci))


def _oracle_requant_search(src, target, trellis):
    # cli/src/pipeline.rs:369-386 with the oracle: reference = reconstruct(ci), make(q) = requantize(ci, q, trellis),
    # metric = f64 SSIMULACRA2, bisection of jpeg/src/target.rs:37-55.
    ref = O.reconstruct(src)
    cache = {}

    def score(q):
        if q not in cache:
            cache[q] = O.ssimulacra2(ref, O.reconstruct(O.requantize(src, q, trellis)))
        return cache[q]

    lo, hi = 1, 100
    best = hi
    if score(hi) >= target:
        while lo < hi:
            mid = lo + (hi - lo) // 2
            if score(mid) >= target:
                hi = mid
                best = mid
            else:
                lo = mid + 1
    return best, cache


@pytest.mark.parametrize("w,h,q0,subs,gray,target,trellis", [
    (96, 72, 92, O.S420, False, 75.0, False), (96, 72, 92, O.S420, False, 60.0, True),
    (130, 66, 95, O.S444, False, 80.0, False), (64, 48, 90, O.S444, True, 70.0, True),
    (200, 120, 85, O.S422, False, 65.0, False)])
def test_requant_search_chooses_the_oracles_quality(ctx, rc, w, h, q0, subs, gray, target, trellis):
    img = rc.synth_rgb8(900 + w, w, h)
    if gray:
        img = np.ascontiguousarray(img[:, :, 2])
    ci = O.build_coefs(img, q0, subs)
    src = rc.CoefImage(ci.width, ci.height, [rc.CoefComponent(c.id, c.h, c.v, c.tq, c.blocks_w, c.coefs)
                                             for c in ci.comps], list(ci.qtables) + [None] * (4 - len(ci.qtables)))
    found = ctx.requant_find_quality(src, target, 100, trellis)
    want, scores = _oracle_requant_search(src, target, trellis)
    assert found.quality == want, (found.quality, want, sorted(scores.items()))
    assert abs(found.score - scores[want]) <= TOL32
    ref = O.requantize(src, want, trellis)
    for a, b in zip(found.coefs.comps, ref.comps):
        np.testing.assert_array_equal(a.coefs, b.coefs)


def test_device_requantisation_matches_the_host(ctx, rc):
    # The requantisation fused into the device IDCT reconstructs exactly what the host requantisation's coefficients
    # reconstruct to (checked through the search's winner: its score at q equals a fresh score of the host result).
    img = rc.synth_rgb8(41, 160, 96)
    ci = O.build_coefs(img, 93, O.S420)
    src = rc.CoefImage(ci.width, ci.height, [rc.CoefComponent(c.id, c.h, c.v, c.tq, c.blocks_w, c.coefs)
                                             for c in ci.comps], list(ci.qtables) + [None] * 2)
    found = ctx.requant_find_quality(src, 70.0, 100, False)
    host = rc.requantize(src, found.quality)
    np.testing.assert_array_equal(ctx.reconstruct(host), O.reconstruct(host))
    s = ctx.ssimulacra2(O.reconstruct(src), ctx.reconstruct(host), rc.FP32)
    assert abs(s - found.score) <= TOL32


@pytest.mark.parametrize("w,h", [(65535, 9), (9, 4099), (4097, 17)])
def test_extreme_aspect_ratios(ctx, rc, w, h):
    # The longest side JPEG allows (encode.rs:417) and very tall strips: encode and reconstruct bit-exact, the score
    # within the FP32 bound (the 8-pixel scale floor leaves one or two scales).
    img = rc.synth_rgb8(60 + h, w, h)
    for subs in (O.S420, O.S444):
        dev = ctx.build_coefs(img, 70, subs)
        ora = O.build_coefs(img, 70, subs)
        for a, b in zip(dev.comps, ora.comps):
            np.testing.assert_array_equal(a.coefs, b.coefs)
        rec = ctx.reconstruct(dev)
        np.testing.assert_array_equal(rec, O.reconstruct(ora))
        assert abs(ctx.ssimulacra2(img, rec, rc.FP32) - O.ssimulacra2(img, rec)) <= TOL32


# ------------------------------------------------------------------------------------ four-component sources

CMYK = ["cmyk.jpg", "ycck.jpg", "ycck_444.jpg"]


def _read_gold(name):
    with open(os.path.join(GOLD, name), "rb") as f:
        return f.read()


@pytest.mark.parametrize("name", CMYK)
def test_four_component_reconstruct_is_bit_exact(ctx, rc, name):
    # reconstruct of CMYK (Adobe transform 0: inverted ink) and YCCK (Adobe 2, 4:2:0-style and 4:4:4) coefficient
    # images: IDCT of four planes, upsampling, cmyk_to_rgb (jpeg/src/decode.rs:1446-1478).
    ci = rc.read_jpeg(_read_gold(name))
    assert len(ci.comps) == 4
    got = ctx.reconstruct(ci)
    assert got.shape == (ci.height, ci.width, 3)
    np.testing.assert_array_equal(got, O.reconstruct(ci))
    # The same planes with no Adobe marker (plain ink amounts) take the other branch of cmyk_to_rgb.
    ci.adobe_transform = -1
    np.testing.assert_array_equal(ctx.reconstruct(ci), O.reconstruct(ci))


@pytest.mark.parametrize("name,target,trellis", [("cmyk.jpg", 70.0, False), ("ycck.jpg", 60.0, True),
                                                 ("ycck_444.jpg", 75.0, False), ("ycck.jpg", 80.0, False)])
def test_four_component_requant_search_chooses_the_oracles_quality(ctx, rc, name, target, trellis):
    # cli/src/pipeline.rs:369-377 searches any non-resampled JPEG source, four-component ones included: candidates
    # are requantised, reconstructed through cmyk_to_rgb and scored as pixels.
    ci = rc.read_jpeg(_read_gold(name))
    found = ctx.requant_find_quality(ci, target, 100, trellis)
    want, scores = _oracle_requant_search(ci, target, trellis)
    assert found.quality == want, (found.quality, want, sorted(scores.items()))
    assert abs(found.score - scores[want]) <= TOL32
    ref = O.requantize(ci, want, trellis)
    for a, b in zip(found.coefs.comps, ref.comps):
        np.testing.assert_array_equal(a.coefs, b.coefs)


def test_rgb_coded_requant_search_takes_the_pixel_route(ctx, rc):
    # Three components stored as R, G, B (Adobe transform 0, decode.rs:1423-1426): no colour conversion, which the
    # fused loader does not do, so the candidates are scored as pixels.
    img = rc.synth_rgb8(77, 120, 90)
    ci = O.build_coefs(img, 92, O.S444)
    src = rc.CoefImage(ci.width, ci.height, [rc.CoefComponent(c.id, c.h, c.v, c.tq, c.blocks_w, c.coefs)
                                             for c in ci.comps], list(ci.qtables) + [None] * 2, 0)
    np.testing.assert_array_equal(ctx.reconstruct(src), O.reconstruct(src))
    found = ctx.requant_find_quality(src, 70.0, 100, False)
    want, scores = _oracle_requant_search(src, 70.0, False)
    assert found.quality == want, (found.quality, want, sorted(scores.items()))


# ------------------------------------------------------------------------------------- winners' coefficients

def _assert_winner(img, ci, quality, sub):
    # Found.coefs (jpeg/src/target.rs:11-16): build_coefs at the chosen quality and subsampling.
    want = O.build_coefs(img, quality, sub)
    assert ci is not None
    _assert_coefs_equal(ci, want)


def test_batch_returns_the_winners_coefficients(ctx, rc):
    cands = [(q, rc.AUTO) for q in (40, 60, 80, 85, 90, 95)]   # Auto: 4:2:0 below q90, 4:4:4 above
    imgs = [rc.synth_rgb8(800 + i, w, h) for i, (w, h) in enumerate([(301, 203), (97, 131), (640, 480), (64, 64)])]
    imgs.append(np.ascontiguousarray(rc.synth_rgb8(805, 150, 90)[:, :, 0]))
    for target in (60.0, 85.0):
        res, win = ctx.score_batch(imgs, cands, target, winners=True)
        for img, r, ci in zip(imgs, res, win):
            assert r.status == 0 and r.winner_written == 1
            _assert_winner(img, ci, r.chosen_quality, rc.AUTO)
    # Device-resident images: the winners land in device buffers, eagerly and from a replayed plan.
    dev = [ctx.upload(im) for im in imgs]
    ptrs = [(d.ptr, im.shape[1], im.shape[0], 1 if im.ndim == 2 else 3) for d, im in zip(dev, imgs)]
    res, win = ctx.score_batch(None, cands, 70.0, device_ptrs=ptrs, winners=True)
    for img, r, ci in zip(imgs, res, win):
        _assert_winner(img, ci, r.chosen_quality, rc.AUTO)
    plan = ctx.batch_plan(ptrs, cands, 70.0, winners=True)
    for rep in range(2):
        got = plan.run()
        for img, r, ci in zip(imgs, got, plan.winners(got)):
            assert r.winner_written == 1
            _assert_winner(img, ci, r.chosen_quality, rc.AUTO)
    plan.close()
    # Jobs that do not ask get none.
    res = ctx.score_batch(imgs[:2], cands, 70.0)
    assert all(r.winner_written == 0 for r in res)

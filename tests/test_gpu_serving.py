"""The continuous-batching server: every request's result equals the one it gets alone, requests that arrive
while others run are admitted and complete, and the HBM budget admits an oversized request by itself."""
import time

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _drain(srv, n, timeout_s=120):
    got = {}
    t0 = time.time()
    while len(got) < n and time.time() - t0 < timeout_s:
        for c in srv.poll(64, 200):
            got[c.id] = c
    assert len(got) == n, f"only {len(got)} of {n} completed"
    return got


def test_results_alone_equal_results_in_the_mix(rc, ctx):
    rng = np.random.default_rng(4)
    reqs = []
    for i in range(10):
        w, h = int(rng.integers(64, 900)), int(rng.integers(64, 900))
        reqs.append((i, rc.synth_rgb8(300 + i, w, h), float([60.0, 70.0, 80.0][i % 3])))
    alone = {}
    for i, img, t in reqs:
        f = ctx.find_quality(img, t, 100, rc.S420)
        alone[i] = (f.quality, f.score)
    for spec in (1, 3):
        srv = rc.Server(0, speculate=spec)
        # Staggered arrivals: half now, half while the first half is in flight.
        for i, img, t in reqs[:5]:
            srv.submit(i, img, t)
        time.sleep(0.05)
        for i, img, t in reqs[5:]:
            srv.submit(i, img, t)
        got = _drain(srv, len(reqs))
        st = srv.stats()
        srv.close()
        for i, _, _ in reqs:
            c = got[i]
            assert c.status == 0
            assert (c.quality, c.score) == alone[i], (spec, i, c.quality, c.score, alone[i])
            assert c.done_s >= c.admit_s >= c.submit_s
        assert st.requests_done == len(reqs)
        assert st.mean_batch > 1.0, "requests were never batched together"


def test_reused_device_blocks_keep_results_alone(rc, ctx):
    # Each wave is drained before the next is submitted, so later waves run in the device blocks that earlier
    # requests of the same sizes left behind (their stale sources and reference states must not leak in).
    sizes = [(700, 500), (640, 560), (700, 500)]
    srv = rc.Server(0, speculate=3)
    for wave in range(3):
        reqs = []
        for j, (w, h) in enumerate(sizes):
            i = wave * len(sizes) + j
            reqs.append((i, rc.synth_rgb8(700 + i, w, h), float([60.0, 70.0, 80.0][(wave + j) % 3])))
        for i, img, t in reqs:
            srv.submit(i, img, t)
        got = _drain(srv, len(reqs))
        for i, img, t in reqs:
            f = ctx.find_quality(img, t, 100, rc.S420)
            assert got[i].status == 0
            assert (got[i].quality, got[i].score) == (f.quality, f.score), (wave, i)
    st = srv.stats()
    srv.close()
    assert st.requests_done == 3 * len(sizes)


def test_budget_admits_oversized_request_alone(rc):
    srv = rc.Server(0, hbm_budget=4096)  # 4 KiB: every request is larger than the whole budget
    imgs = [rc.synth_rgb8(400 + i, 128, 96) for i in range(3)]
    for i, img in enumerate(imgs):
        srv.submit(i, img, 70.0)
    got = _drain(srv, 3)
    st = srv.stats()
    srv.close()
    assert all(c.status == 0 for c in got.values())
    assert st.mean_batch == 1.0  # one at a time


def test_submit_rejects_small(rc):
    srv = rc.Server(0)
    with pytest.raises(rc.RcgError) as e:
        srv.submit(1, np.zeros((7, 40, 3), np.uint8), 70.0)
    assert e.value.status == rc.UNSUPPORTED
    srv.close()


def test_completions_carry_the_winners_coefficients(rc, ctx):
    # Found.coefs (jpeg/src/target.rs:11-16): each completion's coefficients are build_coefs at its chosen quality.
    from oracle import oracle as O
    srv = rc.Server(0, speculate=3)
    reqs = [(i, rc.synth_rgb8(900 + i, w, h), t, sub)
            for i, (w, h, t, sub) in enumerate([(333, 250, 70.0, rc.S420), (128, 96, 80.0, rc.AUTO),
                                                (500, 211, 60.0, rc.S444), (97, 64, 75.0, rc.AUTO)])]
    for i, img, t, sub in reqs:
        srv.submit(i, img, t, 100, sub, coefs=True)
    got = _drain(srv, len(reqs))
    for i, img, t, sub in reqs:
        c = got[i]
        assert c.status == 0 and c.coefs_written == 1
        ci = srv.winner(c)
        want = O.build_coefs(img, c.quality, sub)
        for a, b in zip(ci.comps, want.comps):
            assert (a.h, a.v, a.blocks_w) == (b.h, b.v, b.blocks_w)
            np.testing.assert_array_equal(a.coefs, b.coefs)
        f = ctx.find_quality(img, t, 100, sub)
        assert (c.quality, c.score) == (f.quality, f.score)
    srv.close()


def test_a_request_whose_working_set_cannot_be_allocated_fails_alone(rc, ctx, monkeypatch):
    # Forced from the host: with RCG_TEST_ARENA_LIMIT the workspace refuses working sets above 8 MiB, as a full
    # device would. The 2000x1500 request's working set exceeds it, the small ones' do not: the shared tick fails,
    # is retried request by request, and only the large request fails (RCG_OOM); the others get their alone results.
    small = [(i, rc.synth_rgb8(950 + i, w, h), t) for i, (w, h, t) in
             enumerate([(96, 80, 70.0), (128, 128, 60.0), (200, 150, 80.0)])]
    alone = {i: ctx.find_quality(img, t, 100, rc.S420) for i, img, t in small}
    big = rc.synth_rgb8(960, 2000, 1500)
    monkeypatch.setenv("RCG_TEST_ARENA_LIMIT", str(8 << 20))
    srv = rc.Server(0)
    for i, img, t in small[:2]:
        srv.submit(i, img, t)
    srv.submit(99, big, 70.0)
    srv.submit(*small[2])
    got = _drain(srv, 4)
    srv.close()
    monkeypatch.delenv("RCG_TEST_ARENA_LIMIT")
    assert got[99].status == rc.OOM
    for i, img, t in small:
        assert got[i].status == 0, i
        assert (got[i].quality, got[i].score) == (alone[i].quality, alone[i].score), i


def test_an_exception_in_a_tick_fails_that_tick_and_the_worker_goes_on(rc, ctx, monkeypatch):
    # Forced from the host: RCG_TEST_THROW_TICK=1 makes the worker's first tick throw std::bad_alloc, as a host
    # allocation in the tick would. Its request fails with RCG_OOM; the worker survives and later requests complete
    # with their alone results.
    monkeypatch.setenv("RCG_TEST_THROW_TICK", "1")
    srv = rc.Server(0)
    monkeypatch.delenv("RCG_TEST_THROW_TICK")
    a = rc.synth_rgb8(970, 160, 120)
    srv.submit(1, a, 70.0)
    first = _drain(srv, 1)
    assert first[1].status == rc.OOM
    b = rc.synth_rgb8(971, 200, 100)
    srv.submit(2, b, 70.0)
    srv.submit(3, a, 70.0)
    got = _drain(srv, 2)
    srv.close()
    for i, img in ((2, b), (3, a)):
        f = ctx.find_quality(img, 70.0, 100, rc.S420)
        assert got[i].status == 0 and (got[i].quality, got[i].score) == (f.quality, f.score)


def test_kept_blocks_stay_within_the_budget(rc):
    # Finished requests' device blocks are kept for reuse only while admitted footprints and kept blocks fit the
    # budget: with a budget of about two requests' footprints, a stream of mixed sizes still completes.
    srv = rc.Server(0, hbm_budget=96 << 20)
    imgs = [rc.synth_rgb8(980 + i, w, h) for i, (w, h) in enumerate([(700, 500), (300, 200), (800, 600),
                                                                       (128, 128), (640, 480), (1000, 700)])]
    for i, img in enumerate(imgs):
        srv.submit(i, img, 70.0)
    got = _drain(srv, len(imgs))
    st = srv.stats()
    srv.close()
    assert all(c.status == 0 for c in got.values())
    assert st.peak_footprint <= 96 << 20 or st.mean_batch == 1.0


def test_small_requests_do_not_wait_for_a_large_one(rc, ctx):
    # Requests below 2 MP tick on a lane of their own: submitted just after a 16 MP request, small ones complete
    # before it does, and every result still equals the request's alone result.
    big = rc.synth_rgb8(990, 4000, 4000)
    small = [(i, rc.synth_rgb8(991 + i, 320, 240), t) for i, t in enumerate([60.0, 70.0, 80.0, 70.0])]
    srv = rc.Server(0, speculate=3)
    srv.submit(100, big, 70.0)
    for i, img, t in small:
        srv.submit(i, img, t)
    order = []
    got = {}
    t0 = time.time()
    while len(got) < 5 and time.time() - t0 < 120:
        for c in srv.poll(16, 200):
            got[c.id] = c
            order.append(c.id)
    st = srv.stats()
    srv.close()
    assert len(got) == 5
    assert order.index(100) > 0, order  # at least one small request finished first
    for i, img, t in small:
        f = ctx.find_quality(img, t, 100, rc.S420)
        assert got[i].status == 0 and (got[i].quality, got[i].score) == (f.quality, f.score)
    f = ctx.find_quality(big, 70.0, 100, rc.S420)
    assert got[100].status == 0 and (got[100].quality, got[100].score) == (f.quality, f.score)
    assert st.requests_done == 5


def test_admission_stays_in_arrival_order_across_lanes(rc):
    # A budget smaller than the 2.2 MP request's footprint: it runs alone, once nothing else is admitted. Small
    # requests that arrive after it must not take the budget first (their lane is another one), so it completes
    # before all of them but the one admitted ahead of it.
    srv = rc.Server(0, hbm_budget=32 << 20)
    small = [rc.synth_rgb8(1200 + i, 160, 120) for i in range(5)]
    large = rc.synth_rgb8(1210, 1600, 1400)
    srv.submit(0, small[0], 70.0)
    srv.submit(100, large, 70.0)
    for i in range(1, 5):
        srv.submit(i, small[i], 70.0)
    order = []
    t0 = time.time()
    while len(order) < 6 and time.time() - t0 < 120:
        order += [c.id for c in srv.poll(16, 200) if c.status == 0]
    srv.close()
    assert sorted(order) == [0, 1, 2, 3, 4, 100], order
    assert order.index(100) == 1, order


def test_gray_and_every_subsampling_equal_their_alone_results(rc, ctx):
    # Gray8 requests and RGB8 ones under each subsampling mode, in one server: each completion equals find_quality's
    # result for the same request alone.
    gray = np.ascontiguousarray(rc.synth_rgb8(1300, 300, 211)[:, :, 0])
    reqs = [(0, gray, 70.0, rc.S420), (1, rc.synth_rgb8(1301, 257, 190), 60.0, rc.S444),
            (2, rc.synth_rgb8(1302, 190, 257), 80.0, rc.S422), (3, rc.synth_rgb8(1303, 128, 96), 70.0, rc.AUTO)]
    srv = rc.Server(0, speculate=3)
    for i, img, t, sub in reqs:
        srv.submit(i, img, t, 100, sub)
    got = _drain(srv, len(reqs))
    srv.close()
    for i, img, t, sub in reqs:
        f = ctx.find_quality(img, t, 100, sub)
        assert got[i].status == 0 and (got[i].quality, got[i].score) == (f.quality, f.score), (i, sub)

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

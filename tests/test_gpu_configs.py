"""The batching and serving configurations at their own workload sizes.

Batching: 256 synthetic RGB8 images, W and H uniform in [64, 3000], each swept over the candidates
{40,50,60,70,80,85,90,95} at 4:2:0 through the hipGraph plan. Every replay equals the eager pass bit for bit, every
choice is the reference's bisection (jpeg/src/target.rs:37-55) over the list of returned scores, and on sampled images
(the largest, odd sizes) the scores are within the FP32 bound of the f64 oracle's and the choice is the oracle's.

Serving: Poisson arrivals at 200 requests/s of images whose megapixel count is log-uniform in [0.1, 24] and aspect
uniform in [0.5, 2], targets uniform over {60, 70, 80}: every completion equals the request's alone result, and on
sampled requests the chosen quality is the oracle's bisection over 1..100.
"""
import time

import numpy as np
import pytest

from oracle import oracle as O

pytestmark = pytest.mark.gpu
CANDS_Q = (40, 50, 60, 70, 80, 85, 90, 95)
TOL32 = 0.02
TOL64 = 1e-6


def _bisect(qs, scores, target):
    """jpeg/src/target.rs:37-55 over a candidate list sorted by quality: the highest first, then bisection."""
    lo, hi = 0, len(qs) - 1
    best = hi
    if scores[hi] >= target:
        while lo < hi:
            mid = lo + (hi - lo) // 2
            if scores[mid] >= target:
                hi = mid
                best = mid
            else:
                lo = mid + 1
    return qs[best]


def test_batched_configuration_256_images(ctx, rc):
    rng = np.random.default_rng(2024)
    sizes = [(int(rng.integers(64, 3001)), int(rng.integers(64, 3001))) for _ in range(256)]
    imgs = [rc.synth_rgb8(1000 + i, w, h) for i, (w, h) in enumerate(sizes)]
    print(f"batch: {sum(w * h for w, h in sizes) / 1e6:.1f} MP in {len(sizes)} images", flush=True)
    dev = [ctx.upload(im) for im in imgs]
    ptrs = [(d.ptr, w, h, 3) for d, (w, h) in zip(dev, sizes)]
    cands = [(q, rc.S420) for q in CANDS_Q]
    eager = ctx.score_batch(None, cands, 70.0, device_ptrs=ptrs)
    plan = ctx.batch_plan(ptrs, cands, 70.0)
    for rep in range(2):
        got = plan.run()
        for i, (g, e) in enumerate(zip(got, eager)):
            assert g.status == 0 and e.status == 0
            assert list(g.scores[:8]) == list(e.scores[:8]), (rep, i, sizes[i])
            assert (g.chosen_quality, g.sweep_quality) == (e.chosen_quality, e.sweep_quality), (rep, i)
    plan.close()
    for i, r in enumerate(eager):
        assert r.chosen_quality == _bisect(CANDS_Q, list(r.scores[:8]), 70.0), (i, sizes[i])
    # The oracle on a sample: the largest image, images with a side of at least 2800 px, odd sizes, and the smallest.
    order = sorted(range(256), key=lambda i: sizes[i][0] * sizes[i][1])
    sample = {order[-1], order[0]}
    sample |= set([i for i in range(256) if max(sizes[i]) >= 2800][:2])
    sample |= set([i for i in range(256) if sizes[i][0] % 2 and sizes[i][1] % 2][:3])
    assert len(sample) >= 6
    for i in sorted(sample):
        img, r = imgs[i], eager[i]
        print(f"oracle on image {i}: {sizes[i]}", flush=True)
        osc = [O.ssimulacra2(img, O.reconstruct(O.build_coefs(img, q, O.S420))) for q in CANDS_Q]
        for k in range(8):
            assert abs(r.scores[k] - osc[k]) <= (TOL64 if r.settled[k] else TOL32), (i, CANDS_Q[k])
        assert r.chosen_quality == _bisect(CANDS_Q, osc, 70.0), (i, sizes[i])
    for d in dev:
        d.free()


def _pool(rc, n, seed):
    rng = np.random.default_rng(seed)
    out = []
    for i in range(n):
        mp = float(np.exp(rng.uniform(np.log(0.1), np.log(24.0))))
        aspect = float(rng.uniform(0.5, 2.0))
        w = max(8, int(round((mp * 1e6 * aspect) ** 0.5)))
        h = max(8, int(round(mp * 1e6 / w)))
        out.append(rc.synth_rgb8(5000 + i, w, h))
    return out


def test_serving_configuration_poisson_200_per_second(ctx, rc):
    pool = _pool(rc, 24, 77)
    rng = np.random.default_rng(78)
    n = 300
    picks = rng.integers(0, len(pool), n)
    targets = rng.choice([60.0, 70.0, 80.0], n)
    gaps = rng.exponential(1.0 / 200.0, n)
    alone = {}
    for p, t in set(zip(picks.tolist(), targets.tolist())):
        f = ctx.find_quality(pool[p], t, 100, rc.S420)
        alone[(p, t)] = (f.quality, f.score)
    print(f"serving: {n} requests over {len(alone)} distinct (image, target) searches", flush=True)
    srv = rc.Server(0, speculate=3)
    got = {}
    t = time.perf_counter()
    for i in range(n):
        t += gaps[i]
        d = t - time.perf_counter()
        if d > 0:
            time.sleep(d)
        srv.submit(i, pool[picks[i]], float(targets[i]))
        for c in srv.poll(64, 0):
            got[c.id] = c
    t0 = time.time()
    while len(got) < n and time.time() - t0 < 120:
        for c in srv.poll(64, 200):
            got[c.id] = c
    st = srv.stats()
    srv.close()
    assert len(got) == n
    for i in range(n):
        c = got[i]
        assert c.status == 0, i
        assert (c.quality, c.score) == alone[(int(picks[i]), float(targets[i]))], i
    assert st.mean_batch > 1.0
    # The oracle's bisection on the smallest images of the pool (f64 SSIMULACRA2 of each step).
    small = sorted(set(zip(picks.tolist(), targets.tolist())), key=lambda k: pool[k[0]].size)[:3]
    for p, tt in small:
        q, _, _ = O.find_quality(pool[p], tt, 100, O.S420)
        assert alone[(p, tt)][0] == q, (pool[p].shape, tt)

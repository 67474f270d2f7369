"""Benchmark of the JPEG target-quality inner loop on one MI355X.

Workload (the plain-drop-in configuration): one 4096x4096 RGB8 synthetic image (seeded generator of
rcg_synth_rgb8), candidates q{50,75,90} x {4:2:0, 4:4:4}, trellis off. One step scores all six candidates: encode
(colour, downsample, FDCT, quantisation) + dequantise/IDCT reconstruction + SSIMULACRA2 (FP32 maps, FP64 pooling)
against the source, in one launch sequence, with the source resident in HBM before the timed region.

value = candidate-evaluation throughput in MP/s (image megapixels x candidates / second). The same run also reports
the plain per-image GPU search (serial bisection over q in 1..100, rcg_find_quality_dev) in search-MP/s, the roofline
of the dominant kernel (k_ssim_scale) from HIP events recorded on the library's own stream around every launch in
the timed region, and the CPU oracle (C restatement of the reference CPU path, OpenMP threads) on a bounded sample.

All device memory and synchronisation go through librcgpu (rcg_device_alloc, rcg_synchronize on the stream every
kernel runs on); PyTorch is not loaded, so the process holds a single HIP runtime (torch bundles its own).
"""
import argparse
import importlib
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "MP/s of JPEG DCT+quant+SSIMULACRA2 search, batched vs plain; p99 latency"
HBM_PEAK_GBS = 8000.0
# Algorithmic bytes per pixel per candidate of the dominant kernel (SSIMULACRA2 scales): read the reference and the
# reconstruction at scale 1 (6), write and read back the reconstruction's pyramid (4 + 4), read the reference's
# cached pyramid (4). The kernel's real fetches are larger: it reads the reference side from the per-image cache
# (9 f32 per pixel per scale) in place of the reference's pixels, trading HBM bytes for VALU work.
SSIM_BYTES_PER_PX = 18.0
# Whole candidate evaluation: + read RGB8 (3), write/read coefficients (3 + 3, or 6 + 6 at 4:4:4), write RGB8 (3).
EVAL_BYTES_PER_PX = {3: 30.0, 1: 36.0}


def progress(msg):
    print(f"bench: {msg}", file=sys.stderr, flush=True)


def _pmc_run(counter, bench_args, outdir, name):
    """One rocprofv3 counter pass over a child run of this script; returns [(kernel, dispatch, KiB)] in dispatch
    order, or None if the pass failed."""
    import shutil
    import sqlite3
    import subprocess
    prof = shutil.which("rocprofv3")
    if prof is None:
        return None
    d = os.path.join(outdir, name)
    cmd = ["timeout", "-s", "KILL", "150", prof, "--pmc", counter, "-d", d, "-o", name, "--",
           sys.executable, os.path.abspath(__file__)] + bench_args
    r = subprocess.run(cmd, stdout=subprocess.DEVNULL, stderr=subprocess.DEVNULL, cwd=ROOT)
    dbs = []
    for root, _, files in os.walk(d):
        dbs += [os.path.join(root, f) for f in files if f.endswith("_results.db")]
    if not dbs:
        progress(f"{counter} pass failed (exit {r.returncode})")
        return None
    con = sqlite3.connect(dbs[0])
    rows = con.execute("select dispatch_id, kernel_name, sum(value) from counters_collection "
                       "where counter_name = ? group by dispatch_id order by dispatch_id", (counter,)).fetchall()
    con.close()
    return [(k.split("(")[0], i, v) for i, k, v in rows]


def tuned_policy(ncand):
    """The headline's tile policy, autotuned in a child process (so that this process has not touched the GPU
    when the counter passes run); None if the child failed."""
    import subprocess
    r = subprocess.run([sys.executable, os.path.abspath(__file__), "--autotune-only", str(ncand)],
                       stdout=subprocess.PIPE, stderr=subprocess.DEVNULL, cwd=ROOT, text=True)
    lines = r.stdout.strip().splitlines()
    if r.returncode != 0 or not lines:
        return None
    return json.loads(lines[-1])


def measure_traffic(size, policy):
    """HBM bytes per launch of the dominant kernel (k_ssim_scale) in one headline step, from FETCH_SIZE and
    WRITE_SIZE (one rocprofv3 pass each over a one-step child run under the same tile policy as the timed steps),
    each scaled by what the counter reads for a 512 MiB device-to-device copy (the same counter's calibration
    factor). Runs before this process touches the GPU: the passes are child processes."""
    import tempfile
    out = tempfile.mkdtemp(prefix="rcg_traffic_")
    step = ["--steps", "1", "--warmup", "0", "--size", str(size), "--no-cpu", "--no-plain", "--no-batch",
            "--no-serve", "--no-cli", "--no-traffic", "--tile-policy", json.dumps(policy)]
    copy = ["--copy-calibration", "512"]
    res = {}
    for counter in ("FETCH_SIZE", "WRITE_SIZE"):
        progress(f"traffic: {counter} passes")
        cal = _pmc_run(counter, copy, out, counter.lower() + "_copy")
        run = _pmc_run(counter, step, out, counter.lower() + "_step")
        if not cal or not run:
            return None
        # Calibration: the copy's largest dispatch moved 512 MiB (read for FETCH_SIZE, written for WRITE_SIZE).
        kib = max(v for _, _, v in cal)
        factor = (512 * 1024) / kib if kib > 0 else None
        # The last headline step: the dispatches after the last reference-cache launch from RGB8 bytes (the autotuner
        # launches it only from planes).
        last_ref = max((i for i, (k, _, _) in enumerate(run) if k.startswith("void rcg::k_ssim_ref<") and
                        k.endswith("0u>")), default=None)
        if last_ref is None or factor is None:
            return None
        vals = [v for k, _, v in run[last_ref:] if "k_ssim_scale<float" in k]
        if not vals:
            return None
        res[counter] = {"kib_per_launch": sum(vals) / len(vals), "factor": factor, "launches": len(vals)}
    fetch = res["FETCH_SIZE"]["kib_per_launch"] * res["FETCH_SIZE"]["factor"] * 1024
    write = res["WRITE_SIZE"]["kib_per_launch"] * res["WRITE_SIZE"]["factor"] * 1024
    return {"bytes_per_launch": fetch + write, "fetch_bytes": fetch, "write_bytes": write,
            "fetch_factor": round(res["FETCH_SIZE"]["factor"], 3), "write_factor": round(res["WRITE_SIZE"]["factor"], 3),
            "launches": res["FETCH_SIZE"]["launches"], "tile_policy": policy}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--size", type=int, default=4096)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU oracle baseline")
    ap.add_argument("--no-plain", action="store_true", help="skip the plain per-image search")
    ap.add_argument("--no-batch", action="store_true", help="skip the 256-image batched search")
    ap.add_argument("--no-serve", action="store_true", help="skip the serving run")
    ap.add_argument("--no-cli", action="store_true", help="skip the command-line end-to-end run")
    ap.add_argument("--cli-files", type=int, default=24)
    ap.add_argument("--batch-images", type=int, default=256)
    ap.add_argument("--serve-requests", type=int, default=10000,
                    help="requests of the serving run (Poisson arrivals at 200 per second)")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="only run a device-to-device copy of MIB mebibytes (HBM traffic counter calibration)")
    ap.add_argument("--no-traffic", action="store_true",
                    help="skip the rocprofv3 FETCH_SIZE/WRITE_SIZE passes that give roofline.traffic")
    ap.add_argument("--tile-policy", default=None, metavar="JSON",
                    help="install this headline tile policy (as the bench line prints it) instead of autotuning")
    ap.add_argument("--autotune-only", type=int, default=0, metavar="NCAND",
                    help="only autotune the tile policy for NCAND candidates and print it as JSON")
    args = ap.parse_args()
    if args.autotune_only:
        rc = importlib.import_module("rust-compress_amd")
        print(json.dumps(rc.Context(0).autotune(args.autotune_only).as_dict()), flush=True)
        return
    if args.copy_calibration:
        # One device-to-device copy of a known size (the traffic counters of a plain copy kernel).
        rc = importlib.import_module("rust-compress_amd")
        ctx = rc.Context(0)
        n = args.copy_calibration << 20
        a = rc.DeviceBuffer(ctx, n)
        b = rc.DeviceBuffer(ctx, n)
        rc.lib().rcg_device_copy(ctx.handle, b.ptr, a.ptr, n)
        ctx.synchronize()
        print(json.dumps({"copy_calibration_bytes": n}), flush=True)
        return
    if args.gpus != 1:
        print(f"bench.py: this build measures one GPU; --gpus {args.gpus} is not supported "
              "(images would shard across GPUs with no collective; only the one-GPU share is built)", file=sys.stderr)
        sys.exit(2)

    cands_n = 6  # q{50,75,90} x {4:2:0, 4:4:4}
    policy = json.loads(args.tile_policy) if args.tile_policy else None
    traffic = None
    if not args.no_traffic:
        # Child processes, before any GPU use here; the timed steps then run under the policy the passes measured.
        policy = policy or tuned_policy(cands_n)
        traffic = measure_traffic(args.size, policy) if policy else None
    rc = importlib.import_module("rust-compress_amd")
    ctx = rc.Context(0)
    W = H = args.size
    img = rc.synth_rgb8(2, W, H)
    d_img = ctx.upload(img)  # resident in HBM before the timed region
    ptr = d_img.ptr
    cands = [(q, s) for s in (rc.S420, rc.S444) for q in (50, 75, 90)]
    assert len(cands) == cands_n
    # Tile heights per scale-height class, measured on this GPU for this workload's candidate count.
    if policy:
        tune = rc.Tune()
        tune.wide_nt, tune.narrow_w, tune.nclass = policy["wide_nt"], policy["narrow_w"], len(policy["classes"])
        for i, (min_h, th) in enumerate(policy["classes"]):
            tune.min_h[i], tune.th[i] = min_h, th
        ctx.set_tune(tune)
        tune = ctx.get_tune()
        progress(f"installed tile policy: {tune.as_dict()}")
    else:
        tune = ctx.autotune(len(cands))
        progress(f"autotuned tile policy: {tune.as_dict()}")
    target = 70.0

    def step():
        return ctx.score_batch(None, cands, target, device_ptrs=[(ptr, W, H, 3)])

    for _ in range(args.warmup):
        res = step()
    ctx.synchronize()
    ctx.profile(True)
    t0 = time.perf_counter()
    for _ in range(args.steps):
        res = step()
    ctx.synchronize()
    t1 = time.perf_counter()
    ssim_ms, ssim_n = ctx.profile_get("k_ssim_scale")
    enc_ms, enc_n = ctx.profile_get("k_encode")
    fin_ms, fin_n = ctx.profile_get("k_ssim_finalize")
    ref_ms, ref_n = ctx.profile_get("k_ssim_ref")
    ctx.profile(False)
    ms_per_step = (t1 - t0) * 1000.0 / args.steps
    progress(f"headline workload timed: {ms_per_step:.3f} ms/step")
    mp = W * H / 1e6
    evals = len(cands)
    value = mp * evals / (ms_per_step / 1000.0)

    # Dominant kernel roofline: algorithmic bytes of all its launches / their summed duration.
    ssim_bytes = SSIM_BYTES_PER_PX * W * H * evals * args.steps
    achieved = ssim_bytes / (ssim_ms / 1000.0) / 1e9 if ssim_ms > 0 else 0.0
    r0 = res[0]
    out = {
        "metric": METRIC,
        "value": round(value, 2),
        "unit": "MP/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": f"plain drop-in: one {W}x{H} RGB8 synthetic image, q{{50,75,90}} x {{4:2:0,4:4:4}}, "
                               "trellis off, encode+reconstruct+SSIMULACRA2 per candidate",
                   "candidates": evals, "megapixels": mp},
        "roofline": {"bound": "hbm", "kernel": "k_ssim_scale", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
                     "traffic": round(traffic["bytes_per_launch"]) if traffic else None,
                     "traffic_detail": traffic,
                     "algorithmic_bytes_per_launch": round(SSIM_BYTES_PER_PX * W * H * evals / max(1, ssim_n / args.steps)),
                     "avg_launch_ms": round(ssim_ms / max(1, ssim_n), 4), "launches": ssim_n,
                     "bytes_per_px_per_candidate": SSIM_BYTES_PER_PX},
        "kernels_ms_per_step": {"k_ssim_scale": round(ssim_ms / args.steps, 3),
                                "k_ssim_ref": round(ref_ms / args.steps, 3),
                                "k_encode": round(enc_ms / args.steps, 3),
                                "k_ssim_finalize": round(fin_ms / args.steps, 3)},
        "tile_policy": tune.as_dict(),
        "scores": [round(x, 4) for x in list(r0.scores)[:evals]],
        "settled_fp64": int(sum(r0.settled[:evals])),
    }
    eval_bytes = sum(EVAL_BYTES_PER_PX[1 if s == rc.S444 else 3] for _, s in cands) * W * H
    out["eval_hbm_frac"] = round(eval_bytes / (ms_per_step / 1000.0) / 1e9 / HBM_PEAK_GBS, 4)
    # The same step from a host buffer (rcg_score_batch copies the source over PCIe on every call): the rate a caller
    # that hands over host memory sees. Not `value`, which starts with the source resident in HBM.
    ctx.score_batch([img], cands, target)
    ctx.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        ctx.score_batch([img], cands, target)
    ctx.synchronize()
    host_ms = (time.perf_counter() - t0) * 1000.0 / args.steps
    out["pcie_inclusive"] = {"ms_per_step": round(host_ms, 3), "value": round(mp * evals / (host_ms / 1000.0), 2),
                             "unit": "MP/s"}

    if not args.no_plain:
        ctx.find_quality_dev(ptr, W, H, 3, target, 100, rc.S420)
        ctx.synchronize()
        t0 = time.perf_counter()
        f = ctx.find_quality_dev(ptr, W, H, 3, target, 100, rc.S420)
        t1 = time.perf_counter()
        out["plain_search"] = {"quality": f.quality, "steps": f.steps, "ms": round((t1 - t0) * 1000, 2),
                               "search_mps": round(mp / (t1 - t0), 2),
                               "eval_mps": round(mp * f.steps / (t1 - t0), 2)}

    if not args.no_cpu:
        progress("timing the CPU oracle sample")
        from oracle import oracle as O  # the CPU baseline is the oracle, timed, never the measured path
        threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
        sample = [(50, O.S420), (90, O.S444)]
        t0 = time.perf_counter()
        for q, s in sample:
            O.ssimulacra2(img, O.reconstruct(O.build_coefs(img, q, s)))
        t1 = time.perf_counter()
        out["cpu_baseline"] = {"value": round(mp * len(sample) / (t1 - t0), 3), "unit": "MP/s", "cores": threads,
                               "kind": "port",
                               "sample": f"{len(sample)} candidate evaluations (q50 4:2:0, q90 4:4:4) of the same "
                                         f"{W}x{H} image: oracle build_coefs + reconstruct + f64 SSIMULACRA2"}
        progress("scores against the f64 oracle")
        out["precision"] = precision_report(rc, ctx, O, img, cands, r0)
    if not args.no_batch:
        progress("batched configuration")
        out["batched"] = bench_batch(rc, ctx, args)
    if not args.no_serve:
        progress("serving configuration")
        out["serving"] = bench_serve(rc, args, tune)
    if not args.no_cli:
        progress("command-line end-to-end run")
        out["cli"] = bench_cli(rc, args)
    print(json.dumps(out), flush=True)
    ctx.close()


CANDS_B = (40, 50, 60, 70, 80, 85, 90, 95)


def precision_report(rc, ctx, O, img, cands, r0):
    """Every score of the headline step, and the precision configuration (2048x2048 and 3000x2001 synthetic images
    against their q60 4:2:0 decode), with its error against the f64 oracle. The distorted images are the device's
    reconstructions (bit-exact with the oracle's); the oracle only scores them."""
    import numpy as np
    head = []
    for (q, s), got in zip(cands, list(r0.scores)[:len(cands)]):
        rec = ctx.reconstruct(ctx.build_coefs(img, q, s))
        want = O.ssimulacra2(img, rec)
        head.append({"quality": q, "subsampling": "4:4:4" if s == rc.S444 else "4:2:0", "fp32": round(got, 6),
                     "oracle_f64": round(want, 6), "abs_err": float(f"{abs(got - want):.3e}")})
    configs = []
    for w, h in ((2048, 2048), (3000, 2001)):
        src = rc.synth_rgb8(11, w, h)
        dist = ctx.reconstruct(ctx.build_coefs(src, 60, rc.S420))
        want, wst = O.ssimulacra2(src, dist, with_stats=True)
        s32, st32 = ctx.ssimulacra2(src, dist, rc.FP32, with_stats=True)
        s64, st64 = ctx.ssimulacra2(src, dist, rc.FP64, with_stats=True)
        configs.append({"size": f"{w}x{h}", "oracle_f64": round(want, 6), "fp32": round(s32, 6),
                        "fp64": round(s64, 9), "abs_err_fp32": float(f"{abs(s32 - want):.3e}"),
                        "abs_err_fp64": float(f"{abs(s64 - want):.3e}"),
                        "per_scale_max_abs_err_fp32": [float(f"{v:.3e}") for v in
                                                       np.abs(st32 - wst).max(axis=(1, 2))],
                        "per_scale_max_abs_err_fp64": [float(f"{v:.3e}") for v in
                                                       np.abs(st64 - wst).max(axis=(1, 2))]})
    return {"bound": 0.02, "headline": head, "max_abs_err_headline": max(x["abs_err"] for x in head),
            "configs": configs}


def bench_batch(rc, ctx, args):
    """256 synthetic RGB8 images, W and H uniform in [64, 3000], each swept over 8 candidate qualities at 4:2:0 in
    one batched pass; choice = lowest q with SSIMULACRA2 >= 70 by the reference's bisection over the sorted list.
    Compared with the plain per-image GPU search (serial bisection over q in 1..100) on the same images."""
    import numpy as np
    rng = np.random.default_rng(2024)
    sizes = [(int(rng.integers(64, 3001)), int(rng.integers(64, 3001))) for _ in range(args.batch_images)]
    t0 = time.perf_counter()
    dev = []
    for i, (w, h) in enumerate(sizes):
        dev.append(ctx.upload(rc.synth_rgb8(1000 + i, w, h)))
        if i % 64 == 63:
            progress(f"generated {i + 1} batch images")
    ctx.synchronize()
    gen_s = time.perf_counter() - t0
    ptrs = [(t.ptr, w, h, 3) for t, (w, h) in zip(dev, sizes)]
    cands = [(q, rc.S420) for q in CANDS_B]
    # A batch launch holds many pairs of every size and runs many rounds of workgroups: tune for throughput (time per
    # pair over launches of many rounds), not for one image's candidates.
    tune_b = ctx.autotune(0)
    total_mp = sum(w * h for w, h in sizes) / 1e6
    ctx.score_batch(None, cands, 70.0, device_ptrs=ptrs)  # warm-up (sizes the workspace)
    ctx.synchronize()
    ctx.profile(True)
    t0 = time.perf_counter()
    res = ctx.score_batch(None, cands, 70.0, device_ptrs=ptrs)
    ctx.synchronize()
    t1 = time.perf_counter()
    ssim_ms, ssim_n = ctx.profile_get("k_ssim_scale")
    ref_ms, _ = ctx.profile_get("k_ssim_ref")
    ctx.profile(False)
    batch_s = t1 - t0
    out = {"images": len(sizes), "megapixels": round(total_mp, 2), "candidates": list(CANDS_B),
           "seconds": round(batch_s, 4), "search_mps": round(total_mp / batch_s, 2),
           "eval_mps": round(total_mp * len(CANDS_B) / batch_s, 2),
           "hbm_frac": round(total_mp * 1e6 * (30.0 * len(CANDS_B) + 7.0) / batch_s / 1e9 / HBM_PEAK_GBS, 4),
           "k_ssim_scale_ms": round(ssim_ms, 2), "k_ssim_scale_launches": ssim_n, "k_ssim_ref_ms": round(ref_ms, 2),
           "tile_policy": tune_b.as_dict(),
           "settled_fp64": int(sum(sum(r.settled[:len(CANDS_B)]) for r in res)),
           "chosen_equals_sweep": int(sum(r.chosen_quality == r.sweep_quality for r in res)),
           "generate_s": round(gen_s, 1)}
    progress(f"batched pass: {batch_s:.3f} s")
    # The same pass captured once as a hipGraph and replayed.
    plan = ctx.batch_plan(ptrs, cands, 70.0)
    plan.run()
    ctx.synchronize()
    times = []
    for _ in range(3):  # the median of three replays
        t0 = time.perf_counter()
        gres = plan.run()
        ctx.synchronize()
        times.append(time.perf_counter() - t0)
    graph_s = sorted(times)[1]
    plan.close()
    out["graph_seconds"] = round(graph_s, 4)
    out["graph_search_mps"] = round(total_mp / graph_s, 2)
    out["graph_matches_eager"] = int(all(g.chosen_quality == r.chosen_quality for g, r in zip(gres, res)))
    if not args.no_plain:
        # The plain search scores one pair per launch: it gets the policy tuned for that, not the batch's.
        out["plain_tile_policy"] = ctx.autotune(1).as_dict()
        for p in ptrs[:4]:
            ctx.find_quality_dev(p[0], p[1], p[2], 3, 70.0, 100, rc.S420)  # warm-up under the new policy
        ctx.synchronize()
        t0 = time.perf_counter()
        steps = 0
        for p in ptrs:
            f = ctx.find_quality_dev(p[0], p[1], p[2], 3, 70.0, 100, rc.S420)
            steps += f.steps
        t1 = time.perf_counter()
        out["plain_search_mps"] = round(total_mp / (t1 - t0), 2)
        out["plain_eval_mps"] = round(total_mp * steps / len(ptrs) / (t1 - t0), 2)
        out["plain_seconds"] = round(t1 - t0, 3)
        out["batched_vs_plain"] = round(out["search_mps"] / out["plain_search_mps"], 2)
        # The batch as one captured hipGraph replay (the form the batched configuration names) against the same plain
        # per-image searches.
        out["graph_vs_plain"] = round(out["graph_search_mps"] / out["plain_search_mps"], 2)
    for d in dev:
        d.free()
    return out


def bench_serve(rc, args, tune=None):
    """Poisson arrivals at 200 req/s; megapixels log-uniform in [0.1, 24], aspect uniform in [0.5, 2]; targets
    uniform over {60, 70, 80}. Images come from a pool of 32 synthetic images drawn from that distribution (the
    host cannot hold one image per request); every request is an independent search. Latency = completion -
    submission."""
    import threading
    import numpy as np
    rng = np.random.default_rng(77)
    pool = []
    for i in range(32):
        mp = float(np.exp(rng.uniform(np.log(0.1), np.log(24.0))))
        aspect = float(rng.uniform(0.5, 2.0))
        w = max(8, int(round((mp * 1e6 * aspect) ** 0.5)))
        h = max(8, int(round(mp * 1e6 / w)))
        pool.append(rc.synth_rgb8(5000 + i, w, h))
    n = args.serve_requests
    picks = rng.integers(0, len(pool), n)
    targets = rng.choice([60.0, 70.0, 80.0], n)
    gaps = rng.exponential(1.0 / 200.0, n)
    srv = rc.Server(0, speculate=3, tune=tune)
    # Warm-up: one request initialises the worker's context before the clock starts.
    srv.submit(10 ** 9, pool[0], 70.0)
    while not srv.poll(8, 500):
        pass
    done = {}

    def submitter():
        t = time.perf_counter()
        for i in range(n):
            t += gaps[i]
            d = t - time.perf_counter()
            if d > 0:
                time.sleep(d)
            srv.submit(i, pool[picks[i]], targets[i])

    t0 = time.perf_counter()
    th = threading.Thread(target=submitter)
    th.start()
    while len(done) < n:
        for c in srv.poll(128, 200):
            done[c.id] = c
    t1 = time.perf_counter()
    th.join()
    st = srv.stats()
    srv.close()
    lat = np.array([(c.done_s - c.submit_s) * 1000.0 for c in done.values()])
    mp = sum(pool[picks[i]].shape[0] * pool[picks[i]].shape[1] for i in range(n)) / 1e6
    return {"requests": n, "offered_req_s": 200.0, "seconds": round(t1 - t0, 2),
            "achieved_req_s": round(n / (t1 - t0), 1), "search_mps": round(mp / (t1 - t0), 1),
            "p50_ms": round(float(np.percentile(lat, 50)), 1), "p99_ms": round(float(np.percentile(lat, 99)), 1),
            "mean_batch": round(st.mean_batch, 1), "ticks": int(st.ticks),
            "peak_footprint_gib": round(st.peak_footprint / 2 ** 30, 2),
            "failed": int(sum(c.status != 0 for c in done.values()))}


def bench_cli(rc, args):
    """The command-line driver end to end (python -m rust-compress_amd): PNG files in a directory, log-uniform
    0.3-4 MP, aspect 0.5-2, one --target-ssimulacra2 70 run with the default trellis and one with --no-trellis.
    Wall time includes reading and decoding the PNGs, the searches, entropy coding and writing the JPEGs."""
    import contextlib
    import io as _io
    import shutil
    import tempfile
    import numpy as np
    from PIL import Image
    import importlib
    cli = importlib.import_module("rust-compress_amd.cli")
    rng = np.random.default_rng(91)
    root = tempfile.mkdtemp(prefix="rcg-cli-", dir=os.environ.get("TMPDIR", "/tmp"))
    try:
        src = os.path.join(root, "in")
        os.makedirs(src)
        total_mp = 0.0
        for i in range(args.cli_files):
            mp = float(np.exp(rng.uniform(np.log(0.3), np.log(4.0))))
            aspect = float(rng.uniform(0.5, 2.0))
            w = max(8, int(round((mp * 1e6 * aspect) ** 0.5)))
            h = max(8, int(round(mp * 1e6 / w)))
            Image.fromarray(rc.synth_rgb8(8000 + i, w, h)).save(os.path.join(src, f"f{i:03d}.png"), compress_level=1)
            total_mp += w * h / 1e6
        res = {"files": args.cli_files, "megapixels": round(total_mp, 2), "target": 70.0}
        for name, extra in (("trellis", []), ("no_trellis", ["--no-trellis"])):
            out = os.path.join(root, name)
            buf = _io.StringIO()
            t0 = time.perf_counter()
            with contextlib.redirect_stdout(buf):
                code = cli.main([src, "-f", "jpeg", "--target-ssimulacra2", "70", "-o", out, "--quiet"] + extra)
            t1 = time.perf_counter()
            res[name] = {"seconds": round(t1 - t0, 3), "files_per_s": round(args.cli_files / (t1 - t0), 2),
                         "mp_per_s": round(total_mp / (t1 - t0), 2), "exit": code}
        return res
    finally:
        shutil.rmtree(root, ignore_errors=True)


if __name__ == "__main__":
    main()

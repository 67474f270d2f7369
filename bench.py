"""Benchmark of the JPEG target-quality search on one MI355X.

Headline workload (the batching configuration): 256 synthetic RGB8 images (seeded generator of rcg_synth_rgb8), width
and height each uniform in [64, 3000], each swept over 8 candidate qualities {40,50,60,70,80,85,90,95} at 4:2:0 with
trellis off; the choice is the lowest quality with SSIMULACRA2 >= 70 by the reference's bisection over the sorted
list. One step is one replay of the whole pass captured as a hipGraph (encode, reconstruction planes, reference
states, six SSIMULACRA2 scale launches, pooling), followed on the host by the FP64 settling of scores near the target
and the selection. The images are resident in HBM before the timed region.

value = search throughput in MP/s (image megapixels searched per second). Beside it, as the metric asks: the plain
per-image GPU search (serial bisection over q in 1..100, rcg_find_quality_dev) over the same images, the batched /
plain ratio, and the serving configuration's p50/p99 latency (Poisson arrivals at 200 requests/s). The roofline is
that of the pass's dominant kernel (k_ssim_scale), from HIP events recorded by the library around every launch of the
same pass run eagerly, on the stream each launch runs on; HBM traffic and VALU busy come from rocprofv3 counter
passes over one eager pass in child processes. The CPU baseline is the oracle (C restatement of the reference CPU
path, OpenMP threads) on a bounded sample of the same images: warm-up, then the median of 5 runs.

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
# Per image, once: build the reference pyramid (read 3 + write 4).
REF_BYTES_PER_PX = 7.0
CANDS_B = (40, 50, 60, 70, 80, 85, 90, 95)
TARGET_B = 70.0
SCALE0 = "k_ssim_scale<float, 256, 3u, 1u>"
MI355X_XCDS, MI355X_SIMDS = 8, 256 * 4  # accelerator dies and SIMDs (4 per CU) of the MI355X


def progress(msg):
    print(f"bench: {msg}", file=sys.stderr, flush=True)


def batch_sizes(n):
    """The batching configuration's image sizes: W and H each uniform in [64, 3000] (seed 2024)."""
    import numpy as np
    rng = np.random.default_rng(2024)
    return [(int(rng.integers(64, 3001)), int(rng.integers(64, 3001))) for _ in range(n)]


def upload_batch(rc, ctx, n):
    sizes = batch_sizes(n)
    dev = []
    for i, (w, h) in enumerate(sizes):
        dev.append(ctx.upload(rc.synth_rgb8(1000 + i, w, h)))
        if i % 64 == 63:
            progress(f"generated {i + 1} batch images")
    ctx.synchronize()
    return sizes, dev


def install_policy(rc, ctx, policy):
    tune = rc.Tune()
    tune.wide_nt, tune.narrow_w, tune.nclass = policy["wide_nt"], policy["narrow_w"], len(policy["classes"])
    for i, (min_h, th) in enumerate(policy["classes"]):
        tune.min_h[i], tune.th[i] = min_h, th
    ctx.set_tune(tune)
    return ctx.get_tune()


def _pmc_run(counters, bench_args, outdir, name):
    """One rocprofv3 pass of the counters named in `counters` (space-separated) over a child run of this script;
    returns [(kernel, dispatch, {counter: value}, duration_ns)] in dispatch order, or None if the pass failed."""
    import shutil
    import sqlite3
    import subprocess
    prof = shutil.which("rocprofv3")
    if prof is None:
        return None
    d = os.path.join(outdir, name)
    cmd = ["timeout", "-s", "KILL", "150", prof, "--pmc"] + counters.split() + ["-d", d, "-o", name, "--",
           sys.executable, os.path.abspath(__file__)] + bench_args
    r = subprocess.run(cmd, stdout=subprocess.DEVNULL, stderr=subprocess.DEVNULL, cwd=ROOT)
    dbs = []
    for root, _, files in os.walk(d):
        dbs += [os.path.join(root, f) for f in files if f.endswith("_results.db")]
    if not dbs:
        progress(f"{counters} pass failed (exit {r.returncode})")
        return None
    con = sqlite3.connect(dbs[0])
    rows = con.execute("select dispatch_id, kernel_name, counter_name, sum(value), max(\"end\" - start) "
                       "from counters_collection group by dispatch_id, counter_name order by dispatch_id").fetchall()
    con.close()
    out = {}
    for i, k, c, v, ns in rows:
        e = out.setdefault(i, (k.split("(")[0].replace("void rcg::", "").replace("rcg::", ""), i, {}, ns))
        e[2][c] = v
    return [out[i] for i in sorted(out)] or None


def tuned_policy(jobs):
    """A tile policy autotuned in a child process (so that this process has not touched the GPU when the counter
    passes run); None if the child failed."""
    import subprocess
    r = subprocess.run([sys.executable, os.path.abspath(__file__), "--autotune-only", str(jobs)],
                       stdout=subprocess.PIPE, stderr=subprocess.DEVNULL, cwd=ROOT, text=True)
    lines = r.stdout.strip().splitlines()
    if r.returncode != 0 or not lines:
        return None
    return json.loads(lines[-1])


def counter_passes(images, policy):
    """rocprofv3 counter passes over one eager pass of the batched configuration in child processes (each generates
    the images, installs `policy`, runs the pass once), before this process touches the GPU:
      traffic: FETCH_SIZE and WRITE_SIZE of every k_ssim_scale FP32 launch, each scaled by what the same counter reads
               for a 512 MiB device-to-device copy (that counter's calibration factor);
      VALU:    rocprofv3's VALUBusy (100 * SQ_ACTIVE_INST_VALU / CU_NUM / GRBM_GUI_ACTIVE) of each launch."""
    import tempfile
    out = tempfile.mkdtemp(prefix="rcg_counters_")
    child = ["--batch-pass-child", json.dumps(policy), "--batch-images", str(images)]
    copy = ["--copy-calibration", "512"]
    res = {}
    for counter in ("FETCH_SIZE", "WRITE_SIZE"):
        progress(f"counters: {counter} passes")
        cal = _pmc_run(counter, copy, out, counter.lower() + "_copy")
        run = _pmc_run(counter, child, out, counter.lower() + "_pass")
        if not cal or not run:
            return None, None
        kib = max(v[counter] for _, _, v, _ in cal)
        factor = (512 * 1024) / kib if kib > 0 else None
        vals = [v[counter] for k, _, v, _ in run if k.startswith("k_ssim_scale<float")]
        s0 = [v[counter] for k, _, v, _ in run if k.startswith(SCALE0)]
        if not vals or factor is None:
            return None, None
        res[counter] = {"kib_per_launch": sum(vals) / len(vals), "scale0_kib": sum(s0), "factor": factor,
                        "launches": len(vals)}
    f, w = res["FETCH_SIZE"], res["WRITE_SIZE"]
    fetch = f["kib_per_launch"] * f["factor"] * 1024
    write = w["kib_per_launch"] * w["factor"] * 1024
    traffic = {"bytes_per_launch": fetch + write, "fetch_bytes": fetch, "write_bytes": write,
               "scale0_bytes": f["scale0_kib"] * f["factor"] * 1024 + w["scale0_kib"] * w["factor"] * 1024,
               "fetch_factor": round(f["factor"], 3), "write_factor": round(w["factor"], 3),
               "launches": f["launches"]}
    progress("counters: VALUBusy pass")
    valu = None
    run = _pmc_run("VALUBusy SQ_INSTS_VALU GRBM_GUI_ACTIVE", child, out, "valu_pass")
    if run:
        def pick(prefix):
            v = [x["VALUBusy"] for k, _, x, _ in run if k.startswith(prefix) and "VALUBusy" in x]
            return round(v[0], 1) if v else None
        valu = {"k_ssim_scale_scale0_pct": pick(SCALE0), "k_encode_pct": pick("k_encode"),
                "k_ssim_ref_scale0_pct": pick("k_ssim_ref<256, 0u>"),
                "counter": "VALUBusy = 100*SQ_ACTIVE_INST_VALU/CU_NUM/GRBM_GUI_ACTIVE (rocprofv3), one eager pass"}
        s0 = [(x, ns) for k, _, x, ns in run if k.startswith(SCALE0)]
        if s0 and s0[0][1] > 0 and s0[0][0].get("GRBM_GUI_ACTIVE") and s0[0][0].get("SQ_INSTS_VALU"):
            # GRBM_GUI_ACTIVE sums the XCDs' busy cycles: per XCD over the launch's duration it is the clock. At that
            # clock every SIMD can issue one wave64 VALU instruction per 4 cycles; issue_frac is the share of that
            # peak the launch's SQ_INSTS_VALU fills.
            x, ns = s0[0]
            cycles = x["GRBM_GUI_ACTIVE"] / MI355X_XCDS
            valu["k_ssim_scale_scale0_clock_ghz"] = round(cycles / ns, 3)
            valu["k_ssim_scale_scale0_issue_frac"] = round(x["SQ_INSTS_VALU"] * 4 / (MI355X_SIMDS * cycles), 3)
            valu["issue_frac_def"] = ("SQ_INSTS_VALU*4/(1024 SIMDs * GRBM_GUI_ACTIVE/8 XCDs): share of one wave64 "
                                      "VALU instruction per SIMD per 4 cycles at the launch's clock")
    return traffic, valu


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20, help="timed replays of the batched pass")
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--batch-images", type=int, default=256)
    ap.add_argument("--size", type=int, default=4096, help="side of the plain drop-in image (extra figures)")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU oracle baseline and the precision report")
    ap.add_argument("--no-plain", action="store_true", help="skip the plain per-image search")
    ap.add_argument("--no-dropin", action="store_true", help="skip the plain drop-in (one 4096x4096 image) figures")
    ap.add_argument("--dropin-only", action="store_true",
                    help="only the plain drop-in figures, printed as {\"dropin\": ...} (for profiling that leg alone)")
    ap.add_argument("--stages-only", action="store_true",
                    help="only the forward and reconstruct stages on the drop-in image, printed as {\"stages\": ...}")
    ap.add_argument("--no-serve", action="store_true", help="skip the serving run")
    ap.add_argument("--no-cli", action="store_true", help="skip the command-line end-to-end run")
    ap.add_argument("--cli-files", type=int, default=24)
    ap.add_argument("--serve-requests", type=int, default=10000,
                    help="requests of the serving run (Poisson arrivals at 200 per second)")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="only run a device-to-device copy of MIB mebibytes (HBM traffic counter calibration)")
    ap.add_argument("--no-counters", action="store_true",
                    help="skip the rocprofv3 counter passes (traffic, VALU busy)")
    ap.add_argument("--tile-policy", default=None, metavar="JSON",
                    help="install this batch tile policy (as the bench line prints it) instead of autotuning")
    ap.add_argument("--autotune-only", type=int, default=-1, metavar="JOBS",
                    help="only autotune the tile policy for launches of JOBS pairs (0: throughput) and print it")
    ap.add_argument("--batch-pass-child", default=None, metavar="JSON",
                    help="only run one eager pass of the batched configuration under this tile policy")
    args = ap.parse_args()
    if args.autotune_only >= 0:
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
    if args.batch_pass_child:
        rc = importlib.import_module("rust-compress_amd")
        ctx = rc.Context(0)
        install_policy(rc, ctx, json.loads(args.batch_pass_child))
        sizes, dev = upload_batch(rc, ctx, args.batch_images)
        ptrs = [(t.ptr, w, h, 3) for t, (w, h) in zip(dev, sizes)]
        ctx.score_batch(None, [(q, rc.S420) for q in CANDS_B], TARGET_B, device_ptrs=ptrs)
        ctx.synchronize()
        return
    if args.gpus != 1:
        print(f"bench.py: this build measures one GPU; --gpus {args.gpus} is not supported "
              "(images would shard across GPUs with no collective; only the one-GPU share is built)", file=sys.stderr)
        sys.exit(2)

    if args.stages_only:
        rc = importlib.import_module("rust-compress_amd")
        ctx = rc.Context(0)
        d_img = ctx.upload(rc.synth_rgb8(2, args.size, args.size))
        print(json.dumps({"stages": bench_stages(rc, ctx, d_img.ptr, args.size, args.size, max(10, args.steps))}),
              flush=True)
        d_img.free()
        ctx.close()
        return
    if args.dropin_only:
        rc = importlib.import_module("rust-compress_amd")
        ctx = rc.Context(0)
        print(json.dumps({"dropin": bench_dropin(rc, ctx, args)}), flush=True)
        ctx.close()
        return

    # A batch launch holds many pairs of every size and runs many rounds of workgroups: its tile policy is tuned for
    # throughput (time per pair over launches of many rounds). Tuned in a child, so that the counter passes (children
    # too) run before this process touches the GPU, under the policy the timed steps use.
    policy = json.loads(args.tile_policy) if args.tile_policy else tuned_policy(0)
    traffic = valu = None
    if not args.no_counters and policy:
        traffic, valu = counter_passes(args.batch_images, policy)
    rc = importlib.import_module("rust-compress_amd")
    ctx = rc.Context(0)
    tune_b = install_policy(rc, ctx, policy) if policy else ctx.autotune(0)
    progress(f"batch tile policy: {tune_b.as_dict()}")
    t0 = time.perf_counter()
    sizes, dev = upload_batch(rc, ctx, args.batch_images)
    gen_s = time.perf_counter() - t0
    ptrs = [(t.ptr, w, h, 3) for t, (w, h) in zip(dev, sizes)]
    cands = [(q, rc.S420) for q in CANDS_B]
    total_px = sum(w * h for w, h in sizes)
    total_mp = total_px / 1e6

    # The timed steps: replays of the captured pass.
    plan = ctx.batch_plan(ptrs, cands, TARGET_B)
    for _ in range(args.warmup):
        plan.run()
    ctx.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        gres = plan.run()
    ctx.synchronize()
    t1 = time.perf_counter()
    plan.close()
    ms_per_step = (t1 - t0) * 1000.0 / args.steps
    value = total_mp / (ms_per_step / 1000.0)
    progress(f"batched pass as a graph: {ms_per_step:.2f} ms per replay")

    # The same pass eagerly, as many times, with HIP events around every launch on the stream it runs on: the
    # per-kernel times and the dominant kernel's roofline.
    res = ctx.score_batch(None, cands, TARGET_B, device_ptrs=ptrs)
    ctx.synchronize()
    ctx.profile(True)
    t0 = time.perf_counter()
    for _ in range(args.steps):
        res = ctx.score_batch(None, cands, TARGET_B, device_ptrs=ptrs)
    ctx.synchronize()
    eager_ms = (time.perf_counter() - t0) * 1000.0 / args.steps
    prof = {k: ctx.profile_get(k) for k in ("k_ssim_scale", "k_ssim_ref", "k_encode", "k_ssim_scale_f64",
                                            "k_ssim_finalize")}
    ctx.profile(False)
    ssim_ms, ssim_n = prof["k_ssim_scale"]
    # k_ssim_scale is the FP32 kernel (the FP64 settling launches are timed apart): every pair, every scale.
    launches_per_pass = ssim_n / args.steps
    alg_per_pass = SSIM_BYTES_PER_PX * total_px * len(CANDS_B)
    achieved = alg_per_pass * args.steps / (ssim_ms / 1000.0) / 1e9 if ssim_ms > 0 else 0.0
    eval_bytes = (EVAL_BYTES_PER_PX[3] * len(CANDS_B) + REF_BYTES_PER_PX) * total_px
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
        "config": {"workload": f"batched target-SSIMULACRA2 search: {len(sizes)} synthetic RGB8 images, W and H "
                               "uniform in [64,3000] (seed 2024), 8 candidates q{40,50,60,70,80,85,90,95} at 4:2:0, "
                               "trellis off, target 70; one hipGraph replay of the whole pass per step",
                   "images": len(sizes), "megapixels": round(total_mp, 2), "candidates": list(CANDS_B),
                   "target": TARGET_B},
        "roofline": {"bound": "hbm", "kernel": "k_ssim_scale", "achieved": round(achieved, 1),
                     "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
                     "traffic": round(traffic["bytes_per_launch"]) if traffic else None,
                     "traffic_detail": traffic,
                     "algorithmic_bytes_per_launch": round(alg_per_pass / max(1.0, launches_per_pass)),
                     "avg_launch_ms": round(ssim_ms / max(1, ssim_n), 4), "launches_per_pass": launches_per_pass,
                     "bytes_per_px_per_candidate": SSIM_BYTES_PER_PX,
                     "timed": f"{args.steps} eager passes of the same batch, HIP events around every launch"},
        "valu": valu,
        "search_mps_batched": round(value, 2),
    }
    out["kernels_ms_per_pass"] = {k: round(v[0] / args.steps, 3) for k, v in prof.items()}
    out["eager_pass_ms"] = round(eager_ms, 3)
    out["eager_search_mps"] = round(total_mp / (eager_ms / 1000.0), 2)
    out["pass_hbm_frac"] = round(eval_bytes / (ms_per_step / 1000.0) / 1e9 / HBM_PEAK_GBS, 4)
    out["batch"] = {"tile_policy": tune_b.as_dict(), "generate_s": round(gen_s, 1),
                    "settled_fp64": int(sum(sum(r.settled[:len(CANDS_B)]) for r in res)),
                    "chosen_equals_sweep": int(sum(r.chosen_quality == r.sweep_quality for r in res)),
                    "graph_matches_eager": int(all(g.chosen_quality == r.chosen_quality and
                                                   list(g.scores[:8]) == list(r.scores[:8])
                                                   for g, r in zip(gres, res)))}
    if not args.no_plain:
        progress("plain per-image search over the same images")
        out.update(bench_plain(rc, ctx, ptrs, total_mp, value))
        install_policy(rc, ctx, tune_b.as_dict())
    if not args.no_serve:
        progress("serving configuration")
        # A serving tick holds a few images with three candidates each (speculate = 3): tuned for that launch shape.
        tune_s = ctx.autotune(3)
        install_policy(rc, ctx, tune_b.as_dict())
        sv = bench_serve(rc, args, tune_s)
        sv["tile_policy"] = tune_s.as_dict()
        out["p50_ms"] = sv["p50_ms"]
        out["p99_ms"] = sv["p99_ms"]
        out["serving"] = sv
    if not args.no_cpu:
        progress("timing the CPU oracle sample")
        out["cpu_baseline"] = bench_cpu(rc, sizes)
    for d in dev:
        d.free()
    if not args.no_dropin:
        progress("plain drop-in figures (one 4096x4096 image, 6 candidates)")
        out["dropin"] = bench_dropin(rc, ctx, args)
    if not args.no_cli:
        progress("command-line end-to-end run")
        out["cli"] = bench_cli(rc, args)
    print(json.dumps(out), flush=True)
    ctx.close()


def bench_plain(rc, ctx, ptrs, total_mp, value):
    """The plain per-image GPU search (serial bisection over 1..100, one candidate per step) over the batch's images,
    under the tile policy tuned for one pair per launch: the median of three full runs."""
    pol = ctx.autotune(1)
    for p in ptrs[:4]:
        ctx.find_quality_dev(p[0], p[1], p[2], 3, TARGET_B, 100, rc.S420)  # warm-up under the new policy
    ctx.synchronize()
    runs = []
    for _ in range(3):
        t0 = time.perf_counter()
        steps = 0
        for p in ptrs:
            steps += ctx.find_quality_dev(p[0], p[1], p[2], 3, TARGET_B, 100, rc.S420).steps
        runs.append((time.perf_counter() - t0, steps))
    secs, steps = sorted(runs)[1]
    plain = total_mp / secs
    return {"plain_search_mps": round(plain, 2), "batched_vs_plain": round(value / plain, 2),
            "plain": {"seconds": round(secs, 3), "eval_mps": round(total_mp * steps / len(ptrs) / secs, 2),
                      "mean_steps": round(steps / len(ptrs), 2), "tile_policy": pol.as_dict(),
                      "runs_s": [round(r[0], 3) for r in runs]}}


def bench_cpu(rc, sizes):
    """The oracle on a bounded sample of the batched workload: the first images of the batch (about 10 MP) swept over
    the same 8 candidates (build_coefs + reconstruct + f64 SSIMULACRA2 per candidate); one warm-up run, then the
    median of 5. value is search-MP/s (image megapixels per second), as the headline."""
    from oracle import oracle as O  # the CPU baseline is the oracle, timed, never the measured path
    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    sample, mp = [], 0.0
    for i, (w, h) in enumerate(sizes):
        sample.append(rc.synth_rgb8(1000 + i, w, h))
        mp += w * h / 1e6
        if mp >= 10.0:
            break

    def run():
        for img in sample:
            for q in CANDS_B:
                O.ssimulacra2(img, O.reconstruct(O.build_coefs(img, q, O.S420)))

    run()
    times = []
    for _ in range(5):
        t0 = time.perf_counter()
        run()
        times.append(time.perf_counter() - t0)
    med = sorted(times)[2]
    return {"value": round(mp / med, 3), "unit": "MP/s", "cores": threads, "kind": "port",
            "eval_mps": round(mp * len(CANDS_B) / med, 3), "runs_s": [round(t, 3) for t in times],
            "sample": f"the first {len(sample)} images of the batched configuration ({mp:.2f} MP) x 8 candidates "
                      "at 4:2:0: oracle build_coefs + reconstruct + f64 SSIMULACRA2, warm-up then median of 5; "
                      "value in search-MP/s"}


def bench_dropin(rc, ctx, args):
    """The plain drop-in configuration: one 4096x4096 image, q{50,75,90} x {4:2:0, 4:4:4}, every candidate scored
    per step (rcg_score_batch_dev), source resident; plus the plain search on it, its scores against the f64 oracle,
    the precision configuration and the PCIe-inclusive rate."""
    W = H = args.size
    img = rc.synth_rgb8(2, W, H)
    d_img = ctx.upload(img)
    cands = [(q, s) for s in (rc.S420, rc.S444) for q in (50, 75, 90)]
    ctx.autotune(len(cands))
    step = lambda: ctx.score_batch(None, cands, TARGET_B, device_ptrs=[(d_img.ptr, W, H, 3)])
    for _ in range(3):
        r0 = step()[0]
    ctx.synchronize()
    n = max(10, args.steps)
    t0 = time.perf_counter()
    for _ in range(n):
        r0 = step()[0]
    ctx.synchronize()
    ms = (time.perf_counter() - t0) * 1000.0 / n
    mp = W * H / 1e6
    out = {"workload": f"one {W}x{H} RGB8 image, q{{50,75,90}} x {{4:2:0,4:4:4}}, every candidate per step",
           "ms_per_step": round(ms, 3), "candidate_mps": round(mp * len(cands) / (ms / 1000.0), 2),
           "scores": [round(x, 4) for x in list(r0.scores)[:len(cands)]]}
    # The same step captured once as a hipGraph and replayed (rcg_batch_plan): what a caller that scores the same
    # device buffers repeatedly pays, without the per-call host staging and launches.
    plan = ctx.batch_plan([(d_img.ptr, W, H, 3)], cands, TARGET_B)
    for _ in range(3):
        plan.run()
    ctx.synchronize()
    t0 = time.perf_counter()
    for _ in range(n):
        g0 = plan.run()[0]
    ctx.synchronize()
    gms = (time.perf_counter() - t0) * 1000.0 / n
    plan.close()
    out["graph"] = {"ms_per_step": round(gms, 3), "candidate_mps": round(mp * len(cands) / (gms / 1000.0), 2),
                    "scores_equal_eager": list(g0.scores)[:len(cands)] == list(r0.scores)[:len(cands)]}
    ctx.score_batch([img], cands, TARGET_B)
    t0 = time.perf_counter()
    for _ in range(n):
        ctx.score_batch([img], cands, TARGET_B)
    ctx.synchronize()
    host_ms = (time.perf_counter() - t0) * 1000.0 / n
    out["pcie_inclusive"] = {"ms_per_step": round(host_ms, 3),
                             "candidate_mps": round(mp * len(cands) / (host_ms / 1000.0), 2)}
    ctx.find_quality_dev(d_img.ptr, W, H, 3, TARGET_B, 100, rc.S420)
    ctx.synchronize()
    t0 = time.perf_counter()
    f = ctx.find_quality_dev(d_img.ptr, W, H, 3, TARGET_B, 100, rc.S420)
    t1 = time.perf_counter()
    out["plain_search"] = {"quality": f.quality, "steps": f.steps, "ms": round((t1 - t0) * 1000, 2),
                           "search_mps": round(mp / (t1 - t0), 2)}
    out["stages"] = bench_stages(rc, ctx, d_img.ptr, W, H, n)
    if not args.no_cpu:
        from oracle import oracle as O
        out["precision"] = precision_report(rc, ctx, O, img, cands, r0)
    d_img.free()
    return out


def bench_stages(rc, ctx, d_src, W, H, n):
    """The forward stage (build_coefs: RGB8 -> quantised coefficients, k_encode) and the reconstruct stage
    (coefficients -> RGB8, k_reconstruct) alone, at quality 75, 4:2:0 and 4:4:4, on the device-resident drop-in
    image: each kernel's average launch time by HIP events over n calls, against its algorithmic bytes (read RGB8 and
    write the int16 coefficients; read the coefficients and write RGB8)."""
    res = {}
    for sub, name in ((rc.S420, "4:2:0"), (rc.S444, "4:4:4")):
        L = rc.coef_layout(W, H, 3, 75, sub)
        grids = [(int(L.blocks_w[c]), int(L.blocks_h[c])) for c in range(3)]
        coef_bytes = sum(bw * bh * 64 * 2 for bw, bh in grids)
        bufs = [rc.DeviceBuffer(ctx, bw * bh * 64 * 2) for bw, bh in grids]
        ptrs = [b.ptr for b in bufs]
        out = rc.DeviceBuffer(ctx, W * H * 3)
        qt = rc.quant_tables(75)
        comps = [(c + 1, 2 if (c == 0 and sub == rc.S420) else 1, 2 if (c == 0 and sub == rc.S420) else 1,
                  0 if c == 0 else 1, grids[c][0], grids[c][1], ptrs[c]) for c in range(3)]
        fwd = lambda: ctx.build_coefs_dev(d_src, W, H, 3, 75, sub, ptrs)
        rec = lambda: ctx.reconstruct_dev(W, H, comps, [qt[0], qt[1], None, None], out.ptr)
        entry = {}
        for stage, call, kernel, alg in (("forward", fwd, "k_encode", 3 * W * H + coef_bytes),
                                         ("reconstruct", rec, "k_reconstruct", coef_bytes + 3 * W * H)):
            for _ in range(3):
                call()
            ctx.synchronize()
            ctx.profile(True)
            for _ in range(n):
                call()
            ctx.synchronize()
            ms, k = ctx.profile_get(kernel)
            ctx.profile(False)
            us = ms * 1000.0 / max(1, k)
            gbs = alg / (us * 1e-6) / 1e9
            entry[stage] = {"kernel": kernel, "avg_launch_us": round(us, 2), "launches": k,
                            "algorithmic_bytes": alg, "achieved_gbs": round(gbs, 1),
                            "frac": round(gbs / HBM_PEAK_GBS, 4), "mp_per_s": round(W * H / us, 1)}
        res[name] = entry
        for b in bufs:
            b.free()
        out.free()
    return res


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

"""Benchmark of the JPEG target-quality inner loop on one MI355X.

Workload (the plain-drop-in configuration): one 4096x4096 RGB8 synthetic image (seeded generator of
rcg_synth_rgb8), candidates q{50,75,90} x {4:2:0, 4:4:4}, trellis off. One step scores all six candidates: encode
(colour, downsample, FDCT, quantisation) + dequantise/IDCT reconstruction + SSIMULACRA2 (FP32 maps, FP64 pooling)
against the source, in one launch sequence, with the source resident in HBM before the timed region.

value = candidate-evaluation throughput in MP/s (image megapixels x candidates / second). The same run also reports
the plain per-image GPU search (serial bisection over q in 1..100, rcg_find_quality_dev) in search-MP/s, the roofline
of the dominant kernel (k_ssim_scale) from HIP events recorded on the library's own stream around every launch in
the timed region, and the CPU oracle (C restatement of the reference CPU path, OpenMP threads) on a bounded sample.
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
# cached pyramid (4).
SSIM_BYTES_PER_PX = 18.0
# Whole candidate evaluation: + read RGB8 (3), write/read coefficients (3 + 3, or 6 + 6 at 4:4:4), write RGB8 (3).
EVAL_BYTES_PER_PX = {3: 30.0, 1: 36.0}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--size", type=int, default=4096)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU oracle baseline")
    ap.add_argument("--no-plain", action="store_true", help="skip the plain per-image search")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this build measures one GPU; --gpus {args.gpus} is not supported "
              "(images shard across GPUs with no collective; see DESIGN.md)", file=sys.stderr)
        sys.exit(2)

    import torch
    rc = importlib.import_module("rust-compress_amd")
    ctx = rc.Context(0)
    W = H = args.size
    img = rc.synth_rgb8(2, W, H)
    d_img = torch.from_numpy(img).to("cuda")
    torch.cuda.synchronize()
    ptr = d_img.data_ptr()
    cands = [(q, s) for s in (rc.S420, rc.S444) for q in (50, 75, 90)]
    target = 70.0

    def step():
        return ctx.score_batch(None, cands, target, device_ptrs=[(ptr, W, H, 3)])

    for _ in range(args.warmup):
        res = step()
    torch.cuda.synchronize()
    ctx.profile(True)
    t0 = time.perf_counter()
    for _ in range(args.steps):
        res = step()
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    ssim_ms, ssim_n = ctx.profile_get("k_ssim_scale")
    enc_ms, enc_n = ctx.profile_get("k_encode")
    fin_ms, fin_n = ctx.profile_get("k_ssim_finalize")
    ctx.profile(False)
    ms_per_step = (t1 - t0) * 1000.0 / args.steps
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
                     "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": None,
                     "avg_launch_ms": round(ssim_ms / max(1, ssim_n), 4), "launches": ssim_n,
                     "bytes_per_px_per_candidate": SSIM_BYTES_PER_PX},
        "kernels_ms_per_step": {"k_ssim_scale": round(ssim_ms / args.steps, 3),
                                "k_encode": round(enc_ms / args.steps, 3),
                                "k_ssim_finalize": round(fin_ms / args.steps, 3)},
        "scores": [round(x, 4) for x in list(r0.scores)[:evals]],
        "settled_fp64": int(sum(r0.settled[:evals])),
    }
    eval_bytes = sum(EVAL_BYTES_PER_PX[1 if s == rc.S444 else 3] for _, s in cands) * W * H
    out["eval_hbm_frac"] = round(eval_bytes / (ms_per_step / 1000.0) / 1e9 / HBM_PEAK_GBS, 4)

    if not args.no_plain:
        ctx.find_quality_dev(ptr, W, H, 3, target, 100, rc.S420)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        f = ctx.find_quality_dev(ptr, W, H, 3, target, 100, rc.S420)
        t1 = time.perf_counter()
        out["plain_search"] = {"quality": f.quality, "steps": f.steps, "ms": round((t1 - t0) * 1000, 2),
                               "search_mps": round(mp / (t1 - t0), 2),
                               "eval_mps": round(mp * f.steps / (t1 - t0), 2)}

    if not args.no_cpu:
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
    print(json.dumps(out), flush=True)
    ctx.close()


if __name__ == "__main__":
    main()

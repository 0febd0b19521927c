#!/usr/bin/env python3
"""Benchmark: batched CONTRAfold folds at 200 nt on one MI355X.

Step = one pass of the fold engine over the resident batch of the
"batched fold, fixed length" workload: 10,000 sequences x 200 nt, GC fraction
per sequence ~ Beta(8, 8), per-base forced-unpaired mask Bernoulli(0.1); FP32
logZ, full bpp via inside-outside, 100 Boltzmann samples per sequence
(Philox streams).  value = sequences folded per second (folds/s).

The same JSON line carries
  * roofline: the dominant kernel's algorithmic FP32 flops per launch
    (9 per log-sum-exp / Fast_Exp term, 2 per max-plus term, counted exactly on
    a sample of the batch) over its average launch time, measured with HIP
    events on the stream the kernels run on;
  * cpu_baseline: the oracle (C restatement of the reference's CPU engine) on
    all host cores over a bounded sample of the same sequences;
  * mea: 200-nt MEA folds/s (gamma 6) of the same batch and its ratio to the
    host CPU's all-core MEA throughput;
  * serving (with --serve): p50/p99 latency and served folds/s of the serving
    workload next to plain batched throughput of the same mix.

    python bench.py [--gpus 1] [--steps 3] [--warmup 1] [--n 10000] [--serve]
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

FP32_PEAK_TFLOPS = 157.3     # MI355X FP32 vector/matrix peak (dense)
HBM_PEAK_GBS = 8000.0


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def cpu_baseline(seqs, masks, P, seconds_target=15.0):
    """Oracle on all host cores over a bounded sample of the workload."""
    from oracle import oracle as O
    O.build()
    cores = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else (os.cpu_count() or 1)
    # calibrate on one fold per core, then size the sample to ~seconds_target
    probe = seqs[:cores]
    t = time.perf_counter()
    O.run_batch("bpp32", probe, P, masks=masks[:cores], threads=cores)
    dt = max(time.perf_counter() - t, 1e-3)
    per_s = len(probe) / dt
    n = int(min(len(seqs), max(cores, per_s * seconds_target)))
    t = time.perf_counter()
    O.run_batch("bpp32", seqs[:n], P, masks=masks[:n], threads=cores)
    dt = time.perf_counter() - t
    res = {"value": round(n / dt, 2), "unit": "folds/s", "cores": cores, "kind": "port",
           "sample": f"{n} of the 200-nt sequences, FP32 inside+outside+posterior (logZ + bpp) of the oracle "
                     f"(C restatement of gpu_contrafold/cpu.py:583-820, -O3, no FMA); sampling is not part "
                     f"of the CPU fold", "seconds": round(dt, 2)}
    # MEA (posterior + decode, gamma 6), the headline CPU comparison
    n2 = max(cores, n // 2)
    t = time.perf_counter()
    O.run_batch("mea32", seqs[:n2], P, masks=masks[:n2], gamma=6.0, threads=cores)
    dt2 = time.perf_counter() - t
    res["mea_folds_per_s"] = round(n2 / dt2, 2)
    return res


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--n", type=int, default=10000)
    ap.add_argument("--length", type=int, default=200)
    ap.add_argument("--samples", type=int, default=100)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--serve", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one GPU: every fold is independent and this build runs on a single MI355X; "
              "run with --gpus 1", file=sys.stderr)
        sys.exit(2)

    import contrafold_gpu_amd as cf
    from contrafold_gpu_amd import workcount, workloads
    P = cf.load()
    log(f"[bench] generating {args.n} x {args.length} nt (GC ~ Beta(8,8), mask p=0.1, seed 2)")
    seqs, masks = workloads.gc_masked(args.n, args.length, seed=2)
    eng = cf.default_engine()
    eng.set_params(P)
    batch = eng.batch(seqs, masks)
    flags = cf.CF_BPP | cf.CF_SAMPLE
    log(f"[bench] batch resident: {batch.device_bytes() / 2**20:.0f} MiB before run buffers")
    for w in range(args.warmup):
        t = time.perf_counter()
        batch.run(flags, n_samples=args.samples, seed=w, recompute=True, timing=True)
        log(f"[bench] warmup {w}: {time.perf_counter() - t:.3f} s")
    try:
        import torch
        sync = torch.cuda.synchronize if torch.cuda.is_available() else (lambda: None)
    except Exception:
        sync = lambda: None  # noqa: E731
    kms = []
    dev_ms = []
    sync()
    t0 = time.perf_counter()
    for k in range(args.steps):
        batch.run(flags, n_samples=args.samples, seed=100 + k, recompute=True, timing=True)
        kms.append(batch.kernel_ms())
        dev_ms.append(batch.last_ms())
        log(f"[bench] step {k}: {dev_ms[-1]:.1f} ms device")
    sync()
    wall = time.perf_counter() - t0
    ms_per_step = wall / args.steps * 1e3
    value = args.n / (wall / args.steps)
    kavg = {name: float(np.mean([d[name] for d in kms])) for name in kms[0]}

    # algorithmic work (exact counts on a sample, scaled to the batch)
    wc = workcount.batch_counts(seqs, masks, sample=24)
    flops = {"inside": 9.0 * wc["inside"], "exterior": 9.0 * wc["exterior"],
             "outside": 9.0 * (wc["outside"] + wc["posterior"]), "mea": 2.0 * wc["mea"]}
    dominant = max(kavg, key=kavg.get)
    roof = None
    if dominant in flops and kavg[dominant] > 0:
        ach = flops[dominant] / (kavg[dominant] * 1e-3) / 1e12
        roof = {"kernel": dominant, "bound": "mfma", "achieved": round(ach, 3), "peak": FP32_PEAK_TFLOPS,
                "unit": "TFLOP/s", "frac": round(ach / FP32_PEAK_TFLOPS, 5), "traffic": None,
                "ceiling": "FP32 VALU (no MFMA: log-space DP); peak = MI355X FP32 157.3 TFLOP/s"}
    else:
        # sampling dominates: report it against HBM with the bytes it must read
        roof = {"kernel": dominant, "bound": "hbm", "achieved": None, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": None, "traffic": None}
    per_kernel = {k: {"ms": round(v, 3),
                      "tflops": (round(flops[k] / (v * 1e-3) / 1e12, 3) if k in flops and v > 0 else None)}
                  for k, v in kavg.items()}

    # MEA headline on the same resident batch (inside/outside recomputed)
    batch.run(cf.CF_MEA, gammas=[6.0], recompute=True)
    t = time.perf_counter()
    mea_steps = max(1, args.steps)
    for k in range(mea_steps):
        batch.run(cf.CF_MEA, gammas=[6.0], recompute=True)
    mea_fps = args.n * mea_steps / (time.perf_counter() - t)
    log(f"[bench] MEA folds/s: {mea_fps:.0f}")
    batch.close()

    out = {
        "metric": "folds/s @200 nt (served vs plain batch) + p50/p99 latency, logZ err <=5e-4",
        "value": round(value, 2), "unit": "folds/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 3), "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "fp32", "data": "synthetic",
        "config": {"workload": f"batched fold, fixed length: {args.n} x {args.length} nt, GC~Beta(8,8), "
                               f"mask Bernoulli(0.1); FP32 logZ + bpp (inside-outside) + {args.samples} "
                               f"Boltzmann samples/seq (Philox)",
                   "batch": args.n, "seq_len": args.length, "samples_per_seq": args.samples},
        "roofline": roof,
        "kernels": per_kernel,
        "mea": {"folds_per_s": round(mea_fps, 2), "gamma": 6.0},
    }
    if not args.no_cpu:
        log("[bench] CPU baseline (oracle, all host cores) ...")
        cb = cpu_baseline(seqs, masks, P)
        out["cpu_baseline"] = cb
        out["mea"]["cpu_folds_per_s"] = cb["mea_folds_per_s"]
        out["mea"]["vs_cpu_all_cores"] = round(mea_fps / cb["mea_folds_per_s"], 2)
    if args.serve:
        from contrafold_gpu_amd import serving
        out["serving"] = serving.benchmark(eng, P, log=log)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()

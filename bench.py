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
FP32_UNPACKED_TFLOPS = 78.6  # FP32 VALU issue ceiling without packed ops


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def cpu_baseline(seqs, masks, P, n_samples, seconds_target=15.0):
    """Oracle on all host cores over a bounded sample of the workload."""
    from oracle import oracle as O
    O.build()
    cores = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else (os.cpu_count() or 1)
    # calibrate on one fold per core, then size the sample to ~seconds_target
    probe = seqs[:cores]
    t = time.perf_counter()
    O.run_batch("bpp_sample32", probe, P, masks=masks[:cores], threads=cores, n_samples=n_samples)
    dt = max(time.perf_counter() - t, 1e-3)
    per_s = len(probe) / dt
    n = int(min(len(seqs), max(cores, per_s * seconds_target)))
    t = time.perf_counter()
    O.run_batch("bpp_sample32", seqs[:n], P, masks=masks[:n], threads=cores, n_samples=n_samples)
    dt = time.perf_counter() - t
    # one core, over a few seconds of the same work
    n1 = int(min(len(seqs), max(2, per_s / cores * 3.0)))
    t = time.perf_counter()
    O.run_batch("bpp_sample32", seqs[:n1], P, masks=masks[:n1], threads=1, n_samples=n_samples)
    one_core = n1 / (time.perf_counter() - t)
    res = {"value": round(n / dt, 2), "unit": "folds/s", "cores": cores, "kind": "port",
           "one_core_value": round(one_core, 2),
           "sample": f"first {n} of the same 200-nt sequences and masks: FP32 inside + outside + posterior "
                     f"(logZ + bpp) and {n_samples} Boltzmann samples per sequence, by the oracle (C restatement "
                     f"of gpu_contrafold/cpu.py:583-820 and gpu.py:196-477, -O3, no FMA), one fold per "
                     f"task on a pthread pool of {cores} threads", "seconds": round(dt, 2)}
    # MEA (posterior + decode, gamma 6), the headline CPU comparison
    n2 = max(cores, n // 2)
    t = time.perf_counter()
    O.run_batch("mea32", seqs[:n2], P, masks=masks[:n2], gamma=6.0, threads=cores)
    dt2 = time.perf_counter() - t
    res["mea_folds_per_s"] = round(n2 / dt2, 2)
    return res


def run_config(name, args):
    """Secondary configurations (parity-scale cases with their own measurements)."""
    import contrafold_gpu_amd as cf
    from contrafold_gpu_amd import workloads, autotune
    from contrafold_gpu_amd._lib import CF_GRAPH
    P = cf.load()
    eng = cf.default_engine()
    eng.set_params(P)
    out = {"config": name}
    if name == "c1":
        # 1,000 x 128 nt uniform: FP32 logZ + MEA (gamma 6) + Viterbi; oracle compares on the CPU
        seqs = workloads.uniform(1000, 128, seed=0)
        with eng.batch(seqs) as b:
            b.run(cf.CF_MEA | cf.CF_MFE, gammas=[6.0])
            t = time.perf_counter()
            for _ in range(3):
                b.run(cf.CF_MEA | cf.CF_MFE, gammas=[6.0], recompute=True)
            dt = (time.perf_counter() - t) / 3
            z = b.logz()
            mea = b.mea(0)
            vit, _ = b.mfe()
        from oracle import oracle as O
        O.build()
        k = min(args.check, len(seqs))
        z64, _ = O.run_batch("logz64", seqs[:k], P)
        zz, pairs = O.run_batch("mea32", seqs[:k], P, gamma=6.0)
        _v, vp = O.run_batch("viterbi64", seqs[:k], P)
        out.update({"folds_per_s": round(len(seqs) / dt, 1),
                    "logz_max_abs_err_vs_fp32_oracle": float(np.abs(z[:k] - zz.astype(np.float32)).max()),
                    "logz_max_abs_err_vs_fp64": float(np.abs(z[:k].astype(np.float64) - z64).max()),
                    "mea_identical": int(sum(mea[i] == O.dot_bracket(pairs[i], 128) for i in range(k))),
                    "viterbi_identical": int(sum(vit[i] == O.dot_bracket(vp[i], 128) for i in range(k))),
                    "checked": k})
    elif name == "c3":
        # 20,000 x U[30,500] nt with 2% N: MEA at 3 gammas + Viterbi, length buckets, hipGraph, autotuned
        seqs = workloads.uneven(args.n if args.n != 10000 else 20000, seed=3)
        flags = cf.CF_MEA | cf.CF_MFE
        tuned = autotune.tune(eng, seqs, flags=flags, gammas=(0.5, 6.0, 16.0), sample=1500, repeats=1, log=log)
        with eng.batch(seqs) as b:
            b.run(flags | CF_GRAPH, gammas=[0.5, 6.0, 16.0])
            ms = []
            t = time.perf_counter()
            for _ in range(args.steps):
                b.run(flags | CF_GRAPH, gammas=[0.5, 6.0, 16.0])
                ms.append(b.last_ms())
            dt = (time.perf_counter() - t) / args.steps
        out.update({"folds_per_s": round(len(seqs) / dt, 1), "device_ms": round(float(np.mean(ms)), 2),
                    "tuning": {"edges": tuned["edges"], "threads": tuned["threads"]}, "graph": True})
        eng.set_tuning(*autotune.DEFAULT)
    elif name == "c4":
        # 512 x 150 nt: gradient of mean logZ w.r.t. the 708 raw weights vs FP64 central differences
        from contrafold_gpu_amd import params
        names, raw, pmap = params.load_raw()
        seqs = workloads.uniform(512, 150, seed=4)
        w = np.full(len(seqs), 1.0 / len(seqs), np.float32)
        with eng.batch(seqs) as b:
            b.run(cf.CF_GRAD, weights=w)
            t = time.perf_counter()
            for _ in range(3):
                b.run(cf.CF_GRAD, weights=w, recompute=True)
            dt = (time.perf_counter() - t) / 3
            g = pmap.raw_grad(b.grad())
        from oracle import oracle as O
        O.build()
        rng = np.random.default_rng(4)
        idx = rng.choice(len(raw), 32, replace=False)
        sub = seqs[:args.check]
        h = 1e-4
        rows = []
        for kk in idx:
            e = np.zeros_like(raw)
            e[kk] = h
            fp = O.run_batch("logz64", sub, params.unflatten(pmap.tables(raw + e)))[0].mean()
            fm = O.run_batch("logz64", sub, params.unflatten(pmap.tables(raw - e)))[0].mean()
            rows.append((names[kk], float((fp - fm) / (2 * h))))
        # the GPU gradient over the same subset for the comparison
        with eng.batch(sub) as b:
            b.run(cf.CF_GRAD, weights=np.full(len(sub), 1.0 / len(sub), np.float32))
            gs = pmap.raw_grad(b.grad())
        err = [abs(gs[names.index(n)] - fd) for n, fd in rows]
        out.update({"grad_seconds": round(dt, 4), "seqs": len(seqs),
                    "fd_checked_params": len(rows), "fd_seqs": len(sub),
                    "fd_nonzero": int(sum(abs(fd) > 1e-9 for _, fd in rows)),
                    "max_abs_err": float(max(err)),
                    "max_rel_err_nonzero": float(max([e / abs(fd) for e, (_, fd) in zip(err, rows) if abs(fd) > 1e-3]
                                                     or [0.0]))})
    elif name == "latency":
        # single-fold latency floor: one uniform-ACGU sequence alone on the GPU,
        # one wave (packed layout) vs a 1,024-thread workgroup, device ms per kind
        rng = np.random.default_rng(1)
        rows = []
        for L in (100, 200, 400, 600, 1000):
            seq = "".join(rng.choice(list("ACGU"), L))
            for w in (64, 1024):
                eng.set_tuning([1 << 30], [w])
                with eng.batch([seq]) as b:
                    row = {"L": L, "threads": w}
                    for flags, kw, kind in ((cf.CF_MEA, dict(gammas=[6.0]), "mea"), (cf.CF_BPP, {}, "bpp"),
                                            (cf.CF_SAMPLE, dict(n_samples=10), "sample10")):
                        b.run(flags, recompute=True, **kw)
                        b.run(flags, recompute=True, **kw)
                        row[f"{kind}_ms"] = round(b.last_ms(), 2)
                    rows.append(row)
                    log(f"[latency] {row}")
        eng.set_tuning(*autotune.DEFAULT)
        out["rows"] = rows
    elif name == "c5":
        from contrafold_gpu_amd import serving
        out.update(serving.benchmark(eng, P, n=args.serve_n, log=log))
    print(json.dumps(out), flush=True)


def published_workloads(cf, workloads, P, log, reps=3):
    """The two workloads the reference publishes GPU figures for, run through
    the drop-in API end to end (host strings in, results on the host, like the
    reference's timed calls): logZ of 10,000 x 200 nt uniform ACGU, seed 0
    (Benchmark.md:14-20, 34: 9,830 folds/s on an RTX 5090) and 100 x 200 nt
    x 100 Boltzmann samples (Benchmark.md:11, 68: 5,627 samples/s).  Other
    hardware, so a comparison, never vs_baseline."""
    seqs = workloads.uniform(10000, 200, seed=0)
    cf.logZ_batch(seqs, P)  # warm: library, parameters, code objects
    t = time.perf_counter()
    for _ in range(reps):
        cf.logZ_batch(seqs, P)
    logz_fps = len(seqs) * reps / (time.perf_counter() - t)
    log(f"[bench] logZ_batch 10000 x 200 nt uniform: {logz_fps:.0f} folds/s end to end")
    sseqs = seqs[:100]
    cf.sample_batch(sseqs, P, 100, seed=1)
    t = time.perf_counter()
    for k in range(reps):
        cf.sample_batch(sseqs, P, 100, seed=2 + k)
    sps = len(sseqs) * 100 * reps / (time.perf_counter() - t)
    log(f"[bench] sample_batch 100 x 200 nt x 100: {sps:.0f} samples/s end to end")
    return {
        "logz_10000x200nt_uniform": {"folds_per_s": round(logz_fps, 1), "reference_rtx5090": 9830,
                                      "ratio": round(logz_fps / 9830, 2)},
        "sample_100x200nt_x100": {"samples_per_s": round(sps, 1), "reference_rtx5090": 5627,
                                  "ratio": round(sps / 5627, 2)},
        "note": "drop-in API calls timed end to end (encode, upload, fold, download); reference figures are "
                "the RTX 5090 numbers its Benchmark.md publishes",
    }


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--config", default="c2", choices=["c1", "c2", "c3", "c4", "c5", "latency"],
                    help="c2 (default) is the bench line; c1/c3/c4/c5 print their own measurements")
    ap.add_argument("--check", type=int, default=64, help="sequences the oracle re-folds for c1/c4")
    ap.add_argument("--serve-n", type=int, default=12000)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--n", type=int, default=10000)
    ap.add_argument("--length", type=int, default=200)
    ap.add_argument("--samples", type=int, default=100)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--serve", action="store_true")
    ap.add_argument("--probe", action="store_true",
                    help="for rocprofv3 counter passes: a copy-like calibration kernel of known size "
                         "(torch add: 256 MiB read, 256 MiB written), then exactly one step of the workload")
    args = ap.parse_args()
    if args.config == "c5" or args.serve:
        # the server runs one stream per slot; HIP maps streams onto this many
        # hardware queues (default 4), and streams that share a queue run in
        # order, so short folds would wait behind long ones.  Must be set
        # before the HIP runtime starts (32 is the largest value accepted here).
        try:
            cur = int(os.environ.get("GPU_MAX_HW_QUEUES", "0"))
        except ValueError:
            cur = 0
        if cur < 32:
            os.environ["GPU_MAX_HW_QUEUES"] = "32"
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one GPU: every fold is independent and this build runs on a single MI355X; "
              "run with --gpus 1", file=sys.stderr)
        sys.exit(2)
    if args.config != "c2":
        run_config(args.config, args)
        return

    calib = None
    if args.probe:
        # the calibration kernel runs before the engine starts: under
        # rocprofv3 --pmc, torch.ones crashed when torch started its GPU
        # context after the engine had
        import torch
        x = torch.ones(64 << 20, device="cuda")
        y = x + 1.0
        torch.cuda.synchronize()
        calib = float(y[0])
    import contrafold_gpu_amd as cf
    from contrafold_gpu_amd import workcount, workloads
    P = cf.load()
    log(f"[bench] generating {args.n} x {args.length} nt (GC ~ Beta(8,8), mask p=0.1, seed 2)")
    seqs, masks = workloads.gc_masked(args.n, args.length, seed=2)
    eng = cf.default_engine()
    eng.set_params(P)
    batch = eng.batch(seqs, masks)
    flags = cf.CF_BPP | cf.CF_SAMPLE
    if args.probe:
        batch.run(flags, n_samples=args.samples, seed=1, recompute=True)
        print(json.dumps({"probe": f"{args.n} x {args.length} nt, one step", "calibration": calib}), flush=True)
        return
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
             "outside": 9.0 * wc["outside"], "post": 9.0 * wc["posterior"], "mea": 2.0 * wc["mea"]}
    # The roofline kernel is the longest stage of the batch stream.  The sampler
    # runs on a lane stream beside the outside sweep and the posterior (it needs
    # only the inside tables), so its marks span their time too and do not
    # measure it alone.
    dominant = max((k for k in kavg if k in flops), key=kavg.get)
    roof = None
    if dominant in flops and kavg[dominant] > 0:
        ach = flops[dominant] / (kavg[dominant] * 1e-3) / 1e12
        roof = {"kernel": dominant, "bound": "mfma", "achieved": round(ach, 3), "peak": FP32_PEAK_TFLOPS,
                "unit": "TFLOP/s", "frac": round(ach / FP32_PEAK_TFLOPS, 5), "traffic": None,
                "ceiling": "FP32 VALU (no MFMA: log-space DP); peak = MI355X FP32 157.3 TFLOP/s",
                # the issue ceiling without packed FP32 (256 CUs x 64 lanes x 2 x 2.4 GHz)
                "frac_of_unpacked_fp32": round(ach / FP32_UNPACKED_TFLOPS, 5)}
    else:
        # sampling dominates: report it against HBM with the bytes it must read
        roof = {"kernel": dominant, "bound": "hbm", "achieved": None, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": None, "traffic": None}
    # HBM bytes per launch of the dominant kernel, measured with rocprofv3
    # FETCH_SIZE / WRITE_SIZE on this workload and calibrated on a copy-like
    # kernel of known size (profiles/c2_10000x200nt_traffic.json)
    tpath = os.path.join(ROOT, "profiles", "c2_10000x200nt_traffic.json")
    if os.path.exists(tpath) and args.n == 10000:
        with open(tpath) as fh:
            tdoc = json.load(fh)
        for name, rec in tdoc.get("kernels", {}).items():
            if name.startswith(f"cf_{dominant}_kernel"):
                roof["traffic"] = rec["bytes_per_launch"]
                roof["traffic_unit"] = "bytes/launch"
                # the same launch seen from HBM: measured bytes over its duration
                if kavg.get(dominant, 0) > 0:
                    gbs = rec["bytes_per_launch"] / (kavg[dominant] * 1e-3) / 1e9
                    roof["traffic_gbs"] = round(gbs, 1)
                    roof["traffic_frac_of_hbm_peak"] = round(gbs / HBM_PEAK_GBS, 4)
    # the stages this step runs (the other event pairs only time an empty gap)
    per_kernel = {k: {"ms": round(v, 3),
                      "tflops": (round(flops[k] / (v * 1e-3) / 1e12, 3) if k in flops and v > 0 else None)}
                  for k, v in kavg.items() if k in ("inside", "exterior", "outside", "post", "sample")}

    # MEA headline on the same resident batch (inside/outside recomputed)
    batch.run(cf.CF_MEA, gammas=[6.0], recompute=True)
    t = time.perf_counter()
    mea_steps = max(1, args.steps)
    for k in range(mea_steps):
        batch.run(cf.CF_MEA, gammas=[6.0], recompute=True)
    mea_fps = args.n * mea_steps / (time.perf_counter() - t)
    log(f"[bench] MEA folds/s: {mea_fps:.0f}")
    batch.close()
    published = published_workloads(cf, workloads, P, log)

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
        "cells_per_s": round(args.n * (args.length + 1) * (args.length + 2) / 2 / (wall / args.steps), 1),
        "kernels": per_kernel,
        "kernels_note": "device ms per stage (HIP events); the sampler runs on a lane stream beside the "
                        "outside sweep and the posterior, so those three overlap",
        "mea": {"folds_per_s": round(mea_fps, 2), "gamma": 6.0},
        "published_workloads": published,
    }
    if not args.no_cpu:
        log("[bench] CPU baseline (oracle, all host cores) ...")
        cb = cpu_baseline(seqs, masks, P, args.samples)
        out["cpu_baseline"] = cb
        out["mea"]["cpu_folds_per_s"] = cb["mea_folds_per_s"]
        out["mea"]["vs_cpu_all_cores"] = round(mea_fps / cb["mea_folds_per_s"], 2)
    if args.serve:
        from contrafold_gpu_amd import serving
        out["serving"] = serving.benchmark(eng, P, n=args.serve_n, log=log)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()

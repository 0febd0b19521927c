"""GPU feature tests: sampling statistics and constraints, gradients against
FP64 finite differences, graph replay, launch-shape invariance, serving."""
import math
import os
import time

import numpy as np
import pytest

from conftest import rand_seq

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def pairs_of(db):
    st, out = [], []
    for k, ch in enumerate(db):
        if ch == "(":
            st.append(k)
        elif ch == ")":
            out.append((st.pop(), k))
    assert not st
    return out


CANON = {("A", "U"), ("U", "A"), ("C", "G"), ("G", "C"), ("G", "U"), ("U", "G")}


def test_sample_frequencies_match_bpp(P):
    """Reference test 3 (tests/test_validation.py:44-64) against our own bpp:
    20,000 samples of GCGCGCAAAAGCGCGCAAAAGCGC, seed 1."""
    import contrafold_gpu_amd as cf
    s = "GCGCGCAAAAGCGCGCAAAAGCGC"
    L = len(s)
    n = 20000
    samples = cf.sample_batch([s], P, n, seed=1)[0]
    assert len(samples) == n
    freq = np.zeros((L, L))
    for db in samples:
        for a, b in pairs_of(db):
            freq[a, b] += 1
    freq /= n
    post = cf.bpp(s, P)
    assert np.abs(freq - post).max() < 0.02
    sigma = np.sqrt(post * (1 - post) / n)
    assert (np.abs(freq - post) <= 5 * sigma + 2e-3).all()


def test_sample_masks_and_validity(P):
    """Reference test 4 (tests/test_validation.py:66-76): forced bases never
    pair; every sample is a nested structure of canonical pairs."""
    import contrafold_gpu_amd as cf
    rng = np.random.default_rng(0)
    seqs, masks = [], []
    for t in range(12):
        L = int(rng.integers(10, 90))
        seqs.append(rand_seq(rng, L))
        masks.append((rng.random(L) < 0.3).astype(np.int8))
    res = cf.sample_batch(seqs, P, 300, forced_list=masks, seed=2)
    for s, m, smp in zip(seqs, masks, res):
        for db in smp:
            assert len(db) == len(s)
            for a, b in pairs_of(db):
                assert not m[a] and not m[b]
                assert (s[a], s[b]) in CANON


def test_sample_long_sequences(P):
    """Sequences past the sampler's LDS-resident form (a batch whose longest
    sequence does not fit the per-wave LDS slice reads F5 rows and codes from
    HBM): a 2,338-nt repeat, a masked 700-nt sequence and a short one.  Samples
    are nested canonical structures that respect the mask, and the masked
    sequence's pair frequencies follow bpp within binomial noise.  (Stacks
    deeper than the 128 LDS-resident entries continue in HBM; sampled
    structures of even 1,700 nt need only 10-15 entries, so no test reaches
    that spill.)"""
    import contrafold_gpu_amd as cf
    rng = np.random.default_rng(11)
    rep = "GGGAAACCCAAAAA" * 167
    long_masked = rand_seq(rng, 700)
    mask = (rng.random(700) < 0.1).astype(np.int8)
    seqs = [rep, long_masked, rand_seq(rng, 60)]
    masks = [np.zeros(len(rep), np.int8), mask, np.zeros(60, np.int8)]
    n = 400
    res = cf.sample_batch(seqs, P, n, forced_list=masks, seed=5)
    for s, m, smp in zip(seqs, masks, res):
        assert len(smp) == n
        for db in smp:
            assert len(db) == len(s)
            for a, b in pairs_of(db):
                assert not m[a] and not m[b]
                assert (s[a], s[b]) in CANON
    L = len(long_masked)
    freq = np.zeros((L, L))
    for db in res[1]:
        for a, b in pairs_of(db):
            freq[a, b] += 1
    freq /= n
    post = cf.bpp(long_masked, P, forced=mask)
    sigma = np.sqrt(post * (1 - post) / n)
    assert (np.abs(freq - post) <= 5 * sigma + 0.02).all()


def test_sampling_batch_invariant(P):
    import contrafold_gpu_amd as cf
    rng = np.random.default_rng(4)
    s = rand_seq(rng, 80)
    others = [rand_seq(rng, int(rng.integers(20, 150))) for _ in range(7)]
    alone = cf.sample_batch([s], P, 50, seed=7)[0]
    mixed = cf.sample_batch(others[:3] + [s] + others[3:], P, 50, seed=7)[3]
    assert alone == mixed
    assert cf.sample_batch([s], P, 50, seed=8)[0] != alone


def test_results_independent_of_input_order(P):
    """Equal-length folds are launched heaviest first (most possible base
    pairs of the free bases), so a permutation of the input changes which
    wave folds which sequence but no result: logZ and bpp bit for bit, samples
    equal (1,200 x 80 nt, masked: one packed bucket with fold queues)."""
    from contrafold_gpu_amd import CF_BPP, CF_SAMPLE, Engine, workloads
    seqs, masks = workloads.gc_masked(1200, 80, seed=11)
    perm = np.random.default_rng(5).permutation(len(seqs))
    eng = Engine()
    eng.set_params(P)
    out = []
    for order in (np.arange(len(seqs)), perm):
        with eng.batch([seqs[k] for k in order], masks=[masks[k] for k in order]) as b:
            b.run(CF_BPP | CF_SAMPLE, n_samples=3, seed=9)
            z = np.asarray(b.logz())
            inv = np.empty_like(order)
            inv[order] = np.arange(len(order))
            smp = b.samples()
            out.append((z[inv], [b.bpp(int(inv[k])).copy() for k in (0, 1, 599, 1199)],
                        [smp[int(inv[k])] for k in range(0, 1200, 37)]))
    (z0, p0, s0), (z1, p1, s1) = out
    assert np.array_equal(z0.view(np.uint32), z1.view(np.uint32))
    for a, c in zip(p0, p1):
        assert np.array_equal(np.asarray(a).view(np.uint32), np.asarray(c).view(np.uint32))
    assert s0 == s1


def test_fold_tasks_gpu_order(P):
    import contrafold_gpu_amd as cf
    rng = np.random.default_rng(1)
    seqs = [rand_seq(rng, int(rng.integers(5, 60))) for _ in range(30)]
    out = cf.fold_tasks_gpu(seqs, None, P, chunk=7, seed=3)
    assert [len(x) for x in out] == [len(s) for s in seqs]


def test_gradient_vs_fp64_finite_differences(P, oracle):
    """d mean logZ / d raw weight from the outside-pass adjoint against central
    differences (step 1e-4) of the oracle's FP64 inside pass."""
    import contrafold_gpu_amd as cf
    from contrafold_gpu_amd import params
    names, raw, pmap = params.load_raw()
    rng = np.random.default_rng(4)
    seqs = [rand_seq(rng, 40) for _ in range(6)]
    w = np.full(len(seqs), 1.0 / len(seqs), np.float32)
    _z, tg = cf.logZ_grad(seqs, P, weights=w)
    g = pmap.raw_grad(tg)
    # canonical-reachable params have nonzero gradient somewhere; pick both kinds
    idx = list(rng.choice(len(raw), 24, replace=False))
    idx += [names.index(k) for k in ("multi_base", "multi_unpaired", "external_unpaired", "base_pair_CG",
                                      "hairpin_length_at_least_3", "helix_stacking_CGCG")]
    h = 1e-4

    def mean_logz(r):
        Q = params.unflatten(pmap.tables(r))
        return np.mean([oracle.inside64(s, Q) for s in seqs])
    for k in idx:
        e = np.zeros_like(raw)
        e[k] = h
        fd = (mean_logz(raw + e) - mean_logz(raw - e)) / (2 * h)
        assert abs(g[k] - fd) <= 2e-3 * max(1.0, abs(fd)), (names[k], g[k], fd)


def test_gradient_closed_form_unpairable(P):
    """No pair possible: logZ = L * eu, so d logZ / d eu = L and nothing else moves."""
    import contrafold_gpu_amd as cf
    from contrafold_gpu_amd import params
    _z, tg = cf.logZ_grad(["AAAAAAAAAAAA", "NNNNNNN"], P)
    assert abs(tg[params.OFFSETS["eu"]] - 19.0) < 1e-4
    tg[params.OFFSETS["eu"]] = 0
    assert np.abs(tg).max() < 1e-6


def test_torch_autograd():
    """torch.autograd.Function over the outside-pass adjoint.  Runs in a child
    process that imports torch first, so torch and the engine share one HIP
    runtime (this process started the engine without torch)."""
    pytest.importorskip("torch")
    import subprocess
    import sys
    code = r"""
import sys, numpy as np, torch
sys.path.insert(0, %r)
import contrafold_gpu_amd as cf
from contrafold_gpu_amd import autograd, params, _lib
from conftest import rand_seq
assert len(_lib.hip_runtimes()) <= 1
P = cf.load()
rng = np.random.default_rng(9)
seqs = [rand_seq(rng, 50) for _ in range(4)]
theta = autograd.raw_theta()
z = autograd.logz(theta, seqs)
z.mean().backward()
_z, tg = cf.logZ_grad(seqs, P, weights=np.full(4, 0.25, np.float32))
want = params.param_map().raw_grad(tg)
assert np.allclose(theta.grad.numpy(), want, rtol=1e-5, atol=1e-7)
assert np.allclose(z.detach().numpy(), cf.logZ_batch(seqs, P))
assert len(_lib.hip_runtimes()) == 1
print("autograd ok")
""" % (ROOT,)
    r = subprocess.run([sys.executable, "-c", code], cwd=os.path.join(ROOT, "tests"), capture_output=True,
                       text=True, timeout=110)
    assert r.returncode == 0 and "autograd ok" in r.stdout, (r.stdout[-2000:], r.stderr[-2000:])


def test_graph_replay_matches_eager(engine, P):
    from contrafold_gpu_amd import CF_BPP
    from contrafold_gpu_amd._lib import CF_GRAPH
    rng = np.random.default_rng(12)
    seqs = [rand_seq(rng, int(rng.integers(30, 130))) for _ in range(40)]
    with engine.batch(seqs) as b:
        b.run(CF_BPP, recompute=True)
        z0 = b.logz().copy()
        p0 = b.bpp(3).copy()
        for _ in range(3):
            b.run(CF_BPP | CF_GRAPH)
            assert (b.logz().view(np.uint32) == z0.view(np.uint32)).all()
            assert (b.bpp(3).view(np.uint32) == p0.view(np.uint32)).all()
            b.run(CF_BPP, recompute=True)


def test_bucket_streams_match_serial(engine, P):
    """A batch spanning three length buckets runs each bucket's stages on its own
    stream (eager and under graph capture); the results equal the one-stream
    timed run bit for bit."""
    from contrafold_gpu_amd import CF_MEA, CF_SAMPLE
    from contrafold_gpu_amd._lib import CF_GRAPH
    rng = np.random.default_rng(21)
    seqs = [rand_seq(rng, int(rng.integers(20, 400))) for _ in range(60)]
    with engine.batch(seqs) as b:
        b.run(CF_MEA | CF_SAMPLE, gammas=[2.0, 6.0], n_samples=3, seed=5, recompute=True, timing=True)
        ref = (b.logz().copy(), b.mea(0), b.mea(1), b.samples(), [b.bpp(k).copy() for k in (0, 31, 59)])
        for flags in (0, CF_GRAPH, CF_GRAPH):
            b.run(CF_MEA | CF_SAMPLE | flags, gammas=[2.0, 6.0], n_samples=3, seed=5, recompute=True)
            assert (b.logz().view(np.uint32) == ref[0].view(np.uint32)).all()
            assert b.mea(0) == ref[1] and b.mea(1) == ref[2] and b.samples() == ref[3]
            for k, m in zip((0, 31, 59), ref[4]):
                assert (b.bpp(k).view(np.uint32) == m.view(np.uint32)).all()


def test_bucket_parts_match_serial(P):
    """A large bucket of uneven lengths (2,100 sequences of 20-60 nt, one
    bucket) is cut into parts that run on separate streams (eager and under
    graph capture); the results equal the one-stream timed run bit for bit."""
    from contrafold_gpu_amd import CF_MEA, CF_SAMPLE, Engine
    from contrafold_gpu_amd._lib import CF_GRAPH
    rng = np.random.default_rng(23)
    seqs = [rand_seq(rng, int(rng.integers(20, 61))) for _ in range(2100)]
    eng = Engine()
    eng.set_params(P)
    eng.set_tuning([1 << 30], [64])
    picks = (0, 700, 1400, 2099)
    with eng.batch(seqs) as b:
        b.run(CF_MEA | CF_SAMPLE, gammas=[6.0], n_samples=2, seed=9, recompute=True, timing=True)
        ref = (b.logz().copy(), b.mea(0), b.samples(), [b.bpp(k).copy() for k in picks])
        for flags in (0, CF_GRAPH, CF_GRAPH):
            b.run(CF_MEA | CF_SAMPLE | flags, gammas=[6.0], n_samples=2, seed=9, recompute=True)
            assert (b.logz().view(np.uint32) == ref[0].view(np.uint32)).all()
            assert b.mea(0) == ref[1] and b.samples() == ref[2]
            for k, m in zip(picks, ref[3]):
                assert (b.bpp(k).view(np.uint32) == m.view(np.uint32)).all()


def test_sampler_beside_outside_matches_serial(P):
    """A run with logZ, bpp and samples launches the sampler on a lane stream
    beside the outside sweep and the posterior (eager, timed and under graph
    capture); its samples equal those of a logZ + samples run, where it runs
    alone on the batch stream, and its bpp equals a bpp-only run."""
    from contrafold_gpu_amd import CF_BPP, CF_LOGZ, CF_SAMPLE, Engine
    from contrafold_gpu_amd._lib import CF_GRAPH
    rng = np.random.default_rng(29)
    seqs = [rand_seq(rng, 90) for _ in range(300)]
    eng = Engine()
    eng.set_params(P)
    with eng.batch(seqs) as b:
        b.run(CF_LOGZ | CF_SAMPLE, n_samples=4, seed=3, recompute=True)
        ref_s = b.samples()
        b.run(CF_BPP, recompute=True)
        ref_b = [b.bpp(k).copy() for k in (0, 150, 299)]
        for kw in ({}, {"timing": True}, {"graph": True}, {"graph": True}):
            flags = CF_BPP | CF_SAMPLE | (CF_GRAPH if kw.get("graph") else 0)
            b.run(flags, n_samples=4, seed=3, recompute=True, timing=kw.get("timing", False))
            assert b.samples() == ref_s
            for k, m in zip((0, 150, 299), ref_b):
                assert (b.bpp(k).view(np.uint32) == m.view(np.uint32)).all()
        b.run(CF_BPP | CF_SAMPLE, n_samples=4, seed=3, recompute=True, timing=True)
        assert b.kernel_ms()["sample"] > 0


def test_launch_shape_invariance(P):
    """Results are bit-identical across bucket edges and workgroup sizes (a
    cross-lane race would show up as a mismatch)."""
    from contrafold_gpu_amd import CF_BPP, Engine
    rng = np.random.default_rng(13)
    seqs = [rand_seq(rng, int(rng.integers(10, 260))) for _ in range(30)]
    ref = None
    for edges, threads in (([1 << 30], [64]), ([1 << 30], [256]), ([100, 1 << 30], [128, 192]),
                           ([64, 128, 256, 1 << 30], [64, 128, 256, 256])):
        eng = Engine()
        eng.set_params(P)
        eng.set_tuning(edges, threads)
        with eng.batch(seqs) as b:
            b.run(CF_BPP)
            got = (b.logz().copy(), [b.bpp(k) for k in range(0, 30, 7)])
        eng.close()
        if ref is None:
            ref = got
        else:
            assert (got[0].view(np.uint32) == ref[0].view(np.uint32)).all()
            for x, y in zip(got[1], ref[1]):
                assert (x.view(np.uint32) == y.view(np.uint32)).all()


def test_repeat_determinism(engine, P):
    from contrafold_gpu_amd import CF_MEA
    rng = np.random.default_rng(14)
    seqs = [rand_seq(rng, 200) for _ in range(64)]
    with engine.batch(seqs) as b:
        b.run(CF_MEA, gammas=[6.0], recompute=True)
        a = (b.logz().copy(), b.mea(0), b.bpp(5).copy())
        b.run(CF_MEA, gammas=[6.0], recompute=True)
        assert (b.logz().view(np.uint32) == a[0].view(np.uint32)).all()
        assert b.mea(0) == a[1]
        assert (b.bpp(5).view(np.uint32) == a[2].view(np.uint32)).all()


def test_server_mix_equals_alone(engine, P):
    """Each served result equals the same request served alone, and MEA/bpp/
    logZ equal the plain batch API bit for bit."""
    import contrafold_gpu_amd as cf
    from contrafold_gpu_amd.serving import Server
    rng = np.random.default_rng(15)
    reqs = []
    for t in range(40):
        L = int(rng.integers(5, 220))
        s = rand_seq(rng, L)
        m = (rng.random(L) < 0.15).astype(np.uint8) if t % 3 == 0 else None
        kind = ("mea", "mea", "sample", "bpp", "logz")[t % 5]
        reqs.append((s, m, kind))
    srv = Server(engine, n_streams=3, max_batch=16)
    ids = [srv.submit(s, kind=k, mask=m, n_samples=10, seed=11) for s, m, k in reqs]
    mixed = srv.run_all()
    srv.close()
    for rid, (s, m, kind) in zip(ids, reqs):
        r = mixed[rid]
        z = cf.logZ_batch([s], P, forced_list=[m])[0]
        assert np.float32(r["logz"]) == np.float32(z)
        if kind == "mea":
            assert r["structure"] == cf.mea(s, P, 6.0, m)
        elif kind == "bpp":
            assert (r["bpp"] == cf.bpp(s, P, m)).all()
    # alone
    for rid, (s, m, kind) in list(zip(ids, reqs))[::4]:
        srv = Server(engine, n_streams=1)
        rid2 = srv.submit(s, kind=kind, mask=m, n_samples=10, seed=11)
        r2 = srv.run_all()[rid2]
        srv.close()
        for key in ("structure", "samples", "logz"):
            if key in r2:
                assert r2[key] == mixed[rid][key]


def test_server_admission_budget(engine, P):
    """A request larger than the HBM budget fails cleanly; the rest complete."""
    from contrafold_gpu_amd.serving import Server
    from contrafold_gpu_amd import fold_bytes, CF_MEA
    small = fold_bytes(60, CF_MEA)
    srv = Server(engine, hbm_budget=3 * small, n_streams=2, max_batch=8)
    rng = np.random.default_rng(16)
    ids = [srv.submit(rand_seq(rng, 60)) for _ in range(7)]
    big = srv.submit(rand_seq(rng, 400))
    done = {}
    while len(done) < 8:
        for rid in srv.poll():
            try:
                done[rid] = srv.result(rid)
            except MemoryError:
                done[rid] = None
    srv.close()
    assert done[big] is None
    assert all(done[i] is not None for i in ids)


def test_server_stats_and_long_folds(engine, P):
    """Long folds (> 320 nt) hold at most half the slots; short requests
    submitted behind them still complete, and the counters add up."""
    from contrafold_gpu_amd.serving import Server
    srv = Server(engine, n_streams=4, max_batch=4)
    rng = np.random.default_rng(21)
    longs = [srv.submit(rand_seq(rng, 600)) for _ in range(12)]
    shorts = [srv.submit(rand_seq(rng, 40)) for _ in range(12)]
    done = {}
    t0 = time.time()
    while len(done) < 24 and time.time() - t0 < 60:
        for rid in srv.poll():
            done[rid] = srv.result(rid)
    st = srv.stats()
    srv.close()
    assert len(done) == 24
    assert all(len(done[i]["structure"]) == 600 for i in longs)
    assert all(len(done[i]["structure"]) == 40 for i in shorts)
    assert st["folds"] == 24 and st["batches"] >= 6  # max_batch 4: at least 3 long + 3 short batches
    assert st["max_queue"] >= 1 and st["host_launch_ms"] >= 0 and st["batch_residence_ms"] > 0
    # the short requests did not wait for every long batch: some finished first
    assert min(done[i]["latency_ms"] for i in shorts) < max(done[i]["latency_ms"] for i in longs)


def test_autotune_installs_fastest(engine, P, tmp_path, monkeypatch):
    from contrafold_gpu_amd import autotune, CF_BPP
    monkeypatch.setenv("CF_TUNE_CACHE", str(tmp_path / "tune.json"))
    rng = np.random.default_rng(17)
    seqs = [rand_seq(rng, int(rng.integers(30, 300))) for _ in range(60)]
    cands = autotune.CANDIDATES[:3]
    packs = [(8, 4), (2, 16)]
    with engine.batch(seqs) as b:
        b.run(CF_BPP)
        ref = [b.bpp(k) for k in range(len(seqs))]
    res = autotune.tune(engine, seqs, flags=CF_BPP, sample=60, repeats=1, candidates=cands, packings=packs)
    assert not res["cached"] and len(res["tried"]) == 4  # every candidate has a packed bucket
    # the fastest candidate at the default packing, unless the other packing is
    # faster by the tuner's margin
    base = min(t["ms"] for t in res["tried"] if t["packing"] == list(packs[0]))
    if res["packing"] == list(packs[0]):
        assert res["ms"] == base
    else:  # it passed the bar on its first timing, then was confirmed
        assert any(t["packing"] == res["packing"] and t["ms"] < base * (1.0 - autotune.PACKING_MARGIN)
                   for t in res["tried"])
    again = autotune.tune(engine, seqs, flags=CF_BPP, sample=60, repeats=1, candidates=cands, packings=packs)
    assert again["cached"] and again["edges"] == res["edges"] and again["packing"] == res["packing"]
    # results do not depend on the tuning or the packing
    for pk in packs:
        engine.set_packing(*pk)
        with engine.batch(seqs) as b:
            b.run(CF_BPP)
            assert all(np.array_equal(b.bpp(k), ref[k]) for k in range(len(seqs)))
    engine.set_tuning(*autotune.DEFAULT)
    engine.set_packing(*autotune.PACKINGS[0])


def test_cli_modes(P, tmp_path, capsys):
    import json
    from contrafold_gpu_amd import cli
    import contrafold_gpu_amd as cf
    cli.main(["GGGGAAAACCCC"])
    assert capsys.readouterr().out.strip() == "((((....))))"
    cli.main(["GGGGAAAACCCC", "--viterbi"])
    assert capsys.readouterr().out.strip() == "((((....))))"
    p = tmp_path / "in.jsonl"
    p.write_text(json.dumps({"id": "a", "seq": "GCGCGCAAAAGCGCGC"}) + "\n" +
                 json.dumps({"id": "b", "seq": "GGGAAACCC", "constrain": "000010000"}) + "\n")
    o = tmp_path / "out.jsonl"
    cli.main([str(p), "-o", str(o)])
    recs = [json.loads(x) for x in o.read_text().splitlines()]
    assert recs[0]["structure"] == cf.mea("GCGCGCAAAAGCGCGC", P)
    assert recs[1]["structure"][4] == "."
    cli.main([str(p), "--logz", "-o", str(o)])
    recs = [json.loads(x) for x in o.read_text().splitlines()]
    assert abs(recs[0]["logZ"] - cf.logZ_batch(["GCGCGCAAAAGCGCGC"], P)[0]) < 1e-6
    cli.main([str(p), "--sample", "3", "-o", str(o)])
    recs = [json.loads(x) for x in o.read_text().splitlines()]
    assert len(recs[1]["samples"]) == 3


def test_uneven_lengths_buckets_and_n_bases(engine, P, oracle):
    """Uneven-length batch with 2% N (workload C3 shape, small): every bucket's
    MEA at three gammas and Viterbi equal the oracle."""
    from contrafold_gpu_amd import workloads, CF_MEA, CF_MFE
    seqs = workloads.uneven(24, lo=30, hi=300, seed=3)
    gammas = [0.5, 6.0, 16.0]
    with engine.batch(seqs) as b:
        b.run(CF_MEA | CF_MFE, gammas=gammas)
        got = [b.mea(g) for g in range(3)]
        vit, _ = b.mfe()
    for k, s in enumerate(seqs):
        _z, POST = oracle.posterior32(s, P)
        for g_i, g in enumerate(gammas):
            pair, _ = oracle.mea_decode32(POST, s, g)
            assert got[g_i][k] == oracle.dot_bracket(pair, len(s))
        assert vit[k] == oracle.mfe(s, P)


def test_cli_contrafold_formats(P, tmp_path, capsys):
    from contrafold_gpu_amd import cli
    import contrafold_gpu_amd as cf
    s = "GCGCGCAAAAGCGCGCAAAAGCGC"
    cli.main([s, "--posteriors", "1e-4"])
    text = capsys.readouterr().out
    M = cli.parse_posteriors(text, len(s))
    ref = cf.bpp(s, P)
    keep = ref >= 1e-4
    assert np.allclose(M[keep], ref[keep], rtol=1e-5) and not M[~keep].any()
    cli.main([s, "--partition"])
    line = capsys.readouterr().out.strip()
    assert line.startswith('Log partition coefficient for "0":')
    assert abs(float(line.split()[-1]) - cf.logZ_batch([s], P)[0]) < 1e-5
    cli.main([s, "--parens", "-o", str(tmp_path / "p.txt")])
    assert (tmp_path / "p.txt").read_text().splitlines() == [">0", s, cf.mea(s, P)]

"""GPU tests at the configurations' real shapes, exact-tie pins, and
regression tests of fixed host-side bugs.

* C2 (10,000 x 200 nt, GC ~ Beta(8, 8), masks Bernoulli(0.1)): sequences of
  the seeded 200-nt generator, logZ and every posterior bit-equal to the
  oracle, MEA dot-brackets equal.
* C3 (U[30, 500] nt with 2% N): MEA at gamma 0.5 / 6 / 16 and Viterbi under
  hipGraph replay, equal to the oracle.
* C4 (512 x 150 nt): the gradient at 150 nt on 8 sequences against FP64
  central differences of the oracle.
* C5: served requests above 300 nt equal the plain batch API.
"""
import threading

import numpy as np
import pytest

from conftest import rand_seq
from test_gpu_parity import domains, same_bits
from test_oracle_pins import TIE_CASES, tie_params

pytestmark = pytest.mark.gpu


def bits(x):
    return np.asarray(x, np.float32).view(np.uint32)


# ------------------------------------------------------------ configurations
def test_c2_generator_200nt_bitexact(engine, P, oracle):
    from contrafold_gpu_amd import CF_MEA, workloads
    seqs, masks = workloads.gc_masked(10000, 200, seed=2)
    pick = list(range(0, 10000, 625))       # 16 sequences spread over the batch
    sub = [seqs[k] for k in pick]
    msk = [masks[k] for k in pick]
    with engine.batch(sub, msk) as b:
        b.run(CF_MEA, gammas=[6.0])
        z = b.logz()
        mea = b.mea(0)
        for t, (s, m) in enumerate(zip(sub, msk)):
            z32, POST = oracle.posterior32(s, P, m)
            assert bits(z[t]) == bits(z32), (t, z[t], z32)
            assert (bits(b.bpp(t)) == bits(POST[1:201, 1:201])).all(), t
            pair, _ = oracle.mea_decode32(POST, s, 6.0, m)
            assert mea[t] == oracle.dot_bracket(pair, 200), t


def test_c3_graph_mea_viterbi(P, oracle):
    """Sequences of the seeded U[30, 500] workload with 2% N, MEA at three
    gammas and Viterbi, captured as a hipGraph and replayed twice."""
    from contrafold_gpu_amd import CF_MEA, CF_MFE, Engine, workloads
    from contrafold_gpu_amd._lib import CF_GRAPH
    seqs = workloads.uneven(20000, seed=3)
    L = np.array([len(s) for s in seqs])
    # the longest, the shortest and a spread of the rest
    pick = sorted(set([int(L.argmax()), int(L.argmin())] + list(range(7, 20000, 700))))
    sub = [seqs[k] for k in pick]
    gammas = [0.5, 6.0, 16.0]
    eng = Engine()
    eng.set_params(P)
    with eng.batch(sub) as b:
        for rep in range(3):
            b.run(CF_MEA | CF_MFE | CF_GRAPH, gammas=gammas)
            got = [b.mea(g) for g in range(3)]
            vit, sc = b.mfe()
            if rep == 0:
                first = (got, vit, sc.copy())
            else:
                assert (got, vit) == first[:2] and (sc == first[2]).all()
    zz, pairs = None, None
    for g_i, g in enumerate(gammas):
        _v, pairs = oracle.run_batch("mea32", sub, P, gamma=g, threads=8)
        for k, s in enumerate(sub):
            assert got[g_i][k] == oracle.dot_bracket(pairs[k], len(s)), (k, len(s), g)
    vals, vp = oracle.run_batch("viterbi64", sub, P, threads=8)
    for k, s in enumerate(sub):
        assert vit[k] == oracle.dot_bracket(vp[k], len(s)), (k, len(s))
        assert abs(sc[k] - vals[k]) <= 1e-9 * max(1.0, abs(vals[k]))
    eng.close()


@pytest.mark.parametrize("packing", [(1, 1), (3, 5), (16, 16)])
def test_packing_and_fold_queue_bitexact(P, oracle, packing):
    """The packed sweeps' workgroups take folds from a queue, several one-wave
    teams per workgroup: tables, samples and MEA do not depend on how many
    teams share a workgroup (1, odd counts, 16) or on which wave takes which
    fold, and match the oracle."""
    from contrafold_gpu_amd import CF_BPP, CF_LOGZ, CF_MEA, CF_SAMPLE, Engine
    rng = np.random.default_rng(41)
    seqs = [rand_seq(rng, int(n)) for n in rng.integers(1, 384, 700)]
    masks = [(rng.random(len(q)) < 0.1).astype(np.int8) for q in seqs]
    flags = CF_LOGZ | CF_BPP | CF_SAMPLE | CF_MEA
    eng = Engine()
    eng.set_params(P)
    out = []
    for pk in [(8, 4), packing]:
        eng.set_packing(*pk)
        with eng.batch(seqs, masks) as b:
            b.run(flags, gammas=[6.0], n_samples=4, seed=5)
            out.append((b.logz(), [b.bpp(k) for k in range(0, 700, 7)], b.samples(), b.mea(0)))
    eng.close()
    (z0, p0, s0, m0), (z1, p1, s1, m1) = out
    assert (bits(z0) == bits(z1)).all()
    assert all((bits(x) == bits(y)).all() for x, y in zip(p0, p1))
    assert s0 == s1 and m0 == m1
    for k in range(0, 700, 70):
        z32, POST = oracle.posterior32(seqs[k], P, masks[k])
        L = len(seqs[k])
        assert bits(z1[k]) == bits(z32), k
        assert (bits(p1[k // 7]) == bits(POST[1:L + 1, 1:L + 1])).all(), k


def test_c4_gradient_150nt(P, oracle):
    """d mean logZ / d raw weight over 8 sequences of the 150-nt workload
    against FP64 central differences (step 1e-4) of the oracle."""
    import contrafold_gpu_amd as cf
    from contrafold_gpu_amd import params, workloads
    names, raw, pmap = params.load_raw()
    seqs = workloads.uniform(512, 150, seed=4)[:8]
    w = np.full(len(seqs), 1.0 / len(seqs), np.float32)
    _z, tg = cf.logZ_grad(seqs, P, weights=w)
    g = pmap.raw_grad(tg)
    rng = np.random.default_rng(4)
    idx = list(rng.choice(len(raw), 32, replace=False))
    idx += [names.index(k) for k in ("multi_base", "multi_unpaired", "multi_paired", "external_paired",
                                      "internal_length_at_least_2", "bulge_length_at_least_1")]
    h = 1e-4
    nonzero = 0
    for k in idx:
        e = np.zeros_like(raw)
        e[k] = h
        fp = oracle.run_batch("logz64", seqs, params.unflatten(pmap.tables(raw + e)), threads=8)[0].mean()
        fm = oracle.run_batch("logz64", seqs, params.unflatten(pmap.tables(raw - e)), threads=8)[0].mean()
        fd = (fp - fm) / (2 * h)
        nonzero += abs(fd) > 1e-6
        assert abs(g[k] - fd) <= 2e-3 * max(1.0, abs(fd)), (names[k], g[k], fd)
    assert nonzero >= 10


def test_c5_long_requests_served(engine, P):
    """Requests above 300 nt (up to 1,000) of every kind, served in a mix with
    short ones, equal the plain batch API."""
    import contrafold_gpu_amd as cf
    from contrafold_gpu_amd.serving import Server
    rng = np.random.default_rng(31)
    reqs = []
    for t, L in enumerate([320, 450, 610, 777, 1000, 350, 40, 90, 150, 510]):
        s = rand_seq(rng, L)
        m = (rng.random(L) < 0.15).astype(np.uint8) if t % 3 == 0 else None
        reqs.append((s, m, ("mea", "bpp", "sample", "mea", "mea")[t % 5]))
    srv = Server(engine, n_streams=4, max_batch=8)
    ids = [srv.submit(s, kind=k, mask=m, n_samples=10, seed=3) for s, m, k in reqs]
    got = srv.run_all()
    srv.close()
    for rid, (s, m, kind) in zip(ids, reqs):
        r = got[rid]
        assert bits(r["logz"]) == bits(cf.logZ_batch([s], P, forced_list=[m])[0])
        if kind == "mea":
            assert r["structure"] == cf.mea(s, P, 6.0, m)
        elif kind == "bpp":
            assert (bits(r["bpp"]) == bits(cf.bpp(s, P, m).astype(np.float32))).all()
        else:
            assert len(r["samples"]) == 10 and all(len(x) == len(s) for x in r["samples"])


# ------------------------------------------------------------ exact ties
@pytest.mark.parametrize("case", range(len(TIE_CASES)))
def test_viterbi_exact_ties_gpu(P, oracle, case):
    """Scores built so that several structures tie exactly: the GPU returns
    the structure the reference's candidate order reaches first."""
    from contrafold_gpu_amd import CF_MFE, Engine
    over, seq, want = TIE_CASES[case]
    Q = tie_params(P, **over)
    eng = Engine()
    eng.set_params(Q)
    with eng.batch([seq]) as b:
        b.run(CF_MFE)
        db, _sc = b.mfe()
    eng.close()
    assert db[0] == want == oracle.mfe(seq, Q)


def test_zero_params_mea_and_posteriors(P, oracle):
    """All weights zero: every structure scores 0, many posteriors are equal by
    symmetry and MEA faces near-ties everywhere; the GPU matches the oracle."""
    from contrafold_gpu_amd import CF_MEA, Engine
    Q = tie_params(P)
    rng = np.random.default_rng(41)
    seqs = [rand_seq(rng, L) for L in (8, 20, 37, 64, 101)] + ["GCGCGCGCGCGC", "GGGGAAAACCCC"]
    eng = Engine()
    eng.set_params(Q)
    with eng.batch(seqs) as b:
        b.run(CF_MEA, gammas=[0.5, 1.0, 6.0])
        meas = [b.mea(g) for g in range(3)]
        for k, s in enumerate(seqs):
            z32, POST = oracle.posterior32(s, Q)
            L = len(s)
            assert bits(b.logz()[k]) == bits(z32)
            assert (bits(b.bpp(k)) == bits(POST[1:L + 1, 1:L + 1])).all()
            for g_i, g in enumerate((0.5, 1.0, 6.0)):
                pair, _ = oracle.mea_decode32(POST, s, g)
                assert meas[g_i][k] == oracle.dot_bracket(pair, L), (s, g)
    eng.close()


# ------------------------------------------------------------ regressions
def test_graph_after_buffer_growth(P):
    """A captured graph holds the buffer addresses of its capture.  An eager
    run with more gammas / samples reallocates those buffers; a later graph
    run must re-capture, and its results equal the eager ones bit for bit."""
    from contrafold_gpu_amd import CF_MEA, CF_SAMPLE, Engine
    from contrafold_gpu_amd._lib import CF_GRAPH
    rng = np.random.default_rng(43)
    seqs = [rand_seq(rng, int(rng.integers(30, 160))) for _ in range(50)]
    eng = Engine()
    eng.set_params(P)
    with eng.batch(seqs) as b:
        b.run(CF_MEA, gammas=[6.0])
        ref1 = b.mea(0)
        b.run(CF_MEA | CF_GRAPH, gammas=[6.0])
        assert b.mea(0) == ref1
        b.run(CF_MEA, gammas=[0.5, 2.0, 6.0])
        ref3 = [b.mea(g) for g in range(3)]
        assert ref3[2] == ref1
        b.run(CF_MEA | CF_GRAPH, gammas=[6.0])
        assert b.mea(0) == ref1
        b.run(CF_MEA | CF_GRAPH, gammas=[0.5, 2.0, 6.0])
        assert [b.mea(g) for g in range(3)] == ref3
        b.run(CF_SAMPLE, n_samples=2, seed=4)
        s2 = b.samples()
        b.run(CF_SAMPLE | CF_GRAPH, n_samples=2, seed=4)
        assert b.samples() == s2
        b.run(CF_SAMPLE, n_samples=9, seed=4)
        s9 = b.samples()
        b.run(CF_SAMPLE | CF_GRAPH, n_samples=2, seed=4)
        assert b.samples() == s2
        b.run(CF_SAMPLE | CF_GRAPH, n_samples=9, seed=4)
        assert b.samples() == s9
    eng.close()


def test_results_kept_when_stage_not_rerun(engine, P):
    """A run without sampling (or MEA) keeps the batch's earlier samples and
    structures readable, with their counts."""
    from contrafold_gpu_amd import CF_LOGZ, CF_MEA, CF_SAMPLE
    seqs = ["GGGGAAAACCCC", "GCGCAAGCGCAAGCGC"]
    with engine.batch(seqs) as b:
        b.run(CF_MEA | CF_SAMPLE, gammas=[6.0, 2.0], n_samples=3, seed=1)
        smp, mea = b.samples(), [b.mea(0), b.mea(1)]
        b.run(CF_LOGZ, recompute=True)
        assert b.samples() == smp and [b.mea(0), b.mea(1)] == mea
        assert all("\0" not in x for row in b.samples() for x in row)


def test_fold_tasks_duplicates_across_chunks(P):
    """Copies of one task spread over several chunks draw distinct samples
    (keys derived over the whole task list), and the result does not depend
    on the chunk size."""
    import contrafold_gpu_amd as cf
    rng = np.random.default_rng(45)
    s = rand_seq(rng, 150)
    out = cf.fold_tasks_gpu([s] * 48, None, P, chunk=16, seed=7)
    assert out[0:16] != out[16:32] and out[16:32] != out[32:48]
    assert len(set(out)) > 24
    assert cf.fold_tasks_gpu([s] * 48, None, P, chunk=5, seed=7) == out


def test_concurrent_threads(P):
    """Drop-in calls from several threads share the default engine; each
    returns what it returns alone."""
    import contrafold_gpu_amd as cf
    rng = np.random.default_rng(47)
    work = [[rand_seq(rng, int(rng.integers(20, 200))) for _ in range(40)] for _ in range(6)]
    want = [(cf.logZ_batch(w, P), cf.mea_batch(w, P, 6.0)) for w in work]
    got = [None] * len(work)
    errs = []

    def run(k):
        try:
            for _ in range(3):
                got[k] = (cf.logZ_batch(work[k], P), cf.mea_batch(work[k], P, 6.0))
        except Exception as e:  # pragma: no cover - reported below
            errs.append(e)
    th = [threading.Thread(target=run, args=(k,)) for k in range(len(work))]
    for t in th:
        t.start()
    for t in th:
        t.join()
    assert not errs, errs
    assert got == want


def test_gradient_bitwise_deterministic(P):
    """Per-sequence fixed-point counts and an ordered sum: the gradient is the
    same bit for bit across runs, team widths and bucket layouts."""
    from contrafold_gpu_amd import CF_GRAD, Engine
    rng = np.random.default_rng(49)
    seqs = [rand_seq(rng, int(rng.integers(20, 260))) for _ in range(40)]
    w = rng.random(40).astype(np.float32)
    ref = None
    for edges, threads in (([1 << 30], [64]), ([1 << 30], [256]), ([100, 1 << 30], [128, 512])):
        eng = Engine()
        eng.set_params(P)
        eng.set_tuning(edges, threads)
        with eng.batch(seqs) as b:
            for _ in range(2):
                b.run(CF_GRAD, weights=w, recompute=True)
                g = b.grad()
                if ref is None:
                    ref = g.copy()
                assert g.tobytes() == ref.tobytes()
        eng.close()


# ------------------------------------------------------------ split teams
SPLIT_LENS = (37, 150, 333, 420, 701)


@pytest.fixture(scope="module")
def split_ref(P):
    """Unsplit results (one workgroup per fold, 1,024 threads) of the split
    test set: tables, posteriors, MEA and samples."""
    from contrafold_gpu_amd import CF_MEA, CF_SAMPLE, Engine
    rng = np.random.default_rng(61)
    seqs = [rand_seq(rng, L) for L in SPLIT_LENS]
    masks = [None, (rng.random(150) < 0.1).astype(np.uint8), None, None, (rng.random(701) < 0.05).astype(np.uint8)]
    eng = Engine()
    eng.set_params(P)
    eng.set_tuning([1 << 30], [1024])
    with eng.batch(seqs, masks) as b:
        b.run(CF_MEA | CF_SAMPLE, gammas=[6.0], n_samples=4, seed=2)
        ref = {"logz": b.logz().copy(), "mea": b.mea(0), "samples": b.samples(),
               "tables": {(k, t): b.table(k, t) for k in range(len(seqs))
                          for t in ("FC", "FM", "FM1", "FM2", "FCo", "FMo", "FM1o", "POST")}}
    eng.close()
    return seqs, masks, ref


@pytest.mark.parametrize("K", [2, 4, 8])
def test_split_team_bitexact(P, split_ref, K):
    """Folds run by K workgroups (one per CU) that meet at a barrier per
    anti-diagonal: every inside / outside table, the posteriors, MEA and the
    samples equal the one-workgroup team bit for bit."""
    from contrafold_gpu_amd import CF_MEA, CF_SAMPLE, Engine
    seqs, masks, ref = split_ref
    eng = Engine()
    eng.set_params(P)
    eng.set_tuning([1 << 30], [1024])
    eng.set_split(100, K)
    with eng.batch(seqs, masks) as b:
        for rep in range(2):
            b.run(CF_MEA | CF_SAMPLE, gammas=[6.0], n_samples=4, seed=2, recompute=True)
            assert (bits(b.logz()) == bits(ref["logz"])).all()
            assert b.mea(0) == ref["mea"] and b.samples() == ref["samples"]
        for (k, t), want in ref["tables"].items():
            got = b.table(k, t)
            same = (bits(got) == bits(want)) | (np.isnan(got) & np.isnan(want))
            assert same.all(), (SPLIT_LENS[k], t)
    eng.close()


def test_split_team_oracle_420_1100(P, oracle):
    """Split teams against the oracle at 420 and 1,100 nt (every FP32 table
    and posterior bit-exact), under hipGraph replay too."""
    from contrafold_gpu_amd import CF_BPP, Engine
    from contrafold_gpu_amd._lib import CF_GRAPH
    rng = np.random.default_rng(63)
    seqs = [rand_seq(rng, 420), rand_seq(rng, 1100)]
    eng = Engine()
    eng.set_params(P)
    eng.set_tuning([1 << 30], [512])
    eng.set_split(200, 4)
    with eng.batch(seqs) as b:
        b.run(CF_BPP | CF_GRAPH)
        b.run(CF_BPP | CF_GRAPH)
        for k, s in enumerate(seqs):
            L = len(s)
            z32, POST, T = oracle.posterior32(s, P, tables=True)
            assert bits(b.logz()[k]) == bits(z32)
            assert (bits(b.bpp(k)) == bits(POST[1:L + 1, 1:L + 1])).all()
            fc, fm, _tri = domains(s, None)
            for name, cells in (("FC", fc), ("FM", fm), ("FM1", fm), ("FCo", fc), ("FMo", fm), ("FM1o", fm)):
                ii, jj = np.array(cells).T
                ok = same_bits(b.table(k, name)[ii, jj], T[name][ii, jj])
                assert ok.all(), (L, name, np.array(cells)[~ok][:5])
    eng.close()


def test_sampler_local_and_global_forms_agree(P):
    """The sampler draws into the wave's LDS slice when the batch's longest
    fold lets four workgroups share a CU, else straight into global memory;
    a sequence's samples are the same in both forms (and valid)."""
    import contrafold_gpu_amd as cf
    rng = np.random.default_rng(73)
    short = [rand_seq(rng, 120), rand_seq(rng, 200), rand_seq(rng, 37)]
    masks = [None, (rng.random(200) < 0.1).astype(np.uint8), None]
    a = cf.sample_batch(short, P, 20, forced_list=masks, seed=9)
    b = cf.sample_batch(short + [rand_seq(rng, 1800)], P, 20, forced_list=masks + [None], seed=9)
    assert b[:3] == a
    for s, m, smp in zip(short, masks, a):
        for db in smp:
            st = []
            for t, ch in enumerate(db):
                if ch == "(":
                    st.append(t)
                elif ch == ")":
                    x = st.pop()
                    assert (s[x], s[t]) in {("A", "U"), ("U", "A"), ("C", "G"), ("G", "C"), ("G", "U"), ("U", "G")}
                    assert m is None or (m[x] == 0 and m[t] == 0)
            assert not st

"""GPU parity: the gfx950 kernels against the oracle, bit for bit.

Every FP32 DP table (inside FC/FM/FM1/F5, outside FCo/FMo/FM1o/F5o, the
posterior POST) must equal the oracle's restatement of the reference's float32
engine (gpu_contrafold/cpu.py:583-820) to the bit on every cell of its domain;
non-finite cells must be non-finite in both.  MEA (every gamma) and Viterbi
dot-brackets must be identical strings.  FP64 logZ must match the oracle's
exact inside pass to 1e-9.
"""
import numpy as np
import pytest

from conftest import rand_seq

pytestmark = pytest.mark.gpu

HALF = np.float32(-5e29)
CANON = {(0, 3), (3, 0), (1, 2), (2, 1), (2, 3), (3, 2)}


def cases(seed=11):
    rng = np.random.default_rng(seed)
    out = []
    for t in range(36):
        L = int(rng.integers(1, 72))
        g = rng.beta(8, 8)
        seq = rand_seq(rng, L, p=[(1 - g) / 2, g / 2, g / 2, (1 - g) / 2])
        if t % 5 == 4:
            seq = "".join(c if rng.random() > 0.05 else "N" for c in seq)
        mask = (rng.random(L) < 0.1).astype(np.int8) if t % 3 == 0 else None
        out.append((seq, mask))
    out += [("", None), ("A", None), ("AU", None), ("GGGGAAAACCCC", None), ("N" * 9, None),
            ("gcgcaaagcgcuuuugcgcaaagcgc", None), ("GCGCGCAAAAGCGCGCAAAAGCGC", np.ones(24, np.int8)),
            ("GGGAAACCCAGGGAAACCCUGGGAAACCC", None), (rand_seq(np.random.default_rng(5), 160), None)]
    # Windows whose rows are mostly empty: the window walks skip runs of empty
    # rows when they refill their buffered rows.  A-rich with sparse partners;
    # half the bases forced unpaired; long unpairable runs between helices.
    r2 = np.random.default_rng(7)
    sparse = "".join("A" if r2.random() < 0.85 else "UG"[int(r2.integers(0, 2))] for _ in range(140))
    out += [(sparse, None),
            (rand_seq(r2, 120), (r2.random(120) < 0.5).astype(np.int8)),
            ("GGGCGC" + "A" * 40 + "GCGCCC" + "C" * 30 + "GGAUCC" + "A" * 25 + "GGAUCC", None)]
    return out


CASES = cases()


def encode(seq):
    m = {"A": 0, "C": 1, "G": 2, "U": 3, "T": 3}
    return [4] + [m.get(c, 4) for c in seq.upper()] + [4]


def domains(seq, mask):
    s = encode(seq)
    L = len(seq)
    fo = [0] + [int(x != 0) for x in (mask if mask is not None else np.zeros(L))] + [0]

    def allow(a, b):
        return (s[a], s[b]) in CANON and fo[a] == 0 and fo[b] == 0
    fc = [(i, j) for i in range(1, L + 1) for j in range(i, L) if allow(i, j + 1)]
    fm = [(i, j) for i in range(1, L + 1) for j in range(i + 2, L)]
    tri = [(i, j) for i in range(L + 1) for j in range(i, L + 1)]
    return fc, fm, tri


def same_bits(g, c):
    g = np.asarray(g, np.float32)
    c = np.asarray(c, np.float32)
    both_inf = (g <= HALF) & (c <= HALF)
    return both_inf | (g.view(np.uint32) == c.view(np.uint32))


@pytest.fixture(scope="module", params=[64, 1024], ids=["packed", "wide1024"])
def folded(P, oracle, request):
    """The cases folded with one wave per sequence (packed workgroups) and
    with a 1024-thread workgroup per sequence (the latency-mode sweeps)."""
    from contrafold_gpu_amd import CF_BPP, Engine
    eng = Engine()
    eng.set_params(P)
    eng.set_tuning([1 << 30], [request.param])
    seqs = [c[0] for c in CASES]
    masks = [c[1] for c in CASES]
    b = eng.batch(seqs, masks)
    b.run(CF_BPP)
    refs = [oracle.posterior32(s, P, m, tables=True) for s, m in CASES]
    yield b, refs
    b.close()
    eng.close()


@pytest.mark.parametrize("k", range(len(CASES)))
def test_inside_tables_bitexact(folded, k):
    b, refs = folded
    seq, mask = CASES[k]
    L = len(seq)
    z, _POST, T = refs[k]
    fc, fm, tri = domains(seq, mask)
    for name, cells in (("FC", fc), ("FM", fm), ("FM1", fm)):
        g = b.table(k, name)
        if cells:
            ii, jj = np.array(cells).T
            ok = same_bits(g[ii, jj], T[name][ii, jj])
            assert ok.all(), f"{name} differs at {np.array(cells)[~ok][:5]}"
    # cells outside the FC domain stay non-finite
    g = b.table(k, "FC")
    ii, jj = np.array(tri).T
    inside = set(fc)
    off = np.array([(i, j) not in inside for i, j in tri])
    assert (g[ii[off], jj[off]] <= HALF).all()
    f5 = b.table(k, "F5")
    assert same_bits(f5[:L + 1], T["F5"][:L + 1]).all()
    assert b.logz()[k].view(np.uint32) == np.float32(z).view(np.uint32)


@pytest.mark.parametrize("k", range(len(CASES)))
def test_outside_tables_bitexact(folded, k):
    b, refs = folded
    seq, mask = CASES[k]
    L = len(seq)
    _z, _POST, T = refs[k]
    fc, fm, _ = domains(seq, mask)
    for name, cells in (("FCo", fc), ("FMo", fm), ("FM1o", fm)):
        if not cells:
            continue
        g = b.table(k, name)
        ii, jj = np.array(cells).T
        ok = same_bits(g[ii, jj], T[name][ii, jj])
        assert ok.all(), f"{name} differs at {np.array(cells)[~ok][:5]}"
    f5o = b.table(k, "F5o")
    assert same_bits(f5o[:L + 1], T["F5o"][:L + 1]).all()


@pytest.mark.parametrize("k", range(len(CASES)))
def test_posterior_bitexact(folded, k):
    b, refs = folded
    seq, _mask = CASES[k]
    L = len(seq)
    _z, POST, _T = refs[k]
    got = b.bpp(k)
    want = POST[1:L + 1, 1:L + 1]
    assert got.shape == (L, L)
    assert (got.view(np.uint32) == want.view(np.uint32)).all()


def test_mea_all_gammas_bitexact(engine, P, oracle):
    from contrafold_gpu_amd import CF_MEA
    gammas = [0.5, 1.0, 2.0, 4.0, 6.0, 8.0, 16.0]
    seqs = [c[0] for c in CASES]
    masks = [c[1] for c in CASES]
    with engine.batch(seqs, masks) as b:
        b.run(CF_MEA, gammas=gammas)
        for g_i, g in enumerate(gammas):
            got = b.mea(g_i)
            for k, (s, m) in enumerate(CASES):
                assert got[k] == oracle.mea(s, P, g, m), (k, g)


def test_viterbi_bitexact(engine, P, oracle):
    from contrafold_gpu_amd import CF_MFE
    seqs = [c[0] for c in CASES]
    masks = [c[1] for c in CASES]
    with engine.batch(seqs, masks) as b:
        b.run(CF_MFE)
        got, scores = b.mfe()
    for k, (s, m) in enumerate(CASES):
        pair, sc = oracle.viterbi64(s, P, m)
        assert got[k] == oracle.dot_bracket(pair, len(s)), k
        assert scores[k] == sc, k


def test_logz64_matches_oracle(engine, P, oracle):
    from contrafold_gpu_amd import CF_LOGZ64
    seqs = [c[0] for c in CASES]
    masks = [c[1] for c in CASES]
    with engine.batch(seqs, masks) as b:
        b.run(CF_LOGZ64)
        z = b.logz64()
    for k, (s, m) in enumerate(CASES):
        assert abs(z[k] - oracle.inside64(s, P, m)) < 1e-9, k


def test_dropin_api_known_answers(P):
    import contrafold_gpu_amd as cf
    assert cf.mfe("GGGGAAAACCCC", P) == "((((....))))"   # README.md:101
    assert cf.mea("GGGGAAAACCCC", P, gamma=6) == "((((....))))"  # README.md:104
    z = cf.logZ_batch(["GGGGAAAACCCC", "GCGCGCAAGCGCGC"], P)
    assert len(z) == 2 and all(np.isfinite(z))
    assert abs(cf.cpu_logZ("GGGGAAAACCCC", P) - z[0]) < 5e-4
    pr = cf.bpp("GGGGAAAACCCC", P)
    assert pr.shape == (12, 12) and pr.dtype == np.float64
    assert cf.logZ_batch([""], P) == [0.0]
    assert cf.mea("", P) == ""
    with pytest.raises(ValueError):
        cf.logZ_batch([], P)
    with pytest.raises(ValueError):
        cf.logZ_batch(["GGGAAACCC"], P, forced_list=[[0, 1]])


def _wide_engine(P):
    from contrafold_gpu_amd import Engine
    eng = Engine()
    eng.set_params(P)
    eng.set_tuning([1 << 30], [256])
    return eng


def test_long_sequence_parity(P, oracle):
    """One 420-nt sequence on a 256-thread workgroup (the outside sweep keeps
    its FCo window in an LDS ring)."""
    from contrafold_gpu_amd import CF_MEA
    rng = np.random.default_rng(99)
    s = rand_seq(rng, 420)
    engine = _wide_engine(P)
    with engine.batch([s]) as b:
        b.run(CF_MEA, gammas=[6.0])
        z32, POST = oracle.posterior32(s, P)
        assert b.logz()[0].view(np.uint32) == np.float32(z32).view(np.uint32)
        got = b.bpp(0)
        assert (got.view(np.uint32) == POST[1:421, 1:421].view(np.uint32)).all()
        assert b.mea(0)[0] == oracle.mea(s, P, 6.0)
    engine.close()


@pytest.mark.parametrize("width", [256, 1024, 64])
def test_beyond_lds_ring_parity(P, oracle, width):
    """L = 1100 on a 256-thread workgroup (the FCo ring no longer fits in LDS,
    so the window is read from the global tables) and packed (one wave)."""
    from contrafold_gpu_amd import CF_MEA
    rng = np.random.default_rng(1100)
    s = rand_seq(rng, 1100)
    engine = _wide_engine(P)
    engine.set_tuning([1 << 30], [width])
    with engine.batch([s]) as b:
        b.run(CF_MEA, gammas=[6.0])
        z32, POST = oracle.posterior32(s, P)
        assert b.logz()[0].view(np.uint32) == np.float32(z32).view(np.uint32)
        got = b.bpp(0)
        assert (got.view(np.uint32) == POST[1:1101, 1:1101].view(np.uint32)).all()
        pair, _ = oracle.mea_decode32(POST, s, 6.0)
        assert b.mea(0)[0] == oracle.dot_bracket(pair, 1100)
    engine.close()


def test_packed_layout_parity(P, oracle):
    """One wave per sequence (packed workgroups; the single-branch windows are
    read from the global tables): logZ, every posterior and the MEA structure
    bit-exact against the oracle, lengths 300 down to 1, with masks."""
    from contrafold_gpu_amd import CF_MEA, Engine
    rng = np.random.default_rng(64)
    seqs, masks = [], []
    for L in (300, 257, 130, 65, 64, 17, 5, 1):
        seqs.append(rand_seq(rng, L))
        masks.append((rng.random(L) < 0.1).astype(np.int8) if L % 2 else None)
    eng = Engine()
    eng.set_params(P)
    eng.set_tuning([1 << 30], [64])
    try:
        with eng.batch(seqs, masks) as b:
            b.run(CF_MEA, gammas=[6.0])
            z = b.logz()
            got = b.mea(0)
            for k, (s, m) in enumerate(zip(seqs, masks)):
                z32, POST = oracle.posterior32(s, P, m)
                L = len(s)
                assert z[k].view(np.uint32) == np.float32(z32).view(np.uint32), k
                assert (b.bpp(k).view(np.uint32) == POST[1:L + 1, 1:L + 1].view(np.uint32)).all(), k
                pair, _ = oracle.mea_decode32(POST, s, 6.0, m)
                assert got[k] == oracle.dot_bracket(pair, L), k
    finally:
        eng.close()


@pytest.mark.parametrize("width", [64, 256], ids=["packed", "wide256"])
def test_mea_tile_boundaries(P, oracle, width):
    """The tiled MEA decode (16 x 16 tiles by block diagonal, far terms as tile
    products, near terms in an in-tile wavefront, traceback by replay) at
    lengths on both sides of every tile edge up to four tile rows, and a few
    longer ones, with masks and N bases, three gammas in one run (the items
    of one sequence share its posteriors): every structure equals the
    oracle's decode of the oracle's posteriors."""
    from contrafold_gpu_amd import CF_MEA, Engine
    rng = np.random.default_rng(1616)
    lengths = [L for L in range(1, 66)] + [79, 80, 81, 127, 128, 129, 255, 256, 257, 511, 512, 513]
    seqs, masks = [], []
    for k, L in enumerate(lengths):
        s = rand_seq(rng, L)
        if k % 7 == 3:
            s = "".join(c if rng.random() > 0.05 else "N" for c in s)
        seqs.append(s)
        masks.append((rng.random(L) < 0.1).astype(np.int8) if k % 3 == 1 else None)
    gammas = [0.5, 6.0, 16.0]
    eng = Engine()
    eng.set_params(P)
    eng.set_tuning([1 << 30], [width])
    try:
        with eng.batch(seqs, masks) as b:
            b.run(CF_MEA, gammas=gammas)
            got = [b.mea(g) for g in range(len(gammas))]
            for k, (s, m) in enumerate(zip(seqs, masks)):
                _, POST = oracle.posterior32(s, P, m)
                for g_i, g in enumerate(gammas):
                    pair, _ = oracle.mea_decode32(POST, s, g, m)
                    assert got[g_i][k] == oracle.dot_bracket(pair, len(s)), (width, len(s), g)
    finally:
        eng.close()


def test_viterbi_lane_groups_and_stream(P, oracle):
    """Viterbi with lane groups on short diagonals (groups of 2-8 lanes share a
    cell's split terms and window rows; FP64 maximum, order-free) and on its
    own stream beside the FP32 stages: structures and FP64 scores equal the
    oracle whether Viterbi runs alone, beside an MEA run, or in a replayed
    hipGraph of that run."""
    from contrafold_gpu_amd import CF_MEA, CF_MFE, Engine
    from contrafold_gpu_amd._lib import CF_GRAPH
    rng = np.random.default_rng(808)
    seqs, masks = [], []
    for k, L in enumerate(list(range(1, 41)) + [64, 97, 150, 200, 333]):
        seqs.append(rand_seq(rng, L))
        masks.append((rng.random(L) < 0.1).astype(np.int8) if k % 4 == 2 else None)
    want = [oracle.viterbi64(s, P, m) for s, m in zip(seqs, masks)]
    eng = Engine()
    eng.set_params(P)
    try:
        with eng.batch(seqs, masks) as b:
            runs = [(CF_MFE, {}), (CF_MEA | CF_MFE, dict(gammas=[6.0])),
                    (CF_MEA | CF_MFE | CF_GRAPH, dict(gammas=[6.0])), (CF_MEA | CF_MFE | CF_GRAPH, dict(gammas=[6.0]))]
            for flags, kw in runs:
                b.run(flags, recompute=not (flags & CF_GRAPH), **kw)
                got, scores = b.mfe()
                for k, (s, (pair, sc)) in enumerate(zip(seqs, want)):
                    assert got[k] == oracle.dot_bracket(pair, len(s)), (flags, len(s))
                    assert scores[k] == sc, (flags, len(s))
    finally:
        eng.close()

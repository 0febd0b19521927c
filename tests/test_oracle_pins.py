"""Further pins of the oracle: its CPU sampler, the single-branch loop table
``cs`` and the reference's tie-breaking order.

* ``oracle.sample32`` (restating gpu_contrafold/gpu.py:196-477) is checked the
  way the reference checks its GPU sampler (tests/test_validation.py:44-76):
  pair frequencies over many samples agree with the posterior pair
  probabilities of ``oracle.posterior32`` within binomial noise, no sample
  pairs a forced-unpaired base, and every sample is a nested structure of
  canonical pairs.
* ``cs[l1][l2]`` (cpu.py:114-130) is recomputed from the parameter file by the
  meaning of its feature names -- ``<family>_at_least_k`` fires for a loop
  whose length (or asymmetry) is at least k -- instead of the running sums
  the loaders use, both for the bundled weights and for random weights.
* Exact ties: with scores built so that several structures score exactly the
  same, Viterbi (cpu.py:396-502, first candidate within EPS in the order
  unpaired, k ascending / hairpin, (p asc, q desc), multi) and MEA
  (cpu.py:840-871, strict ``>``) must return the structure that the
  reference's candidate order reaches first.
"""
import numpy as np
import pytest

import brute
from conftest import rand_seq

CANON = {("A", "U"), ("U", "A"), ("C", "G"), ("G", "C"), ("G", "U"), ("U", "G")}


def _pairs(db):
    st, out = [], []
    for t, ch in enumerate(db):
        if ch == "(":
            st.append(t)
        elif ch == ")":
            assert st, db
            out.append((st.pop(), t))
        else:
            assert ch == ".", db
    assert not st, db
    return out


# ---------------------------------------------------------------- sampler
SAMPLER_CASES = [
    ("GGGGAAAACCCC", None),
    ("GCGCAAGCGCAAGCGCAAAGGCGCUUGCGCAAUU", None),
    ("ACGUAGCUAGCUAGGCUAUCGAUCGUAGCGCGAUAUUAGCGCUA", None),
    ("GGGAAACCCAGGGAAACCCUAGCGCUUUAGCGCA", [0, 0, 1, 0, 0, 0, 0, 0, 0, 0, 0, 0, 1, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0,
                                          0, 0, 0, 0, 0, 0, 1, 0, 0, 0]),
]


@pytest.mark.parametrize("case", range(len(SAMPLER_CASES)))
def test_sample32_frequencies_match_posterior(oracle, P, case):
    seq, mask = SAMPLER_CASES[case]
    L = len(seq)
    n = 4000
    _z, POST = oracle.posterior32(seq, P, mask)
    post = POST[1:L + 1, 1:L + 1].astype(np.float64)
    freq = np.zeros((L, L))
    for db in oracle.sample32(seq, P, n, seed=11 + case, forced=mask):
        assert len(db) == L
        for a, b in _pairs(db):
            assert (seq[a], seq[b]) in CANON, (db, a, b)
            if mask is not None:
                assert mask[a] == 0 and mask[b] == 0, (db, a, b)
            freq[a, b] += 1
    freq /= n
    # 5 sigma of binomial noise, plus 2e-3 for the posterior's own
    # approximation: Fast_Exp (cpu.py:560-580) returns 0 below -9.91, so a
    # pair whose contributions are each under ~5e-5 has POST = 0, while the
    # sampler weighs candidates with expf.
    sigma = np.sqrt(np.maximum(post * (1 - post), 1.0 / n) / n)
    excess = np.abs(freq - post) - (5.0 * sigma + 2e-3)
    assert excess.max() < 0, (excess.max(), np.unravel_index(excess.argmax(), excess.shape))
    # a mistake that moves probability mass shows as a biased total as well
    assert abs(freq.sum() - post.sum()) < 0.05 * max(1.0, post.sum())


def test_sample32_reference_setup(oracle, P):
    """The reference's check 3 as it stands (tests/test_validation.py:45-65):
    20,000 samples of GCGCGCAAAAGCGCGCAAAAGCGC, pair frequencies within 0.02
    of the posteriors."""
    s = "GCGCGCAAAAGCGCGCAAAAGCGC"
    L = len(s)
    n = 20000
    freq = np.zeros((L, L))
    for db in oracle.sample32(s, P, n, seed=1):
        for a, b in _pairs(db):
            freq[a, b] += 1
    freq /= n
    assert np.abs(freq - oracle.bpp(s, P)).max() < 0.02


def test_sample32_enumeration_small(oracle, P):
    """On a short sequence every structure has its exact Boltzmann weight
    (tests/brute.py); sampled structure frequencies follow it."""
    seq = "GGGAAACCCU"
    B = brute.enumerate_all(seq, P)
    L = len(seq)
    want = {}
    for st, sc, mu in zip(B["structs"], B["scores"], B["mults"]):
        db = ["."] * L
        for a, b in st:
            db[a - 1], db[b - 1] = "(", ")"
        want["".join(db)] = want.get("".join(db), 0.0) + mu * np.exp(sc - B["logZ"])
    n = 20000
    got = {}
    for db in oracle.sample32(seq, P, n, seed=3):
        got[db] = got.get(db, 0) + 1
    assert set(got) <= set(want)
    for db, p in want.items():
        f = got.get(db, 0) / n
        assert abs(f - p) <= 5 * np.sqrt(max(p * (1 - p), 1.0 / n) / n) + 2e-3, (db, f, p)


def test_sample32_seeded(oracle, P):
    seq = "GCGCAAGCGCAAGCGCAAAGGCGCUUGCGCAAUU"
    a = oracle.sample32(seq, P, 50, seed=5)
    assert a == oracle.sample32(seq, P, 50, seed=5)
    assert a != oracle.sample32(seq, P, 50, seed=6)


def test_sample32_unpairable(oracle, P):
    for seq, mask in (("AAAAAAA", None), ("GGGGCCCC", [1] * 8)):
        assert oracle.sample32(seq, P, 5, seed=1, forced=mask) == ["." * len(seq)] * 5


# ---------------------------------------------------------------- cs table
def _cs_by_features(rv, l1, l2):
    """Sum of the raw weights whose feature fires for a single-branch loop with
    l1 unpaired bases on the left and l2 on the right (l1 + l2 >= 1)."""
    def at_least(prefix, value):
        tot = 0.0
        for name, w in rv.items():
            if name.startswith(prefix + "_at_least_"):
                if value >= int(name.rsplit("_", 1)[1]):
                    tot += w
        return tot

    n = l1 + l2
    if l1 == 0 or l2 == 0:
        return at_least("bulge_length", n)
    v = 0.0
    lo, hi = min(l1, l2), max(l1, l2)
    v += rv.get(f"internal_explicit_{lo}_{hi}", 0.0)
    v += at_least("internal_length", n)
    if l1 == l2:
        v += at_least("internal_symmetric_length", l1)
    v += at_least("internal_asymmetry", abs(l1 - l2))
    return v


def _check_cs(rv, cs, tol):
    checked = 0
    for l1 in range(31):
        for l2 in range(31 - l1):
            if l1 == 0 and l2 == 0:
                continue
            want = _cs_by_features(rv, l1, l2)
            assert abs(cs[l1, l2] - want) <= tol * max(1.0, abs(want)), (l1, l2, cs[l1, l2], want)
            checked += 1
    assert checked == 31 * 32 // 2 - 1


def test_cs_from_feature_names(P, oracle):
    from contrafold_gpu_amd import params
    names, raw = params.read_raw()
    rv = dict(zip(names, raw))
    _check_cs(rv, np.asarray(P["cs"]), 1e-12)
    _check_cs(rv, np.asarray(oracle.load_tables()["cs"]), 1e-12)


def test_cs_from_feature_names_random_weights(oracle):
    """Random weights (no coincidental zeros or cancellations)."""
    from contrafold_gpu_amd import params
    names, raw, pmap = params.load_raw()
    rng = np.random.default_rng(17)
    rnd = rng.normal(size=raw.size)
    rv = dict(zip(names, rnd))
    cs = params.unflatten(pmap.tables(rnd))["cs"]
    _check_cs(rv, np.asarray(cs), 1e-10)
    _check_cs(rv, np.asarray(oracle.tables_from_raw(rv)["cs"]), 1e-10)


# ---------------------------------------------------------------- exact ties
def tie_params(P, **over):
    Z = {k: (np.zeros_like(v) if isinstance(v, np.ndarray) else 0.0) for k, v in P.items()}
    Z.update(over)
    return Z


# (params overrides, sequence, structure the reference's order reaches first)
TIE_CASES = [
    # every structure scores 0: F5 tries "unpaired" first
    ({}, "GGGGAAAACCCC", "............"),
    ({}, "GCAUGCAUGCGUACGAUCG", "..................."),
    # unpaired exterior bases cost 1, everything else 0.  Pair (1,7) encloses a
    # hairpin or a stack on (2,6), both 0: the hairpin comes first.
    ({"eu": -1.0}, "GGAAACC", "(.....)"),
    # F5[4]: k = 0 (pair (1,4)) and k = 2 (pair (3,4) after pair (1,2)) both
    # score 0; k ascending takes k = 0, whose FC then takes the hairpin.
    ({"eu": -1.0}, "GCGC", "(..)"),
]


@pytest.mark.parametrize("case", range(len(TIE_CASES)))
def test_viterbi_exact_ties(oracle, P, case):
    over, seq, want = TIE_CASES[case]
    Q = tie_params(P, **over)
    assert oracle.mfe(seq, Q) == want
    # the answer is one of the optimal structures
    B = brute.enumerate_all(seq, Q) if len(seq) <= 13 else None
    if B is not None:
        pair, sc = oracle.viterbi64(seq, Q)
        assert abs(sc - B["vit"]) < 1e-12
        assert sum(1 for s in B["scores"] if abs(s - B["vit"]) < 1e-12) > 1   # a real tie


def test_mea_exact_tie(oracle):
    """'GC' with P(1,2) = 0.5 and gamma = 1: pairing scores 0.5, leaving both
    bases unpaired scores 2 * (1 - 0.5) / 2 = 0.5 exactly.  The unpaired
    candidates come first and the comparison is strict, so '..' wins."""
    POST = np.zeros((4, 4), np.float32)
    POST[1, 2] = 0.5
    pair, obj = oracle.mea_decode32(POST, "GC", 1.0)
    assert oracle.dot_bracket(pair, 2) == ".." and obj == np.float32(0.5)
    POST[1, 2] = np.nextafter(np.float32(0.5), np.float32(1))
    pair, _ = oracle.mea_decode32(POST, "GC", 1.0)
    assert oracle.dot_bracket(pair, 2) == "()"


def test_mea_exact_tie_split(oracle):
    """'GCGC' with P(1,2) = P(3,4) = P(1,4) = 0.5, gamma = 1: up(1) = up(4)
    = 0 and up(2) = up(3) = 0.25, so '(..)' (0.5 + 0.25 + 0.25) and '()()'
    (0.5 + 0.5) both reach 1.0 exactly.  At (0, 4) the pair candidate (3)
    precedes the splits k = 2, 3 and the comparison is strict, so '(..)'."""
    POST = np.zeros((6, 6), np.float32)
    POST[1, 2] = POST[3, 4] = POST[1, 4] = np.float32(0.5)
    pair, obj = oracle.mea_decode32(POST, "GCGC", 1.0)
    assert obj == np.float32(1.0)
    assert oracle.dot_bracket(pair, 4) == "(..)"


def test_printed_partition_coefficient(oracle, P):
    """gpu_contrafold/reference.py:32 shows the CONTRAfold binary's output line
    'Log partition coefficient for "...": 3.05664' (six significant digits; the
    sequence is not named there, the README's example GGGGAAAACCCC is the
    natural candidate).  The Fast_LogExpPlusOne path that the binary and the
    FP32 engine share agrees to the printed digits (3.0566349 in FP32: within
    the print's half unit, 5e-6, plus FP32 rounding of the sums); the exact
    FP64 inside (3.05661) is within the approximation's error of it."""
    z32, _ = oracle.posterior32("GGGGAAAACCCC", P, None)
    assert abs(float(z32) - 3.05664) <= 1e-5
    assert abs(oracle.inside64("GGGGAAAACCCC", P) - 3.05664) <= 5e-4

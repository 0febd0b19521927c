"""Pins of the oracle (CPU restatement of gpu_contrafold/cpu.py).

The reference cannot run here (no numba, no CUDA, no CONTRAfold binary), and
it ships no golden vectors; its only literal known answers are README.md:101,
104.  The oracle is pinned by independent means: explicit enumeration of all
structures (tests/brute.py), closed forms, invariants and known answers --
each chosen so that a dropped term, a wrong index or a misordered loop fails.
"""
import math

import numpy as np
import pytest

import brute
from conftest import rand_seq


def small_cases(n=40, seed=0, Lmax=12):
    rng = np.random.default_rng(seed)
    out = []
    for t in range(n):
        L = int(rng.integers(1, Lmax + 1))
        seq = rand_seq(rng, L)
        mask = (rng.random(L) < 0.2).astype(int) if t % 3 == 0 else None
        out.append((seq, mask))
    # structured cases with many pairs / multiloops
    out += [("GGGGAAAACCCC", None), ("GCGCAAGCGCAAGCGC", None), ("GGACUUCGGUCC", None),
            ("GUGUGUACACAC", None), ("GGGAAACCCAGGGAAACCCU", None)]
    return out


@pytest.mark.parametrize("case", range(45))
def test_inside64_vs_enumeration(oracle, P, case):
    seq, mask = small_cases()[case]
    B = brute.enumerate_all(seq, P, mask)
    z = oracle.inside64(seq, P, mask)
    assert abs(z - B["logZ"]) < 1e-11


@pytest.mark.parametrize("case", range(45))
def test_posterior32_vs_enumeration(oracle, P, case):
    seq, mask = small_cases()[case]
    L = len(seq)
    B = brute.enumerate_all(seq, P, mask)
    z32, POST = oracle.posterior32(seq, P, mask)
    # FP32 with CONTRAfold's polynomial log/exp approximations
    assert abs(float(z32) - B["logZ"]) < 2e-4
    if L > 1:
        assert np.abs(POST[1:L + 1, 1:L + 1] - B["post"][1:L + 1, 1:L + 1]).max() < 2e-3


@pytest.mark.parametrize("case", range(45))
def test_viterbi64_vs_enumeration(oracle, P, case):
    seq, mask = small_cases()[case]
    L = len(seq)
    B = brute.enumerate_all(seq, P, mask)
    pair, sc = oracle.viterbi64(seq, P, mask)
    assert abs(sc - B["vit"]) < 1e-9
    # the traced structure attains the maximum (within the traceback's EPS)
    assert abs(brute.score(P, B["s"], L, brute.pairs_of(pair, L)) - B["vit"]) < 1e-5


@pytest.mark.parametrize("case", range(45))
@pytest.mark.parametrize("gamma", [0.5, 6.0, 16.0])
def test_mea_vs_enumeration(oracle, P, case, gamma):
    seq, mask = small_cases()[case]
    _z, POST = oracle.posterior32(seq, P, mask)
    pair, obj = oracle.mea_decode32(POST, seq, gamma, mask)
    best = brute.mea_best(seq, POST, gamma, mask)
    assert abs(best - float(obj)) < 1e-5


def test_readme_known_answers(oracle, P):
    # README.md:101, 104
    assert oracle.mfe("GGGGAAAACCCC", P) == "((((....))))"
    assert oracle.mea("GGGGAAAACCCC", P, gamma=6) == "((((....))))"


def test_derivation_multiplicity_is_real(oracle, P):
    """The reference grammar counts a multiloop with m >= 3 branches and t
    trailing unpaired bases C(m-2+t, t) times (tests/brute.py).  With every
    score zero, exp(logZ) counts derivations; enumeration without the
    multiplicity must fall short on a sequence that has such loops."""
    zero = {k: (np.zeros_like(v) if isinstance(v, np.ndarray) else 0.0) for k, v in P.items()}
    seq = "UGAUAGCUAUA"
    B = brute.enumerate_all(seq, zero)
    n_struct = len(B["structs"])
    n_deriv = sum(B["mults"])
    assert n_deriv > n_struct
    assert round(math.exp(oracle.inside64(seq, zero))) == n_deriv


@pytest.mark.parametrize("L", [1, 5, 17, 40])
def test_unpairable_closed_form(oracle, P, L):
    """No pair possible (all A, all N, or fully masked): logZ is the sequential
    sum of L external_unpaired terms, and every posterior is 0."""
    eu64 = P["eu"]
    want64 = 0.0
    for _ in range(L):
        want64 = want64 + eu64
    eu32 = np.float32(eu64)
    want32 = np.float32(0.0)
    for _ in range(L):
        want32 = np.float32(want32 + eu32)
    rng = np.random.default_rng(L)
    for seq, mask in (("A" * L, None), ("N" * L, None), (rand_seq(rng, L), np.ones(L, int))):
        assert oracle.inside64(seq, P, mask) == want64
        z32, POST = oracle.posterior32(seq, P, mask)
        assert z32 == want32
        assert not POST.any()


def test_two_base_closed_form(oracle, P):
    """'AU': the unpaired-unpaired structure and the single pair (1,2) with a
    0-length hairpin."""
    s = {"A": 0, "U": 3}
    a, u = s["A"], s["U"]
    # pair (1,2): exterior branch ep + bp + jA(2,0) (hc + no dangles: j = L, k = 0) + FC(1,1)
    fc = P["hc"][a, u] + P["tm"][a, u, u, a] + P["hp_cum"][0]
    pair = P["ep"] + P["bp"][a, u] + P["hc"][u, a] + fc
    want = np.logaddexp(2 * P["eu"], pair)
    assert abs(oracle.inside64("AU", P) - want) < 1e-12


def test_posterior_invariants(oracle, P):
    rng = np.random.default_rng(7)
    for t in range(6):
        L = int(rng.integers(20, 80))
        seq = rand_seq(rng, L)
        mask = (rng.random(L) < 0.15).astype(int)
        _z, POST = oracle.posterior32(seq, P, mask)
        Pm = POST[1:L + 1, 1:L + 1].astype(np.float64)
        assert (Pm >= 0).all() and (Pm <= 1).all()
        assert np.allclose(np.tril(Pm), 0)
        rows = Pm.sum(0) + Pm.sum(1)
        assert rows.max() < 1.0 + 2e-3
        forced = np.nonzero(mask)[0]
        assert not Pm[forced, :].any() and not Pm[:, forced].any()


def test_lse32_is_contrafold_polynomial(oracle):
    """Fast_LogExpPlusOne: |r(d) - log1p(exp(-d))| small on [0, 11.86), r
    monotone, truncation beyond, NEG guard."""
    d = np.linspace(0, 11.86, 4001, dtype=np.float32)
    got = np.array([float(oracle.lse32(0.0, -x)) for x in d])
    exact = np.log1p(np.exp(-d.astype(np.float64)))
    assert np.abs(got - exact).max() < 2e-5
    assert np.all(np.diff(got) <= 1e-6)
    assert float(oracle.lse32(5.0, 5.0 - 12.0)) == 5.0
    assert float(oracle.lse32(-1e30, 3.0)) == 3.0
    assert float(oracle.lse32(3.0, -1e30)) == 3.0


def test_fexp32_is_contrafold_polynomial(oracle):
    x = np.linspace(-9.9, -1e-4, 3001, dtype=np.float32)
    got = np.array([float(oracle.fexp32(v)) for v in x])
    assert np.abs(got - np.exp(x.astype(np.float64))).max() < 1e-4
    assert float(oracle.fexp32(-10.0)) == 0.0
    assert oracle.fexp32(50.0) == np.float32(1e20)
    assert float(oracle.fexp32(0.0)) == 1.0


def test_batch_runner_matches_single(oracle, P):
    rng = np.random.default_rng(3)
    seqs = [rand_seq(rng, int(rng.integers(5, 40))) for _ in range(12)]
    vals, _ = oracle.run_batch("logz64", seqs, P, threads=4)
    for s, v in zip(seqs, vals):
        assert v == oracle.inside64(s, P)
    vals, pairs = oracle.run_batch("mea32", seqs, P, gamma=6.0, threads=3)
    for s, pr in zip(seqs, pairs):
        assert oracle.dot_bracket(pr, len(s)) == oracle.mea(s, P, 6.0)

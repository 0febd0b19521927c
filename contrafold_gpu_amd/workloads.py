"""Seeded synthetic workloads of the benchmark configurations.

* uniform(n, L, seed): i.i.d. uniform ACGU, the reference's generator
  (benchmarks/generate_sequences.py:16-19) -- 1,000 x 128 nt, seed 0.
* gc_masked(n, L, seed): per-sequence GC fraction g ~ Beta(8, 8), bases with
  p(A,C,G,U) = ((1-g)/2, g/2, g/2, (1-g)/2), per-base forced-unpaired mask
  Bernoulli(0.1) -- 10,000 x 200 nt, seed 2.
* uneven(n, lo, hi, seed): lengths uniform in [lo, hi], uniform ACGU, each
  base N with p = 0.02, no masks -- 20,000 x [30, 500] nt, seed 3.
* serving_trace(n, qps, seed): Poisson arrivals at `qps`; lengths
  log-normal (median 150, sigma 0.6) clipped to [20, 1000]; 30% of requests
  carry a Bernoulli(0.15) mask; mix 70% MEA (gamma 6), 20% sample x10, 10% bpp.
"""
import numpy as np

BASES = np.array(list("ACGU"))


def uniform(n, L, seed=0):
    rng = np.random.default_rng(seed)
    idx = rng.integers(0, 4, size=(n, L), dtype=np.int8)
    return ["".join(BASES[row]) for row in idx]


def gc_masked(n, L, seed=2, mask_p=0.1):
    rng = np.random.default_rng(seed)
    seqs, masks = [], []
    for _ in range(n):
        g = rng.beta(8.0, 8.0)
        p = np.array([(1 - g) / 2, g / 2, g / 2, (1 - g) / 2])
        seqs.append("".join(BASES[rng.choice(4, size=L, p=p)]))
        masks.append((rng.random(L) < mask_p).astype(np.uint8))
    return seqs, masks


def uneven(n, lo=30, hi=500, seed=3, n_frac=0.02):
    rng = np.random.default_rng(seed)
    lens = rng.integers(lo, hi + 1, size=n)
    out = []
    for L in lens:
        b = BASES[rng.integers(0, 4, size=int(L))].copy()
        b[rng.random(int(L)) < n_frac] = "N"
        out.append("".join(b))
    return out


def serving_trace(n, qps, seed=5, median=150.0, sigma=0.6, lo=20, hi=1000):
    """Returns dict with seqs, masks (None or uint8), kinds ('mea'|'sample'|'bpp'),
    arrival_s (sorted, from 0)."""
    rng = np.random.default_rng(seed)
    gaps = rng.exponential(1.0 / qps, size=n)
    arrival = np.cumsum(gaps) - gaps[0]
    lens = np.clip(np.round(np.exp(rng.normal(np.log(median), sigma, size=n))), lo, hi).astype(int)
    seqs, masks, kinds = [], [], []
    for L in lens:
        seqs.append("".join(BASES[rng.integers(0, 4, size=int(L))]))
        masks.append((rng.random(int(L)) < 0.15).astype(np.uint8) if rng.random() < 0.3 else None)
        u = rng.random()
        kinds.append("mea" if u < 0.7 else ("sample" if u < 0.9 else "bpp"))
    return dict(seqs=seqs, masks=masks, kinds=kinds, arrival_s=arrival, lengths=lens)

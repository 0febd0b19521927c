"""Algorithmic work of a fold, counted exactly from the sequence and mask.

Counts the log-sum-exp / Fast_Exp / max-plus terms the reference's recurrences
fold (gpu_contrafold/cpu.py:595-895), per kernel of this engine.  Used for the
roofline figures: 9 FP32 flop per LSE or Fast_Exp term (candidate add, a-b,
three multiply-adds counted as 6, b+r) and 2 per max-plus term.
"""
import numpy as np

CANON = np.zeros((5, 5), bool)
for a, b in ((0, 3), (3, 0), (1, 2), (2, 1), (2, 3), (3, 2)):
    CANON[a, b] = True
CODE = {"A": 0, "C": 1, "G": 2, "U": 3, "T": 3}


def allow_matrix(seq, mask=None):
    L = len(seq)
    s = np.array([4] + [CODE.get(c, 4) for c in seq.upper()] + [4])
    fo = np.zeros(L + 2, bool)
    if mask is not None:
        fo[1:L + 1] = np.asarray(mask)[:L] != 0
    A = CANON[s[:, None], s[None, :]] & ~fo[:, None] & ~fo[None, :]
    return A


def counts(seq, mask=None):
    """Term counts per stage for one fold."""
    L = len(seq)
    A = allow_matrix(seq, mask)
    i_idx, j_idx = np.meshgrid(np.arange(L + 2), np.arange(L + 2), indexing="ij")
    fc = np.zeros((L + 2, L + 2), bool)
    if L >= 2:
        sub = A[1:L, 2:L + 1]  # pair (i, j+1), 1 <= i, j <= L-1
        fc[1:L, 1:L] = sub & (j_idx[1:L, 1:L] >= i_idx[1:L, 1:L])
    n_fc = int(fc.sum())
    # single-branch candidates: inner pair (i+l1+1, j-l2), l1+l2 <= 30, q >= p + 2
    sb = 0
    for l1 in range(31):
        for l2 in range(31 - l1):
            ii = i_idx + l1 + 1
            jj = j_idx - l2
            ok = fc & (jj >= ii + 1) & (ii <= L + 1) & (jj >= 0)
            if not ok.any():
                continue
            ip, jp = np.nonzero(ok)
            sb += int(A[ip + l1 + 1, jp - l2].sum())
    dom = [(j - i) for i in range(1, L + 1) for j in range(i + 2, L)]
    fm2 = int(sum(max(0, s - 3) for s in dom))
    n_dom = len(dom)
    ext = int(sum(A[k + 1, j] for j in range(1, L + 1) for k in range(j)))
    mea = int(sum(max(0, s - 1) + 3 for s in range(1, L + 1) for _ in range(L + 1 - s)))
    return dict(
        L=L,
        inside=n_fc * 2 + sb + fm2 + n_dom * 5,          # FC hairpin+multi, window, FM2 split, FM1/FM
        exterior=2 * (ext + L),                          # F5 and F5o chains
        outside=n_fc * 2 + sb + 2 * fm2 + n_dom * 4,     # FCo ext+FM1o, window, FMo/FM1o splits, rest
        posterior=sb + 2 * n_fc,                         # Fast_Exp terms of POST
        mea=mea,                                         # max-plus terms per gamma
        cells=(L + 1) * (L + 2) // 2,
    )


def batch_counts(seqs, masks=None, sample=32, seed=0):
    """Mean per-fold counts over a random sample of the batch, times len(seqs)."""
    rng = np.random.default_rng(seed)
    n = len(seqs)
    pick = rng.choice(n, size=min(sample, n), replace=False)
    acc = None
    for k in pick:
        c = counts(seqs[k], None if masks is None else masks[k])
        if acc is None:
            acc = {key: 0.0 for key in c}
        for key, v in c.items():
            acc[key] += v
    return {key: v / len(pick) * n for key, v in acc.items()}

"""Brute-force structure enumerator with an independent loop-decomposition scorer.

Test infrastructure.  It enumerates every nested secondary structure of a short
sequence explicitly (no dynamic programming), scores each structure loop by loop
with the complementary-model energy terms, and returns exact log partition
function, pair marginals, Viterbi maximum and MEA maximum.  It is the pin for
the oracle's recurrences (gpu_contrafold/cpu.py:165-275, 583-895): a dropped
term, a double-counted structure or a wrong dangle index changes these sums.

Model conventions (gpu_contrafold/cpu.py): 1-based positions, sentinels 4 at 0
and L+1; a pair (a, b) is allowed when both bases form AU/UA/CG/GC/GU/UG and
neither is forced unpaired; no minimum hairpin length; a loop closed by one
pair with exactly one inner pair needs l1 + l2 <= 30; loops with two or more
inner pairs are multiloops.
"""
import math
from functools import lru_cache

import numpy as np

CANON = {(0, 3), (3, 0), (1, 2), (2, 1), (2, 3), (3, 2)}


def allowed(s, fo, a, b):
    return (s[a], s[b]) in CANON and fo[a] == 0 and fo[b] == 0


def structures(s, fo, L):
    """All nested pairings of positions 1..L as tuples of (a, b), a < b."""
    @lru_cache(maxsize=None)
    def rec(i, j):
        if i > j:
            return [()]
        out = [st for st in rec(i + 1, j)]              # i unpaired
        for k in range(i + 1, j + 1):
            if allowed(s, fo, i, k):
                for a in rec(i + 1, k - 1):
                    for b in rec(k + 1, j):
                        out.append(((i, k),) + a + b)
        return out
    return rec(1, L)


def _jA(P, s, L, d, c):
    """Junction term of branch (c, d) seen from its enclosing loop: hc + dangles."""
    v = P["hc"][s[d], s[c]]
    if d < L:
        v += P["dl"][s[d], s[c], s[d + 1]]
    if c - 1 > 0:
        v += P["dr"][s[d], s[c], s[c - 1]]
    return v


def score(P, s, L, pairs, with_mult=False):
    """Loop-decomposition score of one structure; None if the model excludes it.

    With ``with_mult`` returns (score, multiplicity).  The multiplicity is the
    number of derivations of the structure in the reference grammar
    (cpu.py:179-183, 248-257): FM = FM2 | FM[i][j-1]+mu | FM1 with
    FM2 = sum_k FM1[i][k] FM[k][j] lets trailing unpaired bases of a multiloop
    be generated either inside FM[k][j] or by FM[i][j-1]+mu.  Counting the
    derivations N(m, t) of an FM region with m branches and t trailing unpaired
    bases gives N(1, t) = 1, N(m, t) = N(m-1, t) + N(m, t-1), i.e.
    C(m-1+t, t); the multiloop closing rule uses FM2 with the first branch
    fixed, so a multiloop with m inner branches counts C(m-2+t, t) times.
    """
    partner = {}
    for a, b in pairs:
        partner[a] = b
        partner[b] = a
    total = 0.0
    mult = 1

    def inner_pairs(lo, hi):
        """Pairs directly enclosed in open region lo..hi."""
        res = []
        t = lo
        while t <= hi:
            if t in partner and partner[t] > t:
                res.append((t, partner[t]))
                t = partner[t] + 1
            else:
                t += 1
        return res

    # exterior loop
    ext = inner_pairs(1, L)
    paired_ext = sum(b - a + 1 for a, b in ext)
    total += (L - paired_ext) * P["eu"]
    for c, d in ext:
        total += P["ep"] + P["bp"][s[c], s[d]] + _jA(P, s, L, d, c)
    # every closing pair (a, b): FC cell (i, j) = (a, b - 1)
    for a, b in pairs:
        i, j = a, b - 1
        inner = inner_pairs(a + 1, b - 1)
        if len(inner) == 0:
            d = j - i
            total += P["hc"][s[i], s[j + 1]] + P["tm"][s[i], s[j + 1], s[i + 1], s[j]] + P["hp_cum"][min(d, 30)]
        elif len(inner) == 1:
            c, dd = inner[0]
            l1 = c - a - 1
            l2 = b - dd - 1
            if l1 + l2 > 30:
                return (None, 0) if with_mult else None
            if l1 == 0 and l2 == 0:
                total += P["bp"][s[i + 1], s[j]] + P["stack"][s[i], s[j + 1], s[i + 1], s[j]]
            else:
                p, q = c - 1, dd
                jB1 = P["hc"][s[i], s[j + 1]] + P["tm"][s[i], s[j + 1], s[i + 1], s[j]]
                jB2 = P["hc"][s[q], s[p + 1]] + P["tm"][s[q], s[p + 1], s[q + 1], s[p]]
                sn = 0.0
                if l1 == 0 and l2 == 1:
                    sn = P["b01"][s[j]]
                elif l1 == 1 and l2 == 0:
                    sn = P["b01"][s[i + 1]]
                elif l1 == 1 and l2 == 1:
                    sn = P["i11"][s[i + 1], s[j]]
                total += P["cs"][l1, l2] + P["bp"][s[p + 1], s[q]] + jB1 + jB2 + sn
        else:
            jA = P["hc"][s[i], s[j + 1]]
            if i < L:
                jA += P["dl"][s[i], s[j + 1], s[i + 1]]
            if j > 0:
                jA += P["dr"][s[i], s[j + 1], s[j]]
            total += jA + P["mp"] + P["mb"]
            covered = sum(dd - c + 1 for c, dd in inner)
            total += ((b - a - 1) - covered) * P["mu"]
            for c, dd in inner:
                total += P["mp"] + P["bp"][s[c], s[dd]] + _jA(P, s, L, dd, c)
            t = j - inner[-1][1]
            mult *= math.comb(len(inner) - 2 + t, t)
    return (total, mult) if with_mult else total


def enumerate_all(seq, P, forced=None):
    """Returns dict: logZ, post (L+2)^2 marginals, vit (max score), structs, scores."""
    from oracle.oracle import encode, encode_mask
    s, L = encode(seq)
    fo = encode_mask(L, forced)
    sts = structures(tuple(int(x) for x in s), tuple(int(x) for x in fo), L)
    scored = []
    for st in sts:
        sc, mu = score(P, s, L, st, with_mult=True)
        if sc is not None:
            scored.append((st, sc, mu))
    m = max(sc for _, sc, _ in scored)
    Z = sum(mu * math.exp(sc - m) for _, sc, mu in scored)
    logZ = m + math.log(Z)
    post = np.zeros((L + 2, L + 2))
    for st, sc, mu in scored:
        w = mu * math.exp(sc - logZ)
        for a, b in st:
            post[a, b] += w
    return dict(logZ=logZ, post=post, vit=m, structs=[st for st, _, _ in scored],
                scores=[sc for _, sc, _ in scored], mults=[mu for _, _, mu in scored], L=L, s=s, fo=fo)


def mea_best(seq, post32, gamma, forced=None):
    """Max over all nested allowed structures of sum(POST of pairs) + sum(up of unpaired)."""
    from oracle.oracle import encode, encode_mask
    s, L = encode(seq)
    fo = encode_mask(L, forced)
    P = np.asarray(post32, np.float64)
    up = np.zeros(L + 1)
    for i in range(1, L + 1):
        u = 1.0 - P[1:i, i].sum() - P[i, i + 1:L + 1].sum()
        up[i] = u / (2.0 * gamma)
    best = -1e300
    for st in structures(tuple(int(x) for x in s), tuple(int(x) for x in fo), L):
        paired = set()
        v = 0.0
        for a, b in st:
            v += P[a, b]
            paired.add(a)
            paired.add(b)
        v += sum(up[t] for t in range(1, L + 1) if t not in paired)
        best = max(best, v)
    return best


def pairs_of(pair_arr, L):
    return tuple(sorted((t, int(pair_arr[t])) for t in range(1, L + 1) if pair_arr[t] > t))

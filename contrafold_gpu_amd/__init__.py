"""contrafold_gpu_amd -- CONTRAfold fold engine on MI355X (gfx950).

Drop-in for the Python API of the reference package ``gpu_contrafold``
(gpu_contrafold/__init__.py:17-31): same names, arguments and results.  Every
fold runs in hand-written HIP kernels for gfx950 behind the C ABI in
include/contrafold_amd.h; there is no CPU fallback -- without the built
library or a GPU the calls raise.

    from contrafold_gpu_amd import load, logZ_batch, sample_batch, mfe, mea, bpp
    P = load()
    z = logZ_batch(["GGGGAAAACCCC"], P)
    structs = sample_batch(["GGGGAAAACCCC"], P, n_samples=10)
    db = mea("GGGGAAAACCCC", P, gamma=6)
    prob = bpp("GGGGAAAACCCC", P)
"""
import contextlib
import threading

import numpy as np

from . import params
from .engine import Batch, Engine, encode_batch, fold_bytes, sequence_keys
from ._lib import CF_BPP, CF_GRAD, CF_LOGZ, CF_LOGZ64, CF_MEA, CF_MFE, CF_SAMPLE

__version__ = "0.1.0"

load = params.load

_engine = None
_engine_lock = threading.Lock()


def default_engine():
    """The process-wide engine on device 0 (or $CF_DEVICE), created on first use."""
    global _engine
    with _engine_lock:
        if _engine is None:
            _engine = Engine()
        return _engine


@contextlib.contextmanager
def _prep(P):
    """The default engine with ``P`` on the device, locked for one call: a
    call's batch is created, run and read back under the engine's lock, so
    concurrent calls from several threads (with the same or different
    parameters) do not interleave on the shared context."""
    eng = default_engine()
    with eng.lock:
        eng.set_params(P)
        yield eng


def _one(seq, forced):
    if not isinstance(seq, str):
        raise TypeError("seq must be a string")
    return [seq], (None if forced is None else [forced])


def logZ_batch(seqs, P, forced_list=None, threads=128):
    """Log partition functions, FP32 (gpu.py:480-517).  ``threads`` is
    accepted for compatibility; launch shapes come from the tuner."""
    with _prep(P) as eng, eng.batch(seqs, forced_list, sampling=False) as b:
        b.run(CF_LOGZ)
        return [float(z) for z in b.logz()]


def sample_batch(seqs, P, n_samples, forced_list=None, threads=128, seed=0, keys=None):
    """Boltzmann samples, ``n_samples`` dot-brackets per sequence
    (gpu.py:520-583).  Streams are keyed by (seed, sequence content), so a
    sequence's samples do not depend on the rest of the batch; repeats of one
    (sequence, mask) in a call get distinct streams.  ``keys`` (one 64-bit
    value per sequence) overrides the content keys."""
    if int(n_samples) <= 0:
        return [[] for _ in seqs]
    with _prep(P) as eng, eng.batch(seqs, forced_list, keys=keys) as b:
        b.run(CF_SAMPLE, n_samples=int(n_samples), seed=int(seed))
        return b.samples()


def fold_tasks_gpu(seqs, forced_masks, P, chunk=4096, seed=0):
    """One Boltzmann sample per task, results in input order (gpu.py:586-601).
    Tasks are folded in chunks of ``chunk`` sequences, grouped by length.
    Stream keys are derived once over the whole task list, so repeated
    (sequence, mask) tasks draw distinct samples whichever chunks they land in."""
    n = len(seqs)
    if n == 0:
        return []
    out = [None] * n
    buf, offs, mbuf, _lens = encode_batch(seqs, forced_masks)
    keys = sequence_keys(buf, offs, mbuf)
    order = sorted(range(n), key=lambda k: len(seqs[k]))
    step = max(1, int(chunk))
    for c0 in range(0, n, step):
        idx = order[c0:c0 + step]
        sub = [seqs[k] for k in idx]
        fl = [forced_masks[k] for k in idx] if forced_masks is not None else None
        res = sample_batch(sub, P, 1, forced_list=fl, seed=seed, keys=keys[idx])
        for k, r in zip(idx, res):
            out[k] = r[0]
    return out


def bpp_batch(seqs, P, forced_list=None):
    """Base-pair probability matrices (L x L, float64, 0-based upper triangle)."""
    with _prep(P) as eng, eng.batch(seqs, forced_list, sampling=False) as b:
        b.run(CF_BPP)
        return [b.bpp(k).astype(np.float64) for k in range(b.n)]


def bpp(seq, P, forced=None):
    """Posterior pair probabilities of one sequence (cpu.py:926-929)."""
    s, f = _one(seq, forced)
    return bpp_batch(s, P, f)[0]


def mea_batch(seqs, P, gamma=6.0, forced_list=None):
    """MEA structures; ``gamma`` may be a float or a sequence of floats (then
    one list of structures per gamma)."""
    gam = np.atleast_1d(np.asarray(gamma, np.float64)).astype(np.float32)
    with _prep(P) as eng, eng.batch(seqs, forced_list, sampling=False) as b:
        b.run(CF_MEA, gammas=gam)
        res = [b.mea(g) for g in range(gam.size)]
    return res[0] if np.ndim(gamma) == 0 else res


def mea(seq, P, gamma=6.0, forced=None):
    """Maximum-expected-accuracy structure (cpu.py:932-941)."""
    s, f = _one(seq, forced)
    return mea_batch(s, P, float(gamma), f)[0]


def mfe_batch(seqs, P, forced_list=None, return_scores=False):
    """Viterbi (maximum-probability) structures, FP64 (cpu.py:290-519)."""
    with _prep(P) as eng, eng.batch(seqs, forced_list, sampling=False) as b:
        b.run(CF_MFE)
        db, sc = b.mfe()
    return (db, sc) if return_scores else db


def mfe(seq, P, forced=None):
    """Viterbi structure of one sequence (cpu.py:505-519)."""
    s, f = _one(seq, forced)
    return mfe_batch(s, P, f)[0]


def cpu_logZ(seq, P):
    """FP64 log partition function of one sequence without constraints, the
    quantity of the reference's cpu.logZ (cpu.py:282-287).  Computed on the
    GPU by the FP64 exact-log1p inside pass."""
    with _prep(P) as eng, eng.batch([seq], sampling=False) as b:
        b.run(CF_LOGZ64)
        return float(b.logz64()[0])


def logZ_grad(seqs, P, forced_list=None, weights=None):
    """sum_b w_b * d logZ_b / d table (flat 2577 vector) from the outside
    pass: expected feature counts."""
    with _prep(P) as eng, eng.batch(seqs, forced_list, sampling=False) as b:
        b.run(CF_GRAD, weights=weights)
        return b.logz(), b.grad()


__all__ = ["load", "cpu_logZ", "mfe", "mea", "bpp", "logZ_batch", "sample_batch", "fold_tasks_gpu",
           "bpp_batch", "mea_batch", "mfe_batch", "logZ_grad", "Engine", "Batch", "default_engine", "params",
           "fold_bytes", "encode_batch"]

"""Host mirror of the fold engine: device context, batches, result decoding.

An ``Engine`` owns one C-ABI context (one HIP stream + device parameter
tables).  A ``Batch`` owns the device-resident DP tables of its sequences: a
logZ run leaves the inside tables on the GPU, and a later bpp / MEA / sample /
gradient run on the same batch reuses them.

Threads: the C context is not re-entrant (its pinned staging buffer, device
pool and error string are shared by every call on it), and ctypes releases
the GIL during a call.  Every call into the context therefore holds the
engine's lock; the drop-in functions hold it across a whole batch's life.
"""
import ctypes
import hashlib
import os
import threading
import zlib

import numpy as np

from . import params as _params
from ._lib import (CF_BPP, CF_GRAD, CF_LOGZ, CF_MEA, CF_MFE, CF_RECOMPUTE, CF_SAMPLE, CF_TIMING, check,
                   load_library)

_c = ctypes


def encode_batch(seqs, masks=None):
    """Pack sequences (and optional per-base masks) for cf_batch_create.

    Returns (seq_bytes, offsets int64[n+1], mask bytes or None, lengths).
    Masks follow gpu.py:493-494: None means unconstrained, a mask longer than
    its sequence is truncated, any nonzero entry forces the base unpaired."""
    if isinstance(seqs, (str, bytes)):
        raise TypeError("seqs must be a list of sequences")
    seqs = list(seqs)
    if len(seqs) == 0:
        raise ValueError("empty sequence list")
    if all(type(s) is str for s in seqs):
        # one join and one encode: "replace" maps every character to one byte
        lens = np.fromiter(map(len, seqs), np.int64, len(seqs))
        buf = "".join(seqs).encode("ascii", "replace")
    else:
        enc = [s.encode("ascii", "replace") if isinstance(s, str) else bytes(s) for s in seqs]
        lens = np.fromiter((len(e) for e in enc), np.int64, len(enc))
        buf = b"".join(enc)
    offs = np.zeros(len(seqs) + 1, np.int64)
    np.cumsum(lens, out=offs[1:])
    mbuf = None
    if masks is not None:
        if len(masks) != len(seqs):
            raise ValueError(f"forced_list has {len(masks)} masks for {len(seqs)} sequences")
        if all(type(m) is np.ndarray and m.ndim == 1 for m in masks):
            # the common case, one array per sequence of its own length: one
            # concatenation instead of a copy per sequence
            sizes = np.fromiter((m.size for m in masks), np.int64, len(masks))
            if (sizes == lens).all():
                mbuf = (np.concatenate(masks) != 0).view(np.uint8) if len(masks) else None
                return buf, offs, mbuf, lens
        if any(m is not None for m in masks):
            mb = np.zeros(int(offs[-1]), np.uint8)
            for k, m in enumerate(masks):
                if m is None:
                    continue
                L = int(lens[k])
                arr = np.asarray(m).reshape(-1)
                if arr.size < L:
                    raise ValueError(f"mask {k} has {arr.size} entries for a sequence of length {L}")
                mb[offs[k]:offs[k] + L] = (arr[:L] != 0)
            mbuf = mb
    return buf, offs, mbuf, lens


def sequence_keys(buf, offs, mbuf):
    """64-bit sampling-stream keys from sequence (and mask) content, with an
    occurrence counter for repeats: a sequence's samples do not depend on its
    position in the batch or on the other sequences."""
    n = len(offs) - 1
    seen = {}
    crc, adler = zlib.crc32, zlib.adler32
    view = memoryview(buf)
    mask = mbuf.tobytes() if mbuf is not None else None
    bounds = offs.tolist()
    out = [0] * n
    for k in range(n):
        a, b = bounds[k], bounds[k + 1]
        chunk = view[a:b] if mask is None else bytes(view[a:b]) + mask[a:b]
        h = (crc(chunk) << 32) | adler(chunk)
        occ = seen.get(h, 0)
        seen[h] = occ + 1
        out[k] = (h ^ (occ * 0x9E3779B97F4A7C15)) & 0xFFFFFFFFFFFFFFFF
    return np.array(out, dtype=np.uint64)


class Engine:
    """One device context.  All calls on an engine run on its HIP stream."""

    def __init__(self, device=None):
        self.lib = load_library()
        if device is None:
            device = int(os.environ.get("CF_DEVICE", "0"))
        self.device = device
        ctx = _c.c_void_p()
        rc = self.lib.cf_create(device, _c.byref(ctx))
        if rc != 0:
            raise RuntimeError(f"cf_create(device={device}) failed with status {rc}: no usable HIP device")
        self.ctx = ctx
        self._fp = None
        self.lock = threading.RLock()

    def close(self):
        with self.lock:
            if self.ctx:
                self.lib.cf_destroy(self.ctx)
                self.ctx = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_params(self, P):
        """Upload parameter tables (a cpu.load-style dict or a flat vector)
        when they differ from the ones on the device."""
        v = P if isinstance(P, np.ndarray) else _params.flatten(P)
        v = np.ascontiguousarray(v, np.float64)
        fp = hashlib.blake2b(v.tobytes(), digest_size=16).digest()
        with self.lock:
            if fp != self._fp:
                check(self.lib.cf_set_params(self.ctx, v.ctypes.data_as(_c.POINTER(_c.c_double))), self.ctx,
                      "cf_set_params")
                self._fp = fp

    def set_tuning(self, edges, threads):
        e = (_c.c_int * len(edges))(*edges)
        t = (_c.c_int * len(threads))(*threads)
        with self.lock:
            check(self.lib.cf_set_tuning(self.ctx, e, t, len(edges)), self.ctx, "cf_set_tuning")

    def set_split(self, min_length, wgs_per_fold):
        """Split teams: folds of >= min_length nt in wide buckets run on
        ``wgs_per_fold`` workgroups (CUs) each (cf_set_split); 1 turns it off."""
        with self.lock:
            check(self.lib.cf_set_split(self.ctx, int(min_length), int(wgs_per_fold)), self.ctx, "cf_set_split")

    def set_packing(self, teams_inside, teams_outside):
        """One-wave teams per workgroup of the packed inside and outside
        sweeps (cf_set_packing, 1..16 each); results do not depend on it."""
        with self.lock:
            check(self.lib.cf_set_packing(self.ctx, int(teams_inside), int(teams_outside)), self.ctx,
                  "cf_set_packing")

    def batch(self, seqs, masks=None, keys=None, sampling=True):
        return Batch(self, seqs, masks, keys, sampling)


class Batch:
    """Device-resident fold batch (see cf_batch_* in include/contrafold_amd.h)."""

    def __init__(self, engine, seqs, masks=None, keys=None, sampling=True):
        """``keys``: per-sequence sampling-stream keys; by default derived from
        each sequence's content (``sequence_keys``).  ``sampling=False`` (a
        batch that will not sample) skips deriving them: its streams are then
        keyed by index."""
        self.engine = engine
        self.lib = engine.lib
        buf, offs, mbuf, lens = encode_batch(seqs, masks)
        self.lens = lens
        self.offs = offs
        self.n = len(lens)
        if keys is None and sampling:
            keys = sequence_keys(buf, offs, mbuf)
        kptr = None
        if keys is not None:
            keys = np.ascontiguousarray(keys, np.uint64)
            kptr = keys.ctypes.data_as(_c.POINTER(_c.c_uint64))
        h = _c.c_void_p()
        self.lock = engine.lock
        with self.lock:
            check(self.lib.cf_batch_create(engine.ctx, self.n, buf, offs.ctypes.data_as(_c.POINTER(_c.c_int64)),
                                           mbuf.ctypes.data if mbuf is not None else None, kptr, _c.byref(h)),
                  engine.ctx, "cf_batch_create")
        self.h = h
        self.n_gamma = 0
        self.n_samples = 0

    def close(self):
        if getattr(self, "h", None):
            with self.lock:
                self.lib.cf_batch_destroy(self.h)
            self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def run(self, flags, gammas=(), n_samples=0, seed=0, weights=None, recompute=False, timing=False):
        g = np.ascontiguousarray(np.asarray(gammas, np.float32).reshape(-1))
        w = None
        if weights is not None:
            w = np.ascontiguousarray(np.asarray(weights, np.float32).reshape(-1))
            if w.size != self.n:
                raise ValueError(f"{w.size} weights for {self.n} sequences")
        f = int(flags) | (CF_RECOMPUTE if recompute else 0) | (CF_TIMING if timing else 0)
        with self.lock:
            check(self.lib.cf_batch_run(self.h, f, g.ctypes.data_as(_c.POINTER(_c.c_float)) if g.size else None,
                                        int(g.size), int(n_samples), int(seed) & 0xFFFFFFFFFFFFFFFF,
                                        w.ctypes.data_as(_c.POINTER(_c.c_float)) if w is not None else None),
                  self.engine.ctx, "cf_batch_run")
        if flags & CF_MEA:
            self.n_gamma = int(g.size)
        if flags & CF_SAMPLE:
            self.n_samples = int(n_samples)
        return self

    def _call(self, what, fn, *args):
        """One read-back entry point on this batch, under the engine's lock."""
        with self.lock:
            return check(fn(self.h, *args), self.engine.ctx, what)

    def last_ms(self):
        return float(self.lib.cf_batch_last_ms(self.h))

    def kernel_ms(self):
        k = self.lib.cf_kernel_count()
        out = (_c.c_float * k)()
        self._call("cf_batch_kernel_ms", self.lib.cf_batch_kernel_ms, out)
        return {self.lib.cf_kernel_name(i).decode(): float(out[i]) for i in range(k)}

    def device_bytes(self):
        return int(self.lib.cf_batch_device_bytes(self.h))

    def logz(self):
        out = np.empty(self.n, np.float32)
        self._call("cf_batch_logz", self.lib.cf_batch_logz, out.ctypes.data_as(_c.POINTER(_c.c_float)))
        return out

    def bpp(self, index):
        L = int(self.lens[index])
        out = np.zeros((L, L), np.float32)
        self._call("cf_batch_bpp", self.lib.cf_batch_bpp, int(index), out.ctypes.data_as(_c.POINTER(_c.c_float)))
        return out

    def _split(self, raw, per=None):
        """One string per sequence (per=None) or a list of ``per`` strings."""
        s = raw.decode("ascii")
        out = []
        for k in range(self.n):
            a, L = int(self.offs[k]) * (per or 1), int(self.lens[k])
            if per is None:
                out.append(s[a:a + L])
            else:
                out.append([s[a + t * L:a + (t + 1) * L] for t in range(per)])
        return out

    def mea(self, gamma_index=0):
        buf = _c.create_string_buffer(int(self.offs[-1]) + 1)
        self._call("cf_batch_mea", self.lib.cf_batch_mea, int(gamma_index), buf)
        return self._split(buf.raw[:int(self.offs[-1])])

    def mfe(self):
        buf = _c.create_string_buffer(int(self.offs[-1]) + 1)
        sc = np.empty(self.n, np.float64)
        self._call("cf_batch_mfe", self.lib.cf_batch_mfe, buf, sc.ctypes.data_as(_c.POINTER(_c.c_double)))
        return self._split(buf.raw[:int(self.offs[-1])]), sc

    def samples_raw(self):
        nb = int(self.offs[-1]) * self.n_samples
        buf = _c.create_string_buffer(nb + 1)
        self._call("cf_batch_samples", self.lib.cf_batch_samples, buf)
        return buf.raw[:nb]

    def samples(self):
        return self._split(self.samples_raw(), per=self.n_samples)

    TABLES = ("FC", "FM", "FM1", "FM2", "FCo", "FMo", "FM1o", "FM2o", "POST", "F5", "F5o")

    def table(self, index, name):
        """One DP table of sequence ``index``: triangular tables come back
        dense (L+2)x(L+2) with NaN outside the triangle; F5/F5o as (L+2,)."""
        which = self.TABLES.index(name)
        L = int(self.lens[index])
        if which >= 9:
            out = np.empty(L + 2, np.float32)
            self._call("cf_batch_table", self.lib.cf_batch_table, int(index), which,
                       out.ctypes.data_as(_c.POINTER(_c.c_float)))
            return out
        tri = np.empty((L + 1) * (L + 2) // 2, np.float32)
        self._call("cf_batch_table", self.lib.cf_batch_table, int(index), which,
                   tri.ctypes.data_as(_c.POINTER(_c.c_float)))
        dense = np.full((L + 2, L + 2), np.nan, np.float32)
        pos = 0
        for d in range(L + 1):
            i = np.arange(L + 1 - d)
            dense[i, i + d] = tri[pos:pos + L + 1 - d]
            pos += L + 1 - d
        return dense

    def logz64(self):
        out = np.empty(self.n, np.float64)
        self._call("cf_batch_logz64", self.lib.cf_batch_logz64, out.ctypes.data_as(_c.POINTER(_c.c_double)))
        return out

    def grad(self):
        out = np.empty(_params.N_TABLE, np.float64)
        self._call("cf_batch_grad", self.lib.cf_batch_grad, out.ctypes.data_as(_c.POINTER(_c.c_double)))
        return out


def fold_bytes(L, flags, n_gamma=1, n_samples=0):
    return int(load_library().cf_fold_bytes(int(L), int(flags), int(n_gamma), int(n_samples)))


__all__ = ["Engine", "Batch", "encode_batch", "sequence_keys", "fold_bytes", "CF_LOGZ", "CF_BPP", "CF_MEA",
           "CF_MFE", "CF_SAMPLE", "CF_GRAD"]

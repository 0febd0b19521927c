"""The allocation-free fold (cf_fold_noalloc) against the batch API, called
through the C ABI: logZ bit for bit and MEA dot-brackets identical, on two
calls in a row, the second with a smaller input; oversize calls rejected."""
import ctypes

import numpy as np
import pytest

from conftest import rand_seq

pytestmark = pytest.mark.gpu


def noalloc(eng, seqs, masks, flags, gamma):
    from contrafold_gpu_amd import encode_batch
    from contrafold_gpu_amd._lib import check
    buf, offs, mbuf, lens = encode_batch(seqs, masks)
    z = np.zeros(len(seqs), np.float32)
    out = ctypes.create_string_buffer(int(offs[-1]) + 1)
    with eng.lock:
        rc = eng.lib.cf_fold_noalloc(eng.ctx, len(seqs), buf, offs.ctypes.data_as(ctypes.POINTER(ctypes.c_int64)),
                                     mbuf.ctypes.data if mbuf is not None else None, flags, ctypes.c_float(gamma),
                                     z.ctypes.data_as(ctypes.POINTER(ctypes.c_float)), out)
        check(rc, eng.ctx, "cf_fold_noalloc")
    s = out.raw[:int(offs[-1])].decode("ascii")
    return z, [s[int(offs[k]):int(offs[k + 1])] for k in range(len(seqs))]


def test_noalloc_matches_batch_api(P):
    from contrafold_gpu_amd import CF_LOGZ, CF_MEA, Engine
    eng = Engine()
    eng.set_params(P)
    rng = np.random.default_rng(71)
    big = [rand_seq(rng, L) for L in (200, 180, 150, 120, 90, 60, 33, 7)]
    bmask = [None, (rng.random(180) < 0.1).astype(np.uint8), None, None, None, None, None, None]
    small = [rand_seq(rng, 45), rand_seq(rng, 12)]
    for seqs, masks in ((big, bmask), (small, None)):
        with eng.batch(seqs, masks) as b:
            b.run(CF_MEA, gammas=[4.0])
            zref, mref = b.logz().copy(), b.mea(0)
        z, m = noalloc(eng, seqs, masks, CF_MEA, 4.0)
        assert (z.view(np.uint32) == zref.view(np.uint32)).all()
        assert m == mref
        z2, _ = noalloc(eng, seqs, masks, CF_LOGZ, 0.0)
        assert (z2.view(np.uint32) == zref.view(np.uint32)).all()
    # too large for the fixed array: rejected before any device work
    with pytest.raises(ValueError):
        noalloc(eng, [rand_seq(rng, 800)], None, CF_MEA, 6.0)
    # the batch API still works afterwards
    with eng.batch(small) as b:
        b.run(CF_LOGZ)
        assert np.isfinite(b.logz()).all()
    eng.close()

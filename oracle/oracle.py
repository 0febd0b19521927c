"""Oracle: CPU restatement of the reference's CONTRAfold engine (test infrastructure).

Only tests/, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of
``bench.py`` import this module, and only as the checker.  The product package
(``contrafold_gpu_amd``) never imports it.

Contents
--------
* ``load_tables(path)`` -- restatement of ``cpu.load`` (gpu_contrafold/cpu.py:40-135):
  708 raw weights -> flat FP64 table vector (layout in ``cf_oracle.c``).
* ``encode(seq)`` -- ``cpu.encode`` (cpu.py:138-144).
* ctypes wrappers around ``cf_oracle.c``: FP64 ``inside64`` (cpu.py:165-275),
  FP64 ``viterbi64`` (cpu.py:290-502), FP32 ``posterior32`` (cpu.py:583-820),
  FP32 ``mea_decode32`` (cpu.py:823-895), and the pthread batch runner used as
  the host-CPU baseline.

How each function is pinned (see tests/test_oracle.py):
* load_tables: table shapes, symmetries and cumulative-sum identities checked
  against the params text; cross-checked against the product loader.
* inside64 / posterior32 (Z): brute-force enumeration of every structure for
  L <= 12 with an independent loop-decomposition scorer (tests/brute.py);
  closed forms (all-N / fully masked -> sum of external_unpaired).
* posterior32 (POST): brute-force pair marginals; row sums <= 1; masked rows 0.
* viterbi64: brute-force maximum score; README known answer
  GGGGAAAACCCC -> ((((....)))).
* mea_decode32: brute-force maximum of the MEA objective over all nested
  canonical structures; README known answer GGGGAAAACCCC -> ((((....)))).
* Bit-level FP32 values at near-ties are parity unpinned beyond the restated
  operation order (the reference cannot be executed in this environment).
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libcf_oracle.so")
SRC_PATH = os.path.join(HERE, "cf_oracle.c")

BASES = {"A": 0, "C": 1, "G": 2, "U": 3, "T": 3}
N_TABLE = 2577

C_MAX_SINGLE = 30
D_HAIRPIN = 30
D_BULGE = 30
D_INTERNAL = 30
D_ISYM = 15
D_IASYM = 28
D_IEXP = 4


def build(force=False):
    """Compile cf_oracle.c into libcf_oracle.so (portable x86-64-v3, no FMA contraction)."""
    import subprocess
    if not force and os.path.exists(LIB_PATH) and os.path.getmtime(LIB_PATH) >= os.path.getmtime(SRC_PATH):
        return LIB_PATH
    tmp = LIB_PATH + ".tmp"
    cmd = ["gcc", "-O3", "-march=x86-64-v3", "-ffp-contract=off", "-fno-fast-math", "-fPIC",
           "-shared", "-o", tmp, SRC_PATH, "-lm", "-lpthread"]
    subprocess.check_call(cmd)
    os.replace(tmp, LIB_PATH)
    return LIB_PATH


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = ctypes.CDLL(LIB_PATH)
        i32p = ctypes.POINTER(ctypes.c_int32)
        f64p = ctypes.POINTER(ctypes.c_double)
        f32p = ctypes.POINTER(ctypes.c_float)
        L.cfo_inside64.restype = ctypes.c_double
        L.cfo_inside64.argtypes = [i32p, ctypes.c_int, i32p, f64p]
        L.cfo_viterbi64.restype = None
        L.cfo_viterbi64.argtypes = [i32p, ctypes.c_int, i32p, f64p, i32p, f64p]
        L.cfo_posterior32.restype = ctypes.c_float
        L.cfo_posterior32.argtypes = [i32p, ctypes.c_int, i32p, f32p, f32p, ctypes.c_void_p]
        L.cfo_mea_decode32.restype = None
        L.cfo_mea_decode32.argtypes = [f32p, ctypes.c_int, i32p, i32p, ctypes.c_float, i32p, f32p]
        L.cfo_lse32.restype = ctypes.c_float
        L.cfo_lse32.argtypes = [ctypes.c_float, ctypes.c_float]
        L.cfo_fexp32.restype = ctypes.c_float
        L.cfo_fexp32.argtypes = [ctypes.c_float]
        L.cfo_batch.restype = ctypes.c_int
        L.cfo_batch.argtypes = [ctypes.c_int, ctypes.c_int, i32p, i32p, i32p, i32p, f64p, f32p,
                                ctypes.c_float, f64p, i32p, ctypes.c_int, ctypes.c_int]
        L.cfo_sample32.restype = None
        L.cfo_sample32.argtypes = [i32p, ctypes.c_int, i32p, f32p, f32p, f32p, f32p, f32p, ctypes.c_int,
                                   ctypes.c_uint64, ctypes.c_char_p]
        L.cfo_table_size.restype = ctypes.c_int
        assert L.cfo_table_size() == N_TABLE
        _lib = L
    return _lib


def _p(a, t):
    return a.ctypes.data_as(ctypes.POINTER(t))


# ----------------------------------------------------------------- params
def read_raw(path):
    raw = {}
    with open(path) as f:
        for line in f:
            p = line.split()
            if len(p) == 2:
                raw[p[0]] = float(p[1])
    return raw


def tables_from_raw(raw):
    """cpu.load (cpu.py:40-135) as a dict of FP64 numpy tables."""
    M = 5
    bp = np.zeros((M, M))
    for a in "ACGU":
        for b in "ACGU":
            v = raw.get(f"base_pair_{a}{b}", raw.get(f"base_pair_{b}{a}", 0.0))
            bp[BASES[a], BASES[b]] = v
            bp[BASES[b], BASES[a]] = v
    stack = np.zeros((M, M, M, M))
    tm = np.zeros((M, M, M, M))
    for a in "ACGU":
        for b in "ACGU":
            for c in "ACGU":
                for d in "ACGU":
                    stack[BASES[a], BASES[b], BASES[c], BASES[d]] = raw.get(
                        f"helix_stacking_{a}{b}{c}{d}", raw.get(f"helix_stacking_{d}{c}{b}{a}", 0.0))
                    tm[BASES[a], BASES[b], BASES[c], BASES[d]] = raw.get(f"terminal_mismatch_{a}{b}{c}{d}", 0.0)
    hc = np.zeros((M, M))
    for a in "ACGU":
        for b in "ACGU":
            hc[BASES[a], BASES[b]] = raw.get(f"helix_closing_{a}{b}", 0.0)
    dl = np.zeros((M, M, M))
    dr = np.zeros((M, M, M))
    for a in "ACGU":
        for b in "ACGU":
            for c in "ACGU":
                dl[BASES[a], BASES[b], BASES[c]] = raw.get(f"dangle_left_{a}{b}{c}", 0.0)
                dr[BASES[a], BASES[b], BASES[c]] = raw.get(f"dangle_right_{a}{b}{c}", 0.0)

    def cum(prefix, dmax, start):
        al = np.zeros(dmax + 1)
        for k in range(start, dmax + 1):
            al[k] = raw.get(f"{prefix}_at_least_{k}", 0.0)
        return np.cumsum(al)

    hp_cum = cum("hairpin_length", D_HAIRPIN, 0)
    bulge_cum = cum("bulge_length", D_BULGE, 1)
    il_cum = cum("internal_length", D_INTERNAL, 2)
    isym_cum = cum("internal_symmetric_length", D_ISYM, 1)
    iasym_cum = cum("internal_asymmetry", D_IASYM, 1)
    iexp = np.zeros((D_IEXP + 1, D_IEXP + 1))
    for l1 in range(1, D_IEXP + 1):
        for l2 in range(1, D_IEXP + 1):
            a, b = min(l1, l2), max(l1, l2)
            iexp[l1, l2] = raw.get(f"internal_explicit_{a}_{b}", 0.0)
    i11 = np.zeros((M, M))
    for a in "ACGU":
        for b in "ACGU":
            v = raw.get(f"internal_1x1_nucleotides_{a}{b}", raw.get(f"internal_1x1_nucleotides_{b}{a}", 0.0))
            i11[BASES[a], BASES[b]] = v
            i11[BASES[b], BASES[a]] = v
    b01 = np.zeros(M)
    for a in "ACGU":
        b01[BASES[a]] = raw.get(f"bulge_0x1_nucleotides_{a}", 0.0)
    cs = np.zeros((C_MAX_SINGLE + 1, C_MAX_SINGLE + 1))
    for l1 in range(0, C_MAX_SINGLE + 1):
        for l2 in range(0, C_MAX_SINGLE + 1 - l1):
            if l1 == 0 and l2 == 0:
                continue
            if l1 == 0 or l2 == 0:
                cs[l1, l2] = bulge_cum[min(D_BULGE, l1 + l2)]
            else:
                v = 0.0
                if l1 <= D_IEXP and l2 <= D_IEXP:
                    v += iexp[l1, l2]
                v += il_cum[min(D_INTERNAL, l1 + l2)]
                if l1 == l2:
                    v += isym_cum[min(D_ISYM, l1)]
                v += iasym_cum[min(D_IASYM, abs(l1 - l2))]
                cs[l1, l2] = v
    return dict(bp=bp, stack=stack, tm=tm, hc=hc, dl=dl, dr=dr, hp_cum=hp_cum, cs=cs, i11=i11, b01=b01,
                mb=raw["multi_base"], mu=raw["multi_unpaired"], mp=raw["multi_paired"],
                eu=raw["external_unpaired"], ep=raw["external_paired"])


ORDER = ("bp", "stack", "tm", "hc", "dl", "dr", "hp_cum", "cs", "i11", "b01")
SCALARS = ("mb", "mu", "mp", "eu", "ep")


def flat(P):
    v = np.concatenate([np.asarray(P[k], np.float64).ravel() for k in ORDER] +
                       [np.array([P[k] for k in SCALARS], np.float64)])
    assert v.size == N_TABLE
    return np.ascontiguousarray(v)


def flat32(P):
    """FP32 copy made like cpu._params_f32 (cpu.py:898-910): FP64 tables cast to float32."""
    return np.ascontiguousarray(flat(P).astype(np.float32))


def default_params_path():
    return os.path.join(os.path.dirname(HERE), "contrafold_gpu_amd", "data", "contrafold.params.complementary")


def load_tables(path=None):
    return tables_from_raw(read_raw(path or default_params_path()))


# ----------------------------------------------------------------- encode
def encode(seq):
    L = len(seq)
    s = np.full(L + 2, 4, dtype=np.int32)
    for i, c in enumerate(seq.upper()):
        s[i + 1] = BASES.get(c, 4)
    return s, L


def encode_mask(L, forced):
    fo = np.zeros(L + 2, dtype=np.int32)
    if forced is not None:
        for k in range(L):
            fo[k + 1] = 1 if int(forced[k]) != 0 else 0
    return fo


# ----------------------------------------------------------------- engines
def inside64(seq, P, forced=None):
    s, L = encode(seq)
    fo = encode_mask(L, forced)
    v = flat(P)
    return lib().cfo_inside64(_p(s, ctypes.c_int32), L, _p(fo, ctypes.c_int32), _p(v, ctypes.c_double))


def viterbi64(seq, P, forced=None):
    """Returns (pair array L+2, score)."""
    s, L = encode(seq)
    fo = encode_mask(L, forced)
    v = flat(P)
    pair = np.empty(L + 2, np.int32)
    sc = ctypes.c_double(0.0)
    lib().cfo_viterbi64(_p(s, ctypes.c_int32), L, _p(fo, ctypes.c_int32), _p(v, ctypes.c_double),
                        _p(pair, ctypes.c_int32), ctypes.byref(sc))
    return pair, sc.value


class _Tables32(ctypes.Structure):
    _fields_ = [(n, ctypes.POINTER(ctypes.c_float)) for n in
                ("FC", "FM", "FM1", "F5", "FCo", "FMo", "FM1o", "F5o")]


def posterior32(seq, P, forced=None, tables=False, P32=None):
    """Returns (Z, POST dense (L+2)^2 float32[, dict of dense tables])."""
    s, L = encode(seq)
    fo = encode_mask(L, forced)
    v = P32 if P32 is not None else flat32(P)
    n = L + 2
    POST = np.zeros((n, n), np.float32)
    out = None
    keep = {}
    if tables:
        out = _Tables32()
        for name in ("FC", "FM", "FM1", "FCo", "FMo", "FM1o"):
            keep[name] = np.empty((n, n), np.float32)
            setattr(out, name, _p(keep[name], ctypes.c_float))
        for name in ("F5", "F5o"):
            keep[name] = np.empty(n, np.float32)
            setattr(out, name, _p(keep[name], ctypes.c_float))
    z = lib().cfo_posterior32(_p(s, ctypes.c_int32), L, _p(fo, ctypes.c_int32), _p(v, ctypes.c_float),
                              _p(POST, ctypes.c_float), ctypes.byref(out) if out is not None else None)
    z = np.float32(z)
    if tables:
        return z, POST, keep
    return z, POST


def mea_decode32(POST, seq, gamma, forced=None):
    """Returns (pair array L+2, objective score[0][L])."""
    s, L = encode(seq)
    fo = encode_mask(L, forced)
    pair = np.empty(L + 2, np.int32)
    obj = ctypes.c_float(0.0)
    P = np.ascontiguousarray(POST, np.float32)
    lib().cfo_mea_decode32(_p(P, ctypes.c_float), L, _p(fo, ctypes.c_int32), _p(s, ctypes.c_int32),
                           ctypes.c_float(np.float32(gamma)), _p(pair, ctypes.c_int32), ctypes.byref(obj))
    return pair, np.float32(obj.value)


def dot_bracket(pair, L):
    ch = ["."] * L
    for t in range(1, L + 1):
        if pair[t] > t:
            ch[t - 1] = "("
            ch[pair[t] - 1] = ")"
    return "".join(ch)


def bpp(seq, P, forced=None):
    _z, POST = posterior32(seq, P, forced)
    L = len(seq)
    return POST[1:L + 1, 1:L + 1].astype(np.float64)


def mea(seq, P, gamma=6.0, forced=None):
    _z, POST = posterior32(seq, P, forced)
    pair, _ = mea_decode32(POST, seq, gamma, forced)
    return dot_bracket(pair, len(seq))


def mfe(seq, P, forced=None):
    pair, _ = viterbi64(seq, P, forced)
    return dot_bracket(pair, len(seq))


def sample32(seq, P, n_samples, seed=0, forced=None):
    """CPU Boltzmann samples (restated 3-pass traceback of gpu.py:196-477)."""
    s, L = encode(seq)
    fo = encode_mask(L, forced)
    _z, _POST, T = posterior32(seq, P, forced, tables=True)
    buf = ctypes.create_string_buffer(max(1, n_samples * L))
    v = flat32(P)
    lib().cfo_sample32(_p(s, ctypes.c_int32), L, _p(fo, ctypes.c_int32), _p(v, ctypes.c_float),
                       _p(T["FC"], ctypes.c_float), _p(T["FM"], ctypes.c_float), _p(T["FM1"], ctypes.c_float),
                       _p(T["F5"], ctypes.c_float), int(n_samples), int(seed), buf)
    raw = buf.raw[:n_samples * L].decode("ascii")
    return [raw[k * L:(k + 1) * L] for k in range(n_samples)]


def lse32(a, b):
    return np.float32(lib().cfo_lse32(ctypes.c_float(np.float32(a)), ctypes.c_float(np.float32(b))))


def fexp32(x):
    return np.float32(lib().cfo_fexp32(ctypes.c_float(np.float32(x))))


# ----------------------------------------------------------------- batch runner
KIND = {"logz64": 0, "bpp32": 1, "mea32": 2, "viterbi64": 3, "bpp_sample32": 4}


def pack(seqs, masks=None):
    n = len(seqs)
    lens = np.array([len(s) for s in seqs], np.int32)
    off = np.zeros(n, np.int32)
    if n:
        off[1:] = np.cumsum(lens[:-1] + 2)
    total = int((lens + 2).sum())
    codes = np.full(total, 4, np.int32)
    fo = np.zeros(total, np.int32)
    for b, sq in enumerate(seqs):
        s, L = encode(sq)
        codes[off[b]:off[b] + L + 2] = s
        if masks is not None and masks[b] is not None:
            fo[off[b]:off[b] + L + 2] = encode_mask(L, masks[b])
    return codes, fo, off, lens


def run_batch(kind, seqs, P, masks=None, gamma=6.0, threads=None, n_samples=0):
    """Fold ``seqs`` on ``threads`` host threads.  Returns (values float64[n], pairs list or None)."""
    codes, fo, off, lens = pack(seqs, masks)
    n = len(seqs)
    v64 = flat(P)
    v32 = flat32(P)
    out_val = np.zeros(n, np.float64)
    out_pair = np.full(codes.size, -1, np.int32)
    threads = threads or os.cpu_count() or 1
    lib().cfo_batch(KIND[kind], n, _p(codes, ctypes.c_int32), _p(fo, ctypes.c_int32), _p(off, ctypes.c_int32),
                    _p(lens, ctypes.c_int32), _p(v64, ctypes.c_double), _p(v32, ctypes.c_float),
                    ctypes.c_float(np.float32(gamma)), _p(out_val, ctypes.c_double), _p(out_pair, ctypes.c_int32),
                    int(threads), int(n_samples))
    pairs = None
    if kind in ("mea32", "viterbi64"):
        pairs = [out_pair[off[b]:off[b] + lens[b] + 2].copy() for b in range(n)]
    return out_val, pairs

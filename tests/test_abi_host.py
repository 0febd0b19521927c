"""CPU-side checks: the C ABI library loads and exports every entry point
declared in include/*.h; host logic (packing, masks, keys, parameter map,
workload generators, work counts) behaves."""
import os
import re

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    names = set()
    for fn in os.listdir(os.path.join(ROOT, "include")):
        if not fn.endswith(".h"):
            continue
        text = open(os.path.join(ROOT, "include", fn)).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        for m in re.finditer(r"\b(cf_[a-z0-9_]+)\s*\(", text):
            names.add(m.group(1))
    return names


def test_library_exports_every_declared_symbol():
    import ctypes
    from contrafold_gpu_amd import _lib
    if not os.path.exists(_lib.LIB_PATH):
        from contrafold_gpu_amd import build
        build.build(force=True)
    lib = ctypes.CDLL(_lib.LIB_PATH)
    decl = declared_symbols()
    assert len(decl) >= 30
    missing = [n for n in sorted(decl) if not hasattr(lib, n)]
    assert not missing, missing
    bound = {name for name, _, _ in _lib.SIGNATURES}
    assert decl == bound, (decl ^ bound)


def test_abi_constants_without_gpu():
    from contrafold_gpu_amd import _lib, params
    lib = _lib.load_library()
    assert lib.cf_table_size() == params.N_TABLE == 2577
    assert lib.cf_version() >= 1
    names = [lib.cf_kernel_name(k) for k in range(lib.cf_kernel_count())]
    assert names == [b"inside", b"exterior", b"outside", b"post", b"mea", b"viterbi", b"sample", b"grad"]
    # admission bytes grow like L^2 and with the stages requested
    b200 = lib.cf_fold_bytes(200, _lib.CF_LOGZ, 1, 0)
    b400 = lib.cf_fold_bytes(400, _lib.CF_LOGZ, 1, 0)
    assert 3.5 < b400 / b200 < 4.5
    assert lib.cf_fold_bytes(200, _lib.CF_MEA, 1, 0) > lib.cf_fold_bytes(200, _lib.CF_BPP, 1, 0) > b200
    assert lib.cf_fold_bytes(-1, _lib.CF_LOGZ, 1, 0) == -1


def test_null_and_bad_arguments_return_status():
    import ctypes
    from contrafold_gpu_amd import _lib
    lib = _lib.load_library()
    assert lib.cf_create(0, None) == _lib.CF_EINVAL
    assert lib.cf_set_params(None, None) == _lib.CF_EINVAL
    assert lib.cf_batch_run(None, 1, None, 0, 0, 0, None) == _lib.CF_EINVAL
    assert lib.cf_batch_logz(None, None) == _lib.CF_EINVAL
    assert lib.cf_last_error(None) == b"null context"
    lib.cf_destroy(None)
    lib.cf_batch_destroy(None)
    lib.cf_server_destroy(None)


def test_status_mapping():
    from contrafold_gpu_amd import _lib
    with pytest.raises(ValueError):
        _lib.check(_lib.CF_EINVAL)
    with pytest.raises(MemoryError):
        _lib.check(_lib.CF_ENOMEM)
    with pytest.raises(RuntimeError):
        _lib.check(_lib.CF_EDEVICE)
    assert _lib.check(3) == 3


def test_encode_batch_masks_and_errors():
    from contrafold_gpu_amd.engine import encode_batch, sequence_keys
    buf, offs, mbuf, lens = encode_batch(["ACGU", "", "gg"], [None, None, [1, 0, 5]])
    assert buf == b"ACGUgg" and list(offs) == [0, 4, 4, 6] and list(lens) == [4, 0, 2]
    assert list(mbuf) == [0, 0, 0, 0, 1, 0]          # truncated to L, nonzero -> 1
    assert encode_batch(["AC"], [None])[2] is None
    with pytest.raises(ValueError):
        encode_batch([], None)
    with pytest.raises(ValueError):
        encode_batch(["ACGU"], [[0, 1]])
    with pytest.raises(ValueError):
        encode_batch(["ACGU", "A"], [None])
    with pytest.raises(TypeError):
        encode_batch("ACGU")
    # non-ASCII characters become one placeholder byte each (-> N)
    b2, o2, _, l2 = encode_batch(["AéU"])
    assert l2[0] == 3
    k1 = sequence_keys(*encode_batch(["ACGU", "GGG", "ACGU"])[:3])
    k2 = sequence_keys(*encode_batch(["GGG", "ACGU"])[:3])
    assert k1[0] == k2[1] and k1[2] != k1[0]          # content keys, repeats distinct


def test_encode_batch_array_masks_fast_path():
    """Masks given as one 1-D array per sequence of its own length take one
    concatenation; the bytes equal the per-sequence path's (lists of the same
    values), whatever the integer dtype and nonzero values."""
    from contrafold_gpu_amd.engine import encode_batch
    rng = np.random.default_rng(3)
    seqs = ["ACGU" * int(n) for n in rng.integers(0, 30, 40)]
    arrs = [(rng.random(len(q)) < 0.2).astype(dt) * v
            for q, dt, v in zip(seqs, [np.int8, np.int64, np.uint8, np.int32] * 10, [1, 3, 255, -2] * 10)]
    fast = encode_batch(seqs, arrs)[2]
    slow = encode_batch(seqs, [list(a) for a in arrs])[2]
    assert fast.dtype == np.uint8 and (fast == slow).all() and set(np.unique(fast)) <= {0, 1}
    # a longer array is truncated (per-sequence path), a shorter one refused
    long = [np.ones(len(q) + 2, np.int8) for q in seqs[:3]]
    assert encode_batch(seqs[:3], long)[2].sum() == sum(len(q) for q in seqs[:3])
    with pytest.raises(ValueError):
        encode_batch(["ACGU"], [np.ones(2, np.int8)])


def test_encode_batch_one_join_matches_per_sequence():
    """String batches are encoded with one join: the bytes and lengths equal
    per-sequence ASCII encoding with replacement (non-ASCII characters become
    one '?' byte each), and bytes inputs take the per-item path."""
    from contrafold_gpu_amd.engine import encode_batch
    seqs = ["GGGGAAAACCCC", "acgtNNx\u00e9\u4e2dZ", "", "ACGU"]
    buf, offs, mbuf, lens = encode_batch(seqs)
    ref = [x.encode("ascii", "replace") for x in seqs]
    assert buf == b"".join(ref)
    assert list(lens) == [len(x) for x in ref]
    assert list(offs) == [0, 12, 22, 22, 26]
    buf, offs, _, lens = encode_batch([b"ACGU", "GGCC"])
    assert buf == b"ACGUGGCC" and list(lens) == [4, 4]


def test_sequence_keys_pinned():
    """Sampling-stream keys depend only on content (and mask) plus a repeat
    counter; pinned so that a change to the key derivation, which would change
    every sampled structure, does not pass unnoticed."""
    from contrafold_gpu_amd.engine import encode_batch, sequence_keys
    seqs = ["GGGGAAAACCCC", "ACGU", "ACGU", "", "acgtn"]
    k = sequence_keys(*encode_batch(seqs)[:3])
    assert [int(x) for x in k] == [0x6fe83a4314ea032d, 0xd409af6402b40121, 0x4a3ed6dd7dfe7d34, 0x1,
                                   0xe722dc050601020e]
    masks = [None, [0, 1, 0, 0], None, None, [1, 0, 0, 0, 0]]
    k = sequence_keys(*encode_batch(seqs, masks)[:3])
    assert [int(x) for x in k] == [0x5102b3073b06032d, 0x169a62a6073b0122, 0x1758089107380121, 0x1,
                                   0x75d6a059104c020f]
    # position-independent: the same sequence keys alike wherever it sits
    k2 = sequence_keys(*encode_batch(["ACGU", "GGGGAAAACCCC"])[:3])
    assert int(k2[1]) == 0x6fe83a4314ea032d


def test_param_map_linear_and_gradient_transpose():
    from contrafold_gpu_amd import params
    names, raw, pm = params.load_raw()
    rng = np.random.default_rng(0)
    x = rng.normal(size=raw.size)
    y = rng.normal(size=params.N_TABLE)
    J = pm.dense()
    assert np.allclose(pm.tables(x), J @ x, atol=1e-12)
    assert np.allclose(pm.raw_grad(y), J.T @ y, atol=1e-12)
    assert pm.tables(raw).tobytes() == params.flatten(params.load()).tobytes()


def test_loader_matches_oracle_restatement(oracle):
    from contrafold_gpu_amd import params
    P = params.load()
    Q = oracle.load_tables()
    for k in Q:
        assert np.array_equal(np.asarray(P[k]), np.asarray(Q[k])), k


def test_loader_table_identities(P):
    """Symmetries and running-sum identities of cpu.load (cpu.py:51-130)."""
    from contrafold_gpu_amd import params
    names, raw = params.read_raw()
    rv = dict(zip(names, raw))
    assert np.array_equal(P["bp"], P["bp"].T) and np.array_equal(P["i11"], P["i11"].T)
    assert not P["bp"][4].any() and not P["stack"][4].any()          # N row stays 0
    d = np.diff(P["hp_cum"])
    assert np.allclose(d, [rv[f"hairpin_length_at_least_{k}"] for k in range(1, 31)])
    assert P["hp_cum"][0] == rv["hairpin_length_at_least_0"]
    assert P["cs"][0, 5] == P["cs"][5, 0] == sum(rv[f"bulge_length_at_least_{k}"] for k in range(1, 6))
    # stacking falls back to the reversed key
    for w in ("CGCG", "AUGC", "GCUA"):
        a, b, c, dd = (params.IDX[ch] for ch in w)
        want = rv.get("helix_stacking_" + w, rv.get("helix_stacking_" + w[::-1], 0.0))
        assert P["stack"][a, b, c, dd] == want
    for s in ("mb", "mu", "mp", "eu", "ep"):
        assert isinstance(P[s], float)


def test_workload_generators():
    from contrafold_gpu_amd import workloads
    a = workloads.uniform(5, 30, seed=0)
    assert a == workloads.uniform(5, 30, seed=0) and all(len(s) == 30 for s in a)
    seqs, masks = workloads.gc_masked(50, 100, seed=2)
    frac = np.mean([m.mean() for m in masks])
    assert 0.07 < frac < 0.13
    u = workloads.uneven(200, seed=3)
    L = np.array([len(s) for s in u])
    assert L.min() >= 30 and L.max() <= 500
    nfrac = np.mean([s.count("N") / len(s) for s in u])
    assert 0.01 < nfrac < 0.03
    tr = workloads.serving_trace(2000, 100.0, seed=5)
    assert np.all(np.diff(tr["arrival_s"]) >= 0)
    assert tr["lengths"].min() >= 20 and tr["lengths"].max() <= 1000
    assert 140 <= np.median(tr["lengths"]) <= 160
    kinds = np.array(tr["kinds"])
    assert abs((kinds == "mea").mean() - 0.7) < 0.04


def test_work_counts_closed_forms():
    from contrafold_gpu_amd import workcount
    c = workcount.counts("A" * 50)        # nothing pairs: no FC cells, no windows
    assert c["posterior"] == 0 and c["exterior"] == 100
    c = workcount.counts("GGGGAAAACCCC")
    assert c["cells"] == 13 * 14 // 2 and c["inside"] > 0


def test_serving_steady_rate():
    """The steady-state served rate: a server that keeps up reads the offered
    rate whatever its latency; one that falls behind reads its own capacity;
    failed requests never count."""
    from contrafold_gpu_amd.serving import steady_rate
    n, qps = 10000, 1000.0
    arr = np.arange(n) / qps
    assert abs(steady_rate(arr, np.full(n, 50.0)) - qps) / qps < 0.01      # constant 50 ms latency
    assert abs(steady_rate(arr, np.full(n, 900.0)) - qps) / qps < 0.01     # long but steady
    # capacity 500/s: request k completes at k / 500 s
    lat = (np.arange(n) / 500.0 - arr) * 1e3
    assert abs(steady_rate(arr, lat) - 500.0) / 500.0 < 0.02
    failed = np.full(n, 50.0)
    failed[::2] = -1.0
    assert abs(steady_rate(arr, failed) - qps / 2) / qps < 0.01


def test_autotune_with_a_fake_engine(tmp_path, monkeypatch):
    """Host logic of the tuner: it times every candidate on at least a third of
    the workload (up to 8,000 sequences), then the packings on the fastest one
    when it has a packed bucket, on the workload itself (up to 25,000
    sequences), installs the fastest and caches it."""
    from contrafold_gpu_amd import autotune
    monkeypatch.setenv("CF_TUNE_CACHE", str(tmp_path / "tune.json"))

    class FakeBatch:
        def __init__(self, eng, seqs):
            self.eng, self.n = eng, len(seqs)
            eng.sizes.append(self.n)
        def __enter__(self):
            return self
        def __exit__(self, *a):
            return False
        def run(self, flags, **kw):
            pass
        def last_ms(self):  # a 256-wide last bucket is fastest, and packing (4, 4) on it
            return (5.0 if self.eng.cur[1][-1] == 256 else 9.0) - (1.0 if self.eng.pack == (4, 4) else 0.0)

    class FakeEngine:
        def __init__(self):
            self.sizes, self.cur, self.installed, self.pack = [], None, [], None
        def set_tuning(self, edges, threads):
            self.cur = (list(edges), list(threads))
            self.installed.append(self.cur)
        def set_packing(self, inside, outside):
            self.pack = (inside, outside)
        def batch(self, seqs, masks=None):
            return FakeBatch(self, seqs)

    eng = FakeEngine()
    seqs = ["ACGU" * 10] * 9000
    res = autotune.tune(eng, seqs, sample=1500, repeats=1)
    n_cand, n_pack = len(autotune.CANDIDATES), len(autotune.PACKINGS)
    assert not res["cached"] and len(res["tried"]) == n_cand + n_pack  # + the default on the whole workload
    assert set(eng.sizes) == {3000, 9000}  # candidates on a third of 9,000 (not 1,500), packings on all
    assert res["threads"][-1] == 256 and eng.cur == (res["edges"], res["threads"])
    assert 64 in res["threads"] and res["packing"] == [4, 4] and eng.pack == (4, 4) and res["ms"] == 4.0
    assert all(t["packing"] == list(autotune.PACKINGS[0]) for t in res["tried"][:n_cand])
    eng.pack = None
    again = autotune.tune(eng, seqs, sample=1500, repeats=1)
    assert again["cached"] and again["threads"] == res["threads"] and eng.pack == (4, 4)
    fe = FakeEngine()
    big = autotune.tune(fe, ["ACGU" * 30] * 60000, sample=1500, repeats=1, use_cache=False)
    assert big["ms"] == 4.0 and set(fe.sizes) == {8000, 25000}  # capped at 8,000 and 25,000
    # a packing within the margin of the default does not replace it
    fm = FakeEngine()
    fm_batch_ms = FakeBatch.last_ms
    FakeBatch.last_ms = lambda self: (5.0 if self.eng.cur[1][-1] == 256 else 9.0) - (
        0.1 if self.eng.pack == (4, 4) else 0.0)
    near = autotune.tune(fm, seqs, sample=1500, repeats=1, use_cache=False)
    FakeBatch.last_ms = fm_batch_ms
    assert near["packing"] == list(autotune.PACKINGS[0]) and near["ms"] == 5.0
    # a packing whose first timing passes the bar but whose confirmation runs
    # do not (a lucky sample timing) does not replace the default
    fl = FakeEngine()
    seen = {"n": 0}

    def lucky(self):
        base = 5.0 if self.eng.cur[1][-1] == 256 else 9.0
        if self.eng.pack == (2, 2):
            seen["n"] += 1
            return base - 1.0 if seen["n"] == 1 else base + 0.5
        return base
    FakeBatch.last_ms = lucky
    once = autotune.tune(fl, seqs, sample=1500, repeats=1, use_cache=False)
    FakeBatch.last_ms = fm_batch_ms
    assert seen["n"] == 3 and once["packing"] == list(autotune.PACKINGS[0]) and once["ms"] == 5.0
    # no packed bucket among the candidates: no packing phase, default packing
    fw = FakeEngine()
    wide = autotune.tune(fw, seqs, sample=1500, repeats=1, use_cache=False,
                         candidates=[([autotune.BIG], [128]), ([autotune.BIG], [256])])
    assert len(wide["tried"]) == 2 and wide["packing"] == list(autotune.PACKINGS[0]) and wide["ms"] == 5.0


def test_one_hip_runtime_per_process():
    """torch bundles its own HIP runtime; loading the engine first and torch
    after would map two runtimes into one process.  Importing torch first
    binds the engine to torch's copy; the other order makes the autograd
    module refuse to import (no GPU call is made: loading only)."""
    import subprocess
    import sys
    pytest.importorskip("torch")
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    first = ("import sys; sys.path.insert(0, %r); import torch; from contrafold_gpu_amd import _lib; "
             "_lib.load_library(); assert len(_lib.hip_runtimes()) == 1; print('ok')" % root)
    r = subprocess.run([sys.executable, "-c", first], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0 and "ok" in r.stdout, r.stderr[-1500:]
    later = ("import sys; sys.path.insert(0, %r); from contrafold_gpu_amd import _lib; _lib.load_library()\n"
             "try:\n    import contrafold_gpu_amd.autograd\nexcept ImportError as e:\n"
             "    print('refused', 'two HIP runtimes' in str(e))\n" % root)
    r = subprocess.run([sys.executable, "-c", later], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0 and "refused True" in r.stdout, (r.stdout, r.stderr[-1500:])


def test_write_raw_round_trip(tmp_path):
    """Parameter write-back: load -> write -> load gives the same raw weights
    and tables bit for bit, for the bundled weights and for perturbed
    (trained-like) ones; the gradient's raw order matches the file's."""
    from contrafold_gpu_amd import params
    names, raw, pmap = params.load_raw()
    p1 = tmp_path / "a.params"
    params.write_raw(str(p1), names, raw)
    n2, r2, m2 = params.load_raw(str(p1))
    assert n2 == names and r2.tobytes() == raw.tobytes()
    assert params.flatten(params.load(str(p1))).tobytes() == params.flatten(params.load()).tobytes()
    rng = np.random.default_rng(5)
    trained = raw + rng.normal(scale=1e-3, size=raw.size) * np.pi
    p2 = tmp_path / "b.params"
    params.write_raw(str(p2), names, trained)
    n3, r3, m3 = params.load_raw(str(p2))
    assert r3.tobytes() == trained.tobytes()
    assert m3.tables(r3).tobytes() == pmap.tables(trained).tobytes()
    # a step along the raw gradient direction written and read back unchanged
    g = pmap.raw_grad(rng.normal(size=params.N_TABLE))
    step = raw - 0.01 * g
    params.write_raw(str(p2), names, step)
    assert params.load_raw(str(p2))[1].tobytes() == step.tobytes()
    with pytest.raises(ValueError):
        params.write_raw(str(p2), names[:-1], raw)

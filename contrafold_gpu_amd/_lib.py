"""ctypes binding of libcontrafold_amd.so (include/contrafold_amd.h).

The library is the only compute path: if it is missing or no GPU is visible,
loading raises instead of falling back to anything else.
"""
import ctypes
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libcontrafold_amd.so")

CF_OK, CF_EINVAL, CF_ENOMEM, CF_EDEVICE, CF_ENOPARAMS, CF_ESTATE = 0, -1, -2, -3, -4, -5
CF_LOGZ, CF_BPP, CF_MEA, CF_MFE, CF_SAMPLE, CF_GRAD, CF_LOGZ64 = 1, 2, 4, 8, 16, 32, 64
CF_GRAPH, CF_TIMING, CF_RECOMPUTE = 256, 512, 1024

_c = ctypes
_vp = _c.c_void_p
_i64p = _c.POINTER(_c.c_int64)
_u8p = _c.c_void_p
_f32p = _c.POINTER(_c.c_float)
_f64p = _c.POINTER(_c.c_double)

# (name, restype, argtypes) for every entry point of the C ABI
SIGNATURES = [
    ("cf_version", _c.c_int, []),
    ("cf_table_size", _c.c_int, []),
    ("cf_create", _c.c_int, [_c.c_int, _c.POINTER(_vp)]),
    ("cf_destroy", None, [_vp]),
    ("cf_last_error", _c.c_char_p, [_vp]),
    ("cf_set_params", _c.c_int, [_vp, _f64p]),
    ("cf_set_tuning", _c.c_int, [_vp, _c.POINTER(_c.c_int), _c.POINTER(_c.c_int), _c.c_int]),
    ("cf_set_split", _c.c_int, [_vp, _c.c_int, _c.c_int]),
    ("cf_set_packing", _c.c_int, [_vp, _c.c_int, _c.c_int]),
    ("cf_batch_create", _c.c_int, [_vp, _c.c_int, _c.c_char_p, _i64p, _u8p, _c.POINTER(_c.c_uint64),
                                   _c.POINTER(_vp)]),
    ("cf_batch_run", _c.c_int, [_vp, _c.c_uint, _f32p, _c.c_int, _c.c_int, _c.c_uint64, _f32p]),
    ("cf_batch_logz", _c.c_int, [_vp, _f32p]),
    ("cf_batch_bpp", _c.c_int, [_vp, _c.c_int, _f32p]),
    ("cf_batch_mea", _c.c_int, [_vp, _c.c_int, _c.c_void_p]),
    ("cf_batch_mfe", _c.c_int, [_vp, _c.c_void_p, _f64p]),
    ("cf_batch_samples", _c.c_int, [_vp, _c.c_void_p]),
    ("cf_batch_grad", _c.c_int, [_vp, _f64p]),
    ("cf_batch_logz64", _c.c_int, [_vp, _f64p]),
    ("cf_batch_table", _c.c_int, [_vp, _c.c_int, _c.c_int, _f32p]),
    ("cf_batch_last_ms", _c.c_float, [_vp]),
    ("cf_kernel_count", _c.c_int, []),
    ("cf_kernel_name", _c.c_char_p, [_c.c_int]),
    ("cf_batch_kernel_ms", _c.c_int, [_vp, _f32p]),
    ("cf_batch_device_bytes", _c.c_int64, [_vp]),
    ("cf_batch_destroy", None, [_vp]),
    ("cf_fold_bytes", _c.c_int64, [_c.c_int, _c.c_uint, _c.c_int, _c.c_int]),
    ("cf_copy_probe", _c.c_int, [_vp, _c.c_int64]),
    ("cf_fold_noalloc", _c.c_int, [_vp, _c.c_int, _c.c_char_p, _i64p, _u8p, _c.c_uint, _c.c_float, _f32p,
                                   _c.c_void_p]),
    ("cf_server_create", _c.c_int, [_vp, _c.c_int64, _c.c_int, _c.c_int, _c.POINTER(_vp)]),
    ("cf_server_submit", _c.c_int64, [_vp, _c.c_char_p, _c.c_int, _u8p, _c.c_uint, _c.c_float, _c.c_int,
                                      _c.c_uint64, _c.c_double]),
    ("cf_server_poll", _c.c_int, [_vp, _i64p, _c.c_int]),
    ("cf_server_result", _c.c_int, [_vp, _c.c_int64, _c.c_void_p, _f32p, _f64p]),
    ("cf_server_pending", _c.c_int, [_vp]),
    ("cf_server_stats", _c.c_int, [_vp, _f64p, _c.c_int]),
    ("cf_server_replay", _c.c_int, [_vp, _c.c_int, _c.c_char_p, _i64p, _u8p, _c.POINTER(_c.c_uint), _f32p,
                                    _c.POINTER(_c.c_int), _f64p, _c.c_uint64, _f64p, _f64p, _f64p]),
    ("cf_server_destroy", None, [_vp]),
]

_lib = None


def hip_runtimes():
    """Distinct HIP runtime library files mapped into this process."""
    found = set()
    try:
        with open("/proc/self/maps") as fh:
            for line in fh:
                parts = line.split()
                if len(parts) >= 6 and "libamdhip64" in os.path.basename(parts[-1]):
                    found.add(os.path.realpath(parts[-1]))
    except OSError:
        pass
    return found


def check_single_runtime(what="the fold engine"):
    """Raise if two HIP runtimes share this process.

    The engine library links libamdhip64.so.7.  PyTorch's wheel bundles its
    own copy whose soname is the same, but its libraries ask for it by the
    file name libamdhip64.so, which the dynamic loader does not match against
    an already loaded libamdhip64.so.7.  So importing torch *before* the
    engine gives one runtime (the engine binds to torch's copy by soname),
    while importing torch *after* the engine has started loads a second HIP
    runtime with its own HSA runtime into the process: two runtimes then
    drive one GPU, and GPU work of the one started second has crashed
    (notably under rocprofv3 counter collection, which instruments both)."""
    rts = hip_runtimes()
    if len(rts) > 1:
        raise ImportError(f"{what}: two HIP runtimes are loaded ({', '.join(sorted(rts))}); import torch "
                          f"before contrafold_gpu_amd starts its engine, or not at all")


def load_library(path=None):
    """Load and bind the shared library (raises OSError if it is not built)."""
    global _lib
    if _lib is not None and path is None:
        return _lib
    p = path or LIB_PATH
    if not os.path.exists(p):
        raise OSError(f"{p} is not built; run `python -m contrafold_gpu_amd.build` (needs hipcc)")
    lib = ctypes.CDLL(p)
    check_single_runtime()
    for name, res, args in SIGNATURES:
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    if path is None:
        _lib = lib
    return lib


class FoldError(RuntimeError):
    pass


def check(rc, ctx=None, what=""):
    """Map a C status to the reference's exception classes."""
    if rc >= 0:
        return rc
    msg = ""
    if ctx is not None:
        raw = load_library().cf_last_error(ctx)
        msg = raw.decode() if raw else ""
    text = f"{what}: {msg}" if what else msg
    if rc == CF_EINVAL:
        raise ValueError(text)
    if rc == CF_ENOMEM:
        raise MemoryError(text)
    raise FoldError(f"{text} (status {rc})")

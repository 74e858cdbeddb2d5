"""ctypes binding of libholohip.so (the C-ABI declared in include/holohip.h).

The library is loaded lazily on first use.  There is no fallback: if the
library is missing, or no HIP device is present, every compute call raises.
"""
import ctypes as C
import os
import threading

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libholohip.so")

_lock = threading.Lock()
_lib = None

c_dbl_p = C.c_void_p   # device pointers are passed as integers
_I64 = C.c_int64
_I32 = C.c_int
_D = C.c_double
_P = C.c_void_p
_SZ = C.c_size_t

# name -> (restype, argtypes)
SIGNATURES = {
    "hd_last_error": (C.c_char_p, []),
    "hd_version": (_I32, []),
    "hd_timing_enable": (_I32, [_I32]),
    "hd_timing_read": (_I32, [C.c_char_p, _P, _P]),
    "hd_cosmo_build_dcom_table": (_I32, [_D, _I32, _P, _P, _P]),
    "hd_cosmo_eval": (_I32, [_D, _D, _P, _P, _I32, _P, _I64, _P, _P, _P, _P]),
    "hd_static_density": (_I32, [_P, _I32, _P, _I32, _P, _I32, _P, _P, _P, _I32, _P, _P, _P, _P]),
    "hd_zero_stalled": (_I32, [_P, _P, _I64, _P]),
    "hd_mass_stellar": (_I32, [_P, _I32, _P, _I32, _I32, _P, _P, _P, _P, _P]),
    "hd_find_2pwl_norm": (_I32, [_D, _P, _P, _I64, _D, _D, _D, _D, _I32, _P, _P, _P]),
    "hd_hard_func_2pwl_gw": (_I32, [_P, _P, _P, _P, _I64, _D, _D, _D, _P, _P]),
    "hd_lifetime_2pwl": (_I32, [_P, _P, _P, _I64, _D, _D, _D, _D, _I32, _P, _P]),
    "hd_dbn_gw": (_I32, [_P, _I32, _P, _P, _I32, _P, _I32, _I32, _P, _P, _P, _I32, _P, _P, _P]),
    "hd_dbn_2pwl": (_I32, [_P, _I32, _I32, _D, _I32, _P, _D, _D, _D, _P, _P, _I32, _P, _I32, _P, _I32, _P,
                           _P, _P, _P, _I32, _P, _P, _P]),
    "hd_integrate_3dx1d": (_I32, [_P, _I32, _P, _I32, _P, _I32, _P, _I32, _P, _P, _P]),
    "hd_h2fdf_redz": (_I32, [_P, _I32, _P, _I32, _I32, _P, _P, _I32, _P, _D, _D, _D, _P, _P, _I32, _P, _P]),
    "hd_scatter_workspace_size": (_SZ, [_I32, _I32, _I32, _I32, _I32, _I32, _I32]),
    "hd_scatter": (_I32, [_P, _I32, _I32, _I32, _I32, _P, _P, _P, _P, _P, _I32, _I32, _I32, _I32, _I32, _P, _P, _P,
                          _P, _P, _P, _P, _P, _D, _I32, _P, _SZ, _P, _P]),
    "hd_gemm_f64": (_I32, [_P, _P, _P, _I32, _I32, _I32, _P]),
    "hd_ct_locate": (_I32, [_P, _I32, _P, _P, _I32, _P, _I32, _P, _P, _P, _P]),
    "hd_ct_levels": (_I32, [_P, _P, _I32, _P]),
    "hd_scatter_weights_T": (_I32, [_P, _I32, _P, _P]),
    "hd_bin_params_redz":(_I32, [_P, _I32, _P, _I32, _P, _I32, _P, _I32, _P, _D, _D, _P, _P, _I32, _P, _P]),
    "hd_order_workspace_size": (_SZ, [_I64]),
    "hd_order_desc": (_I32, [_P, _I64, _I64, _I64, _P, _P, _SZ, _P]),
    "hd_realize_counts_workspace_size": (_SZ, [_I64, _I32, _I32, _I32, _I32]),
    "hd_realize_counts": (_I32, [_P, _P, _P, _I64, _I32, _I32, _I32, _I32, _I32, _P, _I32, _P, _P, _P, _P, _P,
                                 _SZ, _P]),
    "hd_realize_workspace_size": (_SZ, [_I64, _I32, _I32, _I32, _I32]),
    "hd_realize": (_I32, [_P, _P, _P, _I64, _I32, _I32, _I32, _I32, _I32, _D, C.c_uint64, _P, _I32, _P, _P, _P,
                          _P, _P, _SZ, _P]),
    "hd_realize_record_counts": (_I32, [_P, _P, _P, _I64, _I32, _I32, _I32, _I32, _I32, _D, C.c_uint64, _P, _I32,
                                        _P, _P, _P, _P, _P, _P, _SZ, _P]),
    "hd_realize_noalloc": (_I32, [_P, _P, _P, _I64, _I32, _I32, _I32, _I32, _I32, _D, C.c_uint64, _P, _I32, _P, _P,
                                  _P, _P, _P]),
    "hd_philox4x32_10": (None, [_P, _P, _P]),
    "hd_xoshiro128ss": (None, [_P, _I64, _P]),
    "hd_poisson_cdf_table": (_I32, [C.c_double, _I32, _P, _P, _P]),
    "hd_poisson_window": (_I32, [C.c_double, _I32, _P, _P, _P]),
    "hd_table_draw_words": (_I32, [_P, _I32, _D, _I32, _I32, _P, _I64, _P, _P, _P]),
    "hd_table_draw_pair_words": (_I32, [_P, _I32, _D, _I32, _I32, _P, _I32, _D, _I32, _I32, _P, _I64, _P, _P, _P]),
    "hd_mlp_gemm_bf16": (_I32, [_P, _I64, _I64, _P, _I64, _I64, _I32, _I32, _I32, _I32, _P, _P, _P, _P, _I64, _P]),
    "hd_mlp_colsum_bf16": (_I32, [_P, _I32, _I32, _P, _P]),
    "hd_mlp_mse": (_I32, [_P, _P, _I64, _P, _P, _P]),
    "hd_mlp_adam": (_I32, [_I64, _P, _P, _P, _P, C.c_float, C.c_float, C.c_float, C.c_float, _I32, _P, _P]),
    "hd_gwb_ideal_workspace_size": (_SZ, []),
    "hd_gwb_ideal": (_I32, [_P, _P, _P, _I32, _P, _I32, _P, _I32, _I32, _P, _I32, _D, _D, _D, _I32, _P, _P, _SZ, _P]),
    "hd_eccen_workspace_size": (_SZ, [_I32, _I32, _I32]),
    "hd_gwb_eccen": (_I32, [_P, _P, _I32, _P, _I32, _P, _P, _I32, _P, _I32, _P, _P, _I32, _I32, _P, _P, _SZ, _P]),
    "hd_eccen_number": (_I32, [_P, _P, _I32, _P, _I32, _P, _P, _I32, _P, _I32, _P, _P, _I32, _I32, _I32, _P, _P, _P,
                               _SZ, _P]),
    "hd_bessel_jpair": (_I32, [_P, _P, _I64, _P, _P]),
    "hd_gw_freq_dist": (_I32, [_P, _P, _I64, _P, _P]),
    "hd_f32_to_bf16": (_I32, [_P, _I64, _P, _P]),
}

HD_SEL_SORTED = 0
HD_SEL_ARGMAX = 1
HD_SEL_NONE = 2
HD_MLP_EPI_STORE = 0
HD_MLP_EPI_BIAS_SILU = 1
HD_MLP_EPI_BIAS_BF16 = 2
HD_MLP_EPI_SILU_BWD = 3
HD_FLAG_SKIP_ZERO_H = 1
HD_DENS_NPAR = 48


class HoloHipError(RuntimeError):
    pass


def load():
    """Load libholohip.so and bind every exported entry point (raises if absent)."""
    global _lib
    if _lib is not None:
        return _lib
    with _lock:
        if _lib is not None:
            return _lib
        if not os.path.exists(LIB_PATH):
            raise HoloHipError(
                f"{LIB_PATH} not found: build it with `python -m holodeck_amd.build` (hipcc, gfx950)")
        lib = C.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(lib, name, None)
            if fn is None:
                continue
            fn.restype = res
            fn.argtypes = args
        _lib = lib
    return _lib


def call(name, *args):
    """Call an int-status entry point; raise HoloHipError with hd_last_error() on failure."""
    lib = load()
    fn = getattr(lib, name)
    rc = fn(*args)
    if rc != 0:
        msg = lib.hd_last_error().decode(errors="replace")
        raise HoloHipError(f"{name} failed ({rc}): {msg}")
    return rc


def size(name, *args):
    return int(getattr(load(), name)(*args))


def philox4x32_10(ctr, key):
    """Host Philox4x32-10 block (known-answer tests)."""
    lib = load()
    c = (C.c_uint32 * 4)(*ctr)
    k = (C.c_uint32 * 2)(*key)
    o = (C.c_uint32 * 4)()
    lib.hd_philox4x32_10(C.addressof(c), C.addressof(k), C.addressof(o))
    return [int(v) for v in o]


def timing_enable(on=True):
    """Start (clearing earlier records) or stop launch timing in the library."""
    call("hd_timing_enable", 1 if on else 0)


def timing_read(name):
    """(total milliseconds, launches) recorded for one timed launch sequence since timing_enable."""
    ms = C.c_double(0.0)
    n = C.c_int(0)
    call("hd_timing_read", name.encode(), C.addressof(ms), C.addressof(n))
    return ms.value, n.value


def poisson_window(lam, cap):
    """(kb, kt, n) of the realization kernels' CDF table for Poisson(lam) with cap entries (host):
    support from kb, table counts kt .. kt + n - 1."""
    kb, kt, n = C.c_int(0), C.c_int(0), C.c_int(0)
    call("hd_poisson_window", float(lam), int(cap), C.addressof(kb), C.addressof(kt), C.addressof(n))
    return kb.value, kt.value, n.value


def xoshiro128ss(state, n):
    """Host xoshiro128**: (state after n steps, the n outputs as uint32)."""
    import numpy as np
    lib = load()
    s = (C.c_uint32 * 4)(*[int(v) & 0xFFFFFFFF for v in state])
    out = np.zeros(max(int(n), 0), dtype=np.uint32)
    lib.hd_xoshiro128ss(C.addressof(s), int(n), out.ctypes.data if n > 0 else None)
    return [int(v) for v in s], out

"""Python (ctypes) binding of libhwyblas_amd, the MI355X Level-3 BLAS.

The product is the C-ABI shared library built from csrc/ (see
include/hwyblas_amd.h); this module only loads it and passes pointers, for
tests and benchmarks. It never computes anything itself: when the library is
missing, loading fails with an error instead of falling back to any other
implementation.
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libhwyblas_amd.so")
HEADER = os.path.join(os.path.dirname(HERE), "include", "hwyblas_amd.h")

_lib = None

_c = ctypes.c_char
_i = ctypes.c_int
_i64 = ctypes.c_int64
_u64 = ctypes.c_uint64
_d = ctypes.c_double
_f = ctypes.c_float
_p = ctypes.c_void_p
_pi = ctypes.POINTER(ctypes.c_int)
_sz = ctypes.c_size_t

# CBLAS enum values (include/hwyblas_amd.h)
RowMajor, ColMajor = 101, 102
NoTrans, Trans, ConjTrans = 111, 112, 113
Upper, Lower = 121, 122
NonUnit, Unit = 131, 132
Left, Right = 141, 142

ERROR_HANDLER = ctypes.CFUNCTYPE(None, ctypes.c_char_p, ctypes.c_int)


class LibraryMissing(ImportError):
    pass


def _declare(L):
    def fn(name, res, args):
        f = getattr(L, name)
        f.restype = res
        f.argtypes = args
    for pre, s in (("d", _d), ("s", _f)):
        fn(f"{pre}gemm_", None, [ctypes.c_char_p, ctypes.c_char_p, _pi, _pi, _pi, _p, _p, _pi,
                                 _p, _pi, _p, _p, _pi])
        fn(f"{pre}trsm_", None, [ctypes.c_char_p] * 4 + [_pi, _pi, _p, _p, _pi, _p, _pi])
        fn(f"cblas_{pre}gemm", None, [_i, _i, _i, _i, _i, _i, s, _p, _i, _p, _i, s, _p, _i])
        fn(f"cblas_{pre}trsm", None, [_i, _i, _i, _i, _i, _i, _i, s, _p, _i, _p, _i])
        for name in ("gemmtr", "gemmt"):
            fn(f"cblas_{pre}{name}", None, [_i, _i, _i, _i, _i, _i, s, _p, _i, _p, _i, s, _p, _i])
        fn(f"cblas_{pre}gemv", None, [_i, _i, _i, _i, s, _p, _i, _p, _i, s, _p, _i])
        fn(f"cblas_{pre}symm", None, [_i, _i, _i, _i, _i, s, _p, _i, _p, _i, s, _p, _i])
        fn(f"cblas_{pre}syrk", None, [_i, _i, _i, _i, _i, s, _p, _i, s, _p, _i])
        fn(f"cblas_{pre}syr2k", None, [_i, _i, _i, _i, _i, s, _p, _i, _p, _i, s, _p, _i])
        fn(f"cblas_{pre}trmm", None, [_i, _i, _i, _i, _i, _i, _i, s, _p, _i, _p, _i])
        fn(f"hwyblas_{pre}gemm_dev", _i, [_c, _c, _i, _i, _i, s, _p, _i, _p, _i, s, _p, _i, _p])
        fn(f"hwyblas_{pre}trsm_dev", _i, [_c, _c, _c, _c, _i, _i, s, _p, _i, _p, _i, _p])
        fn(f"hwyblas_{pre}gemm_dev_cfg", _i, [_i, _c, _c, _i, _i, _i, s, _p, _i, _p, _i, s, _p,
                                             _i, _p])
    fn("LAPACKE_set_nancheck", None, [_i])
    fn("LAPACKE_get_nancheck", _i, [])
    for pre in ("d", "s"):
        for w in ("", "_work"):
            fn(f"LAPACKE_{pre}potrf{w}", _i, [_i, ctypes.c_char, _i, _p, _i])
            fn(f"LAPACKE_{pre}potrs{w}", _i, [_i, ctypes.c_char, _i, _i, _p, _i, _p, _i])
            fn(f"LAPACKE_{pre}posv{w}", _i, [_i, ctypes.c_char, _i, _i, _p, _i, _p, _i])
    fn("dgetrf_", None, [_pi, _pi, _p, _pi, _p, _pi])
    fn("dgetrs_", None, [ctypes.c_char_p, _pi, _pi, _p, _pi, _p, _p, _pi, _pi])
    fn("dgesv_", None, [_pi, _pi, _p, _pi, _p, _p, _pi, _pi])
    fn("dgerfs_", None, [ctypes.c_char_p, _pi, _pi, _p, _pi, _p, _pi, _p, _p, _pi, _p, _pi, _p,
                         _p, _p, _p, _pi])
    for w in ("", "_work"):
        fn(f"LAPACKE_dgetrf{w}", _i, [_i, _i, _i, _p, _i, _p])
        fn(f"LAPACKE_dgetrs{w}", _i, [_i, ctypes.c_char, _i, _i, _p, _i, _p, _p, _i])
        fn(f"LAPACKE_dgesv{w}", _i, [_i, _i, _i, _p, _i, _p, _p, _i])
    for pre in ("z", "c"):
        for r in ("syrk_", "herk_"):
            fn(f"{pre}{r}", None, [ctypes.c_char_p, ctypes.c_char_p, _pi, _pi, _p, _p, _pi, _p,
                                   _p, _pi])
        fn(f"{pre}gemv_", None, [ctypes.c_char_p, _pi, _pi, _p, _p, _pi, _p, _pi, _p, _p, _pi])
        fn(f"cblas_{pre}gemv", None, [_i, _i, _i, _i, _p, _p, _i, _p, _i, _p, _p, _i])
        fn(f"{pre}gemm_", None, [ctypes.c_char_p, ctypes.c_char_p, _pi, _pi, _pi, _p, _p, _pi,
                                 _p, _pi, _p, _p, _pi])
        fn(f"cblas_{pre}gemm", None, [_i, _i, _i, _i, _i, _i, _p, _p, _i, _p, _i, _p, _p, _i])
        fn(f"hwyblas_{pre}gemm_dev", _i, [_c, _c, _i, _i, _i, _p, _p, _i, _p, _i, _p, _p, _i,
                                          _p])
        fn(f"{pre}trsm_", None, [ctypes.c_char_p] * 4 + [_pi, _pi, _p, _p, _pi, _p, _pi])
        fn(f"cblas_{pre}trsm", None, [_i, _i, _i, _i, _i, _i, _i, _p, _p, _i, _p, _i])
        fn(f"hwyblas_{pre}trsm_dev", _i, [_c, _c, _c, _c, _i, _i, _p, _p, _i, _p, _i, _p])
    fn("hwyblas_dgemm_emulated_dev", _i, [_c, _c, _i, _i, _i, _d, _p, _i, _p, _i, _d, _p, _i,
                                          _i, _p, _pi])
    fn("hwyblas_dgemm_emulated_bf16_dev", _i, [_c, _c, _i, _i, _i, _d, _p, _i, _p, _i, _d, _p,
                                               _i, _i, _p, _pi])
    fn("hwyblas_dgemm_emulated_noalloc_dev", _i, [_c, _c, _i, _i, _i, _d, _p, _i, _p, _i, _d,
                                                  _p, _i, _i, _p, _pi])
    fn("hwyblas_gemm_plan_create", _p, [_i, _i, _p, _p, _pi, _pi])
    fn("hwyblas_trsm_plan_create", _p, [_i, _i, _p, _pi, _pi])
    fn("hwyblas_plan_launch", None, [_p, _p])
    fn("hwyblas_plan_launch_graph", None, [_p, _p])
    fn("hwyblas_plan_work_items", _i, [_p])
    fn("hwyblas_plan_problem_config", _i, [_p, _i])
    fn("hwyblas_plan_destroy", None, [_p])
    fn("hwyblas_grouped_config_count", _i, [])
    fn("hwyblas_grouped_config_info", _i, [_i, _pi, _pi, _pi, _pi])
    fn("hwyblas_grouped_default_config", _i, [_i, _i, _i, _i])
    fn("hwyblas_amd_fill_pow2", None, [_p, ctypes.c_int64, ctypes.c_int64, ctypes.c_int64,
                                       ctypes.c_uint64, _i, _i, _p])
    fn("hwyblas_amd_scale_rows_cols", None, [_p, ctypes.c_int64, ctypes.c_int64, ctypes.c_int64,
                                             _p, _p, _p])
    fn("hwyblas_amd_capture_begin", None, [_p])
    fn("hwyblas_amd_capture_end", _p, [_p])
    fn("hwyblas_amd_graph_launch", None, [_p, _p])
    fn("hwyblas_amd_graph_destroy", None, [_p])
    _pf = ctypes.POINTER(ctypes.c_float)
    fn("hwyblas_grouped_tune", _i, [_i, _i, _p, _i, _p, _p])
    fn("hwyblas_grouped_tune_copies", _i, [_i, _i, _i, _i])
    fn("hwyblas_grouped_features", None, [_i, _i, _i, _i, _i, _i, _p])
    fn("hwyblas_grouped_choose_configs", _i, [_i, _i, _p, _i, _p, _p])
    fn("hwyblas_mlp_create", _p, [_i, ctypes.c_uint64])
    fn("hwyblas_mlp_destroy", None, [_p])
    fn("hwyblas_mlp_num_params", _i, [_p])
    fn("hwyblas_mlp_outputs", _i, [_p])
    fn("hwyblas_mlp_get_params", None, [_p, _p])
    fn("hwyblas_mlp_set_params", None, [_p, _p])
    fn("hwyblas_mlp_predict", _i, [_p, _i, _p, _p, _p])
    fn("hwyblas_mlp_gradients", _i, [_p, _i, _p, _p, _p, _p, _pf, _p])
    fn("hwyblas_mlp_train", _i, [_p, _i, _p, _p, _p, _i, _i, ctypes.c_float, ctypes.c_uint64,
                                 _p, _p])
    fn("hwyblas_grouped_set_model", None, [_p])
    fn("hwyblas_dgetrf_dev", _i, [_i, _i, _p, _i, _p, _i, _p])
    fn("hwyblas_dgetrs_dev", _i, [ctypes.c_char, _i, _i, _p, _i, _p, _p, _i, _p])
    fn("hwyblas_set_dgemm_mode", _i, [_i])
    fn("hwyblas_get_dgemm_mode", _i, [])
    fn("dgeqrf_", None, [_pi, _pi, _p, _pi, _p, _p, _pi, _pi])
    fn("hwyblas_dgeqrf_dev", _i, [_i, _i, _p, _i, _p, _p])
    fn("hwyblas_dgerfs_dev", _i, [ctypes.c_char, _i, _i, _p, _i, _p, _i, _p, _p, _i, _p, _i, _p,
                                  _p, _p])
    fn("hwyblas_dgesv_ir_dev", _i, [_i, _i, _p, _i, _p, _i, _p, _p, _i, _p, _i, _i, _i, _p, _p,
                                    _p, _p, _p])
    fn("hwyblas_gemm_config_count", _i, [])
    fn("hwyblas_gemm_config_info", _i, [_i, _pi, _pi, _pi, _pi])
    fn("hwyblas_gemm_config_stages", _i, [_i])
    fn("hwyblas_gemm_default_config", _i, [_i, _c, _c, _i, _i, _i])
    fn("xerbla_", None, [ctypes.c_char_p, _pi, _sz])
    fn("hwyblas_set_error_handler", _p, [_p])
    fn("hwyblas_amd_device_count", _i, [])
    fn("hwyblas_amd_malloc", _p, [_sz])
    fn("hwyblas_amd_free", None, [_p])
    fn("hwyblas_amd_memcpy", None, [_p, _p, _sz, _i])
    fn("hwyblas_amd_memset", None, [_p, _i, _sz])
    fn("hwyblas_amd_synchronize", None, [])
    fn("hwyblas_amd_thread_stream", _p, [])
    fn("hwyblas_amd_stream_create", _p, [])
    fn("hwyblas_amd_stream_destroy", None, [_p])
    fn("hwyblas_amd_stream_synchronize", None, [_p])
    fn("hwyblas_amd_event_create", _p, [])
    fn("hwyblas_amd_event_destroy", None, [_p])
    fn("hwyblas_amd_event_record", None, [_p, _p])
    fn("hwyblas_amd_event_elapsed_ms", _f, [_p, _p])
    fn("hwyblas_amd_fill_uniform", None, [_p, _i, _i64, _i64, _i64, _u64, _d, _d, _p])
    fn("hwyblas_amd_version", ctypes.c_char_p, [])
    fn("hwyblas_amd_profile_enable", None, [_i])
    fn("hwyblas_amd_profile_ms", _f, [_i])


def lib():
    """Loads libhwyblas_amd.so (raises LibraryMissing if it is not built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise LibraryMissing(
                f"{LIB_PATH} is not built; run `python -m hwyblas_amd.build`")
        L = ctypes.CDLL(LIB_PATH)
        _declare(L)
        _lib = L
    return _lib


def exported_symbols_in_header():
    """Names of every function the public header declares."""
    import re
    text = "\n".join(l for l in open(HEADER).read().splitlines()
                     if not l.lstrip().startswith("#"))
    return sorted(set(re.findall(r"HWYBLAS_AMD_API\s+[^;(]*?\b(\w+)\s*\(", text)))


# ---------------------------------------------------------------------------
# Helpers for host arrays (numpy, Fortran order) through the Fortran ABI.

def _ptr(x):
    if x is None:
        return None
    if isinstance(x, DeviceArray):
        return ctypes.c_void_p(x.ptr)
    if isinstance(x, int):  # a raw device address
        return ctypes.c_void_p(x)
    return x.ctypes.data_as(ctypes.c_void_p)


def _ip(v):
    return ctypes.byref(ctypes.c_int(int(v)))


def _sp(v, dtype):
    return ctypes.byref(ctypes.c_double(v) if np.dtype(dtype) == np.float64 else ctypes.c_float(v))


def _pre(dtype):
    return "d" if np.dtype(dtype) == np.float64 else "s"


def gemm(ta, tb, m, n, k, alpha, a, lda, b, ldb, beta, c, ldc, dtype=None):
    """Fortran ?gemm_ on numpy (host) arrays or DeviceArrays; C in place."""
    dtype = dtype or c.dtype
    f = getattr(lib(), f"{_pre(dtype)}gemm_")
    f(ta.encode(), tb.encode(), _ip(m), _ip(n), _ip(k), _sp(alpha, dtype), _ptr(a), _ip(lda),
      _ptr(b), _ip(ldb), _sp(beta, dtype), _ptr(c), _ip(ldc))


def trsm(side, uplo, ta, diag, m, n, alpha, a, lda, b, ldb, dtype=None):
    dtype = dtype or b.dtype
    f = getattr(lib(), f"{_pre(dtype)}trsm_")
    f(side.encode(), uplo.encode(), ta.encode(), diag.encode(), _ip(m), _ip(n),
      _sp(alpha, dtype), _ptr(a), _ip(lda), _ptr(b), _ip(ldb))


def symm(side, uplo, m, n, alpha, a, lda, b, ldb, beta, c, ldc, dtype=None):
    """Fortran ?symm_ (host arrays or DeviceArrays); C in place."""
    dtype = dtype or c.dtype
    f = getattr(lib(), f"{_pre(dtype)}symm_")
    f(side.encode(), uplo.encode(), _ip(m), _ip(n), _sp(alpha, dtype), _ptr(a), _ip(lda),
      _ptr(b), _ip(ldb), _sp(beta, dtype), _ptr(c), _ip(ldc))


def syrk(uplo, trans, n, k, alpha, a, lda, beta, c, ldc, dtype=None):
    dtype = dtype or c.dtype
    f = getattr(lib(), f"{_pre(dtype)}syrk_")
    f(uplo.encode(), trans.encode(), _ip(n), _ip(k), _sp(alpha, dtype), _ptr(a), _ip(lda),
      _sp(beta, dtype), _ptr(c), _ip(ldc))


def syr2k(uplo, trans, n, k, alpha, a, lda, b, ldb, beta, c, ldc, dtype=None):
    dtype = dtype or c.dtype
    f = getattr(lib(), f"{_pre(dtype)}syr2k_")
    f(uplo.encode(), trans.encode(), _ip(n), _ip(k), _sp(alpha, dtype), _ptr(a), _ip(lda),
      _ptr(b), _ip(ldb), _sp(beta, dtype), _ptr(c), _ip(ldc))


def gemmtr(uplo, ta, tb, n, k, alpha, a, lda, b, ldb, beta, c, ldc, dtype=None, name="gemmtr"):
    """Fortran ?gemmtr_ (or, with name="gemmt", ?gemmt_)."""
    dtype = dtype or c.dtype
    f = getattr(lib(), f"{_pre(dtype)}{name}_")
    f(uplo.encode(), ta.encode(), tb.encode(), _ip(n), _ip(k), _sp(alpha, dtype), _ptr(a),
      _ip(lda), _ptr(b), _ip(ldb), _sp(beta, dtype), _ptr(c), _ip(ldc))


def trmm(side, uplo, ta, diag, m, n, alpha, a, lda, b, ldb, dtype=None):
    dtype = dtype or b.dtype
    f = getattr(lib(), f"{_pre(dtype)}trmm_")
    f(side.encode(), uplo.encode(), ta.encode(), diag.encode(), _ip(m), _ip(n),
      _sp(alpha, dtype), _ptr(a), _ip(lda), _ptr(b), _ip(ldb))


def cblas_gemm(layout, ta, tb, m, n, k, alpha, a, lda, b, ldb, beta, c, ldc, dtype=None):
    dtype = dtype or c.dtype
    f = getattr(lib(), f"cblas_{_pre(dtype)}gemm")
    f(layout, ta, tb, m, n, k, alpha, _ptr(a), lda, _ptr(b), ldb, beta, _ptr(c), ldc)


def gemv(trans, m, n, alpha, a, lda, x, incx, beta, y, incy, dtype=None):
    """Fortran ?gemv_ (host arrays or DeviceArrays); y in place."""
    dtype = dtype or y.dtype
    f = getattr(lib(), f"{_pre(dtype)}gemv_")
    f(trans.encode(), _ip(m), _ip(n), _sp(alpha, dtype), _ptr(a), _ip(lda), _ptr(x), _ip(incx),
      _sp(beta, dtype), _ptr(y), _ip(incy))


def potrf(uplo, n, a, lda, dtype=None):
    """Fortran ?potrf_ (host array or DeviceArray) in place; returns info."""
    dtype = dtype or a.dtype
    info = ctypes.c_int(0)
    getattr(lib(), f"{_pre(dtype)}potrf_")(uplo.encode(), _ip(n), _ptr(a), _ip(lda),
                                          ctypes.byref(info))
    return info.value


def potrs(uplo, n, nrhs, a, lda, b, ldb, dtype=None):
    dtype = dtype or b.dtype
    info = ctypes.c_int(0)
    getattr(lib(), f"{_pre(dtype)}potrs_")(uplo.encode(), _ip(n), _ip(nrhs), _ptr(a), _ip(lda),
                                          _ptr(b), _ip(ldb), ctypes.byref(info))
    return info.value


def posv(uplo, n, nrhs, a, lda, b, ldb, dtype=None):
    dtype = dtype or b.dtype
    info = ctypes.c_int(0)
    getattr(lib(), f"{_pre(dtype)}posv_")(uplo.encode(), _ip(n), _ip(nrhs), _ptr(a), _ip(lda),
                                         _ptr(b), _ip(ldb), ctypes.byref(info))
    return info.value


def cblas_level3(routine, *args, dtype):
    """cblas_{s,d}{routine} with its CBLAS arguments (enums as ints)."""
    f = getattr(lib(), f"cblas_{_pre(dtype)}{routine}")
    f(*[_ptr(x) if isinstance(x, (np.ndarray, DeviceArray)) else x for x in args])


def cblas_trsm(layout, side, uplo, ta, diag, m, n, alpha, a, lda, b, ldb, dtype=None):
    dtype = dtype or b.dtype
    f = getattr(lib(), f"cblas_{_pre(dtype)}trsm")
    f(layout, side, uplo, ta, diag, m, n, alpha, _ptr(a), lda, _ptr(b), ldb)


def _zpre(dtype):
    return "z" if np.dtype(dtype) == np.complex128 else "c"


def _zscalar(v, dtype):
    """(re, im) of a complex scalar as a real array of the type's precision."""
    rt = np.float64 if np.dtype(dtype) == np.complex128 else np.float32
    v = complex(v)
    return np.array([v.real, v.imag], dtype=rt)


def zgemm(ta, tb, m, n, k, alpha, a, lda, b, ldb, beta, c, ldc, dtype=None):
    """Fortran cgemm_/zgemm_ on numpy complex (host) arrays or DeviceArrays;
    C in place."""
    dtype = dtype or c.dtype
    al, be = _zscalar(alpha, dtype), _zscalar(beta, dtype)
    getattr(lib(), f"{_zpre(dtype)}gemm_")(
        ta.encode(), tb.encode(), _ip(m), _ip(n), _ip(k), _ptr(al), _ptr(a), _ip(lda), _ptr(b),
        _ip(ldb), _ptr(be), _ptr(c), _ip(ldc))


def cblas_zgemm(layout, ta, tb, m, n, k, alpha, a, lda, b, ldb, beta, c, ldc, dtype=None):
    """cblas_cgemm/cblas_zgemm (enums as ints, scalars as Python complex)."""
    dtype = dtype or c.dtype
    al, be = _zscalar(alpha, dtype), _zscalar(beta, dtype)
    getattr(lib(), f"cblas_{_zpre(dtype)}gemm")(layout, ta, tb, m, n, k, _ptr(al), _ptr(a), lda,
                                                _ptr(b), ldb, _ptr(be), _ptr(c), ldc)


def zgemm_dev(ta, tb, m, n, k, alpha, a, lda, b, ldb, beta, c, ldc, stream=None, dtype=None):
    """hwyblas_{c,z}gemm_dev on device operands; returns the info."""
    dtype = dtype or c.dtype
    al, be = _zscalar(alpha, dtype), _zscalar(beta, dtype)
    return getattr(lib(), f"hwyblas_{_zpre(dtype)}gemm_dev")(
        ta.encode(), tb.encode(), m, n, k, _ptr(al), _ptr(a), lda, _ptr(b), ldb, _ptr(be),
        _ptr(c), ldc, stream)


def zrank_k(routine, uplo, trans, n, k, alpha, a, lda, beta, c, ldc, dtype=None):
    """Fortran {c,z}syrk_ (complex alpha, beta) or {c,z}herk_ (real alpha,
    beta) on numpy complex arrays or DeviceArrays; C in place."""
    dtype = dtype or c.dtype
    if routine == "herk":
        rt = np.float64 if np.dtype(dtype) == np.complex128 else np.float32
        al, be = np.array([alpha], dtype=rt), np.array([beta], dtype=rt)
    else:
        al, be = _zscalar(alpha, dtype), _zscalar(beta, dtype)
    getattr(lib(), f"{_zpre(dtype)}{routine}_")(
        uplo.encode(), trans.encode(), _ip(n), _ip(k), _ptr(al), _ptr(a), _ip(lda), _ptr(be),
        _ptr(c), _ip(ldc))


def zgemv(trans, m, n, alpha, a, lda, x, incx, beta, y, incy, dtype=None):
    """Fortran cgemv_/zgemv_ (numpy complex arrays or DeviceArrays); y in place."""
    dtype = dtype or y.dtype
    al, be = _zscalar(alpha, dtype), _zscalar(beta, dtype)
    getattr(lib(), f"{_zpre(dtype)}gemv_")(
        trans.encode(), _ip(m), _ip(n), _ptr(al), _ptr(a), _ip(lda), _ptr(x), _ip(incx),
        _ptr(be), _ptr(y), _ip(incy))


def cblas_zgemv(layout, trans, m, n, alpha, a, lda, x, incx, beta, y, incy, dtype=None):
    dtype = dtype or y.dtype
    al, be = _zscalar(alpha, dtype), _zscalar(beta, dtype)
    getattr(lib(), f"cblas_{_zpre(dtype)}gemv")(layout, trans, m, n, _ptr(al), _ptr(a), lda,
                                                _ptr(x), incx, _ptr(be), _ptr(y), incy)


def ztrsm(side, uplo, ta, diag, m, n, alpha, a, lda, b, ldb, dtype=None):
    """Fortran ctrsm_/ztrsm_ (numpy complex arrays or DeviceArrays); B in place."""
    dtype = dtype or b.dtype
    al = _zscalar(alpha, dtype)
    getattr(lib(), f"{_zpre(dtype)}trsm_")(
        side.encode(), uplo.encode(), ta.encode(), diag.encode(), _ip(m), _ip(n), _ptr(al),
        _ptr(a), _ip(lda), _ptr(b), _ip(ldb))


def cblas_ztrsm(layout, side, uplo, ta, diag, m, n, alpha, a, lda, b, ldb, dtype=None):
    dtype = dtype or b.dtype
    al = _zscalar(alpha, dtype)
    getattr(lib(), f"cblas_{_zpre(dtype)}trsm")(layout, side, uplo, ta, diag, m, n, _ptr(al),
                                                _ptr(a), lda, _ptr(b), ldb)


def ztrsm_dev(side, uplo, ta, diag, m, n, alpha, a, lda, b, ldb, stream=None, dtype=None):
    dtype = dtype or b.dtype
    al = _zscalar(alpha, dtype)
    return getattr(lib(), f"hwyblas_{_zpre(dtype)}trsm_dev")(
        side.encode(), uplo.encode(), ta.encode(), diag.encode(), m, n, _ptr(al), _ptr(a), lda,
        _ptr(b), ldb, stream)


def getrf(m, n, a, lda, ipiv):
    """Fortran dgetrf_ (host arrays or DeviceArrays); returns info."""
    info = ctypes.c_int(0)
    lib().dgetrf_(_ip(m), _ip(n), _ptr(a), _ip(lda), _ptr(ipiv), ctypes.byref(info))
    return info.value


def getrs(trans, n, nrhs, a, lda, ipiv, b, ldb):
    info = ctypes.c_int(0)
    lib().dgetrs_(trans.encode(), _ip(n), _ip(nrhs), _ptr(a), _ip(lda), _ptr(ipiv), _ptr(b),
                  _ip(ldb), ctypes.byref(info))
    return info.value


def gesv(n, nrhs, a, lda, ipiv, b, ldb):
    info = ctypes.c_int(0)
    lib().dgesv_(_ip(n), _ip(nrhs), _ptr(a), _ip(lda), _ptr(ipiv), _ptr(b), _ip(ldb),
                 ctypes.byref(info))
    return info.value


def geqrf(m, n, a, lda, tau, lwork=None):
    """Fortran dgeqrf_ (host arrays or DeviceArrays for A and tau); returns
    (info, work[0])."""
    info = ctypes.c_int(0)
    work = np.zeros(1)
    lw = max(1, n) if lwork is None else lwork
    lib().dgeqrf_(_ip(m), _ip(n), _ptr(a), _ip(lda), _ptr(tau), _ptr(work), _ip(lw),
                  ctypes.byref(info))
    return info.value, float(work[0])


def geqrf_dev(a, tau, stream=None):
    """hwyblas_dgeqrf_dev on a DeviceArray (m = a.ld rows); tau a DeviceArray."""
    return lib().hwyblas_dgeqrf_dev(a.ld, a.cols, a.ptr, a.ld, tau.ptr, stream)


def gerfs(trans, n, nrhs, a, lda, af, ldaf, ipiv, b, ldb, x, ldx):
    """Fortran dgerfs_ (host arrays or DeviceArrays for the matrices and
    pivots); X refined in place. Returns (info, ferr, berr)."""
    info = ctypes.c_int(0)
    ferr = np.zeros(max(1, nrhs))
    berr = np.zeros(max(1, nrhs))
    work = np.zeros(max(1, 3 * n))
    iwork = np.zeros(max(1, n), dtype=np.int32)
    lib().dgerfs_(trans.encode(), _ip(n), _ip(nrhs), _ptr(a), _ip(lda), _ptr(af), _ip(ldaf),
                  _ptr(ipiv), _ptr(b), _ip(ldb), _ptr(x), _ip(ldx), _ptr(ferr), _ptr(berr),
                  _ptr(work), _ptr(iwork), ctypes.byref(info))
    return info.value, ferr[:nrhs], berr[:nrhs]


def lapacke_lu(routine, layout, *args, work=False):
    """LAPACKE_d{getrf,getrs,gesv}[_work] on host arrays; returns info. Args
    after the layout as in C (chars as str)."""
    f = getattr(lib(), f"LAPACKE_d{routine}{'_work' if work else ''}")
    conv = [_ptr(x) if isinstance(x, np.ndarray) else (x.encode() if isinstance(x, str) else x)
            for x in args]
    return f(layout, *conv)


def lapacke(routine, layout, uplo, *args, dtype=None, work=False):
    """LAPACKE_{s,d}{potrf,potrs,posv}[_work] on host arrays; returns info.
    args after uplo as in C (n, [nrhs,] a, lda, [b, ldb])."""
    dtype = np.dtype(dtype or next(x.dtype for x in args if isinstance(x, np.ndarray)))
    f = getattr(lib(), f"LAPACKE_{_pre(dtype)}{routine}{'_work' if work else ''}")
    return f(layout, uplo.encode(), *[_ptr(x) if isinstance(x, np.ndarray) else x for x in args])


def gemm_dev(ta, tb, m, n, k, alpha, a, lda, b, ldb, beta, c, ldc, stream=None, dtype=None):
    dtype = dtype or c.dtype
    f = getattr(lib(), f"hwyblas_{_pre(dtype)}gemm_dev")
    return f(ta.encode(), tb.encode(), m, n, k, alpha, _ptr(a), lda, _ptr(b), ldb, beta,
             _ptr(c), ldc, stream)


def gemm_dev_cfg(cfg, ta, tb, m, n, k, alpha, a, lda, b, ldb, beta, c, ldc, stream=None,
                 dtype=None):
    dtype = dtype or c.dtype
    f = getattr(lib(), f"hwyblas_{_pre(dtype)}gemm_dev_cfg")
    return f(cfg, ta.encode(), tb.encode(), m, n, k, alpha, _ptr(a), lda, _ptr(b), ldb, beta,
             _ptr(c), ldc, stream)


def dgemm_emulated_dev(ta, tb, m, n, k, alpha, a, lda, b, ldb, beta, c, ldc, nmoduli=0,
                       stream=None, report=True):
    """Emulated DGEMM (INT8 Ozaki II); returns (info, moduli used; 0 = native FP64).

    report=False passes no count pointer: the call then returns without
    synchronizing and the count is None."""
    used = ctypes.c_int(0)
    info = lib().hwyblas_dgemm_emulated_dev(ta.encode(), tb.encode(), m, n, k, alpha, _ptr(a),
                                            lda, _ptr(b), ldb, beta, _ptr(c), ldc, nmoduli,
                                            stream, ctypes.byref(used) if report else None)
    return info, (used.value if report else None)


def dgemm_emulated_noalloc_dev(ta, tb, m, n, k, alpha, a, lda, b, ldb, beta, c, ldc, nmoduli=0,
                               stream=None, report=True):
    """The INT8 emulated DGEMM on the library's fixed workspace (no
    allocation); returns (info, moduli used), info -1 when the problem needs
    more workspace than the library holds."""
    used = ctypes.c_int(0)
    info = lib().hwyblas_dgemm_emulated_noalloc_dev(ta.encode(), tb.encode(), m, n, k, alpha,
                                                    _ptr(a), lda, _ptr(b), ldb, beta, _ptr(c),
                                                    ldc, nmoduli, stream,
                                                    ctypes.byref(used) if report else None)
    return info, (used.value if report else None)


def dgemm_emulated_bf16_dev(ta, tb, m, n, k, alpha, a, lda, b, ldb, beta, c, ldc, slices=0,
                            stream=None, report=True):
    """Emulated DGEMM (BF16 slices, Ozaki I); returns (info, slices used; 0 = native FP64).

    report=False: asynchronous, the count is None."""
    used = ctypes.c_int(0)
    info = lib().hwyblas_dgemm_emulated_bf16_dev(ta.encode(), tb.encode(), m, n, k, alpha,
                                                 _ptr(a), lda, _ptr(b), ldb, beta, _ptr(c), ldc,
                                                 slices, stream,
                                                 ctypes.byref(used) if report else None)
    return info, (used.value if report else None)


def gemm_configs():
    """[(bm, bn, bk, waves)] of the plain GEMM kernel's configurations."""
    out = []
    for i in range(lib().hwyblas_gemm_config_count()):
        v = [ctypes.c_int() for _ in range(4)]
        lib().hwyblas_gemm_config_info(i, *[ctypes.byref(x) for x in v])
        out.append(tuple(x.value for x in v))
    return out


def trsm_dev(side, uplo, ta, diag, m, n, alpha, a, lda, b, ldb, stream=None, dtype=None):
    dtype = dtype or b.dtype
    f = getattr(lib(), f"hwyblas_{_pre(dtype)}trsm_dev")
    return f(side.encode(), uplo.encode(), ta.encode(), diag.encode(), m, n, alpha,
             _ptr(a), lda, _ptr(b), ldb, stream)


class ErrorCapture:
    """Installs an error handler that records (name, position) pairs."""

    def __init__(self):
        self.calls = []
        self._cb = ERROR_HANDLER(lambda name, pos: self.calls.append((name.decode(), pos)))

    def __enter__(self):
        self._prev = lib().hwyblas_set_error_handler(ctypes.cast(self._cb, ctypes.c_void_p))
        return self

    def __exit__(self, *exc):
        lib().hwyblas_set_error_handler(self._prev)
        return False


# ---------------------------------------------------------------------------
# Device memory

class DeviceArray:
    """A column-major device matrix (ld x cols) owned by this object."""

    def __init__(self, ld, cols, dtype=np.float64):
        self.ld, self.cols, self.dtype = int(ld), int(cols), np.dtype(dtype)
        self.nbytes = max(self.ld * self.cols * self.dtype.itemsize, 1)
        self.ptr = lib().hwyblas_amd_malloc(self.nbytes)

    @classmethod
    def from_host(cls, x):
        x = np.asfortranarray(x)
        d = cls(x.shape[0], x.shape[1] if x.ndim > 1 else 1, x.dtype)
        lib().hwyblas_amd_memcpy(d.ptr, x.ctypes.data, x.nbytes, 1)
        return d

    def to_host(self):
        out = np.empty((self.ld, self.cols), dtype=self.dtype, order="F")
        lib().hwyblas_amd_memcpy(out.ctypes.data, self.ptr, out.nbytes, 2)
        return out

    def fill_uniform(self, rows, cols, seed, lo=-1.0, hi=1.0, stream=None):
        lib().hwyblas_amd_fill_uniform(self.ptr, self.dtype.itemsize, rows, cols, self.ld,
                                       seed, lo, hi, stream)

    def free(self):
        if self.ptr:
            lib().hwyblas_amd_free(self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class Event:
    def __init__(self):
        self.ev = lib().hwyblas_amd_event_create()

    def record(self, stream=None):
        lib().hwyblas_amd_event_record(self.ev, stream)

    def elapsed_ms(self, later):
        return lib().hwyblas_amd_event_elapsed_ms(self.ev, later.ev)

    def __del__(self):
        try:
            lib().hwyblas_amd_event_destroy(self.ev)
        except Exception:
            pass


def synchronize():
    lib().hwyblas_amd_synchronize()


def device_count():
    return lib().hwyblas_amd_device_count()


# ---------------------------------------------------------------------------
# Batched (grouped) plans

class GemmProblem(ctypes.Structure):
    """hwyblas_gemm_problem (include/hwyblas_amd.h)."""
    _fields_ = [("transa", ctypes.c_char), ("transb", ctypes.c_char), ("m", _i), ("n", _i),
                ("k", _i), ("alpha", _d), ("beta", _d), ("a", _p), ("lda", _i), ("b", _p),
                ("ldb", _i), ("c", _p), ("ldc", _i)]


class TrsmProblem(ctypes.Structure):
    """hwyblas_trsm_problem (include/hwyblas_amd.h)."""
    _fields_ = [("side", ctypes.c_char), ("uplo", ctypes.c_char), ("transa", ctypes.c_char),
                ("diag", ctypes.c_char), ("m", _i), ("n", _i), ("alpha", _d), ("a", _p),
                ("lda", _i), ("b", _p), ("ldb", _i)]


class PlanError(ValueError):
    def __init__(self, info, index):
        super().__init__(f"problem {index}: invalid argument {info}")
        self.info, self.index = info, index


class Plan:
    """A grouped GEMM or TRSM plan over device-resident problems."""

    def __init__(self, kind, dtype, problems, configs=None):
        self.kind, self.dtype = kind, np.dtype(dtype)
        self.count = len(problems)
        arr = (GemmProblem if kind == "gemm" else TrsmProblem) * max(1, self.count)
        self._problems = arr(*problems)
        info, bad = ctypes.c_int(0), ctypes.c_int(-1)
        if kind == "gemm":
            cfg = None
            if configs is not None:
                self._cfg = np.ascontiguousarray(configs, dtype=np.int32)
                cfg = self._cfg.ctypes.data_as(ctypes.c_void_p)
            self.ptr = lib().hwyblas_gemm_plan_create(self.dtype.itemsize, self.count,
                                                      self._problems, cfg, ctypes.byref(info),
                                                      ctypes.byref(bad))
        else:
            self.ptr = lib().hwyblas_trsm_plan_create(self.dtype.itemsize, self.count,
                                                      self._problems, ctypes.byref(info),
                                                      ctypes.byref(bad))
        if not self.ptr:
            raise PlanError(info.value, bad.value)

    def launch(self, stream=None, graph=False):
        f = lib().hwyblas_plan_launch_graph if graph else lib().hwyblas_plan_launch
        f(self.ptr, stream)

    @property
    def work_items(self):
        return lib().hwyblas_plan_work_items(self.ptr)

    def problem_config(self, i):
        return lib().hwyblas_plan_problem_config(self.ptr, i)

    def destroy(self):
        if getattr(self, "ptr", None):
            lib().hwyblas_plan_destroy(self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.destroy()
        except Exception:
            pass


def grouped_configs():
    """[(bm, bn, unroll, block)] of the grouped GEMM tile configurations."""
    out = []
    for i in range(lib().hwyblas_grouped_config_count()):
        v = [ctypes.c_int() for _ in range(4)]
        lib().hwyblas_grouped_config_info(i, *[ctypes.byref(x) for x in v])
        out.append(tuple(x.value for x in v))
    return out


class Arena:
    """One device allocation holding many column-major matrices (offsets in
    elements), filled from host arrays with one copy."""

    def __init__(self, host_blocks, dtype):
        self.dtype = np.dtype(dtype)
        self.offsets, total = [], 0
        for b in host_blocks:
            self.offsets.append(total)
            total += (b.size + 31) // 32 * 32
        self.total = max(total, 1)
        host = np.zeros(self.total, dtype=self.dtype)
        for off, b in zip(self.offsets, host_blocks):
            host[off:off + b.size] = np.asfortranarray(b).ravel(order="F")
        self.dev = DeviceArray(self.total, 1, self.dtype)
        lib().hwyblas_amd_memcpy(self.dev.ptr, host.ctypes.data, host.nbytes, 1)

    def ptr(self, i):
        return self.dev.ptr + self.offsets[i] * self.dtype.itemsize

    def download(self):
        out = np.empty(self.total, dtype=self.dtype)
        lib().hwyblas_amd_memcpy(out.ctypes.data, self.dev.ptr, out.nbytes, 2)
        return out

    def block(self, flat, i, shape):
        off = self.offsets[i]
        return flat[off:off + shape[0] * shape[1]].reshape(shape, order="F")


# ---------------------------------------------------------------------------
# LU consumer

def getrf_dev(a, update=0, stream=None):
    """LU of the device matrix `a` (DeviceArray, m = a.ld rows unless given)
    in place; returns (ipiv DeviceArray of int32, info)."""
    m, n = a.ld, a.cols
    ipiv = DeviceArray(max(1, min(m, n)), 1, np.int32)
    info = lib().hwyblas_dgetrf_dev(m, n, a.ptr, a.ld, ipiv.ptr, update, stream)
    return ipiv, info


def gesv_ir_dev(a, b, update=0, itmax=5, stream=None, with_ferr=False, factors=False):
    """Mixed-precision solve of A X = B (device arrays); returns
    (X DeviceArray, berr, nbe, iterations, info), with ferr (the forward
    error bound per column) appended when with_ferr, and the factors
    (af DeviceArray, ipiv DeviceArray) appended when factors."""
    n, nrhs = a.ld, b.cols
    af = DeviceArray(n, n)
    x = DeviceArray(n, nrhs)
    ipiv = DeviceArray(max(1, n), 1, np.int32)
    ferr = np.zeros(max(1, nrhs))
    berr = np.zeros(max(1, nrhs))
    nbe = np.zeros(max(1, nrhs))
    iters = np.zeros(max(1, nrhs), dtype=np.int32)
    info = lib().hwyblas_dgesv_ir_dev(n, nrhs, a.ptr, a.ld, af.ptr, n, ipiv.ptr, b.ptr, b.ld,
                                      x.ptr, n, update, itmax,
                                      ferr.ctypes.data if with_ferr else None, berr.ctypes.data,
                                      nbe.ctypes.data, iters.ctypes.data, stream)
    out = (x, berr[:nrhs], nbe[:nrhs], iters[:nrhs], info)
    if with_ferr:
        out += (ferr[:nrhs],)
    if factors:
        out += (af, ipiv)
    return out

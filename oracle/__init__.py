"""ctypes binding of the CPU oracle (oracle/liboracle.so).

TEST INFRASTRUCTURE ONLY. The oracle is a single-threaded restatement of
hwyBLAS's Level-3 algorithms plus double-double references (see oracle.h).
Only the tests, the smoke check and the CPU-baseline leg of bench.py import
this module, as the checker; the product library never does.
"""
import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle.so")
_lib = None

_c_char = ctypes.c_char
_i = ctypes.c_int
_d = ctypes.c_double
_f = ctypes.c_float
_p = ctypes.c_void_p


def build(force=False):
    """Compile liboracle.so with the Makefile beside this file."""
    if force or not os.path.exists(_LIB_PATH):
        subprocess.run(["make", "-B" if force else "-s", "-C", _HERE],
                       check=True, stdout=subprocess.DEVNULL)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = ctypes.CDLL(_LIB_PATH)
        L.orc_splitmix64.restype = ctypes.c_uint64
        L.orc_splitmix64.argtypes = [ctypes.c_uint64]
        L.orc_uniform01.restype = _d
        L.orc_uniform01.argtypes = [ctypes.c_uint64, ctypes.c_uint64]
        for name, real in (("orc_fill_uniform_f64", None), ("orc_fill_uniform_f32", None)):
            fn = getattr(L, name)
            fn.restype = None
            fn.argtypes = [_p, ctypes.c_int64, ctypes.c_int64, ctypes.c_int64,
                           ctypes.c_uint64, _d, _d]
        for pre, s in (("d", _d), ("s", _f)):
            g = getattr(L, f"orc_{pre}gemm")
            g.restype = None
            g.argtypes = [_c_char, _c_char, _i, _i, _i, s, _p, _i, _p, _i, s, _p, _i]
            t = getattr(L, f"orc_{pre}trsm")
            t.restype = None
            t.argtypes = [_c_char, _c_char, _c_char, _c_char, _i, _i, s, _p, _i, _p, _i]
            r = getattr(L, f"orc_{pre}gemm_ref")
            r.restype = None
            r.argtypes = [_c_char, _c_char, _i, _i, _i, s, _p, _i, _p, _i, s, _p, _i, _p, _p]
            e = getattr(L, f"orc_{pre}gemm_ref_entries")
            e.restype = None
            e.argtypes = [_c_char, _c_char, _i, s, _p, _i, _p, _i, s, _p, _i, _p, _p, _i, _p, _p]
            q = getattr(L, f"orc_{pre}trsm_residual")
            q.restype = None
            q.argtypes = [_c_char, _c_char, _c_char, _c_char, _i, _i, s, _p, _i, _p, _i,
                          _p, _i, _p, _i, _p]
        L.orc_ozaki_modulus.restype = _i
        L.orc_ozaki_modulus.argtypes = [_i]
        L.orc_ozaki_total_bits.restype = _i
        L.orc_ozaki_total_bits.argtypes = [_i, ctypes.c_int64]
        L.orc_ozaki_crt.restype = _d
        L.orc_ozaki_crt.argtypes = [_i, _p]
        L.orc_dgemm_ozaki.restype = None
        L.orc_dgemm_ozaki.argtypes = [_c_char, _c_char, _i, _i, _i, _d, _p, _i, _p, _i, _d, _p,
                                      _i, _i]
        L.orc_dgemm_ozaki_bf16.restype = None
        L.orc_dgemm_ozaki_bf16.argtypes = [_c_char, _c_char, _i, _i, _i, _d, _p, _i, _p, _i,
                                           _d, _p, _i, _i]
        L.orc_bf16_digits.restype = None
        L.orc_bf16_digits.argtypes = [_d, _i, _i, _p]
        L.orc_dgeqrf.restype = _i
        L.orc_dgeqrf.argtypes = [_i, _i, _p, _i, _p]
        L.orc_dgeqr2.restype = None
        L.orc_dgeqr2.argtypes = [_i, _i, _p, _i, _p]
        L.orc_dlarft.restype = None
        L.orc_dlarft.argtypes = [_i, _i, _p, _i, _p, _p, _i]
        L.orc_emulated_choose.restype = _i
        L.orc_emulated_choose.argtypes = [_i, _i, _c_char, _c_char, _i, _i, _i, _p, _i, _p, _i,
                                          _p, _p]
        L.orc_emulated_bound.restype = None
        L.orc_emulated_bound.argtypes = [_i, _i, ctypes.c_int64, _d, _i, _p, _p]
        L.orc_emulated_coarse.restype = None
        L.orc_emulated_coarse.argtypes = [_c_char, _c_char, _i, _i, _i, _p, _i, _p, _i, _p, _p,
                                          _p]
        L.orc_fill_pow2.restype = None
        L.orc_fill_pow2.argtypes = [_p, ctypes.c_int64, ctypes.c_int64, ctypes.c_int64,
                                    ctypes.c_uint64, _i, _i]
        L.orc_dgetrf.restype = _i
        L.orc_dgetrf.argtypes = [_i, _i, _p, _i, _p]
        L.orc_dlaswp.restype = None
        L.orc_dlaswp.argtypes = [_i, _p, _i, _i, _i, _p, _i]
        L.orc_dgetrs.restype = None
        L.orc_dgetrs.argtypes = [_c_char, _i, _i, _p, _i, _p, _p, _i]
        L.orc_dgerfs_berr.restype = None
        L.orc_dgerfs_berr.argtypes = [_i, _i, _p, _i, _p, _i, _p, _i, _p]
        L.orc_dlacn2.restype = None
        L.orc_dlacn2.argtypes = [_i, _p, _p, _p, _p, _p, _p]
        L.orc_dpotrf2.restype = _i
        L.orc_dpotrf2.argtypes = [_c_char, _i, _p, _i]
        L.orc_dgerfs.restype = _i
        L.orc_dgerfs.argtypes = [_c_char, _i, _i, _p, _i, _p, _i, _p, _p, _i, _p, _i, _p, _p]
        L.orc_plant_mistake.restype = None
        L.orc_plant_mistake.argtypes = [_i]
        _lib = L
    return _lib


def _ptr(x):
    return None if x is None else x.ctypes.data_as(ctypes.c_void_p)


def _b(c):
    return c.encode() if isinstance(c, str) else c


def _pre(dtype):
    return "d" if np.dtype(dtype) == np.float64 else "s"


# ---- generator ----------------------------------------------------------

def uniform01(seed, index):
    return lib().orc_uniform01(seed, index)


def fill_uniform(rows, cols, seed, lo=-1.0, hi=1.0, dtype=np.float64, ld=None):
    """Column-major rows x cols matrix (Fortran-ordered array of shape
    (ld, cols)) filled by the generator shared with the device fill."""
    ld = rows if ld is None else ld
    x = np.zeros((max(ld, 1), cols), dtype=dtype, order="F")
    fn = lib().orc_fill_uniform_f64 if np.dtype(dtype) == np.float64 else lib().orc_fill_uniform_f32
    fn(_ptr(x), rows, cols, ld, seed, lo, hi)
    return x


# ---- restatement ------------------------------------------------------------

def gemm(ta, tb, m, n, k, alpha, a, lda, b, ldb, beta, c, ldc):
    """In-place C update with hwyBLAS's blocked algorithm (Fortran args)."""
    fn = getattr(lib(), f"orc_{_pre(c.dtype)}gemm")
    fn(_b(ta), _b(tb), m, n, k, alpha, _ptr(a), lda, _ptr(b), ldb, beta, _ptr(c), ldc)


def trsm(side, uplo, ta, diag, m, n, alpha, a, lda, b, ldb):
    fn = getattr(lib(), f"orc_{_pre(b.dtype)}trsm")
    fn(_b(side), _b(uplo), _b(ta), _b(diag), m, n, alpha, _ptr(a), lda, _ptr(b), ldb)


# ---- references ---------------------------------------------------------

def gemm_ref(ta, tb, m, n, k, alpha, a, lda, b, ldb, beta, c, ldc):
    """Double-double expected C (compact m x n, float64) and error scale."""
    exp = np.zeros((max(m, 1), n), dtype=np.float64, order="F")
    bnd = np.zeros((max(m, 1), n), dtype=np.float64, order="F")
    fn = getattr(lib(), f"orc_{_pre(a.dtype)}gemm_ref")
    fn(_b(ta), _b(tb), m, n, k, alpha, _ptr(a), lda, _ptr(b), ldb, beta,
       _ptr(c) if c is not None else None, ldc, _ptr(exp), _ptr(bnd))
    return exp[:m], bnd[:m]


def _threads():
    return max(1, min(16, os.cpu_count() or 1))


def _in_parallel(count, work, per_item_cost):
    """Runs work(lo, hi) over [0, count) in slices on a thread pool (the
    oracle's C functions release the GIL through ctypes) when the job is
    large, else in one call. Results do not depend on the split."""
    threads = _threads()
    if count == 0:
        return
    if threads == 1 or count * per_item_cost < 5e7:
        work(0, count)
        return
    from concurrent.futures import ThreadPoolExecutor
    step = (count + 4 * threads - 1) // (4 * threads)
    with ThreadPoolExecutor(threads) as ex:
        list(ex.map(lambda lo: work(lo, min(count, lo + step)), range(0, count, step)))


def gemm_ref_entries(ta, tb, k, alpha, a, lda, b, ldb, beta, c, ldc, rows, cols):
    rows = np.ascontiguousarray(rows, dtype=np.int32)
    cols = np.ascontiguousarray(cols, dtype=np.int32)
    exp = np.zeros(len(rows), dtype=np.float64)
    bnd = np.zeros(len(rows), dtype=np.float64)
    fn = getattr(lib(), f"orc_{_pre(a.dtype)}gemm_ref_entries")
    cptr = _ptr(c) if c is not None else None

    def work(lo, hi):
        fn(_b(ta), _b(tb), k, alpha, _ptr(a), lda, _ptr(b), ldb, beta, cptr, ldc,
           _ptr(rows[lo:]), _ptr(cols[lo:]), hi - lo, _ptr(exp[lo:]), _ptr(bnd[lo:]))
    _in_parallel(len(rows), work, 20 * max(k, 1))
    return exp, bnd


def trsm_residual(side, uplo, ta, diag, m, n, alpha, a, lda, b0, ldb, x, ldx, which):
    which = np.ascontiguousarray(which, dtype=np.int32)
    ratio = np.zeros(len(which), dtype=np.float64)
    fn = getattr(lib(), f"orc_{_pre(b0.dtype)}trsm_residual")
    ka = m if side in "Ll" else n
    other = m if side in "Ll" else n

    def work(lo, hi):
        fn(_b(side), _b(uplo), _b(ta), _b(diag), m, n, alpha, _ptr(a), lda, _ptr(b0), ldb,
           _ptr(x), ldx, _ptr(which[lo:]), hi - lo, _ptr(ratio[lo:]))
    _in_parallel(len(which), work, 20 * ka * other)
    return ratio


def ozaki_moduli(nmod):
    return [lib().orc_ozaki_modulus(l) for l in range(nmod)]


def ozaki_total_bits(nmod, k):
    return lib().orc_ozaki_total_bits(nmod, k)


def ozaki_crt(nmod, residues):
    r = np.ascontiguousarray(residues, dtype=np.int32)
    return lib().orc_ozaki_crt(nmod, _ptr(r))


def dgemm_ozaki(ta, tb, m, n, k, alpha, a, lda, b, ldb, beta, c, ldc, nmod):
    """In-place C update by the CPU restatement of the emulated DGEMM."""
    lib().orc_dgemm_ozaki(_b(ta), _b(tb), m, n, k, alpha, _ptr(a), lda, _ptr(b), ldb, beta,
                          _ptr(c), ldc, nmod)


def dgemm_ozaki_bf16(ta, tb, m, n, k, alpha, a, lda, b, ldb, beta, c, ldc, slices):
    """In-place C update by the CPU restatement of the BF16-slice emulated DGEMM."""
    lib().orc_dgemm_ozaki_bf16(_b(ta), _b(tb), m, n, k, alpha, _ptr(a), lda, _ptr(b), ldb,
                               beta, _ptr(c), ldc, slices)


def geqrf(a, blocked=True):
    """Householder QR of a copy of `a`: (factored matrix, tau), as xGEQRF
    (blocked=True) or xGEQR2 stores them."""
    h = np.array(a, dtype=np.float64, order="F", copy=True)
    m, n = h.shape
    tau = np.zeros(max(1, min(m, n)))
    if blocked:
        info = lib().orc_dgeqrf(m, n, _ptr(h), max(1, m), _ptr(tau))
        assert info == 0
    else:
        lib().orc_dgeqr2(m, n, _ptr(h), max(1, m), _ptr(tau))
    return h, tau[:min(m, n)]


def larft(v, tau):
    """The upper triangular T of the block reflector I - V T V^T (forward,
    columnwise; V unit lower trapezoidal, entries on and above its diagonal
    not read)."""
    v = np.asfortranarray(v, dtype=np.float64)
    n, k = v.shape
    t = np.zeros((k, k), order="F")
    lib().orc_dlarft(n, k, _ptr(v), max(1, n), _ptr(np.ascontiguousarray(tau, dtype=np.float64)),
                     _ptr(t), max(1, k))
    return t


def emulated_choose(kind, ta, tb, m, n, k, a, lda, b, ldb, forced=0):
    """The count the emulated DGEMM picks (kind "int8" or "bf16"; 0 = native
    FP64) and its statistics: (count, X, gmin)."""
    x = ctypes.c_double(0)
    g = ctypes.c_int(0)
    c = lib().orc_emulated_choose(0 if kind == "int8" else 1, forced, _b(ta), _b(tb), m, n, k,
                                  _ptr(a), lda, _ptr(b), ldb, ctypes.byref(x), ctypes.byref(g))
    return c, x.value, g.value


def emulated_bound(kind, count, k, xmax, gmin):
    """(typical, worst) error bound of `count`, in units of |A||B|(i, j)."""
    t = ctypes.c_double(0)
    w = ctypes.c_double(0)
    lib().orc_emulated_bound(0 if kind == "int8" else 1, count, k, xmax, gmin,
                             ctypes.byref(t), ctypes.byref(w))
    return t.value, w.value


def emulated_coarse(ta, tb, m, n, k, a, lda, b, ldb):
    """(lower bound of |A||B| (m x n), upper row sums of |op(A)|, upper column
    sums of |op(B)|) from the coarse statistics."""
    lower = np.zeros((m, n), order="F")
    rs = np.zeros(m)
    cs = np.zeros(n)
    lib().orc_emulated_coarse(_b(ta), _b(tb), m, n, k, _ptr(a), lda, _ptr(b), ldb, _ptr(lower),
                              _ptr(rs), _ptr(cs))
    return lower, rs, cs


def bf16_digits(x, sc, slices):
    """The integer digits d_1..d_s of x 2^sc (|x 2^sc| < 1)."""
    d = np.zeros(slices, dtype=np.int32)
    lib().orc_bf16_digits(float(x), sc, slices, _ptr(d))
    return d


def plant_mistake(which):
    lib().orc_plant_mistake(which)


def getrf(a):
    """LU of a copy of `a` (Fortran order): (lu, ipiv 1-based, info)."""
    lu_ = np.array(a, dtype=np.float64, order="F", copy=True)
    m, n = lu_.shape
    ipiv = np.zeros(max(1, min(m, n)), dtype=np.int32)
    info = lib().orc_dgetrf(m, n, _ptr(lu_), max(1, m), _ptr(ipiv))
    return lu_, ipiv[:min(m, n)], info


def getrs(trans, lu_, ipiv, b):
    x = np.array(b, dtype=np.float64, order="F", copy=True)
    if x.ndim == 1:
        x = x.reshape(-1, 1, order="F")
    n, nrhs = lu_.shape[0], x.shape[1]
    ipiv = np.ascontiguousarray(ipiv, dtype=np.int32)
    lib().orc_dgetrs(_b(trans), n, nrhs, _ptr(lu_), max(1, n), _ptr(ipiv), _ptr(x), max(1, n))
    return x


def gerfs_berr(a, b, x):
    a = np.asfortranarray(a, dtype=np.float64)
    b = np.asfortranarray(b, dtype=np.float64)
    x = np.asfortranarray(x, dtype=np.float64)
    n, nrhs = b.shape
    out = np.zeros(nrhs)
    lib().orc_dgerfs_berr(n, nrhs, _ptr(a), a.shape[0], _ptr(b), b.shape[0], _ptr(x),
                          x.shape[0], _ptr(out))
    return out


def lacn2(n, matvec, rmatvec):
    """Drives the oracle's xLACN2 with the callbacks matvec(x) = M x and
    rmatvec(x) = M^T x; returns (estimate of ||M||_1, v, number of products)."""
    v = np.zeros(max(1, n))
    x = np.zeros(max(1, n))
    isgn = np.zeros(max(1, n), dtype=np.int32)
    est = np.zeros(1)
    kase = np.zeros(1, dtype=np.int32)
    isave = np.zeros(3, dtype=np.int32)
    calls = 0
    while True:
        lib().orc_dlacn2(n, _ptr(v), _ptr(x), _ptr(isgn), _ptr(est), _ptr(kase), _ptr(isave))
        if kase[0] == 0:
            return float(est[0]), v[:n].copy(), calls
        x[:n] = matvec(x[:n].copy()) if kase[0] == 1 else rmatvec(x[:n].copy())
        calls += 1


def gerfs(trans, a, lu_, ipiv, b, x):
    """xGERFS: (refined copy of x, ferr, berr) for op(A) X = B, with the
    oracle's LU (lu_, 1-based ipiv) of A."""
    a = np.asfortranarray(a, dtype=np.float64)
    lu_ = np.asfortranarray(lu_, dtype=np.float64)
    b = np.asfortranarray(b, dtype=np.float64)
    if b.ndim == 1:
        b = b.reshape(-1, 1, order="F")
    xr = np.array(x, dtype=np.float64, order="F", copy=True).reshape(b.shape, order="F")
    n, nrhs = b.shape
    ipiv = np.ascontiguousarray(ipiv, dtype=np.int32)
    ferr, berr = np.zeros(max(1, nrhs)), np.zeros(max(1, nrhs))
    info = lib().orc_dgerfs(_b(trans), n, nrhs, _ptr(a), max(1, n), _ptr(lu_), max(1, n),
                            _ptr(ipiv), _ptr(b), max(1, n), _ptr(xr), max(1, n), _ptr(ferr),
                            _ptr(berr))
    if info != 0:
        raise ValueError(f"gerfs: invalid argument {-info}")
    return xr, ferr[:nrhs], berr[:nrhs]


def potrf2(uplo, a):
    """xPOTRF2 of a copy of `a` (Fortran order): (factored copy, info)."""
    f = np.array(a, dtype=np.float64, order="F", copy=True)
    n = f.shape[0]
    info = lib().orc_dpotrf2(_b(uplo), n, _ptr(f), max(1, n))
    return f, info


def lu_factors(lu_, ipiv):
    """(P, L, U) with A = P L U from the packed factorization."""
    m, n = lu_.shape
    k = min(m, n)
    L = np.tril(lu_[:, :k], -1) + np.eye(m, k)
    U = np.triu(lu_[:k, :])
    perm = np.arange(m)
    for i, p in enumerate(ipiv):
        perm[[i, p - 1]] = perm[[p - 1, i]]
    P = np.eye(m)[:, perm]
    return P, L, U


def fill_pow2(rows, cols, seed, emin=-30, emax=30):
    """+-u 2^e, u uniform in [1, 2), e uniform in [emin, emax] (Fortran order)."""
    x = np.empty((rows, cols), dtype=np.float64, order="F")
    lib().orc_fill_pow2(_ptr(x), rows, cols, rows, seed, emin, emax)
    return x


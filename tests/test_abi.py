"""The C-ABI boundary, on the CPU (no GPU needed): the library loads and
exports every symbol the public header declares; invalid arguments are
reported at hwyBLAS's info positions (Fortran and CBLAS, both layouts) through
the handler or xerbla_/cblas_xerbla, with outputs untouched; quick returns
need no device. Restates tests/gemm_test.cc:191-212 and
tests/cblas_test.cc:1087-1183 of the reference.
"""
import ctypes
import os
import subprocess
import sys

import numpy as np
import pytest

import hwyblas_amd as H

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_library_exports_every_header_symbol():
    names = H.exported_symbols_in_header()
    assert len(names) >= 30
    out = subprocess.run(["nm", "-D", "--defined-only", H.LIB_PATH], capture_output=True,
                         text=True, check=True).stdout
    exported = {line.split()[-1] for line in out.splitlines() if line.strip()}
    missing = [n for n in names if n not in exported]
    assert not missing, missing
    for n in names:
        assert hasattr(H.lib(), n)


def test_only_c_abi_is_exported():
    out = subprocess.run(["nm", "-D", "--defined-only", H.LIB_PATH], capture_output=True,
                         text=True, check=True).stdout
    funcs = [l.split()[-1] for l in out.splitlines() if " T " in l]
    header = set(H.exported_symbols_in_header())
    extra = [f for f in funcs if f not in header and not f.startswith("_")]
    assert not extra, extra


def _gemm_args(**kw):
    args = dict(ta="N", tb="N", m=4, n=3, k=2, lda=4, ldb=2, ldc=4)
    args.update(kw)
    return args


GEMM_CASES = [  # (changes, expected Fortran info)  (src/gemm.cc:89-106)
    (dict(ta="X"), 1), (dict(tb="?"), 2), (dict(m=-1), 3), (dict(n=-1), 4), (dict(k=-1), 5),
    (dict(lda=3), 8), (dict(ta="T", lda=1), 8), (dict(ldb=1), 10), (dict(tb="T", ldb=2), 10),
    (dict(ldc=3), 13), (dict(ta="X", m=-1), 1), (dict(m=-1, ldc=0), 3),
]


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
@pytest.mark.parametrize("change,info", GEMM_CASES)
def test_gemm_invalid_arguments(dtype, change, info):
    g = _gemm_args(**change)
    a = np.ones((8, 8), dtype=dtype, order="F")
    c = np.full((8, 8), 5.0, dtype=dtype, order="F")
    c0 = c.copy()
    with H.ErrorCapture() as cap:
        H.gemm(g["ta"], g["tb"], g["m"], g["n"], g["k"], 1.0, a, g["lda"], a, g["ldb"], 0.0, c,
               g["ldc"])
    name = "DGEMM" if dtype == np.float64 else "SGEMM"
    assert cap.calls == [(name, info)]
    assert np.array_equal(c, c0)


TRSM_CASES = [  # (src/trsm.cc:151-168)
    (dict(side="X"), 1), (dict(uplo="X"), 2), (dict(ta="X"), 3), (dict(diag="X"), 4),
    (dict(m=-1), 5), (dict(n=-1), 6), (dict(lda=3), 9), (dict(side="R", lda=2), 9),
    (dict(ldb=3), 11),
]


@pytest.mark.parametrize("change,info", TRSM_CASES)
def test_trsm_invalid_arguments(change, info):
    t = dict(side="L", uplo="L", ta="N", diag="N", m=4, n=3, lda=4, ldb=4)
    t.update(change)
    a = np.ones((8, 8), order="F")
    b = np.full((8, 8), 2.0, order="F")
    b0 = b.copy()
    with H.ErrorCapture() as cap:
        H.trsm(t["side"], t["uplo"], t["ta"], t["diag"], t["m"], t["n"], 1.0, a, t["lda"], b,
               t["ldb"])
    assert cap.calls == [("DTRSM", info)]
    assert np.array_equal(b, b0)


def test_cblas_error_positions():
    """CBLAS positions count the layout as argument 1; row-major remaps the
    Fortran positions of the swapped call (src/cblas.cc:529-535, :681-684)."""
    a = np.ones((8, 8), order="F")
    c = np.zeros((8, 8), order="F")
    with H.ErrorCapture() as cap:
        H.cblas_gemm(H.ColMajor, H.NoTrans, H.NoTrans, 4, 3, 2, 1.0, a, 3, a, 2, 0.0, c, 4)
        H.cblas_gemm(H.RowMajor, H.NoTrans, H.NoTrans, 4, 3, 2, 1.0, a, 1, a, 3, 0.0, c, 3)
        H.cblas_gemm(H.RowMajor, H.NoTrans, H.NoTrans, 4, 3, 2, 1.0, a, 2, a, 2, 0.0, c, 3)
        H.cblas_gemm(H.RowMajor, H.NoTrans, H.NoTrans, 4, 3, 2, 1.0, a, 2, a, 3, 0.0, c, 2)
        H.cblas_gemm(77, H.NoTrans, H.NoTrans, 4, 3, 2, 1.0, a, 2, a, 3, 0.0, c, 3)
        H.cblas_gemm(H.ColMajor, 9, H.NoTrans, 4, 3, 2, 1.0, a, 4, a, 2, 0.0, c, 4)
        H.cblas_gemm(H.RowMajor, 9, H.NoTrans, 4, 3, 2, 1.0, a, 2, a, 3, 0.0, c, 3)
        H.cblas_gemm(H.RowMajor, H.NoTrans, H.NoTrans, -1, 3, 2, 1.0, a, 2, a, 3, 0.0, c, 3)
        H.cblas_trsm(H.ColMajor, H.Left, H.Lower, H.NoTrans, H.NonUnit, 4, 3, 1.0, a, 3, c, 4)
        H.cblas_trsm(H.RowMajor, H.Left, H.Lower, H.NoTrans, H.NonUnit, 4, 3, 1.0, a, 4, c, 2)
        H.cblas_trsm(H.RowMajor, H.Left, H.Lower, H.NoTrans, H.NonUnit, -2, 3, 1.0, a, 4, c, 3)
        H.cblas_trsm(H.RowMajor, 5, H.Lower, H.NoTrans, H.NonUnit, 4, 3, 1.0, a, 4, c, 3)
    assert cap.calls == [
        ("cblas_dgemm", 9),    # lda < m
        ("cblas_dgemm", 9),    # row-major: lda < k (A becomes the Fortran B, info 10 -> 9)
        ("cblas_dgemm", 11),   # row-major: ldb < n (Fortran A, info 8 -> 11)
        ("cblas_dgemm", 14),   # row-major: ldc < n
        ("cblas_dgemm", 1),    # invalid layout
        ("cblas_dgemm", 2),    # transa
        ("cblas_dgemm", 2),    # row-major transa is Fortran transb (info 2 -> 2)
        ("cblas_dgemm", 4),    # row-major m < 0 is Fortran n (info 4 -> 4)
        ("cblas_dtrsm", 10),   # lda < m
        ("cblas_dtrsm", 12),   # row-major ldb < n
        ("cblas_dtrsm", 6),    # row-major m < 0 (Fortran n, info 6 -> 6)
        ("cblas_dtrsm", 2),    # side
    ]
    assert np.all(c == 0)


def test_default_reporters_print_and_return():
    """Without a handler, xerbla_ / cblas_xerbla print the reference's messages
    (src/common.cc:71-79, src/cblas.cc:714-726) and the process continues."""
    code = (
        "import numpy as np, hwyblas_amd as H\n"
        "a = np.ones((4, 4), order='F')\n"
        "H.gemm('N', 'N', 2, 2, 2, 1.0, a, 1, a, 4, 0.0, a, 4)\n"
        "H.cblas_trsm(H.ColMajor, H.Left, 0, H.NoTrans, H.NonUnit, 2, 2, 1.0, a, 4, a, 4)\n"
        "print('still running')\n")
    r = subprocess.run([sys.executable, "-c", code], capture_output=True, text=True, cwd=ROOT)
    assert r.returncode == 0, r.stderr
    assert " ** On entry to DGEMM parameter number 8 had an illegal value" in r.stderr
    assert "Parameter 3 to routine cblas_dtrsm was incorrect" in r.stderr
    assert "still running" in r.stdout


def test_quick_returns_need_no_device():
    """m == 0 or n == 0, and (k == 0 or alpha == 0) with beta == 1, return
    before any device work (src/gemm.cc:114-116); TRSM likewise for m or
    n == 0 (src/trsm.cc:171)."""
    a = np.full((4, 4), np.nan, order="F")
    c = np.full((4, 4), 3.0, order="F")
    H.gemm("N", "N", 0, 4, 4, 1.0, a, 4, a, 4, 0.0, c, 4)
    H.gemm("N", "N", 4, 0, 4, 1.0, a, 4, a, 4, 0.0, c, 4)
    H.gemm("N", "N", 4, 4, 0, 1.0, a, 4, a, 4, 1.0, c, 4)
    H.gemm("N", "N", 4, 4, 4, 0.0, a, 4, a, 4, 1.0, c, 4)
    H.trsm("L", "U", "T", "U", 0, 4, 2.0, a, 4, c, 4)
    H.trsm("R", "U", "T", "U", 4, 0, 2.0, a, 4, c, 4)
    assert np.all(c == 3.0)


def test_handler_is_replaced_and_restored():
    calls = []
    cb = H.ERROR_HANDLER(lambda n, p: calls.append((n.decode(), p)))
    prev = H.lib().hwyblas_set_error_handler(ctypes.cast(cb, ctypes.c_void_p))
    a = np.ones((4, 4), order="F")
    H.gemm("N", "N", -1, 2, 2, 1.0, a, 4, a, 4, 0.0, a, 4)
    back = H.lib().hwyblas_set_error_handler(prev)
    assert back == ctypes.cast(cb, ctypes.c_void_p).value
    assert calls == [("DGEMM", 3)]


def test_version_and_device_count_without_gpu():
    assert b"gfx950" in H.lib().hwyblas_amd_version()
    assert H.device_count() >= 0


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
@pytest.mark.parametrize("change,info", GEMM_CASES)
def test_grouped_gemm_plan_rejects_invalid_problem(dtype, change, info):
    """A plan is refused, with the problem's index and the single-call info
    position, before any device work; the handler sees ?GEMM and the info."""
    g = _gemm_args(**change)
    good = H.GemmProblem(b"N", b"N", 4, 3, 2, 1.0, 0.0, None, 4, None, 2, None, 4)
    bad = H.GemmProblem(g["ta"].encode(), g["tb"].encode(), g["m"], g["n"], g["k"], 1.0, 0.0,
                        None, g["lda"], None, g["ldb"], None, g["ldc"])
    with H.ErrorCapture() as cap:
        with pytest.raises(H.PlanError) as e:
            H.Plan("gemm", dtype, [good, good, bad, good])
    assert (e.value.info, e.value.index) == (info, 2)
    assert cap.calls == [("DGEMM" if dtype == np.float64 else "SGEMM", info)]


@pytest.mark.parametrize("change,info", TRSM_CASES)
def test_grouped_trsm_plan_rejects_invalid_problem(change, info):
    t = dict(side="L", uplo="L", ta="N", diag="N", m=4, n=3, lda=4, ldb=4)
    t.update(change)
    bad = H.TrsmProblem(t["side"].encode(), t["uplo"].encode(), t["ta"].encode(),
                        t["diag"].encode(), t["m"], t["n"], 1.0, None, t["lda"], None, t["ldb"])
    with H.ErrorCapture() as cap:
        with pytest.raises(H.PlanError) as e:
            H.Plan("trsm", np.float64, [bad])
    assert (e.value.info, e.value.index) == (info, 0)
    assert cap.calls == [("DTRSM", info)]


def test_grouped_config_table():
    cfgs = H.grouped_configs()
    assert len(cfgs) >= 8
    for bm, bn, uk, block in cfgs:
        assert bm % 16 == 0 and bn % 16 == 0 and uk >= 1 and block in (0, 1)
    assert any(c[3] for c in cfgs) and not all(c[3] for c in cfgs)
    assert H.lib().hwyblas_grouped_config_info(len(cfgs), None, None, None, None) == -1
    for m, n, k in [(8, 8, 8), (100, 70, 300), (512, 512, 512), (1, 512, 1)]:
        assert 0 <= H.lib().hwyblas_grouped_default_config(8, m, n, k) < len(cfgs)


def test_lu_entry_points_reject_invalid_arguments():
    """Argument checks come before any device work (positions as LAPACK's)."""
    with H.ErrorCapture() as cap:
        assert H.lib().hwyblas_dgetrf_dev(-1, 4, None, 4, None, 0, None) == -1
        assert H.lib().hwyblas_dgetrf_dev(4, 4, None, 3, None, 0, None) == -4
        assert H.lib().hwyblas_dgetrf_dev(4, 4, None, 4, None, 7, None) == -6
        assert H.lib().hwyblas_dgetrs_dev(b"X", 4, 1, None, 4, None, None, 4, None) == -1
        assert H.lib().hwyblas_dgetrs_dev(b"N", 4, 1, None, 4, None, None, 3, None) == -8
        assert H.lib().hwyblas_dgesv_ir_dev(4, 17, None, 4, None, 4, None, None, 4, None, 4, 0,
                                            5, None, None, None, None, None) == -2
    assert cap.calls == [("DGETRF", 1), ("DGETRF", 4), ("DGETRF", 6), ("DGETRS", 1),
                         ("DGETRS", 8), ("DGESVIR", 2)]
    # Empty problems return at once.
    assert H.lib().hwyblas_dgetrf_dev(0, 5, None, 1, None, 0, None) == 0


# SYMM, SYRK, SYR2K, GEMMTR/GEMMT, TRMM: the info position of each invalid
# argument, in hwyBLAS's order (src/level3.cc), reported before any device
# work (no GPU needed). Valid defaults are overridden one at a time.
LEVEL3_CASES = [
    ("symm", dict(side="X"), 1), ("symm", dict(uplo="X"), 2), ("symm", dict(m=-1), 3),
    ("symm", dict(n=-1), 4), ("symm", dict(lda=3), 7), ("symm", dict(ldb=3), 9),
    ("symm", dict(ldc=3), 12), ("symm", dict(side="R", n=5, lda=4), 7),
    ("syrk", dict(uplo="X"), 1), ("syrk", dict(trans="X"), 2), ("syrk", dict(n=-1), 3),
    ("syrk", dict(k=-1), 4), ("syrk", dict(lda=3), 7), ("syrk", dict(trans="T", lda=1), 7),
    ("syrk", dict(ldc=3), 10),
    ("syr2k", dict(uplo="X"), 1), ("syr2k", dict(trans="X"), 2), ("syr2k", dict(n=-1), 3),
    ("syr2k", dict(k=-1), 4), ("syr2k", dict(lda=3), 7), ("syr2k", dict(ldb=3), 9),
    ("syr2k", dict(ldc=3), 12),
    ("gemmtr", dict(uplo="X"), 1), ("gemmtr", dict(ta="X"), 2), ("gemmtr", dict(tb="X"), 3),
    ("gemmtr", dict(n=-1), 4), ("gemmtr", dict(k=-1), 5), ("gemmtr", dict(lda=3), 8),
    ("gemmtr", dict(ldb=1), 10), ("gemmtr", dict(ldc=3), 13),
    ("trmm", dict(side="X"), 1), ("trmm", dict(uplo="X"), 2), ("trmm", dict(ta="X"), 3),
    ("trmm", dict(diag="X"), 4), ("trmm", dict(m=-1), 5), ("trmm", dict(n=-1), 6),
    ("trmm", dict(lda=3), 9), ("trmm", dict(ldb=3), 11),
]


@pytest.mark.parametrize("routine,change,info", LEVEL3_CASES)
@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_level3_invalid_arguments(routine, change, info, dtype):
    """m = n = 4, k = 2: every operand's leading dimension valid at 4 (B of
    GEMMTR N x N at 2), then one argument made invalid."""
    g = dict(side="L", uplo="L", trans="N", ta="N", tb="N", diag="N", m=4, n=4, k=2, lda=4,
             ldb=4, ldc=4)
    if routine == "gemmtr":
        g["ldb"] = 2
    g.update(change)
    x = np.zeros(64, dtype=dtype)
    with H.ErrorCapture() as cap:
        if routine == "symm":
            H.symm(g["side"], g["uplo"], g["m"], g["n"], 1.0, x, g["lda"], x, g["ldb"], 1.0, x,
                   g["ldc"], dtype=dtype)
        elif routine == "syrk":
            H.syrk(g["uplo"], g["trans"], g["n"], g["k"], 1.0, x, g["lda"], 1.0, x, g["ldc"],
                   dtype=dtype)
        elif routine == "syr2k":
            H.syr2k(g["uplo"], g["trans"], g["n"], g["k"], 1.0, x, g["lda"], x, g["ldb"], 1.0,
                    x, g["ldc"], dtype=dtype)
        elif routine == "gemmtr":
            H.gemmtr(g["uplo"], g["ta"], g["tb"], g["n"], g["k"], 1.0, x, g["lda"], x, g["ldb"],
                     1.0, x, g["ldc"], dtype=dtype)
        else:
            H.trmm(g["side"], g["uplo"], g["ta"], g["diag"], g["m"], g["n"], 1.0, x, g["lda"], x,
                   g["ldb"], dtype=dtype)
    name = ("D" if dtype == np.float64 else "S") + routine.upper()
    assert cap.calls == [(name, info)]
    assert np.all(x == 0)


def test_gemmt_reports_its_own_name():
    with H.ErrorCapture() as cap:
        H.gemmtr("X", "N", "N", 4, 2, 1.0, None, 4, None, 2, 1.0, None, 4, dtype=np.float64,
                 name="gemmt")
    assert cap.calls == [("DGEMMT", 1)]


@pytest.mark.parametrize("layout,routine,args,position", [
    # Row-major SYMM swaps m and n: an invalid m is still CBLAS position 4.
    ("row", "symm", dict(m=-1), 4), ("col", "symm", dict(m=-1), 4),
    ("row", "symm", dict(n=-1), 5), ("col", "symm", dict(ldc=1), 13),
    ("row", "syrk", dict(k=-1), 5), ("col", "syrk", dict(ldc=1), 11),
    ("row", "syr2k", dict(ldb=1), 10),
    # Row-major GEMMTR swaps the operands: transb is still CBLAS position 4.
    ("row", "gemmtr", dict(tb=0), 4), ("row", "gemmtr", dict(ldb=1), 11),
    ("col", "gemmtr", dict(lda=1), 9),
    ("row", "trmm", dict(m=-1), 6), ("row", "trmm", dict(n=-1), 7),
    ("col", "trmm", dict(diag=0), 5),
])
def test_cblas_level3_error_positions(layout, routine, args, position):
    g = dict(side=H.Left, uplo=H.Lower, trans=H.NoTrans, ta=H.NoTrans, tb=H.NoTrans,
             diag=H.NonUnit, m=4, n=4, k=4, lda=4, ldb=4, ldc=4)
    g.update(args)
    lay = H.RowMajor if layout == "row" else H.ColMajor
    with H.ErrorCapture() as cap:
        if routine == "symm":
            H.cblas_level3("symm", lay, g["side"], g["uplo"], g["m"], g["n"], 1.0, None,
                           g["lda"], None, g["ldb"], 1.0, None, g["ldc"], dtype=np.float64)
        elif routine == "syrk":
            H.cblas_level3("syrk", lay, g["uplo"], g["trans"], g["n"], g["k"], 1.0, None,
                           g["lda"], 1.0, None, g["ldc"], dtype=np.float64)
        elif routine == "syr2k":
            H.cblas_level3("syr2k", lay, g["uplo"], g["trans"], g["n"], g["k"], 1.0, None,
                           g["lda"], None, g["ldb"], 1.0, None, g["ldc"], dtype=np.float64)
        elif routine == "gemmtr":
            H.cblas_level3("gemmtr", lay, g["uplo"], g["ta"], g["tb"], g["n"], g["k"], 1.0,
                           None, g["lda"], None, g["ldb"], 1.0, None, g["ldc"], dtype=np.float64)
        else:
            H.cblas_level3("trmm", lay, g["side"], g["uplo"], g["ta"], g["diag"], g["m"],
                           g["n"], 1.0, None, g["lda"], None, g["ldb"], dtype=np.float64)
    assert cap.calls == [(f"cblas_d{routine}", position)]


@pytest.mark.parametrize("change,info", [
    (dict(trans="X"), 1), (dict(m=-1), 2), (dict(n=-1), 3), (dict(lda=2), 6),
    (dict(incx=0), 8), (dict(incy=0), 11)])
@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_gemv_invalid_arguments(change, info, dtype):
    g = dict(trans="N", m=3, n=2, lda=3, incx=1, incy=1)
    g.update(change)
    x = np.zeros(8, dtype=dtype)
    with H.ErrorCapture() as cap:
        H.gemv(g["trans"], g["m"], g["n"], 1.0, x, g["lda"], x, g["incx"], 1.0, x, g["incy"],
               dtype=dtype)
    assert cap.calls == [(("D" if dtype == np.float64 else "S") + "GEMV", info)]
    with H.ErrorCapture() as cap:  # CBLAS row-major swaps m and n; positions stay
        H.cblas_level3("gemv", H.RowMajor, H.NoTrans, 3, -1, 1.0, None, 3, None, 1, 1.0, None,
                       1, dtype=dtype)
    assert cap.calls == [(f"cblas_{'d' if dtype == np.float64 else 's'}gemv", 4)]


@pytest.mark.parametrize("routine,change,info", [
    ("potrf", dict(uplo="X"), 1), ("potrf", dict(n=-1), 2), ("potrf", dict(lda=2), 4),
    ("potrs", dict(uplo="X"), 1), ("potrs", dict(n=-1), 2), ("potrs", dict(nrhs=-1), 3),
    ("potrs", dict(lda=2), 5), ("potrs", dict(ldb=2), 7),
    ("posv", dict(uplo="X"), 1), ("posv", dict(nrhs=-1), 3), ("posv", dict(ldb=1), 7)])
@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_cholesky_invalid_arguments(routine, change, info, dtype):
    g = dict(uplo="L", n=3, nrhs=2, lda=3, ldb=3)
    g.update(change)
    x = np.zeros(16, dtype=dtype)
    with H.ErrorCapture() as cap:
        if routine == "potrf":
            got = H.potrf(g["uplo"], g["n"], x, g["lda"], dtype=dtype)
        elif routine == "potrs":
            got = H.potrs(g["uplo"], g["n"], g["nrhs"], x, g["lda"], x, g["ldb"], dtype=dtype)
        else:
            got = H.posv(g["uplo"], g["n"], g["nrhs"], x, g["lda"], x, g["ldb"], dtype=dtype)
    assert got == -info
    assert cap.calls == [(("D" if dtype == np.float64 else "S") + routine.upper(), info)]


ZDTYPES = [np.complex128, np.complex64]
ZGEMM_CASES = GEMM_CASES + [(dict(ta="C", lda=1), 8), (dict(tb="C", ldb=2), 10)]


@pytest.mark.parametrize("dtype", ZDTYPES)
@pytest.mark.parametrize("change,info", ZGEMM_CASES)
def test_zgemm_invalid_arguments(dtype, change, info):
    """cgemm_/zgemm_ check as ?gemm_ (src/gemm.cc:84-106), 'C' accepted."""
    g = _gemm_args(**change)
    a = np.ones((8, 8), dtype=dtype, order="F")
    c = np.full((8, 8), 5.0 + 1.0j, dtype=dtype, order="F")
    c0 = c.copy()
    with H.ErrorCapture() as cap:
        H.zgemm(g["ta"], g["tb"], g["m"], g["n"], g["k"], 1.0, a, g["lda"], a, g["ldb"], 0.0, c,
                g["ldc"])
    name = "ZGEMM" if dtype == np.complex128 else "CGEMM"
    assert cap.calls == [(name, info)]
    assert np.array_equal(c, c0)


@pytest.mark.parametrize("layout,change,pos", [
    (H.ColMajor, dict(m=-1), 4), (H.RowMajor, dict(m=-1), 4), (H.RowMajor, dict(n=-1), 5),
    (H.ColMajor, dict(lda=1), 9), (H.RowMajor, dict(ldb=1), 11), (H.RowMajor, dict(ldc=1), 14),
    (99, dict(), 1)])
def test_cblas_zgemm_positions(layout, change, pos):
    """cblas_zgemm reports CBLAS positions; row-major swaps A/B and m/n
    (src/cblas.cc:517-537)."""
    g = dict(m=4, n=3, k=2, lda=4, ldb=4, ldc=4)
    g.update(change)
    a = np.ones((8, 8), dtype=np.complex128, order="F")
    c = np.full((8, 8), 2.0, dtype=np.complex128, order="F")
    c0 = c.copy()
    with H.ErrorCapture() as cap:
        H.cblas_zgemm(layout, H.NoTrans, H.NoTrans, g["m"], g["n"], g["k"], 1.0, a, g["lda"], a,
                      g["ldb"], 0.0, c, g["ldc"])
    assert cap.calls == [("cblas_zgemm", pos)]
    assert np.array_equal(c, c0)


def test_zgemm_quick_returns_need_no_device():
    """m = 0, n = 0, or (k = 0 or alpha = 0) with beta = 1 return before any
    device work; C is untouched."""
    a = np.ones((4, 4), dtype=np.complex128, order="F")
    c = np.full((4, 4), 3.0 - 2.0j, dtype=np.complex128, order="F")
    c0 = c.copy()
    with H.ErrorCapture() as cap:
        H.zgemm("N", "N", 0, 4, 4, 1.0, a, 4, a, 4, 2.0, c, 4)
        H.zgemm("N", "C", 4, 0, 4, 1.0, a, 4, a, 4, 2.0, c, 4)
        H.zgemm("C", "N", 4, 4, 0, 1.0, a, 4, a, 4, 1.0, c, 4)
        H.zgemm("T", "T", 4, 4, 4, 0.0, a, 4, a, 4, 1.0, c, 4)
    assert cap.calls == []
    assert np.array_equal(c, c0)


ZTRSM_CASES = [  # (src/trsm.cc:151-168), 'C' accepted
    (dict(side="X"), 1), (dict(uplo="X"), 2), (dict(ta="X"), 3), (dict(diag="X"), 4),
    (dict(m=-1), 5), (dict(n=-1), 6), (dict(lda=3), 9), (dict(side="R", lda=2), 9),
    (dict(ldb=3), 11), (dict(ta="C", lda=3), 9),
]


@pytest.mark.parametrize("dtype", ZDTYPES)
@pytest.mark.parametrize("change,info", ZTRSM_CASES)
def test_ztrsm_invalid_arguments(dtype, change, info):
    t = dict(side="L", uplo="L", ta="N", diag="N", m=4, n=3, lda=4, ldb=4)
    t.update(change)
    a = np.ones((8, 8), dtype=dtype, order="F")
    b = np.full((8, 8), 2.0 - 1.0j, dtype=dtype, order="F")
    b0 = b.copy()
    with H.ErrorCapture() as cap:
        H.ztrsm(t["side"], t["uplo"], t["ta"], t["diag"], t["m"], t["n"], 1.0, a, t["lda"], b,
                t["ldb"])
    name = "ZTRSM" if dtype == np.complex128 else "CTRSM"
    assert cap.calls == [(name, info)]
    assert np.array_equal(b, b0)


def test_cblas_ztrsm_row_major_positions():
    """Row-major cblas_ztrsm swaps m and n: a bad ldb is CBLAS position 12,
    m < 0 position 6."""
    a = np.ones((8, 8), dtype=np.complex128, order="F")
    b = np.full((8, 8), 2.0, dtype=np.complex128, order="F")
    with H.ErrorCapture() as cap:
        H.cblas_ztrsm(H.RowMajor, H.Left, H.Lower, H.NoTrans, H.NonUnit, 4, 3, 1.0, a, 4, b, 2)
        H.cblas_ztrsm(H.RowMajor, H.Left, H.Lower, H.NoTrans, H.NonUnit, -1, 3, 1.0, a, 4, b, 3)
    assert cap.calls == [("cblas_ztrsm", 12), ("cblas_ztrsm", 6)]


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_lapacke_checks_before_the_routine(dtype, capfd):
    """LAPACKE layout (-1), NaN (minus the argument's position) and row-major
    leading-dimension checks (src/lapack/lapacke.cc:112-133,
    lapacke_symmetric.cc:22-70, lapacke_util.h:420-445) return before any
    device work; NaN in the unreferenced triangle is not looked at by the
    check."""
    p = "d" if dtype == np.float64 else "s"
    H.lib().LAPACKE_set_nancheck(1)
    a = np.eye(4, dtype=dtype, order="F")
    b = np.ones((4, 2), dtype=dtype, order="F")
    assert H.lapacke("potrf", 99, "L", 4, a, 4) == -1
    assert H.lapacke("posv", 7, "U", 4, 2, a, 4, b, 4) == -1
    out = capfd.readouterr().out
    assert f"Wrong parameter 1 in LAPACKE_{p}potrf" in out
    assert f"Wrong parameter 1 in LAPACKE_{p}posv" in out
    an = a.copy(order="F")
    an[3, 1] = np.nan  # lower triangle, column-major
    assert H.lapacke("potrf", H.ColMajor, "L", 4, an, 4) == -4
    assert H.lapacke("potrs", H.ColMajor, "L", 4, 2, an, 4, b, 4) == -5
    bn = b.copy(order="F")
    bn[2, 1] = np.nan
    assert H.lapacke("posv", H.ColMajor, "U", 4, 2, a, 4, bn, 4) == -7
    assert H.lapacke("potrs", H.RowMajor, "L", 4, 2, a, 4, np.ascontiguousarray(bn), 2) == -7
    # Row-major _work: a leading dimension shorter than the columns.
    assert H.lapacke("potrf", H.RowMajor, "L", 4, a, 3, work=True) == -5
    assert H.lapacke("potrs", H.RowMajor, "L", 4, 2, a, 3, b, 2, work=True) == -6
    assert H.lapacke("posv", H.RowMajor, "L", 4, 2, a, 4, b, 1, work=True) == -8
    out = capfd.readouterr().out
    assert f"Wrong parameter 5 in LAPACKE_{p}potrf_work" in out
    assert f"Wrong parameter 8 in LAPACKE_{p}posv_work" in out


def test_lapacke_nancheck_switch():
    H.lib().LAPACKE_set_nancheck(0)
    assert H.lib().LAPACKE_get_nancheck() == 0
    H.lib().LAPACKE_set_nancheck(5)
    assert H.lib().LAPACKE_get_nancheck() == 1


def test_fortran_lu_argument_checks():
    """dgetrf_/dgetrs_/dgesv_ report LAPACK's positions (src/lapack/getrf.cc,
    :341-367) before any device work; LAPACKE_d{getrf,getrs,gesv} check the
    layout, NaNs and row-major leading dimensions first."""
    a = np.eye(4, order="F")
    b = np.ones((4, 2), order="F")
    piv = np.zeros(4, dtype=np.int32)
    with H.ErrorCapture() as cap:
        assert H.getrf(-1, 4, a, 4, piv) == -1
        assert H.getrf(4, 4, a, 3, piv) == -4
        assert H.getrs("X", 4, 2, a, 4, piv, b, 4) == -1
        assert H.getrs("N", 4, 2, a, 4, piv, b, 3) == -8
        assert H.gesv(4, -1, a, 4, piv, b, 4) == -2
        assert H.gesv(4, 2, a, 4, piv, b, 3) == -7
    assert cap.calls == [("DGETRF", 1), ("DGETRF", 4), ("DGETRS", 1), ("DGETRS", 8),
                         ("DGESV", 2), ("DGESV", 7)]
    # dgerfs_: xGERFS's positions (src/lapack/gesvx.cc:359-376), in order.
    x = np.ones((4, 2), order="F")
    cases = [(("X", 4, 2, 4, 4, 4, 4), -1), (("N", -1, 2, 4, 4, 4, 4), -2),
             (("N", 4, -1, 4, 4, 4, 4), -3), (("T", 4, 2, 3, 4, 4, 4), -5),
             (("N", 4, 2, 4, 3, 4, 4), -7), (("N", 4, 2, 4, 4, 3, 4), -10),
             (("N", 4, 2, 4, 4, 4, 3), -12)]
    with H.ErrorCapture() as cap:
        for (tr, n, nrhs, lda, ldaf, ldb, ldx), want in cases:
            info, _, _ = H.gerfs(tr, n, nrhs, a, lda, a, ldaf, piv, b, ldb, x, ldx)
            assert info == want
        info, ferr, berr = H.gerfs("N", 0, 2, a, 4, a, 4, piv, b, 4, x, 4)
        assert info == 0 and np.all(ferr == 0) and np.all(berr == 0)
    assert cap.calls == [("DGERFS", -w) for _, w in cases]
    # dgeqrf_: xGEQRF's positions (src/lapack/geqrf.cc:82-95) and its query.
    tau = np.zeros(4)
    with H.ErrorCapture() as cap:
        assert H.geqrf(-1, 4, a, 4, tau)[0] == -1
        assert H.geqrf(4, -1, a, 4, tau)[0] == -2
        assert H.geqrf(4, 4, a, 3, tau)[0] == -4
        assert H.geqrf(4, 4, a, 4, tau, lwork=3)[0] == -7
        assert H.geqrf(4, 4, a, 4, tau, lwork=0)[0] == -7
        assert H.geqrf(4, 4, a, 4, tau, lwork=-1) == (0, 4 * 32.0)
        assert H.geqrf(0, 4, a, 1, tau, lwork=0)[0] == -7
        assert H.geqrf(0, 4, a, 1, tau, lwork=1) == (0, 1.0)
    assert cap.calls == [("DGEQRF", 1), ("DGEQRF", 2), ("DGEQRF", 4), ("DGEQRF", 7),
                         ("DGEQRF", 7), ("DGEQRF", 7)]
    H.lib().LAPACKE_set_nancheck(1)
    assert H.lapacke_lu("getrf", 3, 4, 4, a, 4, piv) == -1
    an = a.copy(order="F")
    an[1, 2] = np.nan
    assert H.lapacke_lu("getrf", H.ColMajor, 4, 4, an, 4, piv) == -4
    assert H.lapacke_lu("getrs", H.ColMajor, "N", 4, 2, an, 4, piv, b, 4) == -5
    assert H.lapacke_lu("gesv", H.RowMajor, 4, 2, an, 4, piv, b, 2) == -4
    bn = b.copy(order="F")
    bn[0, 0] = np.nan
    assert H.lapacke_lu("gesv", H.ColMajor, 4, 2, a, 4, piv, bn, 4) == -7
    assert H.lapacke_lu("getrf", H.RowMajor, 4, 4, a, 3, piv, work=True) == -5
    assert H.lapacke_lu("getrs", H.RowMajor, "N", 4, 2, a, 4, piv, b, 1, work=True) == -9
    assert H.lapacke_lu("gesv", H.RowMajor, 4, 2, a, 2, piv, b, 2, work=True) == -5


@pytest.mark.parametrize("routine,change,info", [
    ("herk", dict(uplo="X"), 1), ("herk", dict(trans="T"), 2), ("syrk", dict(trans="C"), 2),
    ("syrk", dict(n=-1), 3), ("herk", dict(k=-1), 4), ("syrk", dict(lda=3), 7),
    ("herk", dict(trans="C", lda=1), 7), ("syrk", dict(ldc=3), 10)])
def test_zsyrk_zherk_invalid_arguments(routine, change, info):
    """Complex SYRK takes N or T, HERK N or C (src/level3.cc:93-115)."""
    g = dict(uplo="L", trans="N", n=4, k=2, lda=4, ldc=4)
    g.update(change)
    a = np.ones((8, 8), dtype=np.complex128, order="F")
    c = np.full((8, 8), 3.0 + 1.0j, dtype=np.complex128, order="F")
    c0 = c.copy()
    with H.ErrorCapture() as cap:
        H.zrank_k(routine, g["uplo"], g["trans"], g["n"], g["k"], 1.0, a, g["lda"], 0.5, c,
                  g["ldc"])
    assert cap.calls == [("Z" + routine.upper(), info)]
    assert np.array_equal(c, c0)


def test_error_positions_under_concurrent_callers():
    """Many threads report invalid arguments at once, some through CBLAS
    row-major calls (remapped positions), some through the Fortran entry:
    each report carries its own caller's name and position (no state is
    shared between calls)."""
    import collections
    import threading

    a = np.ones((8, 8), order="F")
    c = np.zeros((8, 8), order="F")
    # (call, expected report): row-major ldb < n is CBLAS position 11,
    # row-major m < 0 is position 4, Fortran lda < m is 8.
    cases = [
        (lambda: H.cblas_gemm(H.RowMajor, H.NoTrans, H.NoTrans, 4, 3, 2, 1.0, a, 2, a, 2,
                              0.0, c, 3), ("cblas_dgemm", 11)),
        (lambda: H.cblas_gemm(H.RowMajor, H.NoTrans, H.NoTrans, -1, 3, 2, 1.0, a, 2, a, 3,
                              0.0, c, 3), ("cblas_dgemm", 4)),
        (lambda: H.gemm("N", "N", 4, 4, 4, 1.0, a, 2, a, 4, 0.0, c, 4), ("DGEMM", 8)),
        (lambda: H.cblas_trsm(H.RowMajor, H.Left, H.Lower, H.NoTrans, H.NonUnit, 4, 3, 1.0, a,
                              4, c, 2), ("cblas_dtrsm", 12)),
    ]
    rounds = 200
    with H.ErrorCapture() as cap:
        def worker(call):
            for _ in range(rounds):
                call()
        threads = [threading.Thread(target=worker, args=(call,)) for call, _ in cases]
        for t in threads:
            t.start()
        for t in threads:
            t.join()
    got = collections.Counter(cap.calls)
    assert got == collections.Counter({want: rounds for _, want in cases})
    assert np.all(c == 0)


def test_dgemm_mode_switch():
    """hwyblas_set_dgemm_mode: 0 and 1 accepted (returning the previous mode),
    anything else refused without a change."""
    L = H.lib()
    start = L.hwyblas_get_dgemm_mode()
    assert start in (0, 1)
    assert L.hwyblas_set_dgemm_mode(1) == start
    assert L.hwyblas_get_dgemm_mode() == 1
    assert L.hwyblas_set_dgemm_mode(7) == -1
    assert L.hwyblas_get_dgemm_mode() == 1
    assert L.hwyblas_set_dgemm_mode(0) == 1
    assert L.hwyblas_set_dgemm_mode(start) == 0

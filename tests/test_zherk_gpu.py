"""Complex SYRK and HERK (csyrk_, zsyrk_, cherk_, zherk_) on the GPU: the
uplo triangle of alpha op(A) op(A)' + beta C against the complex128 product
with the bound 4 (k + 2) eps (|alpha| |op(A)| |op(A)'| + |beta| |C|); the
other triangle comes back untouched; HERK's diagonal is real; alpha = 0 or
k = 0 only scales the triangle (beta = 0: zeros, C not read)."""
import numpy as np
import pytest

import hwyblas_amd as H
from test_zgemm_gpu import zeps, zrand

pytestmark = pytest.mark.gpu
ZDTYPES = [np.complex128, np.complex64]
SENT = 7.25 - 3.5j


@pytest.mark.parametrize("dtype", ZDTYPES)
@pytest.mark.parametrize("routine", ["syrk", "herk"])
@pytest.mark.parametrize("uplo", ["L", "U"])
@pytest.mark.parametrize("trans", ["N", "X"])
def test_zsyrk_zherk(dtype, routine, uplo, trans):
    herm = routine == "herk"
    tr = trans if trans == "N" else ("C" if herm else "T")
    n, k, pad = 150, 70, 2
    ra, ca = (n, k) if tr == "N" else (k, n)
    a = zrand(ra + pad, ca, 1, dtype)
    c = zrand(n + pad, n, 2, dtype)
    i, j = np.indices((n + pad, n))
    tri = ((i >= j) if uplo == "L" else (i <= j)) & (i < n)
    c[~tri] = SENT
    c0 = c.copy(order="F")
    alpha, beta = (0.75, -0.5) if herm else (0.75 - 0.25j, -0.5 + 1.0j)
    H.zrank_k(routine, uplo, tr, n, k, alpha, a, ra + pad, beta, c, n + pad)
    x = a[:ra, :ca].astype(np.complex128)
    opa = x if tr == "N" else (x.T if tr == "T" else np.conj(x.T))
    opb = np.conj(opa.T) if herm else opa.T
    c0n = c0[:n].astype(np.complex128)
    if herm:
        c0n = c0n.copy()
        c0n[np.diag_indices(n)] = c0n[np.diag_indices(n)].real
    exp = alpha * (opa @ opb) + beta * c0n
    scale = abs(alpha) * (np.abs(opa) @ np.abs(opb)) + abs(beta) * np.abs(c0n) + 1e-300
    got = c[:n].astype(np.complex128)
    m = tri[:n]
    r = np.abs(got - exp)[m] / (4.0 * (k + 2) * zeps(dtype) * scale[m])
    assert float(np.max(r)) <= 1.0, float(np.max(r))
    assert np.array_equal(c[~tri], c0[~tri])
    if herm:
        assert np.all(np.diag(c[:n]).imag == 0)


@pytest.mark.parametrize("routine", ["syrk", "herk"])
def test_zsyrk_zherk_no_product(routine):
    n = 60
    a = zrand(n, 4, 3, np.complex128)
    c = zrand(n, n, 4, np.complex128)
    i, j = np.indices((n, n))
    tri = i >= j
    c[tri] = np.nan
    c0 = c.copy(order="F")
    H.zrank_k(routine, "L", "N", n, 4, 0.0, a, n, 0.0, c, n)
    assert np.all(c[tri] == 0) and np.array_equal(c[~tri], c0[~tri])

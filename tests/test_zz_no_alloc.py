"""The emulated DGEMM entry that allocates nothing, through the C-ABI.

hwyblas_dgemm_emulated_noalloc_dev runs the same kernels as
hwyblas_dgemm_emulated_dev on a fixed workspace array instead of the grow-only
per-thread buffer, so its results must be the allocating entry's bit for bit:
checked on two calls in a row (the second on a smaller problem, over the
contents the first left in the array), with the count chosen on the device and
fixed, and across transpositions. A problem beyond the array's size is
refused before any work, leaving C untouched.
"""
import numpy as np
import pytest

import oracle

pytestmark = pytest.mark.gpu


def run(H, noalloc, ta, tb, m, n, k, a, lda, b, ldb, c, ldc, nmod):
    da, db, dc = H.DeviceArray.from_host(a), H.DeviceArray.from_host(b), H.DeviceArray.from_host(c)
    f = H.dgemm_emulated_noalloc_dev if noalloc else H.dgemm_emulated_dev
    info, used = f(ta, tb, m, n, k, 1.25, da, lda, db, ldb, -0.5, dc, ldc, nmod)
    H.synchronize()
    return info, used, dc.to_host()


@pytest.mark.parametrize("nmod", [0, 11])
def test_noalloc_equals_the_allocating_entry(H, nmod):
    # Two calls in a row, the second smaller: it must not depend on what the
    # first left in the shared array.
    for (ta, tb, m, n, k, seed) in (("N", "T", 700, 600, 900, 1), ("T", "N", 300, 200, 500, 4)):
        ra, ca = (m, k) if ta == "N" else (k, m)
        rb, cb = (k, n) if tb == "N" else (n, k)
        a = oracle.fill_uniform(ra, ca, seed)
        b = oracle.fill_uniform(rb, cb, seed + 1)
        c = oracle.fill_uniform(m, n, seed + 2)
        i0, u0, want = run(H, False, ta, tb, m, n, k, a, ra, b, rb, c, m, nmod)
        i1, u1, got = run(H, True, ta, tb, m, n, k, a, ra, b, rb, c, m, nmod)
        assert i0 == 0 and i1 == 0
        assert u1 == u0 and u1 == (nmod or 14), (u0, u1)
        assert np.array_equal(got, want), (m, n, k)


def test_noalloc_native_fallback_matches(H):
    m, n, k = 128, 96, 256
    a = oracle.fill_uniform(m, k, 51)
    b = oracle.fill_uniform(k, n, 52)
    a[:64, 128:] = 0.0
    a[64:, :128] = 0.0
    b[:128, 48:] = 0.0
    b[128:, :48] = 0.0
    c = oracle.fill_uniform(m, n, 53)
    _, u0, want = run(H, False, "N", "N", m, n, k, a, m, b, k, c, m, 0)
    _, u1, got = run(H, True, "N", "N", m, n, k, a, m, b, k, c, m, 0)
    assert u0 == 0 and u1 == 0 and np.array_equal(got, want)


def test_noalloc_refuses_beyond_its_workspace(H):
    m = n = k = 1024  # needs about 75 MB, more than the 64 MiB array
    a = oracle.fill_uniform(m, k, 61)
    b = oracle.fill_uniform(k, n, 62)
    c = oracle.fill_uniform(m, n, 63)
    info, used, got = run(H, True, "N", "N", m, n, k, a, m, b, k, c, m, 0)
    assert info == -1 and used == 0
    assert np.array_equal(got, c)

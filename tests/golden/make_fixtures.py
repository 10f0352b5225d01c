"""Writes tests/golden/blas_fixtures.npz: small GEMM, TRSM, LU and QR cases
whose expected outputs come from an independent library, numpy's BLAS
(OpenBLAS), scipy.linalg.solve_triangular (LAPACK trtrs),
scipy.linalg.lu_factor / lu_solve (LAPACK getrf / getrs) and
scipy.linalg.qr(mode="raw") (LAPACK geqrf: reflectors and tau), in this
container.

hwyBLAS stores no golden vectors (its tests regenerate inputs from seeds and
check bounds), so these fixtures are the external known answers against
which the oracle is pinned; they are data only (inputs and outputs).

Run: python tests/golden/make_fixtures.py
"""
import os

import numpy as np
import scipy
import scipy.linalg

OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "blas_fixtures.npz")


def main():
    rng = np.random.default_rng(20260101)
    data = {}
    cases = []
    shapes = [(7, 5, 3), (33, 17, 65), (64, 64, 64), (1, 9, 300), (70, 1, 2)]
    idx = 0
    for dt in ("f8", "f4"):
        for ta in "NT":
            for tb in "NT":
                for (m, n, k) in shapes:
                    ra, ca = (m, k) if ta == "N" else (k, m)
                    rb, cb = (k, n) if tb == "N" else (n, k)
                    a = rng.uniform(-1, 1, (ra, ca)).astype(dt)
                    b = rng.uniform(-1, 1, (rb, cb)).astype(dt)
                    c = rng.uniform(-1, 1, (m, n)).astype(dt)
                    alpha, beta = 1.5, -0.75
                    A = a if ta == "N" else a.T
                    B = b if tb == "N" else b.T
                    out = (np.asarray(alpha, dt) * (A @ B) + np.asarray(beta, dt) * c).astype(dt)
                    key = f"gemm{idx}"
                    data[key + "_a"], data[key + "_b"], data[key + "_c"] = a, b, c
                    data[key + "_out"] = out
                    cases.append(f"{key},{dt},{ta},{tb},{m},{n},{k},{alpha},{beta}")
                    idx += 1
    data["gemm_cases"] = np.array(cases)

    tcases = []
    idx = 0
    for dt in ("f8", "f4"):
        for side in "LR":
            for uplo in "LU":
                for ta in "NT":
                    for diag in "NU":
                        m, n = 33, 9
                        ka = m if side == "L" else n
                        a = rng.uniform(-1, 1, (ka, ka)) / ka
                        a[np.arange(ka), np.arange(ka)] = (1.5 + rng.uniform(-0.5, 0.5, ka)) * \
                            rng.choice([-1.0, 1.0], ka)
                        a = (np.tril(a) if uplo == "L" else np.triu(a)).astype(dt)
                        b = rng.uniform(-1, 1, (m, n)).astype(dt)
                        alpha = 0.5
                        lower = uplo == "L"
                        unit = diag == "U"
                        trans = 0 if ta == "N" else 1
                        rhs = (np.asarray(alpha, dt) * b).astype(dt)
                        if side == "L":
                            x = scipy.linalg.solve_triangular(a, rhs, trans=trans, lower=lower,
                                                              unit_diagonal=unit)
                        else:  # X op(A) = R  <=>  op(A)^T X^T = R^T
                            x = scipy.linalg.solve_triangular(a, rhs.T, trans=1 - trans,
                                                              lower=lower, unit_diagonal=unit).T
                        key = f"trsm{idx}"
                        data[key + "_a"], data[key + "_b"] = a, b
                        data[key + "_out"] = np.ascontiguousarray(x).astype(dt)
                        tcases.append(f"{key},{dt},{side},{uplo},{ta},{diag},{m},{n},{alpha}")
                        idx += 1
    data["trsm_cases"] = np.array(tcases)
    # LU with partial pivoting (LAPACK dgetrf over OpenBLAS) and the solve
    # with it (dgetrs), square and rectangular; one matrix with a zero column.
    lcases = []
    for idx, (m, n) in enumerate([(1, 1), (7, 7), (16, 16), (33, 33), (64, 64), (100, 100),
                                  (150, 150), (40, 20), (20, 40), (30, 30)]):
        a = rng.standard_normal((m, n))
        if idx == 9:
            a[:, 4] = 0.0
        lu, piv = scipy.linalg.lu_factor(a, check_finite=False)
        key = f"lu{idx}"
        data[key + "_a"], data[key + "_lu"] = a, lu
        data[key + "_ipiv"] = (piv + 1).astype(np.int32)
        if m == n and idx != 9:
            b = rng.uniform(-1, 1, (n, 3))
            data[key + "_b"] = b
            data[key + "_x"] = scipy.linalg.lu_solve((lu, piv), b)
            data[key + "_xt"] = scipy.linalg.lu_solve((lu, piv), b, trans=1)
        lcases.append(f"{key},{m},{n}")
    data["lu_cases"] = np.array(lcases)
    # Householder QR (LAPACK dgeqrf through scipy): the factored matrix (R on
    # and above the diagonal, reflectors below) and tau; square, tall, wide,
    # large enough for the blocked path (k > 128), a zero column (a reflector
    # with tau = 0) and a column already zero below its diagonal.
    qrng = np.random.default_rng(20261203)
    qcases = []
    for idx, (m, n) in enumerate([(1, 1), (5, 3), (3, 5), (40, 40), (200, 150), (150, 200),
                                  (300, 300), (260, 64), (50, 30), (64, 48)]):
        a = qrng.standard_normal((m, n))
        if idx == 8:
            a[:, 7] = 0.0
        if idx == 9:
            a[1:, 0] = 0.0
        (h, tau), _ = scipy.linalg.qr(a, mode="raw", check_finite=False)
        key = f"qr{idx}"
        data[key + "_a"], data[key + "_h"], data[key + "_tau"] = a, np.asfortranarray(h), tau
        qcases.append(f"{key},{m},{n}")
    data["qr_cases"] = np.array(qcases)
    data["versions"] = np.array([f"numpy {np.__version__}", f"scipy {scipy.__version__}"])
    np.savez_compressed(OUT, **data)
    print(f"wrote {OUT}: {len(cases)} gemm, {len(tcases)} trsm, {len(lcases)} lu, "
          f"{len(qcases)} qr cases")


if __name__ == "__main__":
    main()

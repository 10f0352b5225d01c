"""Householder QR on the GPU (dgeqrf_, hwyblas_dgeqrf_dev) through the C-ABI.

Pinned against (1) LAPACK dgeqrf's own output (scipy, the committed
fixtures) and (2) the oracle's restatement of hwyBLAS's xGEQRF: the same
reflectors, signs and tau up to rounding (the GPU blocks by 64 columns and
sums in another order, so the bits differ); (3) the defining properties:
Q^T Q = I and Q R = A to rounding, R upper triangular. Also: the device
entry equals the Fortran entry on host memory bit for bit, tau = 0 for a
column already zero below the diagonal, and operands near underflow take
xLARFG's rescaling path.
"""
import os

import numpy as np
import pytest

import hwyblas_amd as H
import oracle

pytestmark = pytest.mark.gpu
FIX = np.load(os.path.join(os.path.dirname(__file__), "golden", "blas_fixtures.npz"))
EPS = np.finfo(float).eps


def gpu_qr(a):
    m, n = a.shape
    h = np.asfortranarray(a.copy())
    tau = np.zeros(max(1, min(m, n)))
    info, _ = H.geqrf(m, n, h, m, tau)
    assert info == 0
    return h, tau[:min(m, n)]


def explicit_q(h, tau):
    m, n = h.shape
    q = np.eye(m)
    for i in reversed(range(min(m, n))):
        v = np.zeros(m)
        v[i] = 1.0
        v[i + 1:] = h[i + 1:, i]
        q -= tau[i] * np.outer(v, v @ q)
    return q


def close(h, tau, h_ref, tau_ref, m, n, c=64):
    """Entries within c max(m, n) eps of the reference: R relative to the
    largest column norm, the reflectors and tau absolutely (|v| <= 1-ish,
    tau in [1, 2])."""
    k = min(m, n)
    scale = max(1.0, float(np.max(np.abs(np.triu(h_ref)))))
    tol = c * max(m, n) * EPS
    r_err = np.max(np.abs(np.triu(h) - np.triu(h_ref))) / scale
    v_err = np.max(np.abs(np.tril(h, -1)[:, :k] - np.tril(h_ref, -1)[:, :k]), initial=0.0)
    t_err = np.max(np.abs(tau - tau_ref), initial=0.0)
    return r_err <= tol and v_err <= tol * 16 and t_err <= tol, (r_err, v_err, t_err)


def test_qr_matches_lapack_fixtures():
    for line in FIX["qr_cases"]:
        key, m, n = str(line).split(",")
        m, n = int(m), int(n)
        a = FIX[key + "_a"]
        h, tau = gpu_qr(a)
        ok, errs = close(h, tau, FIX[key + "_h"], FIX[key + "_tau"], m, n)
        assert ok, (key, errs)
    h, tau = gpu_qr(FIX["qr9_a"])  # column 0 already zero below the diagonal
    assert tau[0] == 0.0


@pytest.mark.parametrize("m,n", [(513, 300), (300, 513), (1000, 1000), (2000, 130), (64, 64),
                                 (65, 63), (1, 40), (40, 1)])
def test_qr_against_the_oracle_and_its_properties(m, n):
    a = oracle.fill_uniform(m, n, m + 7 * n)
    h, tau = gpu_qr(a)
    ho, to = oracle.geqrf(a)
    ok, errs = close(h, tau, ho, to, m, n)
    assert ok, errs
    k = min(m, n)
    if m <= 2000:
        q = explicit_q(h, tau)
        r = np.triu(h[:k, :])
        assert np.max(np.abs(q.T @ q - np.eye(m))) <= 16 * m * EPS
        assert np.linalg.norm(q[:, :k] @ r - a, 1) <= 16 * m * EPS * np.linalg.norm(a, 1)


def test_device_entry_equals_host_entry_and_is_asynchronous():
    m, n = 700, 400
    a = oracle.fill_uniform(m, n, 5)
    h, tau = gpu_qr(a)
    da = H.DeviceArray.from_host(a)
    dt = H.DeviceArray(min(m, n), 1)
    stream = H.lib().hwyblas_amd_stream_create()
    try:
        assert H.geqrf_dev(da, dt, stream) == 0
        H.lib().hwyblas_amd_stream_synchronize(stream)
    finally:
        H.lib().hwyblas_amd_stream_destroy(stream)
    assert np.array_equal(da.to_host(), h)
    assert np.array_equal(dt.to_host().ravel(), tau)
    # Device operands through the Fortran entry.
    db = H.DeviceArray.from_host(a)
    dt2 = H.DeviceArray(min(m, n), 1)
    info, _ = H.geqrf(m, n, db, m, dt2)
    assert info == 0 and np.array_equal(db.to_host(), h)


def test_near_underflow_takes_the_rescaling_path():
    """A scaled by 2^-1000 (columns' norms below safmin = tiny / eps): xLARFG
    rescales by 1/safmin; the factors equal the oracle's to rounding and
    are those of the unscaled matrix, scaled."""
    m, n = 150, 90
    a0 = oracle.fill_uniform(m, n, 9)
    a = a0 * 2.0 ** -1000
    h, tau = gpu_qr(a)
    ho, to = oracle.geqrf(a)
    h1, t1 = gpu_qr(a0)
    assert np.max(np.abs(tau - to)) <= 64 * m * EPS
    assert np.max(np.abs(tau - t1)) <= 64 * m * EPS
    assert np.max(np.abs(np.tril(h, -1) - np.tril(h1, -1))) <= 1e-12
    r_scaled = np.triu(h)[:n] * 2.0 ** 1000
    assert np.max(np.abs(r_scaled - np.triu(h1)[:n])) <= 1e-12 * np.max(np.abs(np.triu(h1)))


def test_padding_rows_untouched():
    m, n, lda = 130, 70, 137
    a = np.zeros((lda, n), order="F")
    a[:m] = oracle.fill_uniform(m, n, 3)
    a[m:] = 7.25
    h = a.copy(order="F")
    tau = np.zeros(n)
    info, _ = H.geqrf(m, n, h, lda, tau)
    assert info == 0
    assert np.all(h[m:] == 7.25)
    ho, to = oracle.geqrf(a[:m])
    ok, errs = close(h[:m], tau, ho, to, m, n)
    assert ok, errs


PANEL_CASES = [(300, 200, 1), (1000, 700, 2), (2100, 140, 3), (130, 300, 4), (5000, 64, 5),
               (4097, 100, 6)]


def factor_panel_cases():
    """QR of the cases above (one with column 0 zero below the diagonal,
    one scaled near underflow) through the Fortran entry."""
    out = []
    for m, n, seed in PANEL_CASES:
        a = oracle.fill_uniform(m, n, 900 + seed)
        if seed == 2:
            a[1:, 0] = 0.0
        if seed == 3:
            a *= 2.0 ** -1000
        out.extend(gpu_qr(a))
    return out


def _child(env_extra, code_tail, tmp_path):
    import subprocess
    import sys
    here = os.path.dirname(os.path.abspath(__file__))
    root = os.path.dirname(here)
    code = ("import sys, numpy as np\n"
            f"sys.path[:0] = [{root!r}, {here!r}]\n"
            "import hwyblas_amd as H, oracle, test_qr_gpu as T\n" + code_tail)
    env = dict(os.environ, **env_extra)
    subprocess.run([sys.executable, "-c", code], env=env, cwd=root, check=True, timeout=240)


def test_cooperative_panel_equals_column_launches(tmp_path):
    """A panel whose workgroups all fit is factored in one cooperative launch
    with two grid barriers per column; its bits equal the two launches per
    column it replaces (taken with HWYBLAS_AMD_QR_PANEL=steps in a child
    process), tau = 0 columns and xLARFG's rescaling included."""
    out = tmp_path / "steps.npz"
    _child({"HWYBLAS_AMD_QR_PANEL": "steps"},
           f"np.savez({str(out)!r}, *T.factor_panel_cases())\n", tmp_path)
    got = factor_panel_cases()
    with np.load(out) as ref:
        for i in range(len(got)):
            assert got[i].tobytes() == ref[f"arr_{i}"].tobytes(), (PANEL_CASES[i // 2], i % 2)
    assert got[3][0] == 0.0  # the zero column's tau


def test_timed_out_panel_barrier_is_reported(monkeypatch):
    """The failure path of the cooperative panel, forced from the host
    (HWYBLAS_AMD_QR_FAIL_BARRIER=1 raises the flag before the first panel):
    every bounded wait gives up at its first poll and every workgroup
    returns, the panels' tau are NaN, and dgeqrf_ returns info = -1000,
    reported to the error handler as ("DGEQRF", -1000). A normal call
    afterwards factors correctly."""
    m, n = 500, 70
    a0 = oracle.fill_uniform(m, n, 3)
    monkeypatch.setenv("HWYBLAS_AMD_QR_FAIL_BARRIER", "1")
    a = np.asfortranarray(a0.copy())
    tau = np.zeros(n)
    with H.ErrorCapture() as cap:
        info, _ = H.geqrf(m, n, a, m, tau)
    assert info == -1000 and cap.calls == [("DGEQRF", -1000)]
    assert np.all(np.isnan(tau))
    monkeypatch.delenv("HWYBLAS_AMD_QR_FAIL_BARRIER")
    h, t = gpu_qr(a0)
    ho, to = oracle.geqrf(a0)
    ok, errs = close(h, t, ho, to, m, n)
    assert ok, errs


def test_refused_cooperative_launch_falls_back_to_column_launches(tmp_path):
    """This process has made cooperative launches; a second process on the
    same GPU may be refused one (the device's cooperative queue is held).
    Its factorizations fall back to the column launches and equal this
    process's bit for bit."""
    gpu_qr(oracle.fill_uniform(300, 100, 1))  # a cooperative launch here first
    out = tmp_path / "child.npz"
    _child({}, f"np.savez({str(out)!r}, *T.factor_panel_cases())\n", tmp_path)
    got = factor_panel_cases()
    with np.load(out) as ref:
        for i in range(len(got)):
            assert got[i].tobytes() == ref[f"arr_{i}"].tobytes(), (PANEL_CASES[i // 2], i % 2)


def test_tall_matrix_mixes_column_launches_and_cooperative_panels(tmp_path):
    """m = 65600: the first panel's 257 workgroups of 256 rows exceed what
    can be resident at once on a 256-CU device, so it takes the column
    launches, and the later, shorter panels the cooperative launch. R equals
    numpy's (LAPACK dgeqrf, the same signs) to rounding, and the whole
    factorization equals the column launches' bit for bit."""
    m, n = 65600, 200
    a = oracle.fill_uniform(m, n, 12)
    h, tau = gpu_qr(a)
    r_np = np.linalg.qr(a, mode="r")
    scale = float(np.max(np.abs(r_np)))
    assert np.max(np.abs(np.triu(h[:n]) - r_np)) <= 64 * m * EPS * scale
    out = tmp_path / "tall.npz"
    _child({"HWYBLAS_AMD_QR_PANEL": "steps"},
           f"m, n = {m}, {n}\n"
           "h, tau = T.gpu_qr(oracle.fill_uniform(m, n, 12))\n"
           f"np.savez({str(out)!r}, h=h, tau=tau)\n", tmp_path)
    with np.load(out) as ref:
        assert h.tobytes() == ref["h"].tobytes() and tau.tobytes() == ref["tau"].tobytes()


@pytest.mark.parametrize("m,n", [(4096, 2048), (2560, 2560)])
def test_wide_blocks_take_the_split_dma_products(m, n):
    """Blocks whose V^T A2 product (256 rows, trailing columns a multiple
    of 128, k = the block's height >= 2048) has at most 256 tiles of
    128 x 128 run it on the DMA-fed kernel split over k, the partials added
    in order. R equals numpy's (LAPACK dgeqrf, the same signs) to rounding,
    and the reflectors' taus lie in [1, 2] (a nonzero column)."""
    a = oracle.fill_uniform(m, n, 31)
    h, tau = gpu_qr(a)
    r_np = np.linalg.qr(a, mode="r")
    scale = float(np.max(np.abs(r_np)))
    assert np.max(np.abs(np.triu(h[:n]) - r_np)) <= 64 * m * EPS * scale
    assert np.all((tau[:-1] >= 1.0) & (tau[:-1] <= 2.0))

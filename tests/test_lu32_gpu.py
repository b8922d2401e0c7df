"""The complex float LU on the MFMA units (bqd2_cgetrf / bqd2_cgetrs) and its
trailing-update GEMM (bqd2_cgemm), against numpy on the same seeded inputs.

GEMM: complex float inputs, the product formed in complex double by numpy;
bar 2e-6 relative to |A| |B| per entry (FP32 accumulation over k <= 300 terms),
for ragged shapes that cut every tile edge, both modes (C -= A B, C = A B) and
the in-place form used for U12 = L11^-1 A12 (B is C, m <= 128).

LU: the same diagonally weighted matrices as the FP64 LU test (condition
numbers below ~1e3), sizes inside one panel (n = 50), a multiple of the
128-column panel (n = 512), ragged (n = 1000) and the sphere384 system size
(n = 3072), 1 and 3 right-hand sides, plus 16 right-hand sides (MFMA solves).
Bars: backward error |A x - b| / (|A| |x|) below 2e-6 (complex float, n <= 3072)
and the solution within 1e-4 of LAPACK's FP64 solution; a matrix with a zero
column is reported singular.  Both panel pivotings: block-local (the default,
kept while max |L21| <= 16) and over all rows; the matrix that needs a pivot
from outside the first diagonal block makes the block-local panel fall back."""
import ctypes as C

import numpy as np
import pytest

from bem3_aw_qd2_amd import hip_lib

pytestmark = pytest.mark.gpu


def dev(a):
    import torch
    return torch.tensor(np.ascontiguousarray(a).view(np.float32), device="cuda")


def back(t, shape_t):
    return t.cpu().numpy().view(np.complex64).reshape(shape_t).T


def cmat(rng, m, n):
    return (rng.standard_normal((m, n)) + 1j * rng.standard_normal((m, n))).astype(np.complex64)


@pytest.mark.parametrize("m,nc,kk,setmode", [(200, 130, 77, 0), (128, 300, 128, 0), (257, 33, 300, 1),
                                             (1, 1, 1, 0), (384, 256, 128, 1)])
def test_cgemm_matches_numpy(m, nc, kk, setmode):
    import torch
    rng = np.random.default_rng(m + nc + kk)
    A, B, C0 = cmat(rng, m, kk), cmat(rng, kk, nc), cmat(rng, m, nc)
    dA, dB, dC = dev(A.T), dev(B.T), dev(C0.T)
    rc = hip_lib().bqd2_cgemm(setmode, m, nc, kk, C.c_void_p(dA.data_ptr()), m, C.c_void_p(dB.data_ptr()), kk,
                              C.c_void_p(dC.data_ptr()), m, None)
    torch.cuda.synchronize()
    assert rc == 0
    prod = A.astype(np.complex128) @ B.astype(np.complex128)
    ref = prod if setmode else C0.astype(np.complex128) - prod
    got = back(dC, (nc, m))
    scale = np.abs(A).max() * np.abs(B).max() * kk
    err = np.abs(got - ref).max() / scale
    print("m %d nc %d k %d set %d: max err %.2e of |A||B|k" % (m, nc, kk, setmode, err))
    assert err < 2e-6


def test_cgemm_in_place_set():
    """C = T C for a 128 x 128 T, as the panel's U12 = L11^-1 A12."""
    import torch
    rng = np.random.default_rng(5)
    T, X = cmat(rng, 128, 128), cmat(rng, 128, 700)
    dT, dX = dev(T.T), dev(X.T)
    rc = hip_lib().bqd2_cgemm(1, 128, 700, 128, C.c_void_p(dT.data_ptr()), 128, C.c_void_p(dX.data_ptr()), 128,
                              C.c_void_p(dX.data_ptr()), 128, None)
    torch.cuda.synchronize()
    assert rc == 0
    ref = T.astype(np.complex128) @ X.astype(np.complex128)
    err = np.abs(back(dX, (700, 128)) - ref).max() / (np.abs(T).max() * np.abs(X).max() * 128)
    assert err < 2e-6


def matrix(n, seed):
    rng = np.random.default_rng(seed)
    A = rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))
    A += np.diag(np.full(n, 2.0 * np.sqrt(n)))
    return A


def solve_gpu(A, B):
    import torch
    hip = hip_lib()
    n, nrhs = B.shape
    dA = dev(A.astype(np.complex64).T)
    dB = dev(B.astype(np.complex64).T)
    ipiv = torch.zeros(n, dtype=torch.int32, device="cuda")
    ws = torch.zeros(hip.bqd2_cgetrf_wsize(n), dtype=torch.uint8, device="cuda")
    tmp = torch.zeros(2 * n * nrhs, dtype=torch.float32, device="cuda")
    rc = hip.bqd2_cgetrf(n, C.c_void_p(dA.data_ptr()), C.c_void_p(ipiv.data_ptr()), C.c_void_p(ws.data_ptr()), None)
    if rc != 0:
        return rc, None
    rc = hip.bqd2_cgetrs(n, nrhs, C.c_void_p(dA.data_ptr()), C.c_void_p(ws.data_ptr()), C.c_void_p(dB.data_ptr()),
                         C.c_void_p(tmp.data_ptr()), None)
    torch.cuda.synchronize()
    return rc, back(dB, (nrhs, n)).astype(np.complex128)


@pytest.mark.parametrize("n,nrhs", [(50, 1), (512, 3), (1000, 3), (3072, 1), (1000, 16)])
def test_lu32_solve(n, nrhs):
    A = matrix(n, n)
    rng = np.random.default_rng(7)
    B = rng.standard_normal((n, nrhs)) + 1j * rng.standard_normal((n, nrhs))
    rc, X = solve_gpu(A, B)
    assert rc == 0
    Xr = np.linalg.solve(A, B)
    err = np.linalg.norm(X - Xr) / np.linalg.norm(Xr)
    res = np.linalg.norm(A @ X - B) / (np.linalg.norm(A) * np.linalg.norm(X))
    print("n", n, "nrhs", nrhs, "err %.2e res %.2e" % (err, res))
    assert res < 2e-6 and err < 1e-4


def test_lu32_pivots_like_lapack():
    """A matrix whose first column is tiny off one row: the factorisation must
    pivot (without interchanges the float solution would be garbage)."""
    n = 300
    A = matrix(n, 11)
    A[:, 0] *= 1e-6
    A[137, 0] = 40.0
    b = np.ones((n, 1), complex)
    rc, X = solve_gpu(A, b)
    assert rc == 0
    res = np.linalg.norm(A @ X - b) / (np.linalg.norm(A) * np.linalg.norm(X))
    assert res < 2e-6


def test_lu32_reports_singular():
    A = matrix(200, 3)
    A[:, 77] = 0.0
    rc, _ = solve_gpu(A, np.ones((200, 1), complex))
    assert rc != 0
    msg = hip_lib().bqd2_last_error()
    print(msg)
    assert b"zero pivot" in msg


def panel_counts():
    c = np.zeros(2, np.int64)
    assert hip_lib().bqd2_lu32_panel_counts(c.ctypes.data) == 0
    return c


@pytest.mark.parametrize("mode", [0, 1])
def test_lu32_pivoting_modes(mode):
    hip = hip_lib()
    n, nrhs = 1000, 3
    A = matrix(n, 21)
    rng = np.random.default_rng(8)
    B = rng.standard_normal((n, nrhs)) + 1j * rng.standard_normal((n, nrhs))
    c0 = panel_counts()
    assert hip.bqd2_set_lu32_pivoting(mode) == 0
    try:
        rc, X = solve_gpu(A, B)
    finally:
        hip.bqd2_set_lu32_pivoting(1)
    c1 = panel_counts()
    assert rc == 0
    res = np.linalg.norm(A @ X - B) / (np.linalg.norm(A) * np.linalg.norm(X))
    print("mode", mode, "panels (block-local, all rows):", c1 - c0, "res %.2e" % res)
    assert res < 2e-6
    npan = (n + 127) // 128
    assert (c1 - c0).tolist() == ([0, npan] if mode == 0 else [npan, 0])


def test_lu32_local_panel_falls_back():
    """the pivot of column 0 lies outside the first diagonal block: that panel's
    block-local factorisation grows beyond the bound and is redone over all rows"""
    n = 300
    A = matrix(n, 11)
    A[:, 0] *= 1e-6
    A[137, 0] = 40.0
    c0 = panel_counts()
    rc, X = solve_gpu(A, np.ones((n, 1), complex))
    c1 = panel_counts()
    assert rc == 0
    assert (c1 - c0)[1] >= 1
    res = np.linalg.norm(A @ X - 1.0) / (np.linalg.norm(A) * np.linalg.norm(X))
    assert res < 2e-6


def test_lu32_unchecked_block_local_matches_checked():
    """growth bound >= 1e30: the block-local panels go unchecked and the panels
    run beside the trailing updates on their own CUs; on a matrix whose panels
    all stay block-local under the default bound, the factors (and solution)
    are bit-identical to the checked path's"""
    hip = hip_lib()
    n, nrhs = 1000, 3
    A = matrix(n, 21)
    rng = np.random.default_rng(8)
    B = rng.standard_normal((n, nrhs)) + 1j * rng.standard_normal((n, nrhs))
    c0 = panel_counts()
    rc, X1 = solve_gpu(A, B)
    assert rc == 0
    assert (panel_counts() - c0)[1] == 0
    assert hip.bqd2_set_lu32_growth_bound(C.c_double(1e30)) == 0
    try:
        rc, X2 = solve_gpu(A, B)
    finally:
        hip.bqd2_set_lu32_growth_bound(C.c_double(16.0))
    assert rc == 0
    assert np.array_equal(X1, X2)


def test_lu32_unchecked_reports_block_zero_pivot():
    hip = hip_lib()
    A = matrix(200, 3)
    A[:, 77] = 0.0
    assert hip.bqd2_set_lu32_growth_bound(C.c_double(1e30)) == 0
    try:
        rc, _ = solve_gpu(A, np.ones((200, 1), complex))
    finally:
        hip.bqd2_set_lu32_growth_bound(C.c_double(16.0))
    assert rc != 0
    assert b"diagonal block" in hip_lib().bqd2_last_error()
    assert hip.bqd2_set_lu32_growth_bound(C.c_double(0.5)) != 0

"""The dense complex LU (bqd2_zgetrf) and its solves (bqd2_zgetrs) on device
buffers, against numpy's LAPACK on the same seeded matrices.

Sizes cover a single partial block (n = 50), a multiple of the 64-column panel
(n = 512), a ragged last panel (n = 1000) and the system size of the sphere384
workload (n = 3072), with 1 and 3 right-hand sides; on an unweighted random
matrix the interchanges equal LAPACK's and the factors scipy's.  Bar: the solution within
1e-11 normwise of LAPACK's (both are backward stable; the matrices are
diagonally weighted so that their condition numbers stay below ~1e3), and the
residual |A x - b| / (|A| |x|) below 1e-14.  A matrix with a zero column is
reported singular."""
import ctypes as C

import numpy as np
import pytest

from bem3_aw_qd2_amd import hip_lib

pytestmark = pytest.mark.gpu


def matrix(n, seed):
    rng = np.random.default_rng(seed)
    A = rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))
    A += np.diag(np.full(n, 2.0 * np.sqrt(n)))
    return A


def solve_gpu(A, B):
    import torch
    n, nrhs = B.shape
    # column-major device copies (the library's layout)
    dA = torch.tensor(np.ascontiguousarray(A.T).view(np.float64), device="cuda")
    dB = torch.tensor(np.ascontiguousarray(B.T).view(np.float64), device="cuda")
    ipiv = torch.zeros(n, dtype=torch.int32, device="cuda")
    hip = hip_lib()
    rc = hip.bqd2_zgetrf(n, C.c_void_p(dA.data_ptr()), C.c_void_p(ipiv.data_ptr()), None)
    if rc != 0:
        return rc, None
    rc = hip.bqd2_zgetrs(n, nrhs, C.c_void_p(dA.data_ptr()), C.c_void_p(ipiv.data_ptr()),
                         C.c_void_p(dB.data_ptr()), None)
    torch.cuda.synchronize()
    X = dB.cpu().numpy().view(np.complex128).reshape(nrhs, n).T
    return rc, X


@pytest.mark.parametrize("n,nrhs", [(50, 1), (512, 3), (1000, 3), (3072, 1)])
def test_lu_solve_matches_lapack(n, nrhs):
    A = matrix(n, n)
    rng = np.random.default_rng(7)
    B = rng.standard_normal((n, nrhs)) + 1j * rng.standard_normal((n, nrhs))
    rc, X = solve_gpu(A, B)
    assert rc == 0
    Xr = np.linalg.solve(A, B)
    err = np.linalg.norm(X - Xr) / np.linalg.norm(Xr)
    res = np.linalg.norm(A @ X - B) / (np.linalg.norm(A) * np.linalg.norm(X))
    print("n", n, "nrhs", nrhs, "err %.2e res %.2e" % (err, res))
    assert err < 1e-11 and res < 1e-14


def test_lu_reports_singular():
    A = matrix(200, 3)
    A[:, 77] = 0.0
    rc, _ = solve_gpu(A, np.ones((200, 1), complex))
    assert rc != 0
    msg = hip_lib().bqd2_last_error()
    print(msg)
    assert b"zero pivot" in msg


def test_lu_interchanges_match_lapack():
    """a matrix without diagonal weight, so that nearly every column pivots:
    the interchanges equal LAPACK's zgetrf's (max |Re|+|Im|, first of ties) and
    the factors agree with scipy's to rounding"""
    import scipy.linalg
    import torch
    n = 700
    rng = np.random.default_rng(12)
    A = rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))
    dA = torch.tensor(np.ascontiguousarray(A.T).view(np.float64), device="cuda")
    ipiv = torch.zeros(n, dtype=torch.int32, device="cuda")
    assert hip_lib().bqd2_zgetrf(n, C.c_void_p(dA.data_ptr()), C.c_void_p(ipiv.data_ptr()), None) == 0
    torch.cuda.synchronize()
    LU = dA.cpu().numpy().view(np.complex128).reshape(n, n).T
    lu_ref, piv_ref = scipy.linalg.lu_factor(A)
    assert np.array_equal(ipiv.cpu().numpy(), piv_ref)
    err = np.abs(LU - lu_ref).max() / np.abs(lu_ref).max()
    print("interchanges equal; max factor difference %.2e" % err)
    assert err < 1e-12

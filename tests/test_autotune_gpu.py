"""The online autotuner (bqd2_set_autotune / _force / _info / _reset).

Its candidates are launch shapes that give the same bits, so:
* the complex float LU factors and solution are bit-identical for both
  trailing-update tiles, pinned with bqd2_autotune_force, at a ragged size
  (n = 1000) and a multiple of the panel (n = 1536);
* the FP64 GMRES solution of the reference's sample (4800 unknowns) is
  bit-identical for every matrix-vector workgroup size;
* tuning on (the default) settles every size class that was called often
  enough: after twelve factorisations of n = 2048 each class of the LU site,
  and after one GMRES solve the class of its matrix-vector product, holds a valid choice and a timing for every candidate, the choice is the
  fastest candidate, and the tuned runs still give the default candidate's bits;
* mode 0 creates no classes."""
import ctypes as C
import os

import numpy as np
import pytest

from bem3_aw_qd2_amd import exdat, hip_lib, host_lib

pytestmark = pytest.mark.gpu
NCAND = {0: 2, 1: 4}
SAMPLE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "sample_ex.dat")


def dev(a):
    import torch
    return torch.tensor(np.ascontiguousarray(a).view(np.float32), device="cuda")


def matrix(n, seed):
    rng = np.random.default_rng(seed)
    A = rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))
    A += np.diag(np.full(n, 2.0 * np.sqrt(n)))
    return A.astype(np.complex64)


def lu_bits(A, b):
    import torch
    hip = hip_lib()
    n = A.shape[0]
    dA, dB = dev(A.T), dev(b.astype(np.complex64))
    ipiv = torch.zeros(n, dtype=torch.int32, device="cuda")
    ws = torch.zeros(hip.bqd2_cgetrf_wsize(n), dtype=torch.uint8, device="cuda")
    tmp = torch.zeros(2 * n, dtype=torch.float32, device="cuda")
    assert hip.bqd2_cgetrf(n, C.c_void_p(dA.data_ptr()), C.c_void_p(ipiv.data_ptr()), C.c_void_p(ws.data_ptr()),
                           None) == 0
    assert hip.bqd2_cgetrs(n, 1, C.c_void_p(dA.data_ptr()), C.c_void_p(ws.data_ptr()), C.c_void_p(dB.data_ptr()),
                           C.c_void_p(tmp.data_ptr()), None) == 0
    torch.cuda.synchronize()
    return dA.cpu().numpy().tobytes(), ipiv.cpu().numpy().tobytes(), dB.cpu().numpy().tobytes()


def info(site, wait=1):
    hip = hip_lib()
    keys = (C.c_longlong * 64)()
    chosen = (C.c_int * 64)()
    best = (C.c_float * 256)()
    nc = C.c_int()
    cnt = hip.bqd2_autotune_info(site, wait, 64, keys, chosen, best, C.byref(nc))
    assert 0 <= cnt <= 64 and nc.value == NCAND[site]
    return [(keys[i], chosen[i], list(best[4 * i:4 * i + nc.value])) for i in range(cnt)]


@pytest.fixture
def tuner():
    hip = hip_lib()
    yield hip
    for s in NCAND:
        hip.bqd2_autotune_force(s, -1)
    hip.bqd2_set_autotune(1)


@pytest.mark.parametrize("n", [1000, 1536])
def test_lu_candidates_bit_identical(tuner, n):
    A = matrix(n, n)
    b = np.random.default_rng(1).standard_normal(n) + 0j
    ref = None
    for tile in range(NCAND[0]):
        assert tuner.bqd2_autotune_force(0, tile) == 0
        got = lu_bits(A, b)
        if ref is None:
            ref = got
        print("n %d tile candidate %d: factors %s, solution %s" % (n, tile, got[0] == ref[0], got[2] == ref[2]))
        assert got == ref
    assert tuner.bqd2_autotune_force(0, 2) != 0 and tuner.bqd2_autotune_force(2, 0) != 0


def sample_solution():
    m = exdat.Model(host_lib()).dat_read(SAMPLE)
    prob = host_lib().aqd2_problem(m.buf, None)
    hip = hip_lib()
    assert hip.bqd2_set_solver(prob, 0) == 0
    m.solve()
    x = np.zeros(2 * m.nn(), complex)
    assert hip.bqd2_solve_surface(prob, x.ctypes.data) == 0
    m.finalize()
    return x


def test_gmres_candidates_bit_identical(tuner):
    xs = []
    for c in range(NCAND[1]):
        assert tuner.bqd2_autotune_force(1, c) == 0
        xs.append(sample_solution())
    for c, x in enumerate(xs[1:], 1):
        print("rows per workgroup candidate %d: max |x - x0| = %.3e" % (c, np.abs(x - xs[0]).max()))
        assert np.array_equal(x, xs[0])


def test_tuning_settles(tuner):
    assert tuner.bqd2_autotune_reset() == 0
    assert tuner.bqd2_set_autotune(1) == 0
    n = 2048
    A = matrix(n, 3)
    b = np.random.default_rng(2).standard_normal(n) + 0j
    runs = [lu_bits(A, b) for _ in range(12)]
    x = sample_solution()
    for site in NCAND:
        table = info(site)
        print("site", site, table)
        assert table
        # every class of the LU is called at least twelve times; GMRES classes
        # of short solves may still be exploring
        assert any(chosen >= 0 for _, chosen, _ in table)
        for key, chosen, best in table:
            if site == 1 and chosen < 0:
                continue
            assert 0 <= chosen < NCAND[site], (site, key)
            assert all(t > 0.0 for t in best), (site, key, best)
            assert best[chosen] == max(best)
    # the tuned runs give the default candidate's bits
    for s in NCAND:
        assert tuner.bqd2_autotune_force(s, 0) == 0
    ref = lu_bits(A, b)
    assert all(r == ref for r in runs)
    assert np.array_equal(sample_solution(), x)


def test_mode_off_creates_no_classes(tuner):
    assert tuner.bqd2_autotune_reset() == 0
    assert tuner.bqd2_set_autotune(0) == 0
    lu_bits(matrix(700, 4), np.ones(700, complex))
    assert all(len(info(s)) == 0 for s in NCAND)
    assert tuner.bqd2_set_autotune(2) != 0

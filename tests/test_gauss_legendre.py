"""bqd2_gauss_legendre, the n-point Gauss-Legendre rules of the smooth lattice
remainder: nodes and weights against numpy's leggauss (Golub-Welsch), and the
defining property that the n-point rule integrates x^j exactly for j < 2n
(a dropped or mis-weighted node fails it)."""
import ctypes as C

import numpy as np
import pytest

from bem3_aw_qd2_amd import hip_lib


@pytest.mark.parametrize("n", [1, 2, 3, 4, 5, 6, 7, 8, 12, 16, 33, 64])
def test_gauss_legendre(n):
    x, w = np.zeros(n), np.zeros(n)
    assert hip_lib().bqd2_gauss_legendre(n, x.ctypes.data, w.ctypes.data) == 0
    xr, wr = np.polynomial.legendre.leggauss(n)
    assert np.max(np.abs(x - xr)) <= 2e-15
    # numpy's weights lose digits at large n (1e-12 at n = 64; ours integrate
    # the monomials below to 3e-16): the exactness test is the tight pin
    assert np.max(np.abs(w - wr) / wr) <= 5e-12
    assert (np.diff(x) > 0).all()
    for j in range(min(2 * n, 40)):
        exact = 0.0 if j % 2 else 2.0 / (j + 1)
        assert abs(np.sum(w * x ** j) - exact) <= 1e-14


def test_gauss_legendre_rejects_bad_n():
    x, w = np.zeros(70), np.zeros(70)
    assert hip_lib().bqd2_gauss_legendre(0, x.ctypes.data, w.ctypes.data) != 0
    assert hip_lib().bqd2_gauss_legendre(65, x.ctypes.data, w.ctypes.data) != 0

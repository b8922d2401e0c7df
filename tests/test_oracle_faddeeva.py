"""The oracle's Faddeeva function w(z) (oracle/faddeeva_cpu.c, Weideman's rational
approximation) against the known-answer table of the reference's vendored
Faddeeva package (tests/golden/faddeeva_kat.json), against scipy's wofz, and
against the identities the reference's Green's function relies on."""
import ctypes as C
import json
import math
import os

import numpy as np
import pytest

from tests.checkers import oracle

HERE = os.path.dirname(os.path.abspath(__file__))


def w_oracle(z):
    z = np.ascontiguousarray(np.asarray(z, complex).ravel())
    zz = np.ascontiguousarray(np.stack([z.real, z.imag], axis=1))
    out = np.zeros_like(zz)
    oracle().orc_faddeeva_w_batch(C.c_int(len(z)), zz.ctypes.data_as(C.c_void_p), out.ctypes.data_as(C.c_void_p))
    return out[:, 0] + 1j * out[:, 1]


def test_known_answers():
    kat = json.load(open(os.path.join(HERE, "golden", "faddeeva_kat.json")))["cases"]
    z = np.array([complex(*c["z"]) for c in kat])
    w = np.array([complex(*c["w"]) for c in kat])
    # the table also holds |z| ~ 1e160, far outside the arguments of the lattice
    # sums (|z| < 1e3); the rational form overflows there
    use = np.abs(z) < 1e7
    got = w_oracle(z[use])
    ref = w[use]
    err = np.abs(got - ref) / np.where(ref == 0, 1.0, np.abs(ref))
    assert use.sum() >= 44
    assert err.max() < 3e-15, (z[use][err.argmax()], err.max())


def test_against_scipy_wofz():
    special = pytest.importorskip("scipy.special")
    rng = np.random.default_rng(3)
    z = rng.uniform(-30, 30, 4000) + 1j * rng.uniform(-6, 30, 4000)
    got = w_oracle(z)
    ref = special.wofz(z)
    err = np.abs(got - ref) / np.abs(ref)
    assert err.max() < 5e-14


def test_identities():
    rng = np.random.default_rng(4)
    x = rng.uniform(-8, 8, 500)
    # Re w(x) = exp(-x^2) exactly on the real axis (Faddeeva.cc:684-686)
    assert (w_oracle(x + 0j).real == np.array([math.exp(-v * v) for v in x])).all()
    z = rng.uniform(-5, 5, 500) + 1j * rng.uniform(0.01, 5, 500)
    # w(-z) = 2 exp(-z^2) - w(z)
    lhs = w_oracle(-z)
    rhs = 2 * np.exp(-z * z) - w_oracle(z)
    assert np.abs(lhs - rhs).max() / np.abs(rhs).max() < 1e-13
    # w(-conj z) = conj w(z)
    assert np.abs(w_oracle(-np.conj(z)) - np.conj(w_oracle(z))).max() < 1e-15

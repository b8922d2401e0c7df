"""Field and radiation force of the reference's sample on the GPU.

The model and its surface solution are read from tests/golden/sample_ex.dat (the
reference's ex.dat); example1.c of the reference evaluated the total field at
r = (-0.03, 0.01, 0.002) and the force and torque about the origin for both
quadrature settings and printed them to example1.log (14 digits), transcribed
in tests/golden/sample_example1.json.

Bar: p and v within 1e-11 normwise (the field is a sum over 600 elements of
quasi-periodic Green's function values whose series are summed in a different
order); force within 1e-9 of |F|.  The torque about the sphere's centre is
zero analytically and the printed values are rounding residue (|N| ~ 1e-17
against |F| a ~ 2e-14), so the torque is checked in the scale |F| a."""
import json
import os

import numpy as np
import pytest

from bem3_aw_qd2_amd import exdat, host_lib

pytestmark = pytest.mark.gpu
HERE = os.path.dirname(os.path.abspath(__file__))
SAMPLE = os.path.join(HERE, "golden", "sample_ex.dat")
LOG = json.load(open(os.path.join(HERE, "golden", "sample_example1.json")))
RADIUS = 5.0e-3


@pytest.fixture(scope="module")
def model():
    m = exdat.Model(host_lib()).dat_read(SAMPLE)
    yield m
    m.finalize()


def _c(pair):
    return complex(pair[0], pair[1])


@pytest.mark.parametrize("type_", [0, 1])
def test_total_field_matches_example1(model, type_):
    ref = LOG["field"][str(type_)]
    did, p, v = model.pv_t(LOG["r"], type_)
    print("type", type_, "did", did, "p", p, "v", v)
    assert did == 0
    assert abs(p - _c(ref["p"])) <= 1e-11 * abs(_c(ref["p"]))
    vr = np.array([_c(x) for x in ref["v"]])
    assert np.linalg.norm(v - vr) <= 1e-11 * np.linalg.norm(vr)


@pytest.mark.parametrize("type_", [0, 1])
def test_force_matches_example1(model, type_):
    ref = LOG["force"][str(type_)]
    F, N = model.force(LOG["rc"], type_)
    Fr, Nr = np.array(ref["F"]), np.array(ref["N"])
    print("type", type_, "F", F, "N", N)
    assert np.linalg.norm(F - Fr) <= 1e-9 * np.linalg.norm(Fr)
    assert np.linalg.norm(N - Nr) <= 1e-9 * np.linalg.norm(Fr) * RADIUS


def test_batch_equals_single_points(model):
    rng = np.random.default_rng(7)
    pts = rng.uniform(-0.02, 0.02, size=(64, 3))
    p, v, did = model.pv_batch(pts, 1, 1)
    for i in range(0, 64, 9):
        d1, p1, v1 = model.pv_t(pts[i], 1)
        assert d1 == did[i]
        assert p1 == p[i] and (v1 == v[i]).all()


def test_total_is_scattered_plus_incident(model):
    r = LOG["r"]
    _, pt, vt = model.pv_t(r, 0)
    _, ps, vs = model.pv_s(r, 0)
    _, pi, vi = model.pv_i(r, 0)
    assert abs(pt - (ps + pi)) <= 1e-14 * abs(pt)
    assert np.linalg.norm(vt - (vs + vi)) <= 1e-14 * np.linalg.norm(vt)


def test_domain_ids(model):
    # sphere of radius 5 mm at the origin, lattice 15 mm x 15 mm
    assert model.domain_id([0.0, 0.0, 0.0]) == 1
    assert model.domain_id([0.015, -0.03, 0.001]) == 1   # lattice image of the centre
    assert model.domain_id([0.0, 0.0, 0.0049]) == 1
    assert model.domain_id([0.0, 0.0, 0.0051]) == 0
    assert model.domain_id([0.0075, 0.0075, 0.0]) == 0
    assert model.domain_id([0.0, 0.0, 0.5]) == 0


def test_surface_area(model):
    # 600 quads on a sphere of radius 5 mm: close to 4 pi a^2 and both rules agree
    s0 = model.surface_area(1, 0)
    s1 = model.surface_area(1, 1)
    exact = 4.0 * np.pi * RADIUS ** 2
    assert abs(s0 - exact) < 1e-2 * exact
    assert abs(s1 - s0) < 1e-3 * exact

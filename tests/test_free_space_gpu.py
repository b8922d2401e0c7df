"""The free-space problem (periodicity setting 0): the 0.5 mm sphere of the
lattice workloads alone in water.  Both domains then use the free-space
kernels, the open domain's H diagonal carries 1 + F, and the open domain's node
velocities come from the general recovery (solve_pv_bv,
bem3_aw_qd2_solve_bieq.c:429-491) instead of the continuity conditions of the
periodic case.

Against the oracle, on the 24-quadrilateral mesh (bars as for the lattice
workloads): the GPU's surface solution satisfies the oracle's system rows and
dpv rows of both domains (LU 1e-12, GMRES 1e-9), the oracle recomputes the
node velocities of both domains from it (1e-9), and the field of types 2 and 3
on both sides matches orc_field_s (1e-10).

Against the partial-wave series of the fluid sphere (tests/fluid_sphere.py):
on the cube-sphere meshes of 24, 96 and 384 quadrilaterals the scattered
pressure outside (r = 1.5a ... 4a), the pressure inside (r <= 0.6a) and the
radiation force of the 9-point rule converge to the series: the error falls at
least 3x each time the element size halves (flat bilinear elements, O(h^2):
measured 3.4x to 4.2x), to below 2e-2 (scattered pressure, force) and 1e-2
(interior pressure) on the 384-element mesh.  An error that kept falling by 4x
could only be converging to the series itself.  A sign, a wrong 1 + F, a
missing interface factor or a wrong velocity recovery gives an error of order
one that does not fall."""
import ctypes as C
import os
import tempfile

import numpy as np
import pytest

from bem3_aw_qd2_amd import exdat, hip_lib, host_lib, mesh
from tests.checkers import oracle
from tests.fluid_sphere import FluidSphere

pytestmark = pytest.mark.gpu
R = 0.5e-3


def build(name, tri=False):
    tmp = tempfile.mkdtemp()
    paths = mesh.write_config(name, tmp, triangles=tri, periodic=False)
    return exdat.Model(host_lib()).read_inputs(*paths, out=os.path.join(tmp, "ex.dat")).initialize()


def solve(name, mode=0):
    m = build(name)
    prob = host_lib().aqd2_problem(m.buf, None)
    assert hip_lib().bqd2_set_solver(prob, mode) == 0
    m.solve()
    return m


@pytest.fixture(scope="module")
def small():
    out = {mode: solve("sphere24", mode) for mode in (0, 1)}
    yield out
    for m in out.values():
        m.finalize()


ROWS = [(0, 1, 0), (0, 13, 2), (0, 24, 3), (1, 1, 0), (1, 17, 1)]


@pytest.mark.parametrize("mode,bar", [(1, 1e-12), (0, 1e-9)])
@pytest.mark.parametrize("did,t,tn", ROWS)
def test_solution_holds_oracle_rows(small, mode, bar, did, t, tn):
    o = oracle()
    o.orc_row_residual.restype = C.c_double
    o.orc_dpv_row_residual.restype = C.c_double
    m = small[mode]
    r = o.orc_row_residual(m.buf, did, t, tn, None)
    rd = o.orc_dpv_row_residual(m.buf, did, t, tn)
    print("solver", mode, "row", did, t, tn, "residual %.2e dpv %.2e" % (r, rd))
    assert 0.0 <= r < bar and rd < bar


@pytest.mark.parametrize("did,t,tn", [(0, 2, 0), (0, 19, 3), (1, 7, 1)])
def test_velocity_recomputed(small, did, t, tn):
    m = small[1]
    out = os.path.join(tempfile.mkdtemp(), "gpu.dat")
    m.dat_write(out)
    d = exdat.read(out)
    pv = np.zeros(3, complex)
    assert oracle().orc_node_pv_recompute(m.buf, did, t, tn, pv.ctypes.data_as(C.c_void_p)) == 0
    got = d["dom"][did]["pv"][t, tn]
    print("did", did, "t", t, "tn", tn, got, pv)
    assert np.linalg.norm(got - pv) <= 1e-9 * np.linalg.norm(pv)


@pytest.mark.parametrize("type_", [2, 3])
def test_field_matches_oracle(small, type_):
    m = small[1]
    for x, did in (([0.1e-3, -0.2e-3, 0.15e-3], 1), ([0.9e-3, 0.3e-3, -0.5e-3], 0), ([-0.2e-3, 0.1e-3, 1.6e-3], 0)):
        got_did, p, v = m.pv_s(x, type_)
        po, vo, g = (C.c_double * 2)(), (C.c_double * 6)(), (C.c_double * 3)()
        xa = np.array(x)
        form = oracle().orc_field_s(m.buf, did, xa.ctypes.data_as(C.c_void_p), type_, po, vo, g)
        po = complex(po[0], po[1])
        vo = np.array([complex(vo[2 * i], vo[2 * i + 1]) for i in range(3)])
        print("x", x, "did", got_did, "form", form, "p", p, po)
        assert got_did == did and form >= 0
        assert abs(p - po) <= 1e-10 * abs(po)
        assert np.linalg.norm(v - vo) <= 1e-10 * np.linalg.norm(vo)


def series_for(m):
    """the fluid sphere with the model's incident amplitude and direction"""
    _, p0, _ = m.pv_i([0.0, 0.0, 0.0])
    k0 = 2 * np.pi * 1e6 / 1500.0
    h = 1e-5
    d = np.array([np.angle(m.pv_i(h * e)[1] / p0) / (k0 * h) for e in np.eye(3)])
    return FluidSphere(R, 1000.0, 1500.0, 1200.0, 2400.0, 1e6, p0=p0, d=d)


def sample_points():
    u = np.random.default_rng(11).standard_normal((6, 3))
    u /= np.linalg.norm(u, axis=1)[:, None]
    outside = [f * R * v for f in (1.5, 2.0, 4.0) for v in u]
    inside = [f * R * v for f in (0.0, 0.3, 0.6) for v in u[:3]]
    return outside, inside


def series_errors(name):
    m = solve(name)
    s = series_for(m)
    outside, inside = sample_points()
    ps = np.array([m.pv_s(x, 2)[1] for x in outside])
    pw = np.array([s.scattered(x) for x in outside])
    pi = np.array([m.pv_s(x, 2)[1] for x in inside])
    pj = np.array([s.interior(x) for x in inside])
    assert all(m.domain_id(x) == 0 for x in outside) and all(m.domain_id(x) == 1 for x in inside)
    F, _ = m.force((0.0, 0.0, 0.0), 1)
    Fs = s.force()
    m.finalize()
    return (np.abs(ps - pw).max() / np.abs(pw).max(), np.abs(pi - pj).max() / np.abs(pj).max(),
            np.linalg.norm(F - Fs) / np.linalg.norm(Fs))


def test_converges_to_partial_wave_series():
    errs = {}
    for name in ("sphere24", "sphere96", "sphere384"):
        errs[name] = series_errors(name)
        print(name, "scattered %.2e interior %.2e force %.2e" % errs[name])
    for i in range(3):
        assert errs["sphere24"][i] > 3.0 * errs["sphere96"][i] > 9.0 * errs["sphere384"][i], errs
    e = errs["sphere384"]
    assert e[0] < 2e-2 and e[1] < 1e-2 and e[2] < 2e-2, e

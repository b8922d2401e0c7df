"""Linear-triangle meshes on the GPU against the oracle.

Workload: the 24-quadrilateral cube sphere of radius 0.5 mm split into 48
linear triangles (bem3_aw_qd2_amd.mesh, sphere24 with triangles), on the 2 mm
square lattice at 1 MHz.  No reference output with triangles exists: the
oracle's triangle integrals are pinned by tests/test_oracle_lit.py and its
field sums by tests/test_oracle_field.py.  Bars:

* coefficient rows of both domains (G, H with the F diagonal; dG, dH, dF of the
  sphere's domain) against orc_assemble_row: 1e-10 normwise per row;
* the GPU's surface solution satisfies the oracle's system rows: relative row
  residual below 1e-12 with the dense LU, below 1e-9 with GMRES (which stops at
  a scaled residual of 1e-15 over all rows); the oracle recomputes the GPU's
  node velocities from it (1e-9) and the oracle's dpv rows hold to the same bars;
* scattered field of types 2 and 3 in the sphere (p, v) and of type 2 in the
  lattice domain (p; v close to the surface, where both sides use the
  boundary-value form) against orc_field_s, and the boundary field of types 2
  and 3 against orc_field_bd: 1e-10;
* the radiation force of the 4-point and 7-point rules, which no oracle
  function restates, is checked for consistency: the two rules agree with each
  other and with the force on the quadrilateral mesh of the same sphere to the
  discretization error (printed), and the torque about the centre is small."""
import ctypes as C
import os
import tempfile

import numpy as np
import pytest

from bem3_aw_qd2_amd import AssemblyStats, check, exdat, hip_lib, host_lib, mesh  # noqa: F401
from tests.checkers import oracle
from tests.test_oracle_field import element_point

pytestmark = pytest.mark.gpu
R = 0.5e-3


def build(name, tri):
    tmp = tempfile.mkdtemp()
    paths = mesh.write_config(name, tmp, triangles=tri)
    return exdat.Model(host_lib()).read_inputs(*paths, out=os.path.join(tmp, "ex.dat")).initialize(), tmp


def solve(mode):
    m, tmp = build("sphere24", True)
    prob = host_lib().aqd2_problem(m.buf, None)
    assert hip_lib().bqd2_set_solver(prob, mode) == 0
    m.solve()
    out = os.path.join(tmp, "gpu.dat")
    m.dat_write(out)
    return m, exdat.read(out)


@pytest.fixture(scope="module")
def solved():
    m, d = solve(0)
    yield m, d
    m.finalize()


@pytest.fixture(scope="module")
def solved_lu():
    m, d = solve(1)
    yield m, d
    m.finalize()


def oracle_row(buf, did, t, tn, Ne):
    N = 4 * Ne
    G = np.zeros(N, complex); H = np.zeros(N, complex)
    dG = np.zeros(2 * N, complex); dH = np.zeros(2 * N, complex); dF = np.zeros(3)
    oracle().orc_assemble_row.restype = C.c_long
    ptr = lambda a: a.ctypes.data_as(C.c_void_p)
    assert oracle().orc_assemble_row(buf, did, t, tn, ptr(G), ptr(H), ptr(dG), ptr(dH), ptr(dF)) >= 0
    return G, H, dG, dH, dF


def rel(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


@pytest.mark.parametrize("did", [1, 0])
def test_triangle_rows_match_oracle(did):
    m, _ = build("sphere24", True)
    prob = host_lib().aqd2_problem(m.buf, None)
    st = AssemblyStats()
    check(hip_lib().bqd2_assemble(prob, did, C.byref(st)), "assemble")
    Ne = m.ne()   # one sphere: both domains hold every element
    N = 4 * Ne
    G = np.zeros((N, N), complex); H = np.zeros((N, N), complex); dF = np.zeros((N, 3))
    dG = np.zeros((2 * N, N), complex) if did else None
    dH = np.zeros((2 * N, N), complex) if did else None
    ptr = lambda a: a.ctypes.data_as(C.c_void_p) if a is not None else None
    check(hip_lib().bqd2_download_domain(prob, did, ptr(G), ptr(H), ptr(dG), ptr(dH), ptr(dF)), "download")
    print("domain", did, "pairs GL/GH", st.pairs_gl, st.pairs_gh, "split", st.pairs_split)
    for t, tn in [(1, 0), (7, 2), (30, 1), (Ne, 2)]:
        r = (t - 1) * 4 + tn
        oG, oH, odG, odH, odF = oracle_row(m.buf, did, t, tn, Ne)
        eG, eH = rel(G[r], oG), rel(H[r], oH)
        msg = "did=%d t=%d tn=%d G %.2e H %.2e" % (did, t, tn, eG, eH)
        print(msg)
        assert eG < 1e-10 and eH < 1e-10, msg
        assert abs(dF[r, 0] - odF[0]) <= 1e-10 * abs(odF[0]), msg
        if did:
            eg = rel(dG[2 * r:2 * r + 2], odG.reshape(2, N))
            eh = rel(dH[2 * r:2 * r + 2], odH.reshape(2, N))
            parts = np.zeros(3 * Ne)
            oracle().orc_row_fparts(m.buf, did, t, tn, parts.ctypes.data_as(C.c_void_p))
            scale = np.abs(parts.reshape(Ne, 3)[:, 1:]).sum(axis=0)
            ef = (np.abs(dF[r, 1:] - odF[1:]) / scale).max()
            assert eg < 1e-10 and eh < 1e-10 and ef < 1e-12, msg + " dG %.2e dH %.2e dF %.2e" % (eg, eh, ef)
    m.finalize()


ROWS = [(1, 1, 0), (1, 25, 2), (1, 48, 1), (0, 1, 0), (0, 33, 2)]


def row_residuals(m, did, t, tn):
    o = oracle()
    o.orc_row_residual.restype = C.c_double
    o.orc_dpv_row_residual.restype = C.c_double
    r = o.orc_row_residual(m.buf, did, t, tn, None)
    rd = o.orc_dpv_row_residual(m.buf, did, t, tn)
    print("row residual", did, t, tn, r, "dpv", rd)
    return r, rd


@pytest.mark.parametrize("did,t,tn", ROWS)
def test_triangle_lu_solution_holds_oracle_rows(solved_lu, did, t, tn):
    r, rd = row_residuals(solved_lu[0], did, t, tn)
    assert 0.0 <= r < 1e-12 and rd < 1e-12


@pytest.mark.parametrize("did,t,tn", ROWS)
def test_triangle_gmres_solution_holds_oracle_rows(solved, did, t, tn):
    # GMRES stops at a scaled residual of 1e-15 over all rows; a row whose terms
    # cancel can keep a relative residual of its own up to ~1e-10
    r, rd = row_residuals(solved[0], did, t, tn)
    assert 0.0 <= r < 1e-9 and rd < 1e-9


@pytest.mark.parametrize("t,tn", [(2, 0), (17, 1), (48, 2)])
def test_triangle_velocity_recomputed(solved, t, tn):
    m, d = solved
    pv = np.zeros(3, complex)
    assert oracle().orc_node_pv_recompute(m.buf, 1, t, tn, pv.ctypes.data_as(C.c_void_p)) == 0
    got = d["dom"][1]["pv"][t, tn]
    assert np.linalg.norm(got - pv) <= 1e-9 * np.linalg.norm(pv), (got, pv)


def oracle_field(m, did, x, type_):
    p, v, g = (C.c_double * 2)(), (C.c_double * 6)(), (C.c_double * 3)()
    x = np.ascontiguousarray(x, dtype=np.float64)
    form = oracle().orc_field_s(m.buf, did, x.ctypes.data_as(C.c_void_p), type_, p, v, g)
    assert form >= 0
    return form, complex(p[0], p[1]), np.array([complex(v[2 * i], v[2 * i + 1]) for i in range(3)]), np.array(g[:])


PT = np.array([0.3, 0.5, -0.4]) / np.linalg.norm([0.3, 0.5, -0.4])


@pytest.mark.parametrize("type_", [2, 3])
@pytest.mark.parametrize("frac", [0.0, 0.3, 0.6])
def test_triangle_field_in_sphere(solved, type_, frac):
    m, _ = solved
    x = PT * frac * R
    did, p, v = m.pv_s(x, type_)
    form, po, vo, _ = oracle_field(m, 1, x, type_)
    print("frac", frac, "type", type_, "p", p, po)
    assert did == 1 and form == 0
    assert abs(p - po) <= 1e-10 * abs(po)
    assert np.linalg.norm(v - vo) <= 1e-10 * np.linalg.norm(vo)


def test_triangle_field_in_lattice_domain(solved):
    m, d = solved
    for x in ([0.7e-3, 0.2e-3, 0.3e-3], [-0.4e-3, 0.9e-3, -0.8e-3]):
        did, p, v = m.pv_s(x, 2)
        form, po, _, _ = oracle_field(m, 0, x, 2)
        print("x", x, "p", p, po)
        assert did == 0 and form == 2
        assert abs(p - po) <= 1e-10 * abs(po)
    # just outside an element's centre: both sides take the boundary-value form
    s = abs(d["dom"][0]["sid"][5])
    c = element_point(d, s, 0.0, 0.0)
    x = c * (1.0 + 0.01 * R / np.linalg.norm(c))
    did, p, v = m.pv_s(x, 2)
    form, po, vo, g = oracle_field(m, 0, x, 2)
    print("near surface: |grad F|", np.abs(g).max(), "p", p, po)
    assert did == 0 and np.abs(g).max() > 10.0
    assert abs(p - po) <= 1e-10 * abs(po)
    assert np.linalg.norm(v - vo) <= 1e-10 * np.linalg.norm(vo)


@pytest.mark.parametrize("did", [1, 0])
@pytest.mark.parametrize("type_", [2, 3])
def test_triangle_boundary_field(solved, did, type_):
    m, d = solved
    dom = d["dom"][did]
    for t, z, e in [(1, 0.0, 0.0), (9, 0.3, -0.2), (dom["Ne"], -0.1, 0.4)]:
        p, v = m.pv_t_bd(did, t, z, e, type_)
        po, vo = (C.c_double * 2)(), (C.c_double * 6)()
        assert oracle().orc_field_bd(m.buf, did, t, C.c_double(z), C.c_double(e), type_, po, vo) == 0
        po = complex(po[0], po[1])
        vo = np.array([complex(vo[2 * i], vo[2 * i + 1]) for i in range(3)])
        if did == 0:
            _, pi, vi = m.pv_i(element_point(d, abs(dom["sid"][t]), z, e))
            po, vo = po + pi, vo + vi
        print("did", did, "t", t, "p", p, po)
        assert abs(p - po) <= 1e-10 * abs(po)
        assert np.linalg.norm(v - vo) <= 1e-10 * np.linalg.norm(vo)


def test_triangle_force_consistency():
    out = {}
    for tri in (False, True):
        m, _ = build("sphere96", tri)
        m.solve()
        for type_ in (0, 1):
            out[(tri, type_)] = m.force((0.0, 0.0, 0.0), type_)
        m.finalize()
    Fq = out[(False, 1)][0]
    for key, (F, N) in out.items():
        print(key, "F", F, "N", N, "rel to quad 9-point %.2e" % (np.linalg.norm(F - Fq) / np.linalg.norm(Fq)))
    for type_ in (0, 1):
        F, N = out[(True, type_)]
        assert np.linalg.norm(F - Fq) <= 0.05 * np.linalg.norm(Fq)
        assert np.linalg.norm(N) <= 0.05 * np.linalg.norm(Fq) * R

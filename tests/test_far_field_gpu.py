"""Field points and the boundary elements far below or above them: plane-wave
amplitudes.

For a point of the periodic domain, bqd2_field takes the lattice sums over the
bands of boundary elements at least h below or above it from plane-wave
amplitudes of the solution (the pure spectral series of d3_qpgf_d2.c:484-533,
summed once per order over each band's items), and the other bands per
(point, source) pair.  Checked:

* against the per-pair sums of the same solved problem (bqd2_set_field_far(p, 0))
  at points above and below the spheres, from several periods away down to
  points level with them (only the bands far above or below then take the
  amplitudes), on quads and triangles, square and triangular lattices, oblique
  incidence, rule types 0-3: p and v within 1e-10, point by point;
* the pressure against the oracle's (orc_field_s, rule type 2): 1e-10;
* the amplitudes follow the solution: after a new incident amplitude and a new
  solve, far points equal the per-pair sums again and scale with the amplitude;
* a far point alone gives the bits it gives inside a batch;
* boundary points (bqd2_field_bd, the boundary field behind the type-1 force)
  take the bands far below or above them from the same amplitudes: their field
  and the forces of both rules against the per-pair sums.
"""
import ctypes as C
import math
import os
import tempfile

import numpy as np
import pytest

from bem3_aw_qd2_amd import exdat, hip_lib, host_lib, mesh
from tests.checkers import oracle

pytestmark = pytest.mark.gpu

TRI_LATTICE = dict(d1=(1.8e-3, 0.0), d2=(0.9e-3, 0.9e-3 * math.sqrt(3.0)), theta=math.radians(30.0),
                   psi=math.radians(15.0))
CASES = {
    "quads_square": ("sphere96", False, {}),
    "quads_triangular": ("sphere96", False, TRI_LATTICE),
    "triangles_triangular": ("sphere24", True, TRI_LATTICE),
}


def solved(name):
    cfg, tri, over = CASES[name]
    tmp = tempfile.mkdtemp()
    paths = mesh.write_config(cfg, tmp, triangles=tri, **over)
    m = exdat.Model(host_lib()).read_inputs(*paths, out=os.path.join(tmp, "ex.dat")).initialize()
    m.solve()
    return m, host_lib().aqd2_problem(m.buf, None)


def far_info(prob):
    on, orders = C.c_int(), C.c_int()
    zb, zt, h = C.c_double(), C.c_double(), C.c_double()
    assert hip_lib().bqd2_field_far_info(prob, C.byref(on), C.byref(orders), C.byref(zb), C.byref(zt),
                                         C.byref(h)) == 0
    return on.value, orders.value, zb.value, zt.value, h.value


def field(prob, pts, type_, total=0):
    pts = np.ascontiguousarray(pts, dtype=np.float64)
    n = len(pts)
    out = np.zeros((n, 8))
    did = np.zeros(n, np.int32)
    assert hip_lib().bqd2_field(prob, n, pts.ctypes.data, type_, total, out.ctypes.data, did.ctypes.data, None) == 0
    return out, did


def cplx(out):
    p = out[:, 0] + 1j * out[:, 1]
    v = out[:, 2::2] + 1j * out[:, 3::2]
    return p, v


def points(zb, zt, h, rng, n_each=6):
    """far points above and below (heights from 1.0001 h to several periods),
    then points short of the threshold"""
    far, near = [], []
    for f in (1.0001, 1.3, 2.0, 4.0, 12.0):
        for side in (1, -1):
            z = zt + f * h if side > 0 else zb - f * h
            xy = rng.uniform(-3e-3, 3e-3, (n_each, 2))
            far += [[x, y, z] for x, y in xy]
    for f in (0.5, 0.95):
        for side in (1, -1):
            z = zt + f * h if side > 0 else zb - f * h
            xy = rng.uniform(-3e-3, 3e-3, (3, 2))
            near += [[x, y, z] for x, y in xy]
    # level with the spheres, laterally outside them (the lattice cell's edge)
    for z in (0.0, 0.3 * zt, 0.7 * zb):
        xy = rng.uniform(0.8e-3, 1.0e-3, (3, 2)) * rng.choice([-1, 1], (3, 2))
        near += [[x, y, z] for x, y in xy]
    return np.array(far), np.array(near)


def compare(p1, v1, p0, v0, bar):
    for i in range(len(p0)):
        assert abs(p1[i] - p0[i]) <= bar * abs(p0[i]), (i, p1[i], p0[i])
        assert np.linalg.norm(v1[i] - v0[i]) <= bar * np.linalg.norm(v0[i]), (i, v1[i], v0[i])


@pytest.mark.parametrize("name", sorted(CASES))
def test_far_points_match_per_pair_sums(name):
    m, prob = solved(name)
    hip = hip_lib()
    try:
        on, orders, zb, zt, h = far_info(prob)
        print(name, "far field: on", on, "orders", orders, "z extent", zb, zt, "h", h, flush=True)
        assert on == 1 and 0 < orders <= 32768 and h > 0 and zt > zb
        far, near = points(zb, zt, h, np.random.default_rng(11))
        for type_ in (0, 1, 2, 3):
            print("type", type_, flush=True)
            o1, d1 = field(prob, np.vstack([far, near]), type_)
            assert hip.bqd2_set_field_far(prob, 0) == 0
            o0, d0 = field(prob, np.vstack([far, near]), type_)
            assert hip.bqd2_set_field_far(prob, 1) == 0
            assert (d1 == 0).all() and (d0 == 0).all()
            p1, v1 = cplx(o1)
            p0, v0 = cplx(o0)
            err_p = np.max(np.abs(p1 - p0) / np.abs(p0))
            err_v = np.max(np.linalg.norm(v1 - v0, axis=1) / np.linalg.norm(v0, axis=1))
            print("  max rel err p %.2e v %.2e" % (err_p, err_v), flush=True)
            compare(p1, v1, p0, v0, 1e-10)
        # the far field of the counters: points and amplitude work recorded
        c = (C.c_longlong * 19)()
        assert hip.bqd2_counters(prob, c, 19) == 0
    finally:
        m.finalize()


@pytest.mark.parametrize("name", ["quads_triangular", "triangles_triangular"])
def test_far_pressure_matches_oracle(name):
    m, prob = solved(name)
    try:
        _, _, zb, zt, h = far_info(prob)
        for x in ([0.3e-3, -0.2e-3, zt + 1.1 * h], [-0.7e-3, 0.4e-3, zb - 2.5 * h], [0.1e-3, 0.6e-3, zt + 6 * h]):
            did, p, _ = m.pv_s(x, 2)
            po, vo, g = (C.c_double * 2)(), (C.c_double * 6)(), (C.c_double * 3)()
            xa = np.ascontiguousarray(x, dtype=np.float64)
            form = oracle().orc_field_s(m.buf, 0, xa.ctypes.data_as(C.c_void_p), 2, po, vo, g)
            assert form >= 0 and did == 0
            po = complex(po[0], po[1])
            print("x", x, "p", p, "oracle", po, "rel", abs(p - po) / abs(po), flush=True)
            assert abs(p - po) <= 1e-10 * abs(po)
    finally:
        m.finalize()


def test_amplitudes_follow_the_solution():
    m, prob = solved("quads_triangular")
    hip = hip_lib()
    try:
        _, _, zb, zt, h = far_info(prob)
        far, _ = points(zb, zt, h, np.random.default_rng(5), n_each=2)
        pa, va = cplx(field(prob, far, 0)[0])
        c = TRI_LATTICE
        amp = complex(2.0, 0.5)
        pw7 = (C.c_double * 7)(amp.real, amp.imag, c["theta"], c["psi"], 0.0, 0.0, 0.0)
        assert hip.bqd2_set_incidence(prob, pw7) == 0
        assert hip.bqd2_solve_surface(prob, None) == 0
        assert hip.bqd2_solve_velocity(prob) == 0
        pb, vb = cplx(field(prob, far, 0)[0])
        assert hip.bqd2_set_field_far(prob, 0) == 0
        p0, v0 = cplx(field(prob, far, 0)[0])
        assert hip.bqd2_set_field_far(prob, 1) == 0
        compare(pb, vb, p0, v0, 1e-10)
        compare(pb, vb, amp * pa, amp * va, 1e-10)
    finally:
        m.finalize()


def test_far_point_alone_equals_batch():
    m, prob = solved("quads_square")
    try:
        _, _, zb, zt, h = far_info(prob)
        far, near = points(zb, zt, h, np.random.default_rng(3), n_each=1)
        pts = np.vstack([far, near])
        ob, _ = field(prob, pts, 1, total=1)
        for i in range(0, len(pts), 3):
            oi, _ = field(prob, pts[i:i + 1], 1, total=1)
            assert np.array_equal(oi[0], ob[i])
    finally:
        m.finalize()


@pytest.mark.parametrize("name", sorted(CASES))
def test_boundary_points_match_per_pair_sums(name):
    m, prob = solved(name)
    hip = hip_lib()
    tri = CASES[name][1]
    try:
        ne = m.ne()  # one sphere: domain 0 holds every element
        t = np.arange(1, ne + 1, dtype=np.int32)
        ze = np.ascontiguousarray(np.random.default_rng(7).uniform(-0.9, 0.9, (ne, 2)))
        res = {}
        for far in (1, 0):
            assert hip.bqd2_set_field_far(prob, far) == 0
            out = np.zeros((ne, 8))
            if not tri:  # quadrilateral local coordinates
                assert hip.bqd2_field_bd(prob, 0, ne, t.ctypes.data, ze.ctypes.data, 1, 0, out.ctypes.data) == 0
            FN = {}
            for type_ in (0, 1):
                F, N = (C.c_double * 3)(), (C.c_double * 3)()
                assert hip.bqd2_force(prob, (C.c_double * 3)(0.0, 0.0, 0.0), type_, F, N) == 0
                FN[type_] = (np.array(F[:]), np.array(N[:]))
            res[far] = (out, FN)
        assert hip.bqd2_set_field_far(prob, 1) == 0
        if not tri:
            p1, v1 = cplx(res[1][0])
            p0, v0 = cplx(res[0][0])
            print(name, "boundary points: max rel err p %.2e v %.2e" % (
                np.max(np.abs(p1 - p0) / np.abs(p0)),
                np.max(np.linalg.norm(v1 - v0, axis=1) / np.linalg.norm(v0, axis=1))), flush=True)
            compare(p1, v1, p0, v0, 1e-10)
        for type_ in (0, 1):
            (F1, N1), (F0, N0) = res[1][1][type_], res[0][1][type_]
            eF = np.linalg.norm(F1 - F0) / np.linalg.norm(F0)
            eN = np.linalg.norm(N1 - N0) / (np.linalg.norm(N0) + 1e-3 * np.linalg.norm(F0))
            print(name, "force type", type_, "rel err F %.2e N %.2e" % (eF, eN), flush=True)
            assert eF < 1e-10 and eN < 1e-10
    finally:
        m.finalize()

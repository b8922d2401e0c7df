"""The oracle's linear-triangle element integrals (oracle/lit_cpu.c) pinned by
mathematics.  No reference output with triangles exists, so each part is
checked against an independent evaluation written here in numpy on the
physical triangle (r = cr0 + cr1 z + cr2 e over the equilateral reference
triangle, area element |cr1 x cr2| dz de), not against the kite-shaped
quadrature parts the oracle shares with the reference:

* the interpolation functions are 1 on their own node and 0 on the others;
* the solid angle of a closed triangulated sphere (row sums of F) is 1/2 and
  its tangential derivatives vanish;
* regular pair integrals (free-space and periodic, G, H, F, dG/dv, dH/dv,
  dF/dv) against a collapsed (Duffy) Gauss rule of high order;
* the self element: G against a polar rule of high order around the node, the
  principal-value dG/dv against a central difference of that G in the node
  position, and the periodic self element as the free-space one plus the
  smooth remainder of the lattice sum."""
import ctypes as C
import os
import tempfile

import numpy as np
import pytest

from bem3_aw_qd2_amd import exdat, host_lib, mesh
from tests.checkers import oracle, oracle_eval
from tests.lattices import make_qpd2

SQ3 = np.sqrt(3.0)
CORNERS = np.array([[-0.5, -0.5 * SQ3], [1.0, 0.0], [-0.5, 0.5 * SQ3]])
NODES = 0.4 * CORNERS           # collocation nodes at 2/5 towards the corners
OFF_BD = 736


def lit_N(n, z, e):
    ia = 2.5
    if n == 0:
        return (1.0 - ia * (z + SQ3 * e)) / 3.0
    if n == 1:
        return (1.0 + 2.0 * ia * z) / 3.0
    return (1.0 - ia * (z - SQ3 * e)) / 3.0


@pytest.fixture(scope="module")
def tri_model():
    tmp = tempfile.mkdtemp()
    paths = mesh.write_config("sphere24", tmp, triangles=True)
    m = exdat.Model(host_lib()).read_inputs(*paths, out=os.path.join(tmp, "ex.dat")).initialize()
    out = os.path.join(tmp, "m.dat")
    m.dat_write(out)
    d = exdat.read(out)
    yield m, d
    m.finalize()


def geom(d, s):
    n = d["rn"][d["ed"][s][:3]]
    cr0 = n.sum(axis=0) / 3.0
    cr1 = (-n[0] + 2.0 * n[1] - n[2]) / 3.0
    cr2 = (n[2] - n[0]) / SQ3
    return cr0, cr1, cr2


def duffy(n):
    """points (z, e) and weights of a collapsed Gauss rule on the reference triangle"""
    x, w = np.polynomial.legendre.leggauss(n)
    u, wu = 0.5 * (x + 1), 0.5 * w
    U, V = np.meshgrid(u, u, indexing="ij")
    W = np.outer(wu, wu) * U
    A, B, Cc = CORNERS
    P = A[None, None, :] + U[..., None] * (B - A) + (U * V)[..., None] * (Cc - B)
    jac = abs((B - A)[0] * (Cc - B)[1] - (B - A)[1] * (Cc - B)[0])
    return P[..., 0].ravel(), P[..., 1].ravel(), (W * jac).ravel()


def bd_ptr(m):
    return C.c_void_p(C.addressof(m.buf) + OFF_BD)


def test_interpolation_functions_are_nodal():
    for n in range(3):
        for m in range(3):
            assert abs(lit_N(n, *NODES[m]) - (n == m)) < 1e-15
            assert abs(oracle_N(n, *NODES[m]) - (n == m)) < 1e-15


def oracle_N(n, z, e):
    o = oracle()
    o.orc_lit_N.restype = C.c_double
    return o.orc_lit_N(C.c_int(n), C.c_double(z), C.c_double(e))


@pytest.mark.parametrize("t,tn", [(1, 0), (7, 2), (30, 1)])
def test_solid_angle_of_triangulated_sphere(tri_model, t, tn):
    m, d = tri_model
    ne = m.ne()
    for did, sign in ((1, 1.0), (0, -1.0)):
        parts = np.zeros(3 * ne)
        assert oracle().orc_row_fparts(m.buf, did, t, tn, parts.ctypes.data_as(C.c_void_p)) == 0
        P = parts.reshape(ne, 3)
        assert abs(P[:, 0].sum() - 0.5 * sign) < 1e-12
        assert (np.abs(P[:, 1:].sum(axis=0)) < 1e-13 * np.abs(P[:, 1:]).sum(axis=0)).all()


def brute_pair(cr0, cr1, cr2, rt, v0, v1, k, npts, gf=None):
    """G, H, F and their v0/v1 derivatives (target-position derivatives) of one
    pair by a Duffy rule on the physical triangle"""
    z, e, w = duffy(npts)
    V = np.cross(cr1, cr2)
    aV = np.linalg.norm(V)
    nrm = V / aV
    r = cr0[None] + z[:, None] * cr1[None] + e[:, None] * cr2[None]
    R = r - rt[None]
    aR = np.linalg.norm(R, axis=1)
    dA = w * aV
    if gf is None:
        g = np.exp(1j * k * aR) / (4 * np.pi * aR)
        dg = R * ((1j * k * aR - 1) * np.exp(1j * k * aR) / (4 * np.pi * aR ** 3))[:, None]   # d/dR
    else:
        g, dg = gf(R)
    out = {}
    Ns = np.array([lit_N(n, z, e) for n in range(3)])
    out["G"] = (Ns * g[None] * dA[None]).sum(axis=1)
    out["H"] = (Ns * (dg @ nrm)[None] * dA[None]).sum(axis=1)
    out["F"] = np.sum((R @ nrm) / aR ** 3 * dA) / (4 * np.pi)
    return out, (z, e, w, R, aR, nrm, dA, Ns)


def call_pair(m, periodic, rt, v0, v1, s, k, n):
    o = oracle()
    x, w = np.polynomial.legendre.leggauss(n)
    CC = np.zeros(9, complex)
    d0 = np.zeros(9, complex)
    d1 = np.zeros(9, complex)
    a = lambda v: np.ascontiguousarray(v, dtype=np.float64).ctypes.data_as(C.c_void_p)
    if periodic:
        o.orc_q0_lit_pair.restype = C.c_long
        r = o.orc_q0_lit_pair(bd_ptr(m), a(rt), a(v0), a(v1), C.c_int(s), C.c_int(n), a(x), a(w),
                              CC.ctypes.data_as(C.c_void_p), d0.ctypes.data_as(C.c_void_p),
                              d1.ctypes.data_as(C.c_void_p))
        assert r > 0
    else:
        o.orc_lit_pair(bd_ptr(m), a(rt), a(v0), a(v1), C.c_int(s), C.c_double(k), C.c_int(n), a(x), a(w),
                       CC.ctypes.data_as(C.c_void_p), d0.ctypes.data_as(C.c_void_p), d1.ctypes.data_as(C.c_void_p))
    return CC, d0, d1


def target_near(d, s, dist, along=(0.3, 0.2)):
    cr0, cr1, cr2 = geom(d, s)
    V = np.cross(cr1, cr2)
    nrm = V / np.linalg.norm(V)
    h = np.linalg.norm(cr1)
    rt = cr0 + along[0] * cr1 + along[1] * cr2 + dist * h * nrm
    t0 = cr1 / np.linalg.norm(cr1)
    t1 = np.cross(nrm, t0)
    return rt, t0, t1


@pytest.mark.parametrize("dist,n", [(2.0, 12), (0.4, 64)])
def test_free_pair_against_duffy_rule(tri_model, dist, n):
    m, d = tri_model
    k = 2 * np.pi * 1e6 / 2400.0
    s = 5
    cr0, cr1, cr2 = geom(d, s)
    rt, v0, v1 = target_near(d, s, dist)
    CC, d0, d1 = call_pair(m, False, rt, v0, v1, s, k, n)
    ref, _ = brute_pair(cr0, cr1, cr2, rt, v0, v1, k, 300)
    assert np.abs(CC[:3] - ref["G"]).max() < 1e-10 * np.abs(ref["G"]).max()
    assert np.abs(CC[4:7] - ref["H"]).max() < 1e-10 * np.abs(ref["H"]).max()
    assert abs(CC[8].real - ref["F"]) < 1e-10 * abs(ref["F"])
    # target derivatives by central difference of the brute-force integrals
    hs = 1e-6 * np.linalg.norm(cr1)
    for dv, v in ((d0, v0), (d1, v1)):
        rp, _ = brute_pair(cr0, cr1, cr2, rt + hs * v, v0, v1, k, 300)
        rm, _ = brute_pair(cr0, cr1, cr2, rt - hs * v, v0, v1, k, 300)
        fdG = (rp["G"] - rm["G"]) / (2 * hs)
        fdH = (rp["H"] - rm["H"]) / (2 * hs)
        fdF = (rp["F"] - rm["F"]) / (2 * hs)
        assert np.abs(dv[:3] - fdG).max() < 1e-6 * np.abs(fdG).max()
        assert np.abs(dv[4:7] - fdH).max() < 1e-6 * np.abs(fdH).max()
        assert abs(dv[8].real - fdF) < 1e-6 * abs(fdF) + 1e-9


def test_periodic_pair_against_duffy_rule(tri_model):
    m, d = tri_model
    q = make_qpd2(2 * np.pi * 1e6 / 1500.0, 1e-6, np.pi / 4, (2e-3, 0.0), (0.0, 2e-3))
    s = 9
    cr0, cr1, cr2 = geom(d, s)
    rt, v0, v1 = target_near(d, s, 1.5)

    def gf(R):
        G, dG, _, t = oracle_eval(2, q, R)
        assert (t > 0).all()
        return G, dG

    CC, d0, d1 = call_pair(m, True, rt, v0, v1, s, 0.0, 12)
    ref, _ = brute_pair(cr0, cr1, cr2, rt, v0, v1, 0.0, 120, gf=gf)
    assert np.abs(CC[:3] - ref["G"]).max() < 1e-10 * np.abs(ref["G"]).max()
    assert np.abs(CC[4:7] - ref["H"]).max() < 1e-10 * np.abs(ref["H"]).max()
    assert abs(CC[8].real - ref["F"]) < 1e-10 * abs(ref["F"])


def polar_self(cr1, cr2, zt, et, f, nth=240, nr=80):
    """integral over the reference triangle of f(R, z, e) dA by a polar rule about
    (zt, et): three sub-triangles to the edges, Gauss in theta and r"""
    aV = np.linalg.norm(np.cross(cr1, cr2))
    xt, wt = np.polynomial.legendre.leggauss(nth)
    xr, wr = np.polynomial.legendre.leggauss(nr)
    P = np.array([zt, et])
    total = 0.0
    for j in range(3):
        A, B = CORNERS[j], CORNERS[(j + 1) % 3]
        ta, tb = np.arctan2(*(A - P)[::-1]), np.arctan2(*(B - P)[::-1])
        if tb < ta:
            tb += 2 * np.pi
        th = 0.5 * (tb - ta) * xt + 0.5 * (tb + ta)
        c, s = np.cos(th), np.sin(th)
        # distance from P along (c, s) to the edge AB
        ab = B - A
        den = c * ab[1] - s * ab[0]
        rm = ((A[0] - P[0]) * ab[1] - (A[1] - P[1]) * ab[0]) / den
        rr = 0.5 * rm[:, None] * (xr[None] + 1)
        z = P[0] + rr * c[:, None]
        e = P[1] + rr * s[:, None]
        K = c[:, None, None] * cr1[None, None] + s[:, None, None] * cr2[None, None]
        R = rr[..., None] * K
        val = f(R, z, e)
        total = total + np.sum(val * rr * (0.5 * rm[:, None]) * wr[None] * (0.5 * (tb - ta)) * wt[:, None])
    return total * aV


def call_self(m, periodic, s, zt, et, k):
    o = oracle()
    CC = np.zeros(9, complex)
    if periodic:
        o.orc_q0_lit_self.restype = C.c_long
        assert o.orc_q0_lit_self(bd_ptr(m), C.c_int(s), C.c_double(zt), C.c_double(et),
                                 CC.ctypes.data_as(C.c_void_p)) > 0
    else:
        o.orc_lit_self(bd_ptr(m), C.c_int(s), C.c_double(zt), C.c_double(et), C.c_double(k),
                       CC.ctypes.data_as(C.c_void_p))
    return CC


@pytest.mark.parametrize("node", [0, 1, 2])
def test_free_self_against_polar_rule(tri_model, node):
    m, d = tri_model
    k = 2 * np.pi * 1e6 / 2400.0
    s = 11
    cr0, cr1, cr2 = geom(d, s)
    zt, et = NODES[node]
    CC = call_self(m, False, s, zt, et, k)
    for n in range(3):
        ref = polar_self(cr1, cr2, zt, et,
                         lambda R, z, e: np.exp(1j * k * np.linalg.norm(R, axis=-1)) /
                         (4 * np.pi * np.linalg.norm(R, axis=-1)) * lit_N(n, z, e))
        assert abs(CC[n] - ref) < 1e-10 * abs(ref)
    assert np.abs(CC[4:9]).max() == 0.0


@pytest.mark.parametrize("node", [0, 2])
def test_free_self_pv_derivative_by_central_difference(tri_model, node):
    """dG/dv of the self element is the derivative of its G in the node position
    along the unit tangent (cr1 or cr2 direction)"""
    m, d = tri_model
    k = 2 * np.pi * 1e6 / 2400.0
    s = 11
    cr0, cr1, cr2 = geom(d, s)
    zt, et = NODES[node]
    o = oracle()
    for direction, cr in ((0, cr1), (1, cr2)):
        dC = np.zeros(9, complex)
        o.orc_lit_self_dPV(bd_ptr(m), C.c_int(s), C.c_double(zt), C.c_double(et), C.c_double(k), C.c_int(direction),
                           dC.ctypes.data_as(C.c_void_p))
        h = 1e-4
        dz, de = (h, 0.0) if direction == 0 else (0.0, h)
        Gp = call_self(m, False, s, zt + dz, et + de, k)
        Gm = call_self(m, False, s, zt - dz, et - de, k)
        fd = (Gp[:3] - Gm[:3]) / (2 * h) / np.linalg.norm(cr)
        # agreement to 2-3e-4: the oracle restates the reference's split of the
        # principal value (D_bar on r > DEPS by Gauss rules, the disc r < DEPS by
        # its closed form, lit_dcoef.c:461-637); halving the closed form's dN/dr
        # worsens the agreement to 1e-1, so the terms are weighted as intended
        # and the residual is the scheme's own
        assert np.abs(dC[:3] - fd).max() < 1e-3 * np.abs(fd).max(), (dC[:3], fd)


def test_periodic_self_is_free_self_plus_smooth_remainder(tri_model):
    m, d = tri_model
    q = make_qpd2(2 * np.pi * 1e6 / 1500.0, 1e-6, np.pi / 4, (2e-3, 0.0), (0.0, 2e-3))
    s = 3
    cr0, cr1, cr2 = geom(d, s)
    zt, et = NODES[1]
    V = np.cross(cr1, cr2)
    nrm = V / np.linalg.norm(V)
    CC = call_self(m, True, s, zt, et, 0.0)
    free = call_self(m, False, s, zt, et, q.k)

    def rem(R, z, e, n, which):
        flat = R.reshape(-1, 3)
        G, dG, _, t = oracle_eval(2, q, flat)
        aR = np.linalg.norm(flat, axis=1)
        g0 = np.exp(1j * q.k * aR) / (4 * np.pi * aR)
        dg0 = flat * ((1j * q.k * aR - 1) * np.exp(1j * q.k * aR) / (4 * np.pi * aR ** 3))[:, None]
        if which == "G":
            v = G - g0
        else:
            v = (dG - dg0) @ nrm
        return v.reshape(R.shape[:-1]) * lit_N(n, z, e)

    for n in range(3):
        gs = polar_self(cr1, cr2, zt, et, lambda R, z, e: rem(R, z, e, n, "G"), nth=60, nr=24)
        hs = polar_self(cr1, cr2, zt, et, lambda R, z, e: rem(R, z, e, n, "H"), nth=60, nr=24)
        assert abs(CC[n] - (free[n] + gs)) < 1e-9 * abs(CC[n])
        assert abs(CC[n + 4] - hs) < 1e-8 * abs(hs) + 1e-12 * abs(CC[n])

"""Parity at the sizes of the benchmark configurations, GPU against the oracle.

Workloads (bem3_aw_qd2_amd.mesh):
* sphere384: 384 quads, 0.5 mm sphere on the 2 mm square lattice at 1 MHz,
  theta = 1e-6 (the CPU-runnable configuration);
* sphere3456: 3456 quads on the 1.8 mm triangular lattice, theta = 30 deg,
  phi = 15 deg, and the same sphere split into 6912 linear triangles.

For each: the GPU assembles and solves; sampled rows of both domains' G, H
(and, in the sphere's domain, dG, dH, dF) against orc_assemble_row at 1e-10
normwise per row.  The sampled rows include targets on the side faces and the
poles of the cube sphere, whose nodes lie exactly level with nodes of other
elements (exact z == 0 offsets, where the lattice sums stop on a rounding
residue), and every row includes the neighbouring elements that take the GH64
rule or its split.  The GPU's surface solution is put into the oracle's system
rows (A x - b relative to |A||x|: below 1e-9 with GMRES, which stops at a
scaled residual of 1e-15 over all rows).  The type-0 force is compared with
orc_force on the same solution (1e-10) and the boundary field that the type-1
force integrates with orc_field_bd_total at sampled 9- or 7-point rule points
(1e-10)."""
import ctypes as C
import os
import tempfile
import time

import numpy as np
import pytest

from bem3_aw_qd2_amd import exdat, hip_lib, host_lib, mesh
from tests.checkers import oracle

pytestmark = pytest.mark.gpu
CASES = {
    "sphere384": ("sphere384", False),
    "sphere3456": ("sphere3456", False),
    "sphere3456_tri": ("sphere3456", True),
}


def ptr(a):
    return a.ctypes.data_as(C.c_void_p) if a is not None else None


def level_rows(d, did, count, seed=0):
    """rows (t, tn) of domain did whose node has exactly the z of some node of
    another element, spread over the mesh, plus a node on each pole face"""
    dom = d["dom"][did]
    ne = dom["Ne"]
    nl = 3 if d["ed"][abs(dom["sid"][1])][3] == 0 else 4
    zs = {}
    for s in range(1, ne + 1):
        for n in range(nl):
            zs.setdefault(d["ren"][abs(dom["sid"][s])][n][2], set()).add(s)
    rng = np.random.default_rng(seed)
    level = []
    for t in rng.permutation(np.arange(1, ne + 1)):
        for tn in range(nl):
            if len(zs[d["ren"][abs(dom["sid"][t])][tn][2]] - {t}) > 0:
                level.append((int(t), tn))
                break
        if len(level) >= count:
            break
    # the last face of the cube sphere is a pole face
    level.append((ne, nl - 1))
    return level


def field_points(d, count, seed=1):
    rng = np.random.default_rng(seed)
    ne = d["dom"][0]["Ne"]
    tri = d["ed"][abs(d["dom"][0]["sid"][1])][3] == 0
    return [(int(rng.integers(1, ne + 1)), int(rng.integers(0, 7 if tri else 9))) for _ in range(count)], tri


@pytest.fixture(scope="module", params=sorted(CASES))
def solved(request):
    name, tri = CASES[request.param]
    tmp = tempfile.mkdtemp()
    paths = mesh.write_config(name, tmp, triangles=tri)
    m = exdat.Model(host_lib()).read_inputs(*paths, out=os.path.join(tmp, "ex.dat")).initialize()
    t0 = time.time()
    m.solve()
    print("%s: solved on the GPU in %.1f s" % (request.param, time.time() - t0), flush=True)
    out = os.path.join(tmp, "gpu.dat")
    m.dat_write(out)
    d = exdat.read(out)
    yield request.param, m, d
    m.finalize()


def test_rows_and_solution(solved):
    case, m, d = solved
    o = oracle()
    o.orc_assemble_row.restype = C.c_long
    o.orc_set_cr_exp(C.c_int(1))
    prob = host_lib().aqd2_problem(m.buf, None)
    na = 2 * m.nn()
    x = np.zeros(na, complex)
    o.orc_solution_vector(m.buf, ptr(x))
    nrows = {"sphere384": 4, "sphere3456": 2, "sphere3456_tri": 1}[case]
    try:
        for did in (0, 1):
            rows = level_rows(d, did, nrows, seed=did)
            Ne = d["dom"][did]["Ne"]
            N = 4 * Ne
            idx = np.array([(t - 1) * 4 + tn for t, tn in rows], np.int32)
            G = np.zeros((len(rows), N), complex)
            H = np.zeros((len(rows), N), complex)
            dG = np.zeros((len(rows), 2 * N), complex) if did else None
            dH = np.zeros((len(rows), 2 * N), complex) if did else None
            dF = np.zeros((len(rows), 3))
            assert hip_lib().bqd2_download_rows(prob, did, len(rows), ptr(idx), ptr(G), ptr(H), ptr(dG), ptr(dH),
                                                ptr(dF)) == 0
            for i, (t, tn) in enumerate(rows):
                oG, oH = np.zeros(N, complex), np.zeros(N, complex)
                odG = np.zeros(2 * N, complex) if did else None
                odH = np.zeros(2 * N, complex) if did else None
                odF = np.zeros(3)
                t0 = time.time()
                assert o.orc_assemble_row(m.buf, did, t, tn, ptr(oG), ptr(oH), ptr(odG), ptr(odH), ptr(odF)) >= 0
                eG = np.linalg.norm(G[i] - oG) / np.linalg.norm(oG)
                eH = np.linalg.norm(H[i] - oH) / np.linalg.norm(oH)
                msg = "%s did %d row (%d, %d): G %.2e H %.2e (oracle %.1f s)" % (case, did, t, tn, eG, eH,
                                                                                 time.time() - t0)
                if did:
                    eg = np.linalg.norm(dG[i] - odG) / np.linalg.norm(odG)
                    eh = np.linalg.norm(dH[i] - odH) / np.linalg.norm(odH)
                    msg += " dG %.2e dH %.2e" % (eg, eh)
                    assert eg < 1e-10 and eh < 1e-10, msg
                print(msg, flush=True)
                assert eG < 1e-10 and eH < 1e-10, msg
                assert abs(dF[i, 0] - odF[0]) <= 1e-10 * abs(odF[0]), msg
                # the GPU solution in the oracle's system row
                A = np.zeros(na, complex)
                b = np.zeros(1, complex)
                o.orc_system_row(m.buf, did, ptr(oG), ptr(oH), ptr(A), ptr(b))
                res = abs(A @ x - b[0]) / (np.abs(A) @ np.abs(x))
                print("  system row residual %.2e" % res, flush=True)
                assert res < 1e-9, msg
    finally:
        o.orc_set_cr_exp(C.c_int(0))


def test_force_and_boundary_field(solved):
    from concurrent.futures import ThreadPoolExecutor
    case, m, d = solved
    o = oracle()
    rc = np.zeros(3)
    F, N = m.force(rc, 0)
    Fo, No = np.zeros(3), np.zeros(3)
    assert o.orc_force(m.buf, ptr(rc), 0, ptr(Fo), ptr(No)) == 0
    eF = np.linalg.norm(F - Fo) / np.linalg.norm(Fo)
    eN = np.linalg.norm(N - No) / max(np.linalg.norm(No), np.linalg.norm(Fo) * 0.5e-3)
    print(case, "type-0 force F %.2e N %.2e" % (eF, eN))
    assert eF < 1e-10 and eN < 1e-10
    pts, tri = field_points(d, {"sphere384": 8, "sphere3456": 4, "sphere3456_tri": 3}[case])
    off = 736 + (592 if tri else 280)          # BOUD.zt_37 / zt_49 in AQD2
    nq = 7 if tri else 9
    zt = np.array(C.cast(C.addressof(m.buf) + off, C.POINTER(C.c_double * nq)).contents[:])
    et = np.array(C.cast(C.addressof(m.buf) + off + 8 * nq, C.POINTER(C.c_double * nq)).contents[:])
    gpu = [m.pv_t_bd(0, t, zt[q], et[q], 1) for t, q in pts]
    o.orc_set_cr_exp(C.c_int(1))

    def one(tq):
        t, q = tq
        po, vo = (C.c_double * 2)(), (C.c_double * 6)()
        assert o.orc_field_bd_total(m.buf, t, C.c_double(zt[q]), C.c_double(et[q]), 1, po, vo) == 0
        return complex(po[0], po[1]), np.array([complex(vo[2 * i], vo[2 * i + 1]) for i in range(3)])

    try:
        with ThreadPoolExecutor(max_workers=len(pts)) as ex:
            ref = list(ex.map(one, pts))
    finally:
        o.orc_set_cr_exp(C.c_int(0))
    for (t, q), (p, v), (po, vo) in zip(pts, gpu, ref):
        ep, ev = abs(p - po) / abs(po), np.linalg.norm(v - vo) / np.linalg.norm(vo)
        print(case, "boundary point t %d q %d: p %.2e v %.2e" % (t, q, ep, ev))
        assert ep < 1e-10 and ev < 1e-10

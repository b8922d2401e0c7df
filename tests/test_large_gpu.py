"""The system layout (bqd2_create_large) against the default layout and the
reference's own outputs, on the reference's shipped sample (600 quads on a
periodic lattice: one sphere, each element with its own four nodes, which is
the layout's precondition).

* Rows: the lattice domain's rows written straight into the 2N x 2N array and
  the sphere's rows assembled window by window give the same bits as the
  default layout's separate matrices (G, H of both domains, dF).
* Solutions: every solver of the layout (FP64 GMRES on the array as an
  operator; the complex float LU in the memory of the sphere's half with the
  sphere's rows partly resident and partly assembled again in every product of
  the refinement; the FP64 LU of the array in place, after which the velocity
  recovery assembles domain 0 again) gives P and dP within 1e-10 of the
  reference's ex.dat, pv and dpv within 1e-8, the surface solution within 1e-10
  of the default layout's with the same solver, and both force rules within
  1e-9 of the reference's example1.log (the bars of the mixed-precision test of
  the default layout)."""
import ctypes as C
import json
import os

import numpy as np
import pytest

from bem3_aw_qd2_amd import LargeStats, check, exdat, hip_lib, host_lib

pytestmark = pytest.mark.gpu
HERE = os.path.dirname(os.path.abspath(__file__))
SAMPLE = os.path.join(HERE, "golden", "sample_ex.dat")
LOG = json.load(open(os.path.join(HERE, "golden", "sample_example1.json")))


def problem(large, lwin=256, max_resident=-1):
    hip, host = hip_lib(), host_lib()
    m = exdat.Model(host).dat_read(SAMPLE)
    mp = host.aqd2_model_new(m.buf)
    p = C.c_void_p()
    check((hip.bqd2_create_large if large else hip.bqd2_create)(mp, C.byref(p)), "bqd2_create")
    host.aqd2_model_free(mp)
    if large:
        check(hip.bqd2_large_config(p, lwin, max_resident), "bqd2_large_config")
    return m, p


def solve(p, solver, ne, nn):
    hip = hip_lib()
    check(hip.bqd2_set_solver(p, solver), "bqd2_set_solver")
    for d in (0, 1):
        check(hip.bqd2_assemble(p, d, None), "bqd2_assemble")
    x = np.zeros(2 * nn, complex)
    check(hip.bqd2_solve_surface(p, x.ctypes.data), "bqd2_solve_surface")
    check(hip.bqd2_solve_velocity(p), "bqd2_solve_velocity")
    sol = []
    for d in (0, 1):
        P = np.zeros((ne + 1, 4), complex)
        dP = np.zeros((ne + 1, 4), complex)
        pv = np.zeros((ne + 1, 4, 3), complex)
        dpv = np.zeros((ne + 1, 4, 3), complex)
        check(hip.bqd2_download_solution(p, d, P.ctypes.data, dP.ctypes.data, pv.ctypes.data, dpv.ctypes.data),
              "bqd2_download_solution")
        sol.append({"P": P, "dP": dP, "pv": pv, "dpv": dpv})
    forces = {}
    rc = (C.c_double * 3)(*LOG["rc"])
    for t in (0, 1):
        F, N = (C.c_double * 3)(), (C.c_double * 3)()
        check(hip.bqd2_force(p, rc, t, F, N), "bqd2_force")
        forces[t] = np.array(F[:])
    return x, sol, forces


def test_rows_bit_identical():
    hip = hip_lib()
    m0, p0 = problem(False)
    m1, p1 = problem(True, lwin=256)
    ne, nn = m0.ne(), m0.nn()
    N = 4 * ne
    try:
        for p in (p0, p1):
            check(hip.bqd2_set_solver(p, 0), "bqd2_set_solver")
            for d in (0, 1):
                check(hip.bqd2_assemble(p, d, None), "bqd2_assemble")
        # the layout assembles the sphere's rows into its array for FP64 GMRES
        check(hip.bqd2_solve_surface(p1, None), "bqd2_solve_surface")
        rows = np.array([0, 1, 5, 255, 256, 257, 1023, 1500, 2047, 2048, N - 2, N - 1], np.int32)
        for d in (0, 1):
            out = []
            for p in (p0, p1):
                G = np.zeros((len(rows), N), complex)
                H = np.zeros((len(rows), N), complex)
                dF = np.zeros((len(rows), 3))
                check(hip.bqd2_download_rows(p, d, len(rows), rows.ctypes.data, G.ctypes.data, H.ctypes.data,
                                             None, None, dF.ctypes.data), "bqd2_download_rows")
                out.append((G, H, dF))
            for name, a, b in zip(("G", "H", "dF"), out[0], out[1]):
                print("domain %d %s: max |default - layout| %.3e" % (d, name, np.abs(a - b).max()))
                assert np.array_equal(a, b)
    finally:
        hip.bqd2_destroy(p0)
        hip.bqd2_destroy(p1)
        m0.finalize()
        m1.finalize()


@pytest.mark.parametrize("solver", [0, 2, 1])
def test_layout_solvers_match_reference(solver):
    hip = hip_lib()
    ref = exdat.read(SAMPLE)
    # solver 2: 10 windows of 256 rows, 3 of them resident, the others assembled
    # again in every product of the refinement
    m1, p1 = problem(True, lwin=256, max_resident=3)
    m0, p0 = problem(False)
    ne, nn = m0.ne(), m0.nn()
    try:
        x1, s1, f1 = solve(p1, solver, ne, nn)
        st = LargeStats()
        check(hip.bqd2_large_info(p1, C.byref(st)), "bqd2_large_info")
        print("solver %d layout stats %s" % (solver, st.as_dict()))
        x0, s0, f0 = solve(p0, solver, ne, nn)
    finally:
        hip.bqd2_destroy(p0)
        hip.bqd2_destroy(p1)
        m0.finalize()
        m1.finalize()
    if solver == 2:
        assert st.resident == 3 and st.recomputed > 0
    if solver == 1:
        assert st.ms_reassemble > 0.0
    ex = np.linalg.norm(x1 - x0) / np.linalg.norm(x0)
    print("solver %d: surface solution, layout vs default %.2e" % (solver, ex))
    assert ex < 1e-10
    for d in (0, 1):
        for name, bar in (("P", 1e-10), ("dP", 1e-10), ("pv", 1e-8), ("dpv", 1e-8)):
            e = exdat.relerr(s1[d][name], ref["dom"][d][name])
            e0 = exdat.relerr(s1[d][name], s0[d][name])
            print("domain %d %-3s layout-ref %.2e layout-default %.2e" % (d, name, e, e0))
            assert e < bar and e0 < bar
    for t in (0, 1):
        Fr = np.array(LOG["force"][str(t)]["F"])
        e = np.linalg.norm(f1[t] - Fr) / np.linalg.norm(Fr)
        e0 = np.linalg.norm(f1[t] - f0[t]) / np.linalg.norm(f0[t])
        print("force type %d: layout-log %.2e layout-default %.2e" % (t, e, e0))
        assert e <= 1e-9 and e0 <= 1e-9

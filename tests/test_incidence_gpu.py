"""Several incident plane waves on one assembled operator (bqd2_set_incidence).

The periodic operator depends on the incidence only through the in-plane wave
vector modulo the reciprocal lattice.  On the 96-quad sphere lattice (2 mm
square, 1 MHz, base incidence theta = 1e-6, psi = pi/4) two other incidences
share it: the mirrored direction pi - theta, and the direction of the (1, 0)
diffraction order, whose in-plane wave vector is the base one plus g1
(|k_par + g1| = 3142 < k = 4189, a propagating order).  For each, the solution
on the base problem after bqd2_set_incidence must equal the solution of a
problem created and assembled with that incidence: surface P and dP within
1e-10 normwise, force within 1e-9, for the FP64 GMRES solver, the FP64 LU
(whose factors are kept: a later incidence rebuilds the right-hand side only)
and the mixed-precision solver (whose factors are reused, not recomputed).  An
incidence with another in-plane wave vector is rejected."""
import ctypes as C
import math
import os
import tempfile

import numpy as np
import pytest

from bem3_aw_qd2_amd import exdat, hip_lib, host_lib, mesh

pytestmark = pytest.mark.gpu
CFG = mesh.CONFIGS["sphere96"]
K = 2 * math.pi * CFG["f"] / CFG["host"][1]
TH0, PS0 = CFG["theta"], CFG["psi"]


def incidences():
    kx, ky = K * math.sin(TH0) * math.cos(PS0), K * math.sin(TH0) * math.sin(PS0)
    g1 = 2 * math.pi / CFG["d1"][0]
    kx1 = kx + g1
    th1 = math.asin(math.hypot(kx1, ky) / K)
    return {"mirrored": (math.pi - TH0, PS0), "order10": (th1, math.atan2(ky, kx1))}


def model(theta, psi, tmp):
    paths = mesh.write_config("sphere96", tmp, theta=theta, psi=psi)
    m = exdat.Model(host_lib()).read_inputs(*paths, out=os.path.join(tmp, "ex.dat")).initialize()
    return m, host_lib().aqd2_problem(m.buf, None)


def results(prob, mn):
    hip = hip_lib()
    out = {}
    for d in range(mn + 1):
        ne = 96
        P = np.zeros((ne + 1) * 4, complex)
        dP = np.zeros((ne + 1) * 4, complex)
        pv = np.zeros((ne + 1) * 12, complex)
        dpv = np.zeros((ne + 1) * 12, complex)
        assert hip.bqd2_download_solution(prob, d, P.ctypes.data, dP.ctypes.data, pv.ctypes.data,
                                          dpv.ctypes.data) == 0
        out[d] = (P[4:], dP[4:])
    F = (C.c_double * 3)()
    N = (C.c_double * 3)()
    rc = (C.c_double * 3)(0.0, 0.0, 0.0)
    assert hip.bqd2_force(prob, rc, 0, F, N) == 0
    out["F"] = np.array(F[:])
    return out


def rel(a, b):
    return np.linalg.norm(a - b) / np.linalg.norm(b)


@pytest.mark.parametrize("mode", [0, 1, 2])
def test_incidences_share_the_operator(mode):
    hip = hip_lib()
    with tempfile.TemporaryDirectory() as tmp:
        base, prob = model(TH0, PS0, os.path.join(tmp, "base"))
        assert hip.bqd2_set_solver(prob, mode) == 0
        base.solve()
        ms0 = C.c_double()
        assert hip.bqd2_lu32_info(prob, C.byref(ms0)) == 0
        for name, (th, ps) in incidences().items():
            pw7 = (C.c_double * 7)(1.0, 0.0, th, ps, 0.0, 0.0, 0.0)
            assert hip.bqd2_set_incidence(prob, pw7) == 0
            assert hip.bqd2_solve_surface(prob, None) == 0
            assert hip.bqd2_solve_velocity(prob) == 0
            got = results(prob, base.mn())
            ms = C.c_double()
            assert hip.bqd2_lu32_info(prob, C.byref(ms)) == 0
            assert ms.value == ms0.value  # the factors were not recomputed
            fresh, prob2 = model(th, ps, os.path.join(tmp, name))
            assert hip.bqd2_set_solver(prob2, mode) == 0
            fresh.solve()
            ref = results(prob2, fresh.mn())
            fresh.finalize()
            for d in range(base.mn() + 1):
                eP, edP = rel(got[d][0], ref[d][0]), rel(got[d][1], ref[d][1])
                print("mode %d %s domain %d: P %.2e dP %.2e" % (mode, name, d, eP, edP))
                assert eP < 1e-10 and edP < 1e-10
            eF = rel(got["F"], ref["F"])
            print("mode %d %s force %.2e" % (mode, name, eF))
            assert eF < 1e-9
        bad = (C.c_double * 7)(1.0, 0.0, 0.3, 0.2, 0.0, 0.0, 0.0)
        assert hip.bqd2_set_incidence(prob, bad) == -1
        assert b"reciprocal lattice" in hip.bqd2_last_error()
        base.finalize()

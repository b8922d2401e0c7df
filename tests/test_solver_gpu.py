"""The two solvers of the boundary-value system and of the dpv systems on the
GPU: GMRES with Jacobi row scaling (the default) and the dense LU
(bqd2_set_solver(p, 1)), both against the reference's ex.dat solution of its
shipped sample and against each other.

Bars: P and dP within 1e-10 normwise of ex.dat for both solvers (the
reference solved with MKL PARDISO; GMRES stops at a scaled residual of 1e-15 on
a system whose condition number is ~1e5), pv and dpv within 1e-8; the two
solvers agree to 1e-10.  The Krylov solves must converge without the LU
fallback, in fewer than 200 steps."""
import ctypes as C
import os

import numpy as np
import pytest

from bem3_aw_qd2_amd import exdat, hip_lib, host_lib

pytestmark = pytest.mark.gpu
HERE = os.path.dirname(os.path.abspath(__file__))
SAMPLE = os.path.join(HERE, "golden", "sample_ex.dat")


def solve(mode, tmp_path):
    m = exdat.Model(host_lib()).dat_read(SAMPLE)
    prob = host_lib().aqd2_problem(m.buf, None)
    assert hip_lib().bqd2_set_solver(prob, mode) == 0
    m.solve()
    its = (C.c_int * 17)()
    res = (C.c_double * 17)()
    lu = C.c_int()
    assert hip_lib().bqd2_solver_info(prob, its, res, C.byref(lu)) == 0
    out = str(tmp_path / ("ex_%d.dat" % mode))
    m.dat_write(out)
    m.finalize()
    return exdat.read(out), list(its[:3]), list(res[:3]), lu.value


def test_gmres_and_lu_match_reference(tmp_path):
    ref = exdat.read(SAMPLE)
    g, its, res, lu = solve(0, tmp_path)
    print("GMRES steps", its, "residuals", res, "LU fallbacks", lu)
    assert lu == 0 and all(0 < i < 200 for i in its) and all(r <= 1e-15 for r in res)
    l, its_l, _, lu_l = solve(1, tmp_path)
    assert lu_l == 3 and its_l == [0, 0, 0]
    for d in range(ref["MN"] + 1):
        for name, bar in (("P", 1e-10), ("dP", 1e-10), ("pv", 1e-8), ("dpv", 1e-8)):
            eg = exdat.relerr(g["dom"][d][name], ref["dom"][d][name])
            el = exdat.relerr(l["dom"][d][name], ref["dom"][d][name])
            egl = exdat.relerr(g["dom"][d][name], l["dom"][d][name])
            print("domain %d %-3s gmres %.2e lu %.2e gmres-lu %.2e" % (d, name, eg, el, egl))
            assert eg < bar and el < bar and egl < max(bar, 1e-10)

"""Mixed-precision solve (bqd2_set_solver(p, 2)): the row-scaled system factored
in complex float on the MFMA units and used as the right preconditioner of
GMRES on the FP64 system, on the reference's shipped sample (600 quads,
4800 unknowns, not a multiple of the 128-column panel).

Bars, the same as the FP64 solvers': P and dP within 1e-10 normwise of the
reference's ex.dat, pv and dpv within 1e-8, the surface solution within 1e-10
of the FP64 GMRES solution, and the force of both rules within 1e-9 of the
reference's example1.log (the 1e-8 force bar of the mixed form with a margin).
The refinement must reach the FP64 solver's scaled residual (1e-15) in far
fewer steps than unpreconditioned GMRES, and a second solve on the same
assembly reuses the factors.

The triangular sweeps of the factors run as a HIP graph built at the first
application: three solves in a row (each replays the graph at every refinement
step, with the GMRES kernels and host synchronisations in between) must give
the same bits as the sweeps launched one by one."""
import ctypes as C
import json
import os

import numpy as np
import pytest

from bem3_aw_qd2_amd import exdat, hip_lib, host_lib

pytestmark = pytest.mark.gpu
HERE = os.path.dirname(os.path.abspath(__file__))
SAMPLE = os.path.join(HERE, "golden", "sample_ex.dat")
LOG = json.load(open(os.path.join(HERE, "golden", "sample_example1.json")))


def info(prob):
    its = (C.c_int * 17)()
    res = (C.c_double * 17)()
    lu = C.c_int()
    assert hip_lib().bqd2_solver_info(prob, its, res, C.byref(lu)) == 0
    ms = C.c_double()
    assert hip_lib().bqd2_lu32_info(prob, C.byref(ms)) == 0
    return list(its[:3]), list(res[:3]), lu.value, ms.value


def run(mode, tmp_path, repeat=1):
    m = exdat.Model(host_lib()).dat_read(SAMPLE)
    prob = host_lib().aqd2_problem(m.buf, None)
    assert hip_lib().bqd2_set_solver(prob, mode) == 0
    m.solve()
    infos = [info(prob)]
    for _ in range(repeat - 1):
        assert hip_lib().bqd2_solve_surface(prob, None) == 0
        infos.append(info(prob))
    out = str(tmp_path / ("ex_%d.dat" % mode))
    m.dat_write(out)
    forces = {t: m.force(LOG["rc"], t) for t in (0, 1)}
    m.finalize()
    return exdat.read(out), infos, forces


def test_mixed_solution_matches_reference(tmp_path):
    ref = exdat.read(SAMPLE)
    x, infos, forces = run(2, tmp_path, repeat=2)
    (its, res, lu, ms), (its2, res2, _, ms2) = infos
    print("refinement steps", its, "residuals", res, "LU fallbacks", lu, "factor ms", ms, "second solve", its2, ms2)
    assert lu == 0 and 0 < its[0] <= 20 and res[0] <= 1e-15 and ms > 0.0
    # the second solve reuses the factors of the first (same assembly)
    assert its2[0] == its[0] and ms2 == ms
    g, _, gforces = run(0, tmp_path)
    for d in range(ref["MN"] + 1):
        for name, bar in (("P", 1e-10), ("dP", 1e-10), ("pv", 1e-8), ("dpv", 1e-8)):
            e = exdat.relerr(x["dom"][d][name], ref["dom"][d][name])
            eg = exdat.relerr(x["dom"][d][name], g["dom"][d][name])
            print("domain %d %-3s mixed-ref %.2e mixed-gmres %.2e" % (d, name, e, eg))
            assert e < bar and eg < max(bar, 1e-10)
    for t in (0, 1):
        F, N = forces[t]
        Fr = np.array(LOG["force"][str(t)]["F"])
        Fg = gforces[t][0]
        ef, eg = np.linalg.norm(F - Fr) / np.linalg.norm(Fr), np.linalg.norm(F - Fg) / np.linalg.norm(Fg)
        print("force type %d: mixed-ref %.2e mixed-fp64 %.2e" % (t, ef, eg))
        assert ef <= 1e-9 and eg <= 1e-9


def test_graph_replays_match_eager():
    import numpy as np
    m = exdat.Model(host_lib()).dat_read(SAMPLE)
    prob = host_lib().aqd2_problem(m.buf, None)
    hip = hip_lib()
    assert hip.bqd2_set_solver(prob, 2) == 0
    m.solve()
    na = 2 * m.nn()
    xs = []
    try:
        for graphs in (0, 1, 1, 1):
            assert hip.bqd2_set_graphs(graphs) == 0
            x = np.zeros(na, complex)
            assert hip.bqd2_solve_surface(prob, x.ctypes.data) == 0
            assert hip.bqd2_solve_velocity(prob) == 0
            xs.append(x)
    finally:
        hip.bqd2_set_graphs(1)
    m.finalize()
    for i, x in enumerate(xs[1:]):
        print("replay solve %d: max |x - x_eager| = %.3e" % (i, np.abs(x - xs[0]).max()))
        assert np.array_equal(x, xs[0])

"""Remainder rules of the periodic pair assembly against the reference's rules.

In the periodic domain the pair kernel integrates the free-space kernel on the
reference's rule (GL12, or GH64 for near pairs) and the smooth remainder of the
lattice sum on a Gauss rule of 4, 6 or 8 points a side chosen per pair from the
distance to the target's other lattice images (bqd2_set_remainder_rules).  The
reference sums the whole lattice sum on every GL12 / GH64 point
(q0_bil_dcoef_v2_cc.c:129-281, q0_lit_dcoef_v2_cc.c:156-321); its rules are
what bqd2_set_remainder_rules(p, 0) assembles.  Checked here, whole matrices of
domain 0, quads and triangles, square and triangular lattices, oblique
incidence, lattices of 2.5 to 4 sphere radii (all three remainder orders):

* G and H entry by entry within 1e-10 of the reference's rules (relative, with a
  floor of 1e-14 of the row's largest entry);
* the solid-angle terms (dF, from the free-space kernel on the reference's
  points in both) bit for bit;
* the remainder rules are taken by most pairs and lower the lattice-sum
  evaluations.
The samples' and configuration meshes' rows, solutions and forces against the
oracle and the reference's logs run with the remainder rules (the default) in
their own tests."""
import ctypes as C
import math
import os
import tempfile

import numpy as np
import pytest

from bem3_aw_qd2_amd import AssemblyStats, exdat, hip_lib, host_lib, mesh

pytestmark = pytest.mark.gpu

TRI = dict(d1=(1.8e-3, 0.0), d2=(0.9e-3, 0.9e-3 * math.sqrt(3.0)), theta=math.radians(30.0), psi=math.radians(15.0))
TIGHT = dict(d1=(1.25e-3, 0.0), d2=(0.0, 1.25e-3), theta=math.radians(20.0), psi=math.radians(40.0))
CASES = {
    "sphere96_square": ("sphere96", False, {}),
    "sphere96_triangular": ("sphere96", False, TRI),
    "sphere96_tight": ("sphere96", False, TIGHT),
    "sphere96_tri_triangular": ("sphere96", True, TRI),
    "sphere96_tri_tight": ("sphere96", True, TIGHT),
    "sphere384": ("sphere384", False, {}),
}


def ptr(a):
    return a.ctypes.data_as(C.c_void_p)


def assemble0(prob, ne):
    hip = hip_lib()
    st = AssemblyStats()
    assert hip.bqd2_assemble(prob, 0, C.byref(st)) == 0
    N = 4 * ne
    G, H = np.zeros((N, N), complex), np.zeros((N, N), complex)
    dF = np.zeros((N, 3))
    assert hip.bqd2_download_domain(prob, 0, ptr(G), ptr(H), None, None, ptr(dF)) == 0
    return st, G, H, dF


@pytest.mark.parametrize("name", sorted(CASES))
def test_remainder_rules_match_reference_rules(name):
    cfg, tri, over = CASES[name]
    tmp = tempfile.mkdtemp()
    paths = mesh.write_config(cfg, tmp, triangles=tri, **over)
    m = exdat.Model(host_lib()).read_inputs(*paths, out=os.path.join(tmp, "ex.dat")).initialize()
    hip = hip_lib()
    try:
        prob = host_lib().aqd2_problem(m.buf, None)
        ne = m.ne()  # one sphere: every element bounds domain 0
        s1, G1, H1, F1 = assemble0(prob, ne)
        assert hip.bqd2_set_remainder_rules(prob, 0) == 0
        s0, G0, H0, F0 = assemble0(prob, ne)
        assert hip.bqd2_set_remainder_rules(prob, 1) == 0
        print("%s: pairs %d gh %d split %d smooth %d free points %d; lattice evaluations %d -> %d"
              % (name, s1.pairs_gl + s1.pairs_gh, s1.pairs_gh, s1.pairs_split, s1.pairs_smooth, s1.free_points,
                 s0.pair_evals, s1.pair_evals), flush=True)
        assert s0.pairs_smooth == 0 and s1.pairs_smooth > 0.5 * (s1.pairs_gl + s1.pairs_gh)
        assert s1.pair_evals < s0.pair_evals
        assert np.array_equal(F1, F0)
        for A1, A0, what in ((G1, G0, "G"), (H1, H0, "H")):
            rowmax = np.max(np.abs(A0), axis=1, keepdims=True)
            err = np.abs(A1 - A0)
            used = rowmax[:, 0] > 0  # the fourth slot of a triangle's rows is unused (all zero)
            # entries: 1e-10 relative, with a floor of 1e-14 of the row's largest
            # entry (below it the rounding of any summation order dominates)
            big = np.abs(A0) >= 1e-4 * rowmax
            rel = err[big] / np.abs(A0[big])
            print("  %s: max err %.2e of the row's largest entry; entries above 1e-4 of it: max rel err %.2e"
                  % (what, np.max(err[used] / rowmax[used]), rel.max()), flush=True)
            assert (err <= 1e-10 * np.abs(A0) + 1e-14 * rowmax).all(), what
    finally:
        m.finalize()

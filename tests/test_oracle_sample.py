"""The oracle's assembly pinned by the reference's own solution of its shipped
sample (tests/golden/sample_ex.dat, a copy of the reference's ex.dat: 600
bilinear quads, sphere of radius 5 mm, square 15 mm lattice, 100 kHz).

The reference solved its system with MKL PARDISO; its solution therefore
satisfies the reference's matrix rows to rounding.  The oracle's rows (GL12/GH64
pair blocks, the polar self-element blocks, the 1+F / F diagonal, the system
layout and right-hand side) applied to that solution must leave a residual at
rounding level too: an error of 1e-10 in any entry would show as a residual of
that size.  The particle velocity and the dpv system are pinned the same way."""
import ctypes as C
import os

import numpy as np
import pytest

from bem3_aw_qd2_amd import exdat, host_lib
from tests.checkers import oracle

HERE = os.path.dirname(os.path.abspath(__file__))
SAMPLE = os.path.join(HERE, "golden", "sample_ex.dat")


@pytest.fixture(scope="module")
def sample():
    m = exdat.Model(host_lib()).dat_read(SAMPLE)
    yield m, exdat.read(SAMPLE)
    m.finalize()


@pytest.mark.parametrize("did,t,tn", [(1, 1, 0), (1, 300, 2), (1, 600, 3), (0, 1, 0), (0, 457, 2)])
def test_system_rows_hold_reference_solution(sample, did, t, tn):
    m, _ = sample
    o = oracle()
    o.orc_row_residual.restype = C.c_double
    terms = C.c_long()
    r = o.orc_row_residual(m.buf, did, t, tn, C.byref(terms))
    assert 0.0 <= r < 1e-13, r
    if did == 0:
        assert terms.value > 0   # periodic Green's function terms were summed


@pytest.mark.parametrize("t,tn", [(7, 3), (123, 0), (599, 1)])
def test_interior_velocity_recomputed(sample, t, tn):
    m, ref = sample
    pv = np.zeros(3, complex)
    assert oracle().orc_node_pv_recompute(m.buf, 1, t, tn, pv.ctypes.data_as(C.c_void_p)) == 0
    want = ref["dom"][1]["pv"][t][tn]
    assert np.linalg.norm(pv - want) <= 1e-10 * np.linalg.norm(want)


@pytest.mark.parametrize("d,t,tn", [(1, 5, 1), (1, 411, 3)])
def test_dpv_rows_hold_reference_solution(sample, d, t, tn):
    m, _ = sample
    o = oracle()
    o.orc_dpv_row_residual.restype = C.c_double
    assert o.orc_dpv_row_residual(m.buf, d, t, tn) < 1e-12


@pytest.mark.parametrize("fname,did,t,tn", [("analysis_sample1_ex1.dat", 0, 5, 1), ("analysis_sample1_ex1.dat", 1, 9, 2),
                                            ("analysis_sample2_ex2.dat", 2, 11, 0), ("analysis_sample2_ex2.dat", 3, 40, 3),
                                            ("verification_veri.dat", 0, 333, 2)])
def test_other_samples_rows(fname, did, t, tn):
    """The same pin on the reference's other solved samples: a triangular
    lattice, three media (domains 1-3) and a 50 mm lattice."""
    m = exdat.Model(host_lib()).dat_read(os.path.join(HERE, "golden", fname))
    o = oracle()
    o.orc_row_residual.restype = C.c_double
    r = o.orc_row_residual(m.buf, did, t, tn, None)
    m.finalize()
    assert 0.0 <= r < 1e-13, r


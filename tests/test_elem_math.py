"""Element integrals of the oracle pinned by mathematics.

The static double-layer coefficient F (the solid-angle kernel (r.n)/(4 pi R^3))
summed over a closed surface is the solid angle seen from the collocation point
over 4 pi: 1/2 at a point of a smooth part of the boundary (Gauss).  Each row of
the assembled matrices carries that sum (1 + F on the periodic domain's diagonal,
F on an interior domain's).  This exercises every pair rule (GL12, GH64) and the
polar self-element rule together: a dropped or misweighted element, a wrong
sign of the normal or a wrong self contribution moves the sum by far more than
its quadrature error."""
import ctypes as C
import os

import numpy as np
import pytest

from bem3_aw_qd2_amd import exdat, host_lib
from tests.checkers import oracle

HERE = os.path.dirname(os.path.abspath(__file__))
SAMPLE = os.path.join(HERE, "golden", "sample_ex.dat")


@pytest.fixture(scope="module")
def model():
    m = exdat.Model(host_lib()).dat_read(SAMPLE)
    yield m
    m.finalize()


@pytest.mark.parametrize("t,tn", [(1, 0), (77, 1), (300, 2), (600, 3)])
def test_solid_angle_of_closed_surface(model, t, tn):
    o = oracle()
    parts = np.zeros(3 * 600)
    assert o.orc_row_fparts(model.buf, 1, t, tn, parts.ctypes.data_as(C.c_void_p)) == 0
    F = parts.reshape(600, 3)[:, 0].sum()
    # interior domain: the boundary is traversed with the object's outward
    # normal, seen from inside the object: F = -1/2 or 1/2 depending on the
    # orientation convention; its magnitude is fixed by Gauss's theorem
    assert abs(abs(F) - 0.5) < 1e-8, F
    # the tangential derivatives of the total solid angle vanish on a smooth surface
    dF = parts.reshape(600, 3)[:, 1:].sum(axis=0)
    scale = np.abs(parts.reshape(600, 3)[:, 1:]).sum(axis=0)
    assert (np.abs(dF) < 1e-8 * scale).all()


def test_solid_angle_opposite_sides_cancel(model):
    """The periodic domain sees each element from the other side: its F row sum
    has the opposite sign (the reference's diagonal is 1 + F there)."""
    o = oracle()
    p0 = np.zeros(3 * 600)
    p1 = np.zeros(3 * 600)
    assert o.orc_row_fparts(model.buf, 1, 10, 0, p1.ctypes.data_as(C.c_void_p)) == 0
    assert o.orc_row_fparts(model.buf, 0, 10, 0, p0.ctypes.data_as(C.c_void_p)) == 0
    F1 = p1.reshape(600, 3)[:, 0].sum()
    F0 = p0.reshape(600, 3)[:, 0].sum()
    assert abs(F0 + F1) < 1e-8

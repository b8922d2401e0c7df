"""Full boundary-value solve of the reference's shipped sample on the GPU.

The reference's own output for this problem (600 bilinear quads, sphere of
radius 5 mm on a 15 mm square lattice, 100 kHz) is committed as
tests/golden/sample_ex.dat (copied from the reference's ex.dat).  Reading it
gives the model; solve_bieq_aqd2 re-solves it on the GPU and the surface
solution is compared with the reference's.  Bar: P and dP within 1e-10
normwise per domain (the reference solved with MKL PARDISO; a dense partial
pivoting LU differs at rounding level), pv and dpv within 1e-8."""
import ctypes as C
import os

import numpy as np
import pytest

from bem3_aw_qd2_amd import exdat, host_lib

pytestmark = pytest.mark.gpu
HERE = os.path.dirname(os.path.abspath(__file__))
SAMPLE = os.path.join(HERE, "golden", "sample_ex.dat")


def test_sample_surface_solution(tmp_path):
    ref = exdat.read(SAMPLE)
    m = exdat.Model(host_lib()).dat_read(SAMPLE)
    m.solve()
    out = str(tmp_path / "gpu_ex.dat")
    m.dat_write(out)
    m.finalize()
    got = exdat.read(out)
    assert got["size"] == ref["size"] == os.path.getsize(SAMPLE)
    report = {}
    for d in range(ref["MN"] + 1):
        for name in ("P", "dP", "pv", "dpv"):
            report[(d, name)] = exdat.relerr(got["dom"][d][name], ref["dom"][d][name])
    print({"%d:%s" % k: "%.2e" % v for k, v in report.items()})
    for d in range(ref["MN"] + 1):
        assert report[(d, "P")] < 1e-10
        assert report[(d, "dP")] < 1e-10
        assert report[(d, "pv")] < 1e-8
        assert report[(d, "dpv")] < 1e-8

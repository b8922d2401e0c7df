"""Radiation force and torque on the GPU (force_FN_aqd2) against the oracle's
restatement of bem3_aw_qd2_force.c:4-278 (oracle/force_cpu.c), on the same
surface solution.

Bars: F within 1e-10 of |F|; N within 1e-10 of max(|N|, |F| a), a the sphere
radius (the torque about the centre of a sphere is rounding residue).

* The 24-quad cube sphere and its 48-triangle split (radius 0.5 mm) on the
  2 mm square lattice at near-normal incidence and on the 1.8 mm triangular
  lattice at theta = 30 deg, phi = 15 deg (1 MHz), solved on the GPU: types 0
  (node values; 4-point rules) and 1 (boundary field at 9 points per quad, 7 per
  triangle) against orc_force, computed here.
* The boundary field behind type 1 (pv_t_bd_aqd2, types 0 and 1) at points of
  the reference's verification sample against orc_field_bd_total.
* The reference's solved samples: type 1 against the oracle's force committed
  in tests/golden/force_oracle.json (correctly rounded z == 0 exponentials, as
  the GPU evaluates them)."""
import ctypes as C
import json
import os
import tempfile

import numpy as np
import pytest

from bem3_aw_qd2_amd import exdat, host_lib, mesh
from tests.checkers import oracle

pytestmark = pytest.mark.gpu
HERE = os.path.dirname(os.path.abspath(__file__))
G = os.path.join(HERE, "golden")
R = 0.5e-3
LATTICES = {
    "square": {},
    "triangular": dict(d1=(1.8e-3, 0.0), d2=(0.9e-3, 0.9e-3 * np.sqrt(3.0)), theta=np.pi / 6, psi=np.pi / 12),
}


def oracle_force(m, rc, type_, cr=True):
    o = oracle()
    rc = np.ascontiguousarray(rc, dtype=float)
    F, N = np.zeros(3), np.zeros(3)
    o.orc_set_cr_exp(C.c_int(1 if cr else 0))
    try:
        err = o.orc_force(m.buf, rc.ctypes.data_as(C.c_void_p), int(type_), F.ctypes.data_as(C.c_void_p),
                          N.ctypes.data_as(C.c_void_p))
    finally:
        o.orc_set_cr_exp(C.c_int(0))
    assert err == 0
    return F, N


def errors(F, N, Fo, No, radius):
    eF = np.linalg.norm(F - Fo) / np.linalg.norm(Fo)
    eN = np.linalg.norm(N - No) / max(np.linalg.norm(No), np.linalg.norm(Fo) * radius)
    return eF, eN


@pytest.mark.parametrize("lattice", sorted(LATTICES))
@pytest.mark.parametrize("tri", [False, True])
def test_force_matches_oracle(lattice, tri):
    tmp = tempfile.mkdtemp()
    paths = mesh.write_config("sphere24", tmp, triangles=tri, **LATTICES[lattice])
    m = exdat.Model(host_lib()).read_inputs(*paths, out=os.path.join(tmp, "ex.dat")).initialize()
    m.solve()
    rc = np.array([1.0e-4, -0.5e-4, 0.2e-4])
    try:
        for type_ in (0, 1):
            F, N = m.force(rc, type_)
            Fo, No = oracle_force(m, rc, type_)
            eF, eN = errors(F, N, Fo, No, R)
            print(lattice, "triangles" if tri else "quads", type_, "F %.2e N %.2e" % (eF, eN))
            assert eF < 1e-10 and eN < 1e-10
    finally:
        m.finalize()


def test_boundary_field_points_verification():
    """pv_t_bd_aqd2 of types 0 and 1 at points of the reference's verification
    sample (600 quads, 50 mm lattice), the element's own polar rule included"""
    m = exdat.Model(host_lib()).dat_read(os.path.join(G, "verification_veri.dat"))
    d = exdat.read(os.path.join(G, "verification_veri.dat"))
    ne = d["dom"][0]["Ne"]
    o = oracle()
    o.orc_set_cr_exp(C.c_int(1))
    try:
        for t, z, e in [(1, 0.0, 0.0), (77, 0.7745966692414834, -0.7745966692414834), (300, 0.0, 0.7745966692414834),
                        (ne, -0.5773502691896258, 0.5773502691896258)]:
            for type_ in (0, 1):
                p, v = m.pv_t_bd(0, t, z, e, type_)
                po, vo = (C.c_double * 2)(), (C.c_double * 6)()
                assert o.orc_field_bd_total(m.buf, t, C.c_double(z), C.c_double(e), type_, po, vo) == 0
                po = complex(po[0], po[1])
                vo = np.array([complex(vo[2 * i], vo[2 * i + 1]) for i in range(3)])
                ep, ev = abs(p - po) / abs(po), np.linalg.norm(v - vo) / np.linalg.norm(vo)
                print("t", t, "type", type_, "p %.2e v %.2e" % (ep, ev))
                assert ep < 1e-10 and ev < 1e-10
    finally:
        o.orc_set_cr_exp(C.c_int(0))
        m.finalize()


SAMPLES = {
    "sample": "sample_ex.dat",
    "analysis_sample1": "analysis_sample1_ex1.dat",
    "analysis_sample2": "analysis_sample2_ex2.dat",
    "verification": "verification_veri.dat",
}


@pytest.mark.parametrize("name", sorted(SAMPLES))
def test_sample_type1_force_matches_oracle_fixture(name):
    path = os.path.join(G, "force_oracle.json")
    fix = json.load(open(path)) if os.path.exists(path) else {}
    if name not in fix:
        pytest.skip("no oracle type-1 force committed for %s" % name)
    log = json.load(open(os.path.join(G, name + "_example1.json")))
    m = exdat.Model(host_lib()).dat_read(os.path.join(G, SAMPLES[name]))
    try:
        for type_, key in ((0, "type0"), (1, "type1_cr")):
            F, N = m.force(fix[name]["rc"], type_)
            eF, eN = errors(F, N, np.array(fix[name][key]["F"]), np.array(fix[name][key]["N"]), log["radius"])
            print(name, key, "F %.2e N %.2e" % (eF, eN))
            assert eF < 1e-10 and eN < 1e-10
    finally:
        m.finalize()

"""The oracle's radiation force and boundary field (oracle/force_cpu.c) pinned
without the GPU:

* type 0 (node values of the surface solution plus the incident wave, the
  4-point rule) on the reference's four solved samples against their
  example1.log, to rounding;
* the boundary field with the point rules of field types 0 and 1 (4/9 points
  on quads, 4/7 on triangles) fed the boundary values of a plane wave inside a
  closed sphere: the field at boundary points is that wave to the
  discretization error, which falls with refinement;
* the point rules of the periodic domain against the 12 x 12 Gauss-Legendre
  rule of the same lattice sum on a distant element (quadrature error only);
* type 1 on the samples (tests/golden/force_oracle.json, written by
  tests/golden/make_force_fixtures.py): within the reference log's own
  type-0/type-1 difference of the log, and the correctly rounded and libm modes
  of the z == 0 stopping rule within a few parts in 1e8 of each other."""
import ctypes as C
import json
import os

import numpy as np
import pytest

from bem3_aw_qd2_amd import exdat, host_lib
from tests.checkers import oracle
from tests.test_oracle_field import boundary_error, plane_wave_model

HERE = os.path.dirname(os.path.abspath(__file__))
G = os.path.join(HERE, "golden")
SAMPLES = {
    "sample": "sample_ex.dat",
    "analysis_sample1": "analysis_sample1_ex1.dat",
    "analysis_sample2": "analysis_sample2_ex2.dat",
    "verification": "verification_veri.dat",
}
OFF_BD = 736


def force_errors(F, N, Fr, Nr, radius):
    eF = np.linalg.norm(np.asarray(F) - Fr) / np.linalg.norm(Fr)
    eN = np.linalg.norm(np.asarray(N) - Nr) / max(np.linalg.norm(Nr), np.linalg.norm(Fr) * radius)
    return eF, eN


@pytest.mark.parametrize("name", sorted(SAMPLES))
def test_type0_force_matches_log(name):
    log = json.load(open(os.path.join(G, name + "_example1.json")))
    m = exdat.Model(host_lib()).dat_read(os.path.join(G, SAMPLES[name]))
    rc = np.array(log["rc"], float)
    F, N = np.zeros(3), np.zeros(3)
    err = oracle().orc_force(m.buf, rc.ctypes.data_as(C.c_void_p), 0, F.ctypes.data_as(C.c_void_p),
                             N.ctypes.data_as(C.c_void_p))
    m.finalize()
    assert err == 0
    eF, eN = force_errors(F, N, np.array(log["force"]["0"]["F"]), np.array(log["force"]["0"]["N"]), log["radius"])
    # measured <= 2.7e-15 (the log prints 15 significant digits)
    assert eF < 1e-13 and eN < 1e-13, (eF, eN)


@pytest.fixture(scope="module")
def pw_models():
    out = {}
    for name in ("sphere24", "sphere96"):
        for tri in (False, True):
            out[(name, tri)] = plane_wave_model(name, tri)
    yield out
    for m, _, _ in out.values():
        m.finalize()


@pytest.mark.parametrize("tri", [False, True])
@pytest.mark.parametrize("type_", [0, 1])
def test_point_rule_boundary_field_is_the_plane_wave(pw_models, tri, type_):
    """measured on 96 quads 1.8e-3 (type 0) and 4.6e-4 (type 1), on 192
    triangles 3.9e-3 and 2.0e-3; 3.1-5.8 times smaller than on 24 elements"""
    e96 = boundary_error(pw_models[("sphere96", tri)], type_)
    e24 = boundary_error(pw_models[("sphere24", tri)], type_)
    assert e96 < 5e-3, e96
    assert e24 > 2.5 * e96, (e24, e96)


def _periodic_model(tri):
    import tempfile
    from bem3_aw_qd2_amd import mesh
    tmp = tempfile.mkdtemp()
    paths = mesh.write_config("sphere96", tmp, triangles=tri)
    m = exdat.Model(host_lib()).read_inputs(*paths, out=os.path.join(tmp, "ex.dat")).initialize()
    f = os.path.join(tmp, "m.dat")
    m.dat_write(f)
    return m, exdat.parse(open(f, "rb").read())


@pytest.mark.parametrize("tri", [False, True])
def test_periodic_point_rules_against_gauss_legendre(tri):
    """The 9-/7-point rules and the 4-point rule of the lattice sum on an
    element on the far side of the sphere against the 12 x 12 Gauss-Legendre
    rule of the same integrand: agreement to the rules' quadrature error, far
    above rounding (the routines differ) and far below a dropped term or sign."""
    m, d = _periodic_model(tri)
    o = oracle()
    nn = 3 if tri else 4
    centre = lambda s: d["ren"][s][:nn].mean(axis=0)
    rt = np.ascontiguousarray(centre(1))
    far = max(range(2, d["Ne"] + 1), key=lambda s: np.linalg.norm(centre(s) - rt))
    cc = {}
    for typ in (0, 1):
        CC = np.zeros(9, complex)
        c = o.orc_coef_pts(m.buf, 0, rt.ctypes.data_as(C.c_void_p), far, typ, CC.ctypes.data_as(C.c_void_p))
        assert c > 0
        cc[typ] = CC
    GL, d0, d1 = np.zeros(9, complex), np.zeros(9, complex), np.zeros(9, complex)
    v = np.array([1.0, 0.0, 0.0])
    x, w = (np.ascontiguousarray(a) for a in np.polynomial.legendre.leggauss(12))
    fn = o.orc_q0_lit_pair if tri else o.orc_q0_bil_pair
    fn.restype = C.c_long
    bd = C.c_void_p(C.addressof(m.buf) + OFF_BD)
    c = fn(bd, rt.ctypes.data_as(C.c_void_p), v.ctypes.data_as(C.c_void_p), v.ctypes.data_as(C.c_void_p), far, 12,
           x.ctypes.data_as(C.c_void_p), w.ctypes.data_as(C.c_void_p), GL.ctypes.data_as(C.c_void_p),
           d0.ctypes.data_as(C.c_void_p), d1.ctypes.data_as(C.c_void_p))
    m.finalize()
    assert c > 0
    for typ, bar in ((1, 1e-4), (0, 1e-2)):
        for blk in (slice(0, nn), slice(4, 4 + nn)):
            e = np.linalg.norm(cc[typ][blk] - GL[blk]) / np.linalg.norm(GL[blk])
            print(tri, typ, blk, "%.2e" % e)
            assert 1e-12 < e < bar, (typ, blk, e)
        eF = abs(cc[typ][8] - GL[8]) / abs(GL[8])
        assert eF < bar, (typ, eF)


FIX = os.path.join(G, "force_oracle.json")


@pytest.mark.parametrize("name", sorted(SAMPLES))
def test_type1_force_fixture(name):
    """The oracle's type-1 force of the samples (computed once, committed)
    against the log.  Mode 0 (the reference's algorithm with this machine's
    libm) reproduces the root sample's log to 1.3e-10 and the verification
    sample's to 7.4e-8; the log's own type-0/type-1 difference is ~1e-3, so a
    restatement error in the 9-point field would show far above that.  Mode 1
    (exact stopping residues at level offsets, the GPU's) differs from mode 0
    only through the reference's libm sensitivity: 4.6e-10 on the root sample,
    2.2e-7 on the 50 mm lattice of the verification sample."""
    fix = json.load(open(FIX)) if os.path.exists(FIX) else {}
    if name not in fix:
        pytest.skip("no oracle type-1 force committed for %s" % name)
    e = fix[name]
    log = json.load(open(os.path.join(G, name + "_example1.json")))
    Fr, Nr = np.array(log["force"]["1"]["F"]), np.array(log["force"]["1"]["N"])
    F0 = np.array(log["force"]["0"]["F"])
    rule_gap = np.linalg.norm(Fr - F0) / np.linalg.norm(Fr)
    eF, eN = force_errors(e["type1_libm"]["F"], e["type1_libm"]["N"], Fr, Nr, log["radius"])
    print(name, "mode 0 vs log: F %.2e N %.2e (log type 0 vs 1: %.2e)" % (eF, eN, rule_gap))
    assert eF < 1e-7 and eN < 1e-7
    if name != "verification":
        assert eF < 1e-8 and eN < 1e-8
    eF, eN = force_errors(e["type1_cr"]["F"], e["type1_cr"]["N"], np.array(e["type1_libm"]["F"]),
                          np.array(e["type1_libm"]["N"]), log["radius"])
    print(name, "mode 1 vs mode 0: F %.2e N %.2e" % (eF, eN))
    assert eF < 5e-7 and eN < 5e-7
    if name != "verification":
        assert eF < 1e-8 and eN < 1e-8

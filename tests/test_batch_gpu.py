"""Assembly of several problems in one persistent launch (bqd2_assemble_batch)
against assembly one problem at a time (bqd2_assemble): the periodic domains
alone (did = 0), and every domain of every problem (did = -1: the periodic
pairs and self blocks, the free-space pairs with their tangential derivatives
and self blocks, of quadrilateral and triangle meshes, all in one launch).

Three uneven instances (24 quads on the 2 mm square lattice, 96 quads on the
1.8 mm triangular lattice at theta = 30 deg, 108 triangles on a 2.4 mm square
lattice at 0.8 MHz) and a free-space problem in the same batch.
Each pair of a periodic domain is computed the same way whichever workgroup of
the persistent launch takes it, so G, H and dF must be bit-identical to the
one-at-a-time assembly, and so must the counters of Green's function terms."""
import ctypes as C
import math
import os
import tempfile

import numpy as np
import pytest

from bem3_aw_qd2_amd import AssemblyStats, exdat, hip_lib, host_lib, mesh

pytestmark = pytest.mark.gpu
INSTANCES = [
    ("sphere24", {}, False, True),
    ("sphere96", dict(d1=(1.8e-3, 0.0), d2=(0.9e-3, 0.9e-3 * math.sqrt(3.0)), theta=math.pi / 6,
                      psi=math.pi / 12), False, True),
    ("sphere24", dict(n=3, d1=(2.4e-3, 0.0), d2=(0.0, 2.4e-3), f=0.8e6, theta=0.4, psi=0.3), True, True),
    ("sphere24", {}, False, False),
]


def problem(tmp, name, over, tri, periodic):
    paths = mesh.write_config(name, tmp, triangles=tri, periodic=periodic, **over)
    m = exdat.Model(host_lib()).read_inputs(*paths, out=os.path.join(tmp, "ex.dat")).initialize()
    return m, host_lib().aqd2_problem(m.buf, None)


def matrices(prob, n, did=0, derivs=False):
    G = np.zeros(n * n * 2)
    H = np.zeros(n * n * 2)
    dF = np.zeros(n * 3)
    dG = np.zeros(4 * n * n) if derivs else None
    dH = np.zeros(4 * n * n) if derivs else None
    assert hip_lib().bqd2_download_domain(prob, did, G.ctypes.data, H.ctypes.data,
                                          dG.ctypes.data if derivs else None, dH.ctypes.data if derivs else None,
                                          dF.ctypes.data) == 0
    return (G, H, dF) + ((dG, dH) if derivs else ())


def test_batch_equals_one_at_a_time():
    hip = hip_lib()
    with tempfile.TemporaryDirectory() as tmp:
        models, probs = [], []
        for i, (name, over, tri, per) in enumerate(INSTANCES):
            m, p = problem(os.path.join(tmp, str(i)), name, over, tri, per)
            models.append(m)
            probs.append(p)
        single, sstats = [], []
        for m, p in zip(models, probs):
            st = AssemblyStats()
            assert hip.bqd2_assemble(p, 0, C.byref(st)) == 0
            single.append(matrices(p, 4 * m.ne()))
            sstats.append(st)
        arr = (C.c_void_p * len(probs))(*probs)
        bst = (AssemblyStats * len(probs))()
        assert hip.bqd2_assemble_batch(arr, len(probs), 0, bst) == 0
        for i, (m, p) in enumerate(zip(models, probs)):
            G, H, dF = matrices(p, 4 * m.ne())
            same = [np.array_equal(a, b) for a, b in zip((G, H, dF), single[i])]
            print("instance %d (%d elements): G, H, dF identical %s; terms %s / %s" %
                  (i, m.ne(), same, list(bst[i].pair_terms), list(sstats[i].pair_terms)))
            assert all(same)
            assert list(bst[i].pair_terms) == list(sstats[i].pair_terms)
            assert bst[i].pair_evals == sstats[i].pair_evals
        for m in models:
            m.finalize()


def test_batch_all_domains_equals_one_at_a_time():
    hip = hip_lib()
    with tempfile.TemporaryDirectory() as tmp:
        models, probs = [], []
        for i, (name, over, tri, per) in enumerate(INSTANCES):
            m, p = problem(os.path.join(tmp, str(i)), name, over, tri, per)
            models.append(m)
            probs.append(p)
        single, sstats = [], []
        for i, (m, p) in enumerate(zip(models, probs)):
            per_dom = []
            for d in range(m.mn() + 1):
                st = AssemblyStats()
                assert hip.bqd2_assemble(p, d, C.byref(st)) == 0
                sstats.append(st)
            for d in range(m.mn() + 1):
                per_dom.append(matrices(p, 4 * m.ne(), d, derivs=(d == 1 or not INSTANCES[i][3])))
            single.append(per_dom)
        arr = (C.c_void_p * len(probs))(*probs)
        bst = (AssemblyStats * len(sstats))()
        assert hip.bqd2_assemble_batch(arr, len(probs), -1, bst) == 0
        k = 0
        for i, (m, p) in enumerate(zip(models, probs)):
            for d in range(m.mn() + 1):
                got = matrices(p, 4 * m.ne(), d, derivs=(d == 1 or not INSTANCES[i][3]))
                same = [np.array_equal(a, b) for a, b in zip(got, single[i][d])]
                print("instance %d domain %d (%d elements): G, H, dF(, dG, dH) identical %s; self terms %s / %s"
                      % (i, d, m.ne(), same, list(bst[k].self_terms), list(sstats[k].self_terms)))
                assert all(same)
                if d == 0:
                    assert list(bst[k].pair_terms) == list(sstats[k].pair_terms)
                    assert list(bst[k].self_terms) == list(sstats[k].self_terms)
                k += 1
        for m in models:
            m.finalize()



def test_batch_beyond_one_dispatch():
    """Two 3456-quad instances: their free-space pairs (9.6e7 single-wave
    workgroups) exceed what one dispatch can address (the grid's size in
    work-items is a 32-bit number), so the batched launches go in parts; sampled
    rows of every matrix must still equal one instance at a time."""
    hip = hip_lib()
    over2 = dict(d1=(2.2e-3, 0.0), d2=(0.0, 2.2e-3), theta=0.3, psi=0.4)
    with tempfile.TemporaryDirectory() as tmp:
        models, probs = [], []
        for i, over in enumerate(({}, over2)):
            m, p = problem(os.path.join(tmp, str(i)), "sphere3456", over, False, True)
            models.append(m)
            probs.append(p)
        N = 4 * models[0].ne()
        rows = np.array([0, 1, 5, 4000, 9999, N - 1], np.int32)

        def sample(p, d):
            G = np.zeros((len(rows), N), complex)
            H = np.zeros((len(rows), N), complex)
            dF = np.zeros((len(rows), 3))
            dG = np.zeros((len(rows), 2 * N), complex) if d == 1 else None
            dH = np.zeros((len(rows), 2 * N), complex) if d == 1 else None
            assert hip.bqd2_download_rows(p, d, len(rows), rows.ctypes.data, G.ctypes.data, H.ctypes.data,
                                          dG.ctypes.data if d == 1 else None, dH.ctypes.data if d == 1 else None,
                                          dF.ctypes.data) == 0
            return (G, H, dF) + ((dG, dH) if d == 1 else ())

        single = []
        for p in probs:
            for d in (0, 1):
                assert hip.bqd2_assemble(p, d, None) == 0
            single.append([sample(p, d) for d in (0, 1)])
        arr = (C.c_void_p * len(probs))(*probs)
        assert hip.bqd2_assemble_batch(arr, len(probs), -1, None) == 0
        for i, p in enumerate(probs):
            for d in (0, 1):
                same = [np.array_equal(a, b) for a, b in zip(sample(p, d), single[i][d])]
                print("instance %d domain %d: sampled rows identical %s" % (i, d, same))
                assert all(same)
        for m in models:
            m.finalize()

"""The reference's own driver programs, compiled unchanged from its sources and
linked against this repository's libbem3_aw_qd2.so (oracle/Makefile target
`drivers`, built where the reference sources are present; the binaries travel
with the tree), run on the GPU:

* example1 on the reference's ex.dat must print the field and force of the
  reference's example1.log (tests/golden/sample_example1.json);
* aw_d3qd2_bv_solver on generated inputs must write the ex.dat that the
  library's own solve writes;
* example2 drives pv_t_aqd2 per point from OpenMP threads: its intensity grid
  must equal |p|^2 of the batched field call."""
import json
import os
import re
import subprocess

import numpy as np
import pytest

from bem3_aw_qd2_amd import exdat, host_lib, mesh

pytestmark = pytest.mark.gpu
HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
REF = os.path.join(ROOT, "oracle", "_ref")
SAMPLE = os.path.join(HERE, "golden", "sample_ex.dat")
LOG = json.load(open(os.path.join(HERE, "golden", "sample_example1.json")))


def driver(name):
    path = os.path.join(REF, name)
    if not os.path.exists(path):
        pytest.skip("%s not built (make -C oracle drivers where the reference sources are)" % name)
    return path


def run(args, cwd, timeout=240, threads=None):
    env = dict(os.environ)
    if threads:
        env["OMP_NUM_THREADS"] = str(threads)
    r = subprocess.run(args, cwd=cwd, capture_output=True, text=True, timeout=timeout, env=env)
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]
    return r.stdout


def parse_c(line):
    a = re.findall(r"[-+]?\d\.\d+e[-+]\d+", line)
    return complex(float(a[0]), float(a[1]))


def test_example1_prints_the_reference_log(tmp_path):
    out = run([driver("example1"), SAMPLE], str(tmp_path)).split("\n")
    blocks = [i for i, l in enumerate(out) if l.startswith("p  =")]
    assert len(blocks) == 2
    for k, i in enumerate(blocks):
        ref = LOG["field"][str(k)]
        p = parse_c(out[i])
        v = np.array([parse_c(out[i + j]) for j in (1, 2, 3)])
        assert abs(p - complex(*ref["p"])) <= 1e-11 * abs(complex(*ref["p"]))
        vr = np.array([complex(*x) for x in ref["v"]])
        assert np.linalg.norm(v - vr) <= 1e-11 * np.linalg.norm(vr)
    Fl = [l for l in out if l.startswith("F = (")]
    Nl = [l for l in out if l.startswith("N = (")]
    for k in (0, 1):
        F = np.array([float(x) for x in re.findall(r"[-+]?\d\.\d+e[-+]\d+", Fl[k])])
        N = np.array([float(x) for x in re.findall(r"[-+]?\d\.\d+e[-+]\d+", Nl[k])[:3]])
        Fr, Nr = np.array(LOG["force"][str(k)]["F"]), np.array(LOG["force"][str(k)]["N"])
        assert np.linalg.norm(F - Fr) <= 1e-9 * np.linalg.norm(Fr)
        assert np.linalg.norm(N - Nr) <= 1e-9 * np.linalg.norm(Fr) * 5e-3


def test_solver_driver_writes_the_library_solution(tmp_path):
    paths = mesh.write_config("sphere24", str(tmp_path))
    out = str(tmp_path / "drv_ex.dat")
    run([driver("aw_d3qd2_bv_solver")] + paths + [out], str(tmp_path))
    m = exdat.Model(host_lib()).read_inputs(*paths, out=str(tmp_path / "x.dat")).initialize().solve()
    mine = str(tmp_path / "lib_ex.dat")
    m.dat_write(mine)
    m.finalize()
    a, b = exdat.read(out), exdat.read(mine)
    for d in range(a["MN"] + 1):
        for name in ("P", "dP", "pv", "dpv"):
            assert np.array_equal(a["dom"][d][name], b["dom"][d][name]), (d, name)


def test_example2_threads_match_batched_field(tmp_path):
    import time
    t0 = time.perf_counter()
    run([driver("example2"), SAMPLE], str(tmp_path), timeout=170, threads=8)
    print("example2, 40 000 points from 8 threads: %.2f s" % (time.perf_counter() - t0))
    rows = [l.split() for l in open(str(tmp_path / "Ip_xz.txt")) if l.strip() and not l.startswith("#")]
    data = np.array([[float(x) for x in r] for r in rows])
    assert data.shape == (200 * 200, 3)
    # the driver prints coordinates with %g; rebuild them as it computes them
    # (example2.c:16-18: rang = 3 lambda0, dr = 2 rang / (max - 1))
    lam0 = np.frombuffer(exdat.read(SAMPLE)["pw"], np.float64, 1, 160)[0]
    assert abs(lam0 - 0.015) < 1e-12
    rang = 3.0 * lam0
    dr = rang * 2.0 / 199.0
    idx = np.arange(0, 200 * 200, 997)
    i, j = idx // 200, idx % 200
    pts = np.stack([-rang + i * dr, np.zeros(len(idx)), -rang + j * dr], axis=1)
    assert np.allclose(pts[:, 0], data[idx, 0], rtol=1e-5) and np.allclose(pts[:, 2], data[idx, 1], rtol=1e-5)
    m = exdat.Model(host_lib()).dat_read(SAMPLE)
    p, v, did = m.pv_batch(pts, 0, 1)
    m.finalize()
    # intensities printed to 15 digits
    assert np.allclose(data[idx, 2], np.abs(p) ** 2, rtol=1e-13, atol=0)


def test_single_point_calls_are_coalesced():
    """pv_t_aqd2 from 8 host threads at once: the points travel in shared
    bqd2_field batches (fewer batches than points) and each result equals the
    batched call's bit for bit"""
    import ctypes as C
    import threading
    host = host_lib()
    m = exdat.Model(host).dat_read(SAMPLE)
    rng = np.random.default_rng(3)
    pts = rng.uniform(-0.03, 0.03, (8, 30, 3))
    b0, p0 = C.c_longlong(), C.c_longlong()
    host.aqd2_point_batches(C.byref(b0), C.byref(p0))
    res = [[None] * 30 for _ in range(8)]

    def work(t):
        for i in range(30):
            res[t][i] = m.pv_t(pts[t, i], 0)

    th = [threading.Thread(target=work, args=(t,)) for t in range(8)]
    for t in th:
        t.start()
    for t in th:
        t.join()
    b1, p1 = C.c_longlong(), C.c_longlong()
    host.aqd2_point_batches(C.byref(b1), C.byref(p1))
    p, v, did = m.pv_batch(pts.reshape(-1, 3), 0, 1)
    m.finalize()
    nb, npnt = b1.value - b0.value, p1.value - p0.value
    print("%d points in %d batches" % (npnt, nb))
    assert npnt == 240 and nb < 240
    for t in range(8):
        for i in range(30):
            d, pp, vv = res[t][i]
            k = 30 * t + i
            assert d == did[k] and pp == p[k] and np.array_equal(np.asarray(vv), v[k])

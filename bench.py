#!/usr/bin/env python3
"""One-GPU benchmark of the boundary-value solve of a sphere lattice.

A step is one complete solve request on the device-resident problem:
  * assembly of the coefficient matrices of both domains;
  * the dense complex solve of the 2NN system (GMRES, dense LU as fallback);
  * recovery of the particle velocity and its normal derivative on the boundary;
  * pressure and velocity on field grids;
  * radiation force and torque with both quadrature settings (type 0 and 1).
The mesh and media are uploaded before the timed region (inputs resident in HBM).

Workloads (bem3_aw_qd2_amd.mesh.CONFIGS):
  sphere3456 (default)  cube-sphere of 6 x 24 x 24 = 3456 bilinear quads, radius
                0.5 mm (1200 kg/m^3, 2400 m/s) in water (1000 kg/m^3, 1500 m/s)
                on a triangular 1.8 mm lattice (60 degrees), plane wave of 1 Pa at
                1 MHz, theta = 30 deg, phi = 15 deg; field on 256 x 256 points in
                each of the y = 0, z = 0 and x = 0 planes.  --triangles takes the
                6912-triangle mesh of the same sphere.
  sphere384     6 x 8 x 8 = 384 quads, same sphere and media on a square 2.0 mm
                lattice, theta = 1e-6 rad, phi = pi/4 (exact normal incidence
                makes the in-plane wave vector zero, which the reference's model
                rejects); field on 64 x 64 points in the y = 0 plane.

The JSON line also carries:
  roofline      the off-diagonal pair kernel of the periodic domain (k_q0_pairs),
                FP64 vector bound.  Algorithmic flops per launch = sum over the
                Green's function terms it summed, by kind, of a fixed flop count
                per term (F_TERM below); its duration comes from HIP events
                recorded around the launch on the problem's stream.
  cpu_baseline  the C restatement of the algorithm (oracle/) on the host cores:
                a seeded sample of (row, element) coefficient blocks of both
                domains plus self blocks, and Green's function evaluations for the
                field and force, extrapolated to one full request by the counts
                of the GPU run; the dense solve by numpy's LAPACK at n <= 4096,
                scaled by n^3.
"""
import argparse
import contextlib
import ctypes as C
import json
import math
import os
import sys
import tempfile
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

METRIC = "solves/sec + p50/p99 latency at <=1e-8 rel force error; GF evals/sec"
# flops per lattice term of a value + gradient evaluation, counted from the term
# formulas of d3_qpgf_d2.c (:1814-1838 spatial, :1965-2009 spectral) with add/mul 1,
# FMA 2, sqrt/div 8, exp 20, sincos 30, erfc 40, w(z) 120
F_TERM = {"spatial": 256, "propagating": 288, "evanescent": 212, "pure_spectral": 96}
# the same for a term of the Hessian evaluation of the field (:2131-2170,
# :2349-2398): the value-and-gradient count plus the Hessian's own operations in
# the same convention -- spatial: the second radial factor (16), its combination
# with the first (6), six components (28), six more accumulations (12), 62;
# propagating: d2g_zz (12), five products with the order's wave vector (10),
# accumulations (12), 34; evanescent: the exp of d2g_zz (24), its combination
# (8), five products (10), accumulations (12), 54; pure spectral: components and
# accumulations, 32
F_TERM_V2 = {"spatial": 318, "propagating": 322, "evanescent": 266, "pure_spectral": 128}
# flops per GH64 point of a split pair: free-space kernel and gradient (sincos 30,
# 18 mul/add/FMA) and its accumulation into the four shape functions (42)
F_FREE = 90
# flops of the far field's plane-wave sums in the same convention.  Amplitude
# term (one boundary item and one order, k_far_amp): phase argument and height 4;
# exp(i a.y) exp(-b dz) (sincos 30, exp 20) 50; kappa . w 6; the four
# coefficients alpha + i (kappa . w) beta (complex product and sum) 32; their
# products with the source factor accumulated 32: 124.  Point-order term (one
# point and one order, k_far_eval): the order's a_n, b_n (6 products, b_n^2 3,
# sqrt 8) 19; phase argument 3; the target factor 50; its product with A_n 6;
# pressure 2; three gradient components 16; three boundary-value components 24:
# 120
F_FAR = {"amplitude_term": 124, "point_order_term": 120}
PEAK_FP64_TFLOPS = 78.6


def log(*a):
    print(*a, file=sys.stderr, flush=True)


@contextlib.contextmanager
def quiet_stdout():
    """Silence the C library's printf output (model reading and initialisation)."""
    libc = C.CDLL(None)
    sys.stdout.flush()
    libc.fflush(None)
    saved = os.dup(1)
    devnull = os.open(os.devnull, os.O_WRONLY)
    os.dup2(devnull, 1)
    try:
        yield
    finally:
        libc.fflush(None)
        os.dup2(saved, 1)
        os.close(saved)
        os.close(devnull)


def grid_points(n, half, planes=1):
    """n x n points over [-half, half]^2 in the y = 0 plane, then (planes = 3)
    the z = 0 and x = 0 planes"""
    x = np.linspace(-half, half, n)
    A, B = np.meshgrid(x, x, indexing="ij")
    a, b, o = A.ravel(), B.ravel(), np.zeros(n * n)
    sets = [np.stack([a, o, b], axis=1), np.stack([a, b, o], axis=1), np.stack([o, a, b], axis=1)][:planes]
    return np.ascontiguousarray(np.concatenate(sets, axis=0))


class Request:
    """One solve request on a device-resident problem through the C ABI."""

    def __init__(self, hip, prob, mn, pts, mode=0):
        from bem3_aw_qd2_amd import AssemblyStats, check
        self.hip, self.p, self.mn = hip, prob, mn
        self.mode = mode
        check(hip.bqd2_set_solver(prob, mode), "bqd2_set_solver")
        self.pts = pts
        self.n = len(pts)
        self.out = np.zeros((self.n, 8))
        self.did = np.zeros(self.n, np.int32)
        self.stats = [AssemblyStats() for _ in range(mn + 1)]
        self.cnt = np.zeros(16, np.int64)

    def run(self):
        from bem3_aw_qd2_amd import check
        hip, p = self.hip, self.p
        t = {}
        t0 = time.perf_counter()
        for d in range(self.mn + 1):
            check(hip.bqd2_assemble(p, d, C.byref(self.stats[d])), "bqd2_assemble")
        t1 = time.perf_counter()
        check(hip.bqd2_solve_surface(p, None), "bqd2_solve_surface")
        t2 = time.perf_counter()
        check(hip.bqd2_solve_velocity(p), "bqd2_solve_velocity")
        t3 = time.perf_counter()
        ev = C.c_longlong(0)
        check(hip.bqd2_field(p, self.n, self.pts.ctypes.data, 0, 1, self.out.ctypes.data, self.did.ctypes.data,
                             C.byref(ev)), "bqd2_field")
        self.field_counts = self._counters()
        fms = C.c_double(0.0)
        check(hip.bqd2_field_kernel_ms(p, C.byref(fms)), "bqd2_field_kernel_ms")
        self.field_kernel_ms = fms.value
        t4 = time.perf_counter()
        rc = (C.c_double * 3)(0.0, 0.0, 0.0)
        self.F = np.zeros((2, 3))
        self.N = np.zeros((2, 3))
        for typ in (0, 1):
            F = (C.c_double * 3)()
            N = (C.c_double * 3)()
            check(hip.bqd2_force(p, rc, typ, F, N), "bqd2_force")
            self.F[typ], self.N[typ] = F[:], N[:]
        self.force_counts = self._counters()
        t5 = time.perf_counter()
        its = (C.c_int * 17)()
        res = (C.c_double * 17)()
        lu = C.c_int()
        check(hip.bqd2_solver_info(p, its, res, C.byref(lu)), "bqd2_solver_info")
        ms = C.c_double()
        check(hip.bqd2_lu32_info(p, C.byref(ms)), "bqd2_lu32_info")
        self.solver = {"mode": self.mode, "krylov_steps": list(its[:self.mn + 2]),
                       "scaled_residual": list(res[:self.mn + 2]), "lu_fallbacks": lu.value,
                       "lu32_factor_ms": ms.value}
        t.update(assemble=t1 - t0, solve=t2 - t1, velocity=t3 - t2, field=t4 - t3, force=t5 - t4, total=t5 - t0)
        return t

    def _counters(self):
        from bem3_aw_qd2_amd import check
        c = np.zeros(19, np.int64)
        check(self.hip.bqd2_counters(self.p, c.ctypes.data, 19), "bqd2_counters")
        return c

    def gf_evals(self):
        n = sum(s.gf_evals for s in self.stats)
        return n + int(self.field_counts[15]) + int(self.force_counts[15])


def measured_traffic(workload, key="traffic_bytes_per_launch"):
    """HBM bytes per launch of k_q0_pairs (or of the kernel the key names) from
    the committed rocprofv3 summary of this workload (FETCH_SIZE + WRITE_SIZE,
    scaled by a copy kernel of known size), or None when there is none."""
    path = os.path.join(ROOT, "profiles", "k_q0_pairs_%s.md" % workload)
    if not os.path.exists(path):
        return None, None
    for line in open(path):
        if line.startswith(key + ":"):
            return float(line.split(":")[1]), "profiles/k_q0_pairs_%s.md" % workload
    return None, None


def pair_flops(st):
    kinds = ("spatial", "propagating", "evanescent", "pure_spectral")
    return sum(st.pair_terms[i] * F_TERM[k] for i, k in enumerate(kinds)) + st.free_points * F_FREE


def cpu_baseline(model, req, npairs, nself, seed=0):
    """Oracle timings on the host cores, extrapolated to one request."""
    orc_path = os.path.join(ROOT, "oracle", "liboracle.so")
    if not os.path.exists(orc_path):
        return None
    orc = C.CDLL(orc_path)
    orc.orc_pair_sample.restype = C.c_long
    orc.orc_pair_sample.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_void_p, C.c_void_p]
    orc.orc_threads.restype = C.c_int
    want = int(os.environ.get("OMP_NUM_THREADS", "0")) or min(os.cpu_count() or 1, 16)
    threads = int(orc.orc_threads(want))
    rng = np.random.default_rng(seed)
    total = 0.0
    parts = {}
    for d, st in enumerate(req.stats):
        ne = int(st.self_rows) // 4
        nrow = int(st.self_rows)
        rows = rng.integers(0, nrow, size=npairs).astype(np.int32)
        src = rng.integers(1, ne + 1, size=npairs).astype(np.int32)
        # off-diagonal sample: redraw sources that hit the row's own element
        own = rows // 4 + 1
        src = np.where(src == own, (src % ne) + 1, src).astype(np.int32)
        t0 = time.perf_counter()
        r = orc.orc_pair_sample(model.buf, d, npairs, rows.ctypes.data, src.ctypes.data)
        t_off = time.perf_counter() - t0
        if r < 0:
            return None
        srow = rng.choice(nrow, size=nself, replace=False).astype(np.int32)
        ssrc = (srow // 4 + 1).astype(np.int32)
        t0 = time.perf_counter()
        r = orc.orc_pair_sample(model.buf, d, nself, srow.ctypes.data, ssrc.ctypes.data)
        t_self = time.perf_counter() - t0
        if r < 0:
            return None
        est = t_off / npairs * (nrow * (ne - 1)) + t_self / nself * nrow
        parts["assemble_d%d" % d] = est
        total += est
    # Green's function rates of the oracle for the field (value, gradient, Hessian)
    # and the boundary force (value, gradient) at offsets of the workload's size
    orc.orc_qpgf_batch.argtypes = [C.c_int, C.c_void_p, C.c_int, C.c_void_p, C.c_double, C.c_void_p, C.c_void_p,
                                   C.c_void_p, C.c_void_p]
    qd = np.frombuffer(model.buf, dtype=np.float64, count=8, offset=model.qpd2_offset()).copy()
    ns = 32768
    pts = rng.uniform(-1.5e-3, 1.5e-3, size=(ns, 3))
    G = np.zeros(ns, complex)
    dG = np.zeros((ns, 3), complex)
    d2G = np.zeros((ns, 6), complex)
    ret = np.zeros(ns, np.int32)
    rate = {}
    for mode in (2, 3):
        t0 = time.perf_counter()
        orc.orc_qpgf_batch(mode, qd.ctypes.data, ns, pts.ctypes.data, 1e-15, G.ctypes.data, dG.ctypes.data,
                           d2G.ctypes.data, ret.ctypes.data)
        rate[mode] = ns / (time.perf_counter() - t0)
    parts["field"] = req.field_counts[15] / rate[3]
    parts["force"] = req.force_counts[15] / rate[2]
    total += parts["field"] + parts["force"]
    # dense solve (LAPACK through numpy) at up to n = 4096, scaled by n^3 to the system size
    na = 2 * model.nn()
    ns = min(na, 4096)
    A = rng.standard_normal((ns, ns)) + 1j * rng.standard_normal((ns, ns))
    b = rng.standard_normal(ns) + 0j
    t0 = time.perf_counter()
    np.linalg.solve(A, b)
    parts["solve"] = (time.perf_counter() - t0) * (na / ns) ** 3
    total += parts["solve"]
    sample = ("oracle coefficient blocks of %d random off-diagonal (row, element) pairs and %d self blocks per "
              "domain, Green's function rates for the field and force evaluation counts, "
              "numpy LAPACK solve at n=%d scaled by n^3 to n=%d; extrapolated to one request: %s s"
              % (npairs, nself, ns, na, ", ".join("%s %.1f" % (k, v) for k, v in parts.items())))
    return {"value": 1.0 / total, "unit": "solves/s", "cores": threads, "kind": "port", "sample": sample,
            "cpu_model": cpu_model(), "host_cpus": os.cpu_count()}


def cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return None


def force_vs_oracle(host, model, prob, req, npts, seed=0):
    """The benched request's force against the oracle on the same surface
    solution (copied from the device into the model's arrays):
    type 0 (node values and the incident wave) in full; type 1 through the
    boundary field it integrates (the total field at the elements' 9 points),
    recomputed by the oracle at a sample of those points (each point sums every
    element's 9-point rule and the own element's polar rule; correctly rounded
    z == 0 exponentials, as on the GPU).  Relative errors: F normwise, field
    points max over p and |v|."""
    from concurrent.futures import ThreadPoolExecutor
    orc_path = os.path.join(ROOT, "oracle", "liboracle.so")
    if not os.path.exists(orc_path):
        return None
    orc = C.CDLL(orc_path)
    host.aqd2_pull_solution(model.buf, prob)
    F, N = np.zeros(3), np.zeros(3)
    rc = np.zeros(3)
    t0 = time.perf_counter()
    err = orc.orc_force(model.buf, rc.ctypes.data_as(C.c_void_p), 0, F.ctypes.data_as(C.c_void_p),
                        N.ctypes.data_as(C.c_void_p))
    if err:
        return None
    e0 = float(np.linalg.norm(req.F[0] - F) / np.linalg.norm(F))
    eN0 = float(np.linalg.norm(req.N[0] - N) / max(np.linalg.norm(N), np.linalg.norm(F) * 0.5e-3))
    rng = np.random.default_rng(seed)
    ne = model.ne()
    b = model.lib
    # the 9-point rule's (zeta, eta): BOUD.zt_49 and .et_49 at offsets 280 and
    # 352 of BOUD, itself at 736 in AQD2 (include/bem3_aw_qd2.h)
    zt = np.array(C.cast(C.addressof(model.buf) + 736 + 280, C.POINTER(C.c_double * 9)).contents[:])
    et = np.array(C.cast(C.addressof(model.buf) + 736 + 352, C.POINTER(C.c_double * 9)).contents[:])
    picks = [(int(rng.integers(1, ne + 1)), int(rng.integers(0, 9))) for _ in range(npts)]

    def one(tq):
        t, q = tq
        p = (C.c_double * 2)()
        v = (C.c_double * 6)()
        b.pv_t_bd_aqd2(p, v, 0, t, C.c_double(zt[q]), C.c_double(et[q]), 1, model.buf)
        po, vo = (C.c_double * 2)(), (C.c_double * 6)()
        orc.orc_set_cr_exp(1)
        r = orc.orc_field_bd_total(model.buf, t, C.c_double(zt[q]), C.c_double(et[q]), 1, po, vo)
        if r:
            return None
        g = np.array(p[:] + v[:]).view(complex)
        o = np.array(po[:] + vo[:]).view(complex)
        return max(abs(g[0] - o[0]) / abs(o[0]), np.linalg.norm(g[1:] - o[1:]) / np.linalg.norm(o[1:]))

    with ThreadPoolExecutor(max_workers=min(npts, os.cpu_count() or 1, 16)) as ex:
        errs = list(ex.map(one, picks))
    orc.orc_set_cr_exp(0)
    if any(e is None for e in errs):
        return None
    return {"type0_F_vs_oracle": e0, "type0_N_vs_oracle": eN0,
            "type1_boundary_field_vs_oracle": float(max(errs)), "type1_points_checked": npts,
            "seconds": time.perf_counter() - t0}


def lu32_leg(hip, prob, na):
    """The mixed-precision solve on the request's assembled system: the
    row-scaled system factored in complex float on the MFMA units, then FP64
    GMRES preconditioned by the factors; further incidences with the same
    in-plane wave vector reuse the factors.  Compared with the FP64 GMRES
    solution of the timed steps (surface unknowns and type-0 force)."""
    from bem3_aw_qd2_amd import check
    out = {}
    for mode in (0, 2):
        check(hip.bqd2_set_solver(prob, mode), "bqd2_set_solver")
        t0 = time.perf_counter()
        xs = np.zeros(na, np.complex128)
        check(hip.bqd2_solve_surface(prob, xs.ctypes.data), "bqd2_solve_surface")
        t1 = time.perf_counter()
        check(hip.bqd2_solve_velocity(prob), "bqd2_solve_velocity")
        F = (C.c_double * 3)()
        N = (C.c_double * 3)()
        check(hip.bqd2_force(prob, (C.c_double * 3)(0.0, 0.0, 0.0), 0, F, N), "bqd2_force")
        its = (C.c_int * 17)()
        res = (C.c_double * 17)()
        lu = C.c_int()
        check(hip.bqd2_solver_info(prob, its, res, C.byref(lu)), "bqd2_solver_info")
        out[mode] = dict(x=xs, F=np.array(F[:]), t=t1 - t0, its=its[0], res=res[0])
    # the FP64 direct solve (the dense LU on the MFMA units replacing the MKL
    # solve): system build, factorisation and the two triangular solves
    check(hip.bqd2_set_solver(prob, 1), "bqd2_set_solver")
    x1 = np.zeros(na, np.complex128)
    t0 = time.perf_counter()
    check(hip.bqd2_solve_surface(prob, x1.ctypes.data), "bqd2_solve_surface")
    t_lu64 = time.perf_counter() - t0
    lu64 = {"solve_ms": t_lu64 * 1e3, "tflops_8n3_3_over_solve": 8.0 / 3.0 * na ** 3 / t_lu64 / 1e12,
            "peak_fp64_mfma_tflops": 78.6, "pivoting": "partial over all rows (LAPACK's interchanges)",
            "solution_rel_diff_vs_gmres": float(np.linalg.norm(x1 - out[0]["x"]) / np.linalg.norm(out[0]["x"]))}
    check(hip.bqd2_set_solver(prob, 2), "bqd2_set_solver")
    # a second incidence on the same operator: the factors are reused
    t0 = time.perf_counter()
    check(hip.bqd2_solve_surface(prob, None), "bqd2_solve_surface")
    t_reuse = time.perf_counter() - t0
    ms = C.c_double()
    check(hip.bqd2_lu32_info(prob, C.byref(ms)), "bqd2_lu32_info")
    check(hip.bqd2_set_solver(prob, 0), "bqd2_set_solver")
    flops = 8.0 / 3.0 * na ** 3
    x0, x2 = out[0]["x"], out[2]["x"]
    return {"unknowns": na, "factor_ms": ms.value, "factor_tflops_8n3_3": flops / (ms.value * 1e-3) / 1e12,
            "factor_mfma_tflops_6n3_3": 0.75 * flops / (ms.value * 1e-3) / 1e12,
            "trailing_update": "3 real products per complex product (3M): the nominal 8n^3/3 count is 4/3 of the "
                               "MFMA work",
            "panels": "block-local pivots, unchecked (the refinement validates the factors), next panel on 8 "
                      "CUs of its own beside the update",
            "peak_fp32_mfma_tflops": 157.3, "refinement_steps": out[2]["its"], "scaled_residual": out[2]["res"],
            "solve_ms_with_factor": out[2]["t"] * 1e3, "solve_ms_factor_reused": t_reuse * 1e3,
            "fp64_gmres_steps": out[0]["its"], "fp64_gmres_solve_ms": out[0]["t"] * 1e3,
            "solution_rel_diff_vs_fp64": float(np.linalg.norm(x2 - x0) / np.linalg.norm(x0)),
            "fp64_lu": lu64,
            "force_rel_diff_vs_fp64": float(np.linalg.norm(out[2]["F"] - out[0]["F"]) / np.linalg.norm(out[0]["F"]))}


AUTOTUNE_SITES = {0: ("lu32_trailing_update", ["3m_64x64", "3m_128x64"]),
                  1: ("gmres_matvec", ["rows4", "rows2", "rows8", "rows1"])}


def autotune_tables(hip):
    """The online autotuner's classes on this device: per site, each size class
    (half octaves of the dimension; GMRES: 8 class + right-hand sides) with the
    candidate it settled on (null while exploring) and the highest rate seen per
    candidate (trailing update: TFLOP/s; GMRES product: TB/s of the matrix)."""
    out = {}
    for site, (name, cands) in AUTOTUNE_SITES.items():
        keys = (C.c_longlong * 256)()
        chosen = (C.c_int * 256)()
        best = (C.c_float * 1024)()
        nc = C.c_int()
        cnt = hip.bqd2_autotune_info(site, 1, 256, keys, chosen, best, C.byref(nc))
        rows = []
        for i in range(min(max(cnt, 0), 256)):
            rows.append({"class": int(keys[i]), "chosen": cands[chosen[i]] if chosen[i] >= 0 else None,
                         "best_rate": {cands[c]: (round(best[4 * i + c] / 1e9, 3) if best[4 * i + c] >= 0 else None)
                                       for c in range(nc.value)}})
        out[name] = rows
    return out


def batch_leg(args, hip, host):
    """A scaled-down form of the batched workload: `batch_instances` uneven
    instances (6 n^2 quads, n in {8, 12, 16}: 384, 864 and 1536 quads of the
    workload's 384-3456; ka in [0.5, 3]; square or triangular lattice with
    period / radius in [3, 6]; seeded), each solved for 16 incidences that
    share its operator (8 amplitudes and phases, each at theta and at the
    mirrored pi - theta).  Assembly one instance at a time
    against one batched launch of the periodic pairs; the 16 solves (surface,
    velocity, type-0 force) by FP64 GMRES each against one complex float
    factorisation reused by all 16 (mixed precision)."""
    from bem3_aw_qd2_amd import check, exdat, mesh
    rng = np.random.default_rng(1)
    a, c = 0.5e-3, 1500.0
    models, probs, tmp = [], [], tempfile.mkdtemp(prefix="bqd2_batch_")
    for i in range(args.batch_instances):
        n = int(rng.choice([8, 12, 16]))
        period = a * rng.uniform(3.0, 6.0)
        d2 = (0.5 * period, 0.5 * math.sqrt(3.0) * period) if rng.random() < 0.5 else (0.0, period)
        ka = rng.uniform(0.5, 3.0)
        theta, psi = rng.uniform(0.1, 1.0), rng.uniform(0.1, 1.4)
        paths = mesh.write_config("sphere384", os.path.join(tmp, str(i)), n=n, d1=(period, 0.0), d2=d2,
                                  f=ka * c / (2 * math.pi * a), theta=theta, psi=psi)
        with quiet_stdout():
            m = exdat.Model(host).read_inputs(*paths, out=os.path.join(tmp, "ex%d.dat" % i)).initialize()
        models.append((m, theta, psi))
        probs.append(host.aqd2_problem(m.buf, None))
    arr = (C.c_void_p * len(probs))(*probs)

    def assemble(batched):
        t0 = time.perf_counter()
        if batched:
            check(hip.bqd2_assemble_batch(arr, len(probs), 0, None), "bqd2_assemble_batch")
        else:
            for p in probs:
                check(hip.bqd2_assemble(p, 0, None), "bqd2_assemble")
        for p in probs:
            check(hip.bqd2_assemble(p, 1, None), "bqd2_assemble")
        return time.perf_counter() - t0

    t_one = assemble(False)
    t_bat = assemble(True)
    incid = [(amp, mir) for amp in np.exp(1j * np.linspace(0, 2 * math.pi, 8, endpoint=False)) *
             np.linspace(0.5, 2.0, 8) for mir in (False, True)]

    def solves(mode):
        forces = []
        t0 = time.perf_counter()
        for p, (m, theta, psi) in zip(probs, models):
            check(hip.bqd2_set_solver(p, mode), "bqd2_set_solver")
            for amp, mir in incid:
                pw7 = (C.c_double * 7)(amp.real, amp.imag, math.pi - theta if mir else theta, psi, 0.0, 0.0, 0.0)
                check(hip.bqd2_set_incidence(p, pw7), "bqd2_set_incidence")
                check(hip.bqd2_solve_surface(p, None), "bqd2_solve_surface")
                check(hip.bqd2_solve_velocity(p), "bqd2_solve_velocity")
                F = (C.c_double * 3)()
                N = (C.c_double * 3)()
                check(hip.bqd2_force(p, (C.c_double * 3)(0.0, 0.0, 0.0), 0, F, N), "bqd2_force")
                forces.append(np.array(F[:]))
        return time.perf_counter() - t0, forces

    t_plain, f_plain = solves(0)
    t_mixed, f_mixed = solves(2)
    err = max(float(np.linalg.norm(x - y) / np.linalg.norm(y)) for x, y in zip(f_mixed, f_plain))
    nsol = len(probs) * len(incid)
    out = {"instances": len(probs), "elements": [m.ne() for m, _, _ in models], "incidences_per_instance": len(incid),
           "assemble_s_one_at_a_time": t_one, "assemble_s_batched": t_bat,
           "plain_fp64_solves_per_s": nsol / t_plain, "factor_reuse_solves_per_s": nsol / t_mixed,
           "solve_speedup": t_plain / t_mixed, "force_rel_diff": err}
    with quiet_stdout():
        for m, _, _ in models:
            m.finalize()
    return out


def diffraction_incidences(k, d1, d2, theta0, psi0, count=16, max_theta=math.pi / 3):
    """`count` incidences that share one quasi-periodic operator: the plane wave
    (theta0, psi0) and the directions of its propagating diffraction orders with
    theta <= max_theta (in-plane wave vectors k_par + m1 b1 + m2 b2, b the
    reciprocal lattice), nearest orders first; amplitudes and phases vary, and
    directions repeat when there are fewer than `count` orders."""
    det = d1[0] * d2[1] - d1[1] * d2[0]
    b1 = np.array([d2[1], -d2[0]]) * (2 * math.pi / det)
    b2 = np.array([-d1[1], d1[0]]) * (2 * math.pi / det)
    kp = k * math.sin(theta0) * np.array([math.cos(psi0), math.sin(psi0)])
    dirs = []
    for m1 in range(-12, 13):
        for m2 in range(-12, 13):
            q = kp + m1 * b1 + m2 * b2
            sq = float(np.hypot(q[0], q[1])) / k
            if 1e-6 < sq <= math.sin(max_theta):
                dirs.append((m1 * m1 + m2 * m2, m1, m2, math.asin(sq), math.atan2(q[1], q[0])))
    dirs.sort()
    out = []
    for j in range(count):
        _, _, _, th, ps = dirs[j % len(dirs)]
        amp = complex(np.exp(2j * math.pi * j / count) * (0.5 + j / count))
        out.append((amp, th, ps))
    return out, len(dirs)


def batch_config_leg(args):
    """The batched workload at its full size: 64 instances, 6 n^2 quads with n
    drawn from {8, 12, 16, 20, 24} (384-3456 quads), ka uniform in [0.5, 3],
    square or triangular lattice with period / radius uniform in [3, 6], seeded.
    Instances are created in waves that fit the device-memory budget; per wave:
    assembly one instance (and domain) at a time, then every pair and self block
    of every instance in one launch per task kind (bqd2_assemble_batch, did = -1),
    with the GPU time of each from HIP events; then each instance's 16
    incidences (its plane wave and the directions of its propagating
    diffraction orders with theta <= 60 deg, which share its operator) solved
    with FP64 GMRES each (plain) and with the complex float factors reused by
    all 16 and FP64 refinement (reuse): surface solve, velocity recovery and
    type-0 force per incidence."""
    from bem3_aw_qd2_amd import AssemblyStats, check, exdat, hip_lib, host_lib, mesh, serve
    hip, host = serve._bind()
    rng = np.random.default_rng(args.batch_seed)
    a, c = 0.5e-3, 1500.0
    tmp = tempfile.mkdtemp(prefix="bqd2_batchcfg_")
    specs = []
    for i in range(args.batch_instances):
        n = int(rng.choice([8, 12, 16, 20, 24]))
        period = a * rng.uniform(3.0, 6.0)
        tri_lat = bool(rng.random() < 0.5)
        d2 = (0.5 * period, 0.5 * math.sqrt(3.0) * period) if tri_lat else (0.0, period)
        ka = float(rng.uniform(0.5, 3.0))
        theta, psi = float(rng.uniform(0.05, math.pi / 3)), float(rng.uniform(0.0, 2 * math.pi))
        over = dict(n=n, d1=(period, 0.0), d2=d2, f=ka * c / (2 * math.pi * a), theta=theta, psi=psi)
        paths = mesh.write_config("sphere384", os.path.join(tmp, str(i)), **over)
        with quiet_stdout():
            m = exdat.Model(host).read_inputs(*paths, out=os.path.join(tmp, "ex%d.dat" % i)).initialize()
        mp = host.aqd2_model_new(m.buf)
        k = ka / a
        inc, norders = diffraction_incidences(k, (period, 0.0), d2, theta, psi)
        specs.append({"m": m, "mp": mp, "ne": m.ne(), "ka": ka, "ratio": period / a, "tri_lattice": tri_lat,
                      "inc": inc, "orders": norders,
                      "bytes": int(hip.bqd2_footprint(mp, 2, 0, 0))})
    budget = int(args.batch_budget_gib * 2 ** 30)
    waves, cur, used = [], [], 0
    for sp in specs:
        if cur and used + sp["bytes"] > budget:
            waves.append(cur)
            cur, used = [], 0
        cur.append(sp)
        used += sp["bytes"]
    waves.append(cur)
    log("batch (configs): %d instances, %d-%d quads, %d memory waves" %
        (len(specs), min(sp["ne"] for sp in specs), max(sp["ne"] for sp in specs), len(waves)))
    gpu_one = gpu_bat = wall_one = wall_bat = 0.0
    t_plain = t_reuse = 0.0
    nsol, errs, steps = 0, [], []
    for wi, wave in enumerate(waves):
        probs = []
        for sp in wave:
            p = C.c_void_p()
            check(hip.bqd2_create(sp["mp"], C.byref(p)), "bqd2_create")
            probs.append(p)
        arr = (C.c_void_p * len(probs))(*probs)
        # one at a time (first assembly: the rule decisions and pair lists are made here)
        t0 = time.perf_counter()
        for p, sp in zip(probs, wave):
            for d in (0, 1):
                st = AssemblyStats()
                check(hip.bqd2_assemble(p, d, C.byref(st)), "bqd2_assemble")
                gpu_one += (st.ms_pairs + st.ms_self) * 1e-3
        wall_one += time.perf_counter() - t0
        log("  wave %d: %d instances assembled one at a time (%.1f s)" % (wi, len(wave), time.perf_counter() - t0))
        # the pairs and self blocks of the wave, one launch per task kind
        bst = (AssemblyStats * (2 * len(probs)))()
        t0 = time.perf_counter()
        check(hip.bqd2_assemble_batch(arr, len(probs), -1, bst), "bqd2_assemble_batch")
        wall_bat += time.perf_counter() - t0
        gpu_bat += (bst[0].ms_pairs + sum(bst[j].ms_self for j in range(2 * len(probs)))) * 1e-3
        log("  wave %d: batched assembly (%.1f s)" % (wi, time.perf_counter() - t0))
        for ii, (p, sp) in enumerate(zip(probs, wave)):
            forces = {}
            for mode in (0, 2):
                check(hip.bqd2_set_solver(p, mode), "bqd2_set_solver")
                t0 = time.perf_counter()
                for j, (amp, th, ps) in enumerate(sp["inc"]):
                    pw7 = (C.c_double * 7)(amp.real, amp.imag, th, ps, 0.0, 0.0, 0.0)
                    check(hip.bqd2_set_incidence(p, pw7), "bqd2_set_incidence")
                    check(hip.bqd2_solve_surface(p, None), "bqd2_solve_surface")
                    check(hip.bqd2_solve_velocity(p), "bqd2_solve_velocity")
                    F, N = (C.c_double * 3)(), (C.c_double * 3)()
                    check(hip.bqd2_force(p, (C.c_double * 3)(0.0, 0.0, 0.0), 0, F, N), "bqd2_force")
                    forces[(mode, j)] = np.array(F[:])
                    if mode == 2:
                        its = (C.c_int * 17)()
                        check(hip.bqd2_solver_info(p, its, None, None), "bqd2_solver_info")
                        steps.append(its[0])
                dt = time.perf_counter() - t0
                if mode == 0:
                    t_plain += dt
                    dt_plain = dt
                else:
                    t_reuse += dt
                    dt_reuse = dt
            for j in range(len(sp["inc"])):
                f0, f2 = forces[(0, j)], forces[(2, j)]
                errs.append(float(np.linalg.norm(f2 - f0) / np.linalg.norm(f0)))
            nsol += len(sp["inc"])
            log("    instance %d of wave %d (%d quads): 16 incidences, FP64 GMRES %.2f s, factors reused %.2f s"
                % (ii, wi, sp["ne"], dt_plain, dt_reuse))
        for p in probs:
            hip.bqd2_destroy(p)
        log("  wave %d: %d instances done (assembly GPU one at a time %.1f s, batched %.1f s; solves so far "
            "plain %.1f s, reuse %.1f s)" % (wi, len(wave), gpu_one, gpu_bat, t_plain, t_reuse))
    for sp in specs:
        host.aqd2_model_free(sp["mp"])
        with quiet_stdout():
            sp["m"].finalize()
    return {
        "instances": len(specs), "waves": len(waves), "budget_gib": args.batch_budget_gib,
        "elements": [sp["ne"] for sp in specs], "ka": [round(sp["ka"], 4) for sp in specs],
        "period_over_radius": [round(sp["ratio"], 4) for sp in specs],
        "triangular_lattice": [sp["tri_lattice"] for sp in specs],
        "diffraction_orders_within_60deg": [sp["orders"] for sp in specs],
        "incidences_per_instance": 16,
        "assembly_gpu_s_one_at_a_time": gpu_one, "assembly_gpu_s_batched": gpu_bat,
        "assembly_wall_s_one_at_a_time_incl_rules": wall_one, "assembly_wall_s_batched": wall_bat,
        "assembly_speedup_gpu": gpu_one / gpu_bat,
        "plain_fp64_solves_per_s": nsol / t_plain, "factor_reuse_solves_per_s": nsol / t_reuse,
        "solve_speedup": t_plain / t_reuse,
        "refinement_steps_mean": float(np.mean(steps)), "refinement_steps_max": int(np.max(steps)),
        "force_rel_diff_max": max(errs),
        "solves_per_s_end_to_end_batched_reuse": nsol / (gpu_bat + t_reuse),
        "solves_per_s_end_to_end_plain": nsol / (gpu_one + t_plain),
    }


# mesh sizes of the request streams: "small" 6 n^2 quads for n in 4..8 (or
# their triangle splits); "config" 384-6144 elements: 6 n^2 quads for
# n in {8, 12, 16, 20, 24, 28, 32}, or 12 n^2 triangles for n in {6, 9, 12, 15, 18, 22}
STREAM_SIZES = {"small": {False: list(range(4, 9)), True: list(range(4, 9))},
                "config": {False: [8, 12, 16, 20, 24, 28, 32], True: [6, 9, 12, 15, 18, 22]}}


def serving_stream(nreq, rate, repeat, seed=0, args_mirror=0.0, shape="small"):
    """A seeded stream of solve requests of mixed size: Poisson arrivals at
    `rate` per second; each request is, with probability `repeat`, another
    incidence on an operator already requested (a new amplitude and phase of
    the same plane wave, or with probability `args_mirror` the mirrored
    direction pi - theta), else a new operator: quad or triangle cube-sphere
    (sizes of STREAM_SIZES[shape]), square or triangular lattice with
    period / radius in [2.5, 6], ka in [0.3, 4], sphere density ratio in
    [0.8, 3], sound-speed ratio in [0.5, 2], random incidence.  Lattices with a
    diffraction order within 1e-3 of grazing (a Wood anomaly) are redrawn."""
    rng = np.random.default_rng(seed)
    a, rho, c = 0.5e-3, 1000.0, 1500.0
    ops, reqs, t = [], [], 0.0
    while len(reqs) < nreq:
        t += rng.exponential(1.0 / rate)
        if ops and rng.random() < repeat:
            k = int(rng.integers(len(ops)))
            over = dict(ops[k]["over"])
            mirrored = rng.random() < args_mirror
            if mirrored:
                over["theta"] = math.pi - over["theta"]
            amp = complex(np.exp(1j * rng.uniform(0, 2 * math.pi)) * rng.uniform(0.5, 2.0))
            over["p"] = (amp.real, amp.imag)
            reqs.append(dict(t=t, key=k + 1, over=over, tri=ops[k]["tri"], mirrored=mirrored))
            continue
        while True:
            period = a * rng.uniform(2.5, 6.0)
            tri = rng.random() < 0.5
            d1, d2 = (period, 0.0), ((0.5 * period, 0.5 * math.sqrt(3.0) * period) if rng.random() < 0.5
                                     else (0.0, period))
            ka = rng.uniform(0.3, 4.0)
            f = ka * c / (2 * math.pi * a)
            theta, psi = rng.uniform(0.05, 1.2), rng.uniform(0.05, 2 * math.pi - 0.05)
            k = 2 * math.pi * f / c
            kx, ky = k * math.sin(theta) * math.cos(psi), k * math.sin(theta) * math.sin(psi)
            det = d1[0] * d2[1] - d1[1] * d2[0]
            g1 = (2 * math.pi * d2[1] / det, -2 * math.pi * d2[0] / det)
            g2 = (-2 * math.pi * d1[1] / det, 2 * math.pi * d1[0] / det)
            near = min(abs(((kx + i * g1[0] + j * g2[0]) ** 2 + (ky + i * g1[1] + j * g2[1]) ** 2) / k ** 2 - 1.0)
                       for i in range(-12, 13) for j in range(-12, 13))
            if near > 1e-3:
                break
        sizes = STREAM_SIZES[shape][tri]
        over = dict(n=int(sizes[int(rng.integers(len(sizes)))]), d1=d1, d2=d2, f=f, theta=theta, psi=psi,
                    sphere=(rho * rng.uniform(0.8, 3.0), c * rng.uniform(0.5, 2.0)))
        ops.append(dict(over=over, tri=tri))
        reqs.append(dict(t=t, key=len(ops), over=over, tri=tri, mirrored=False))
    return reqs


def run_stream(reqs, models, slots, solver, keyed, budget, pts):
    """Submit each request at its arrival time; returns per-request results and
    the wall time from the first arrival to the last completion."""
    from bem3_aw_qd2_amd import serve
    with serve.Server(budget, slots) as srv:
        t0 = time.perf_counter()
        tickets = []
        for r, m in zip(reqs, models):
            dt = r["t"] - (time.perf_counter() - t0)
            if dt > 0:
                time.sleep(dt)
            tickets.append(srv.submit(m, key=r["key"] if keyed else 0, solver=solver, pts=pts, force_type=0))
        res = []
        for i, t in enumerate(tickets):
            res.append(srv.wait(t))
            log("  request %d of %d done at %.1f s (%d elements)" % (i + 1, len(tickets), time.perf_counter() - t0,
                                                                     models[i].ne))
        wall = time.perf_counter() - t0 - reqs[0]["t"]
        st = srv.stats()
    for r in res:
        if r["status"] != 0:
            raise RuntimeError("serving request failed: " + r["error"])
    return res, wall, st


def serving_bench(args):
    """The plain FP64 path (one request at a time, every operator assembled,
    FP64 GMRES) against the serving form (resident operators, `slots` concurrent
    requests, the mixed-precision solver) on the same request stream."""
    from bem3_aw_qd2_amd import serve
    reqs = serving_stream(args.serve_requests, args.serve_rate, args.serve_repeat, args_mirror=args.serve_mirror,
                          shape=args.serve_shape)
    models = [serve.PackedModel("sphere384", triangles=r["tri"], **r["over"]) for r in reqs]
    pts = grid_points(8, 1.5e-3)
    budget = int(args.serve_budget_gib * 2 ** 30)
    log("serving: %d requests, %d operators, rate %.1f/s" % (len(reqs), len(set(r["key"] for r in reqs)),
                                                             args.serve_rate))
    plain, wall_p, _ = run_stream(reqs, models, 1, 0, False, budget, pts)
    fast, wall_f, st = run_stream(reqs, models, args.serve_slots, args.serve_solver, True, budget, pts)
    lat_p = np.array([r["t_queue"] + r["t_run"] for r in plain]) * 1e3
    lat_f = np.array([r["t_queue"] + r["t_run"] for r in fast]) * 1e3
    errs = [float(np.linalg.norm(a["F"] - b["F"]) / np.linalg.norm(b["F"])) for a, b in zip(fast, plain)]
    ferr = max(errs)
    ferr_fresh = max([e for e, a in zip(errs, fast) if not a["reused"]] or [0.0])
    ferr_reused = max([e for e, a, r in zip(errs, fast, reqs) if a["reused"] and not r["mirrored"]] or [0.0])
    ferr_mirr = max([e for e, a, r in zip(errs, fast, reqs) if a["reused"] and r["mirrored"]] or [0.0])
    elems = [m.ne for m in models]
    return {
        "requests": len(reqs), "operators": len(set(r["key"] for r in reqs)), "mesh_sizes": args.serve_shape,
        "elements_min_max": [int(min(elems)), int(max(elems))], "arrival_rate_per_s": args.serve_rate,
        "repeat_fraction": args.serve_repeat,
        "plain_fp64": {"solves_per_s": len(reqs) / wall_p, "p50_ms": float(np.percentile(lat_p, 50)),
                       "p99_ms": float(np.percentile(lat_p, 99)), "slots": 1, "solver": "FP64 GMRES",
                       "resident_operators": False},
        "serving": {"solves_per_s": len(reqs) / wall_f, "p50_ms": float(np.percentile(lat_f, 50)),
                    "p99_ms": float(np.percentile(lat_f, 99)), "slots": args.serve_slots,
                    "solver": {0: "FP64 GMRES", 2: "complex float LU (MFMA) + FP64 GMRES refinement"}[args.serve_solver],
                    "resident_operators": True, "hits": st["hits"], "misses": st["misses"],
                    "evictions": st["evictions"]},
        "speedup": wall_p / wall_f,
        "force_rel_err_vs_plain": ferr,
        "force_rel_err_vs_plain_fresh_operator": ferr_fresh,
        "force_rel_err_vs_plain_resident_operator": ferr_reused,
        "force_rel_err_vs_plain_resident_mirrored": ferr_mirr,
        "mirrored_fraction": args.serve_mirror,
    }


def large_main(args):
    """The mixed-precision instance in the system layout (bqd2_create_large): one
    request with the complex float LU and GMRES refinement (assembly of both
    domains, surface solve, velocity recovery, force and torque of both rules),
    then the FP64 LU of the same system as the reference solution (its velocity
    recovery assembles domain 0 again), and the differences between the two."""
    import torch
    from bem3_aw_qd2_amd import AssemblyStats, LargeStats, check, exdat, hip_lib, host_lib, mesh
    hip, host = hip_lib(), host_lib()
    hip.bqd2_device_bytes.argtypes = [C.c_void_p, C.POINTER(C.c_size_t)]
    check(hip.bqd2_set_autotune(args.autotune), "bqd2_set_autotune")
    cfg = mesh.CONFIGS[args.workload]
    tmp = tempfile.mkdtemp(prefix="bqd2_bench_")
    paths = mesh.write_config(args.workload, tmp)
    with quiet_stdout():
        model = exdat.Model(host).read_inputs(*paths, out=os.path.join(tmp, "ex.dat")).initialize()
    ne, nn = model.ne(), model.nn()
    mp = host.aqd2_model_new(model.buf)
    p = C.c_void_p()
    check(hip.bqd2_create_large(mp, C.byref(p)), "bqd2_create_large")
    host.aqd2_model_free(mp)
    check(hip.bqd2_large_config(p, args.large_rows, args.large_resident), "bqd2_large_config")
    log("bench: workload %s in the system layout, %d elements, %d unknowns" % (args.workload, ne, 2 * nn))
    from bem3_aw_qd2_amd._native import PROGRESS_FN
    t_start = time.perf_counter()

    def progress(stage, done, total, _ctx):
        log("    %s: %d/%d (%.0f s)" % (stage.decode(), done, total, time.perf_counter() - t_start))
    progress_cb = PROGRESS_FN(progress)  # kept alive until the problem is destroyed
    check(hip.bqd2_set_progress(progress_cb, None), "bqd2_set_progress")
    rc3 = (C.c_double * 3)(0.0, 0.0, 0.0)

    def request(solver, assemble):
        t, st = {}, [AssemblyStats(), AssemblyStats()]
        t0 = time.perf_counter()
        if assemble:
            for d in (0, 1):
                check(hip.bqd2_assemble(p, d, C.byref(st[d])), "bqd2_assemble")
                log("  domain %d assembled: %.1f s" % (d, time.perf_counter() - t0))
        t1 = time.perf_counter()
        check(hip.bqd2_set_solver(p, solver), "bqd2_set_solver")
        x = np.zeros(2 * nn, complex)
        check(hip.bqd2_solve_surface(p, x.ctypes.data), "bqd2_solve_surface")
        t2 = time.perf_counter()
        log("  surface solve (solver %d): %.1f s" % (solver, t2 - t1))
        check(hip.bqd2_solve_velocity(p), "bqd2_solve_velocity")
        t3 = time.perf_counter()
        log("  velocity recovery: %.1f s" % (t3 - t2))
        F, N = np.zeros((2, 3)), np.zeros((2, 3))
        for typ in (0, 1):
            f, n = (C.c_double * 3)(), (C.c_double * 3)()
            check(hip.bqd2_force(p, rc3, typ, f, n), "bqd2_force")
            F[typ], N[typ] = f[:], n[:]
        t4 = time.perf_counter()
        log("  force and torque, both rules: %.1f s" % (t4 - t3))
        its, res, lu = (C.c_int * 17)(), (C.c_double * 17)(), C.c_int()
        check(hip.bqd2_solver_info(p, its, res, C.byref(lu)), "bqd2_solver_info")
        ls = LargeStats()
        check(hip.bqd2_large_info(p, C.byref(ls)), "bqd2_large_info")
        nb = C.c_size_t()
        check(hip.bqd2_device_bytes(p, C.byref(nb)), "bqd2_device_bytes")
        free_b, total_b = torch.cuda.mem_get_info()
        t.update(assemble=t1 - t0, solve=t2 - t1, velocity=t3 - t2, force=t4 - t3, total=t4 - t0)
        return {"x": x, "F": F, "N": N, "t": t, "st": st, "krylov_steps": list(its[:3]),
                "scaled_residual": list(res[:3]), "lu": lu.value,
                "layout": dict(ls.as_dict(), problem_device_bytes=nb.value, device_free_bytes=free_b,
                               device_total_bytes=total_b)}

    log("mixed-precision request")
    mixed = request(2, True)
    st0 = mixed["st"][0]
    flops = pair_flops(st0)
    achieved = flops / (st0.ms_pairs * 1e-3) / 1e12
    n3 = 8.0 / 3.0 * (2 * nn) ** 3
    line = {
        "metric": METRIC, "value": 1.0 / mixed["t"]["total"], "unit": "solves/s", "n_gpus": 1, "steps": 1,
        "warmup": 0, "ms_per_step": mixed["t"]["total"] * 1e3, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "c64 LU + f64 refinement",
        "data": "synthetic (generated cube-sphere mesh, plane-wave incidence)",
        "config": {"workload": "%s: %d-quad cube-sphere a=%.2g m, square lattice %.3g m (2.2 radii), ka=6 "
                               "(f=%.6g Hz), theta=%.3g phi=%.4g; system layout, complex float LU + FP64 GMRES "
                               "refinement, velocity recovery, force/torque type 0 and 1"
                               % (args.workload, ne, cfg["radius"], cfg["d1"][0], cfg["f"], cfg["theta"],
                                  cfg["psi"]),
                   "elements": ne, "unknowns": 2 * nn, "parallelism": "1 GPU"},
        "phases_ms": {k: v * 1e3 for k, v in mixed["t"].items()},
        "layout": mixed["layout"],
        "solver": {"krylov_steps": mixed["krylov_steps"], "scaled_residual": mixed["scaled_residual"],
                   "lu_fallbacks": mixed["lu"],
                   "lu32_tflops_nominal": n3 / (mixed["layout"]["ms_factor"] * 1e-3) / 1e12
                   if mixed["layout"]["ms_factor"] > 0 else None},
        "force": {"type0": mixed["F"][0].tolist(), "type1": mixed["F"][1].tolist()},
        "torque": {"type0": mixed["N"][0].tolist(), "type1": mixed["N"][1].tolist()},
        "roofline": {"bound": "valu", "pipe": "FP64 vector ALU (MI355X FP64 vector peak 78.6 TFLOP/s)",
                     "kernel": "k_q0_pairs", "achieved": achieved, "peak": PEAK_FP64_TFLOPS, "unit": "TFLOP/s",
                     "frac": achieved / PEAK_FP64_TFLOPS, "traffic": None, "ms_per_launch": st0.ms_pairs,
                     "flops_per_launch": flops},
        "cpu_baseline": None,
    }
    if not args.no_lu_reference:
        log("FP64 LU of the same system (reference solution)")
        ref = request(1, False)
        ex = float(np.linalg.norm(mixed["x"] - ref["x"]) / np.linalg.norm(ref["x"]))
        ef = {t: float(np.linalg.norm(mixed["F"][t] - ref["F"][t]) / np.linalg.norm(ref["F"][t])) for t in (0, 1)}
        line["vs_fp64_lu"] = {
            "surface_solution_rel_err": ex, "force_rel_err_type0": ef[0], "force_rel_err_type1": ef[1],
            "fp64_lu_request_ms": {k: v * 1e3 for k, v in ref["t"].items()}, "fp64_lu_layout": ref["layout"],
            "fp64_lu_tflops": n3 / (ref["layout"]["ms_factor"] * 1e-3) / 1e12 if ref["layout"]["ms_factor"] else None,
            "force_fp64_lu": {"type0": ref["F"][0].tolist(), "type1": ref["F"][1].tolist()}}
        log("mixed vs FP64 LU: surface solution %.2e, force type 0 %.2e, type 1 %.2e" % (ex, ef[0], ef[1]))
    hip.bqd2_destroy(p)
    check(hip.bqd2_set_progress(PROGRESS_FN(), None), "bqd2_set_progress")
    with quiet_stdout():
        model.finalize()
    print(json.dumps(line), flush=True)


def main():
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default="sphere3456")
    ap.add_argument("--grid", type=int, default=0, help="field grid per plane (default 64, 256 for sphere3456)")
    ap.add_argument("--planes", type=int, default=0, help="field planes y=0, z=0, x=0 (default 1, 3 for sphere3456)")
    ap.add_argument("--triangles", action="store_true", help="the linear-triangle mesh of the workload")
    ap.add_argument("--solver", type=int, default=0,
                    help="0: FP64 GMRES (default), 1: FP64 LU, 2: complex float LU on MFMA + FP64 GMRES refinement")
    ap.add_argument("--cpu-pairs", type=int, default=384)
    ap.add_argument("--cpu-self", type=int, default=16)
    ap.add_argument("--field-far", type=int, default=1, choices=(0, 1),
                    help="field points far above or below the lattice of spheres from plane-wave amplitudes (1) "
                         "or from per-pair lattice sums (0)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-force-check", action="store_true", help="skip the oracle check of the force")
    ap.add_argument("--check-points", type=int, default=16,
                    help="boundary points of the type-1 force whose field the oracle recomputes")
    ap.add_argument("--no-serving", action="store_true", help="skip the request-stream measurement")
    ap.add_argument("--no-lu32", action="store_true", help="skip the mixed-precision solve of the request's system")
    ap.add_argument("--no-batch", action="store_true", help="skip the batched-instances measurement")
    ap.add_argument("--autotune", type=int, default=1, choices=(0, 1),
                    help="online autotuning of launch shapes (0: default shapes everywhere)")
    ap.add_argument("--batch-instances", type=int, default=None,
                    help="instances of the batch leg (default 6; 64 with --batch-config)")
    ap.add_argument("--batch-config", action="store_true",
                    help="only the batched workload at its full size (64 instances of 384-3456 quads)")
    ap.add_argument("--batch-seed", type=int, default=2)
    ap.add_argument("--batch-budget-gib", type=float, default=240.0)
    ap.add_argument("--lu32-lookahead", type=int, default=1, choices=(0, 1),
                    help="the mixed-precision factorisation's next panel on 8 CUs of its own (1) or not (0)")
    ap.add_argument("--serving-only", action="store_true", help="only the request-stream measurement")
    ap.add_argument("--serve-requests", type=int, default=48)
    ap.add_argument("--serve-rate", type=float, default=20.0, help="Poisson arrival rate, requests per second")
    ap.add_argument("--serve-repeat", type=float, default=0.75,
                    help="probability that a request reuses an earlier request's operator")
    ap.add_argument("--serve-mirror", type=float, default=0.0,
                    help="probability that a repeated request takes the mirrored incidence pi - theta (its "
                         "in-plane wave vector then differs from the resident operator's in the last bits, which "
                         "moves the lattice sums at points level with a source element by up to 1e-6 in the force)")
    ap.add_argument("--serve-slots", type=int, default=4)
    ap.add_argument("--serve-shape", default="small", choices=sorted(STREAM_SIZES),
                    help="mesh sizes of the stream: small (96-768 elements) or config (384-6144 elements)")
    ap.add_argument("--serve-solver", type=int, default=2)
    ap.add_argument("--serve-budget-gib", type=float, default=64.0)
    ap.add_argument("--layout", default="", choices=("", "default", "system"),
                    help="system: the system layout of bqd2_create_large (default for sphere13824)")
    ap.add_argument("--large-rows", type=int, default=2048, help="system layout: rows per window of the sphere")
    ap.add_argument("--large-resident", type=int, default=-1,
                    help="system layout: windows the mixed solver keeps resident (-1: as memory allows)")
    ap.add_argument("--no-lu-reference", action="store_true",
                    help="system layout: skip the FP64 LU reference solution")
    ap.add_argument("--copy-mib", type=int, default=0,
                    help="before the steps, copy a device buffer of this size once (a known-traffic kernel that "
                         "calibrates profiler byte counters)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py: this build measures one GPU; the solve is not sharded across GPUs (requested --gpus %d)"
              % args.gpus, file=sys.stderr)
        sys.exit(2)

    import torch
    if not torch.cuda.is_available():
        print("bench.py: no GPU visible", file=sys.stderr)
        sys.exit(1)
    if args.layout == "system" or (args.layout == "" and args.workload == "sphere13824"):
        large_main(args)
        return
    from bem3_aw_qd2_amd import hip_lib as _hip_lib
    _hip_lib().bqd2_set_lu32_lookahead(args.lu32_lookahead)
    if args.batch_instances is None:
        args.batch_instances = 64 if args.batch_config else 6
    if args.batch_config:
        bc = batch_config_leg(args)
        print(json.dumps({"metric": METRIC, "value": bc["solves_per_s_end_to_end_batched_reuse"], "unit": "solves/s",
                          "n_gpus": 1, "higher_is_better": True, "dtype": "f64 (c64 factors, f64 refinement)",
                          "config": {"workload": "batched instances (64, 384-3456 quads, 16 incidences each)"},
                          "batch": bc}), flush=True)
        return
    if args.serving_only:
        sv = serving_bench(args)
        print(json.dumps({"metric": METRIC, "value": sv["serving"]["solves_per_s"], "unit": "solves/s",
                          "n_gpus": 1, "higher_is_better": True, "dtype": "f64",
                          "config": {"workload": "request stream"}, "serving": sv}), flush=True)
        return
    from bem3_aw_qd2_amd import check, exdat, hip_lib, host_lib, mesh
    hip = hip_lib()
    check(hip.bqd2_set_autotune(args.autotune), "bqd2_set_autotune")
    host = host_lib()
    cfg = mesh.CONFIGS[args.workload]
    big = args.workload == "sphere3456"
    args.grid = args.grid or (256 if big else 64)
    args.planes = args.planes or (3 if big else 1)
    tmp = tempfile.mkdtemp(prefix="bqd2_bench_")
    paths = mesh.write_config(args.workload, tmp, triangles=args.triangles)
    with quiet_stdout():
        model = exdat.Model(host).read_inputs(*paths, out=os.path.join(tmp, "ex.dat")).initialize()
    prob = host.aqd2_problem(model.buf, None)
    lam = cfg["host"][1] / cfg["f"]
    check(hip.bqd2_set_field_far(prob, args.field_far), "bqd2_set_field_far")
    check(hip.bqd2_set_field_timing(prob, 1), "bqd2_set_field_timing")
    req = Request(hip, prob, model.mn(), grid_points(args.grid, 3.0 * lam, args.planes), args.solver)
    log("bench: workload %s, %d elements, %d unknowns" % (args.workload, model.ne(), 2 * model.nn()))

    if args.copy_mib:
        a = torch.ones(args.copy_mib << 17, dtype=torch.float64, device="cuda")
        b = torch.empty_like(a)
        torch.mul(a, 1.0, out=b)  # elementwise kernel: reads and writes copy_mib MiB each
        torch.cuda.synchronize()
        log("copied %d MiB device to device" % args.copy_mib)
        del a, b
    for i in range(args.warmup):
        t = req.run()
        log("warmup %d: %.3f s (assemble %.3f, solve %.3f, velocity %.3f, field %.3f, force %.3f)"
            % (i, t["total"], t["assemble"], t["solve"], t["velocity"], t["field"], t["force"]))
    F_ref = req.F.copy() if args.warmup else None

    torch.cuda.synchronize()
    lat, phases, flops, ms_pairs, field_ms, evals = [], [], [], [], [], 0
    t_start = time.perf_counter()
    for i in range(args.steps):
        t = req.run()
        lat.append(t["total"])
        phases.append(t)
        st0 = req.stats[0]
        flops.append(pair_flops(st0))
        ms_pairs.append(st0.ms_pairs)
        field_ms.append(req.field_kernel_ms)
        evals += req.gf_evals()
        log("step %d: %.3f s (assemble %.3f, solve %.3f, velocity %.3f, field %.3f, force %.3f)"
            % (i, t["total"], t["assemble"], t["solve"], t["velocity"], t["field"], t["force"]))
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t_start

    # repeatability of the force between requests (the FP64 path is deterministic)
    F_last = req.F.copy()
    if F_ref is None:
        F_ref = F_last
    force_rep = float(max(np.linalg.norm(F_last[k] - F_ref[k]) / np.linalg.norm(F_ref[k]) for k in (0, 1)))
    force_check = None
    if not args.no_force_check:
        log("force check: the oracle on the request's surface solution")
        force_check = force_vs_oracle(host, model, prob, req, args.check_points)
        force_check["repeatability"] = force_rep

    achieved = float(np.sum(flops) / (np.sum(ms_pairs) * 1e-3) / 1e12)
    st0 = req.stats[0]
    traffic, traffic_src = measured_traffic(args.workload)
    roof = {"bound": "valu", "pipe": "FP64 vector ALU (MI355X FP64 vector peak 78.6 TFLOP/s)",
            "kernel": "k_q0_pairs", "achieved": achieved, "peak": PEAK_FP64_TFLOPS,
            "unit": "TFLOP/s", "frac": achieved / PEAK_FP64_TFLOPS, "traffic": traffic, "traffic_source": traffic_src,
            "ms_per_launch": float(np.mean(ms_pairs)),
            "flops_per_launch": float(np.mean(flops)),
            "terms_per_launch": {"spatial": int(st0.pair_terms[0]), "propagating": int(st0.pair_terms[1]),
                                 "evanescent": int(st0.pair_terms[2]), "pure_spectral": int(st0.pair_terms[3])},
            "evals_per_launch": int(st0.pair_evals), "split_pairs": int(st0.pairs_split),
            "gh_pairs": int(st0.pairs_gh), "flops_per_term": F_TERM, "flops_per_free_point": F_FREE,
            "remainder_rule_pairs": int(st0.pairs_smooth), "free_points_per_launch": int(st0.free_points),
            # the lattice evaluations the reference's rules take for the same pairs
            # (the whole sum on every GL12 / GH64 point, three parts per triangle)
            "evals_per_launch_reference_rules": int((st0.pairs_gl * 144 + st0.pairs_gh * 4096) *
                                                    (3 if args.triangles else 1))}

    # the field phase: Hessian evaluations of the field points (k_field_part, the
    # phase's dominant kernel; the phase time bounds its duration from above)
    fc = req.field_counts
    kinds = ("spatial", "propagating", "evanescent", "pure_spectral")
    far_flops = float(int(fc[18]) * F_FAR["amplitude_term"] + int(fc[17]) * F_FAR["point_order_term"])
    field_flops = float(sum(int(fc[11 + i]) * F_TERM_V2[k] for i, k in enumerate(kinds)))
    field_s = float(np.mean([p["field"] for p in phases]))
    part_s = float(np.mean(field_ms)) * 1e-3
    fon, forders = C.c_int(), C.c_int()
    check(hip.bqd2_field_far_info(prob, C.byref(fon), C.byref(forders), None, None, None), "bqd2_field_far_info")
    ftraffic, ftraffic_src = measured_traffic(args.workload, "k_field_part_traffic_bytes_per_launch")
    # the Hessian lattice sums of the points near the boundary, k_field_part's
    # own time (HIP events on the call's stream); the far points' plane-wave
    # sums (k_far_*) and the rest of the phase are listed beside it
    roof_field = {"bound": "valu", "kernel": "k_field_part",
                  "achieved": field_flops / part_s / 1e12, "peak": PEAK_FP64_TFLOPS, "unit": "TFLOP/s",
                  "frac": field_flops / part_s / 1e12 / PEAK_FP64_TFLOPS, "ms_per_launch": part_s * 1e3,
                  "traffic": ftraffic, "traffic_source": ftraffic_src,
                  "field_phase_ms": field_s * 1e3, "far_flops_per_phase": far_flops,
                  "flops_per_launch": field_flops,
                  "terms_per_launch": {k: int(fc[11 + i]) for i, k in enumerate(kinds)},
                  "evals_per_launch": int(fc[15]), "flops_per_term": F_TERM_V2,
                  "far_field": {"on": fon.value, "orders": forders.value, "points": int(fc[16]),
                                "point_order_terms": int(fc[17]), "amplitude_terms": int(fc[18]),
                                "flops_per_term": F_FAR}}

    cpu = None
    if not args.no_cpu:
        log("cpu baseline: sampling the oracle")
        cpu = cpu_baseline(model, req, args.cpu_pairs, args.cpu_self)

    lat_ms = np.array(lat) * 1e3
    ph = {k: float(np.mean([p[k] for p in phases]) * 1e3) for k in phases[0]}
    line = {
        "metric": METRIC,
        "value": args.steps / elapsed,
        "unit": "solves/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": elapsed / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (generated cube-sphere mesh, plane-wave incidence)",
        "config": {"workload": "%s: %d-%s cube-sphere a=%.2g m, lattice d1=%s d2=%s m, f=%.3g Hz, "
                               "theta=%.3g phi=%.4g; surface solve + %d x %dx%d field grid (%s, |.|<=3 lambda) "
                               "+ force/torque type 0 and 1"
                               % (args.workload, model.ne(), "triangle" if args.triangles else "quad",
                                  cfg["radius"], cfg["d1"], cfg["d2"], cfg["f"], cfg["theta"], cfg["psi"],
                                  args.planes, args.grid, args.grid, ", ".join(["y=0", "z=0", "x=0"][:args.planes])),
                   "elements": model.ne(), "unknowns": 2 * model.nn(), "field_points": len(req.pts),
                   "parallelism": "1 GPU"},
        "p50_ms": float(np.percentile(lat_ms, 50)),
        "p99_ms": float(np.percentile(lat_ms, 99)),
        "gf_evals_per_s": evals / elapsed,
        "force_check": force_check,
        "phases_ms": ph,
        "solver": req.solver,
        "force": {"type0": req.F[0].tolist(), "type1": req.F[1].tolist()},
        "torque": {"type0": req.N[0].tolist(), "type1": req.N[1].tolist()},
        "roofline": roof,
        "roofline_field": roof_field,
        "cpu_baseline": cpu,
    }
    if not args.no_lu32:
        log("mixed-precision solve of the request's system")
        line["lu32"] = lu32_leg(hip, prob, 2 * model.nn())
    with quiet_stdout():
        model.finalize()
    if not args.no_serving:
        log("serving: request stream")
        line["serving"] = serving_bench(args)
    if not args.no_batch:
        log("batch: instances with shared factorisations")
        line["batch"] = batch_leg(args, hip, host)
    line["autotune"] = autotune_tables(hip)
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()

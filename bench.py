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
# flops per GH64 point of a split pair: free-space kernel and gradient (sincos 30,
# 18 mul/add/FMA) and its accumulation into the four shape functions (42)
F_FREE = 90
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
        c = np.zeros(16, np.int64)
        check(self.hip.bqd2_counters(self.p, c.ctypes.data, 16), "bqd2_counters")
        return c

    def gf_evals(self):
        n = sum(s.gf_evals for s in self.stats)
        return n + int(self.field_counts[15]) + int(self.force_counts[15])


def measured_traffic(workload):
    """HBM bytes per launch of k_q0_pairs from the committed rocprofv3 summary of
    this workload (FETCH_SIZE + WRITE_SIZE, scaled by a copy kernel of known
    size), or None when there is none."""
    path = os.path.join(ROOT, "profiles", "k_q0_pairs_%s.md" % workload)
    if not os.path.exists(path):
        return None, None
    for line in open(path):
        if line.startswith("traffic_bytes_per_launch:"):
            return float(line.split(":")[1]), "profiles/k_q0_pairs_%s.md" % workload
    return None, None


def pair_flops(st):
    kinds = ("spatial", "propagating", "evanescent", "pure_spectral")
    return sum(st.pair_terms[i] * F_TERM[k] for i, k in enumerate(kinds)) + st.pairs_split * 4096 * F_FREE


def cpu_baseline(model, req, npairs, nself, seed=0):
    """Oracle timings on the host cores, extrapolated to one request."""
    orc_path = os.path.join(ROOT, "oracle", "liboracle.so")
    if not os.path.exists(orc_path):
        return None
    orc = C.CDLL(orc_path)
    orc.orc_pair_sample.restype = C.c_long
    orc.orc_pair_sample.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_void_p, C.c_void_p]
    threads = int(os.environ.get("OMP_NUM_THREADS", "0")) or min(os.cpu_count() or 1, 16)
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
    return {"value": 1.0 / total, "unit": "solves/s", "cores": threads, "kind": "port", "sample": sample}


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
    ap.add_argument("--no-cpu", action="store_true")
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
    from bem3_aw_qd2_amd import exdat, hip_lib, host_lib, mesh
    hip = hip_lib()
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
    lat, phases, flops, ms_pairs, evals = [], [], [], [], 0
    t_start = time.perf_counter()
    for i in range(args.steps):
        t = req.run()
        lat.append(t["total"])
        phases.append(t)
        st0 = req.stats[0]
        flops.append(pair_flops(st0))
        ms_pairs.append(st0.ms_pairs)
        evals += req.gf_evals()
        log("step %d: %.3f s (assemble %.3f, solve %.3f, velocity %.3f, field %.3f, force %.3f)"
            % (i, t["total"], t["assemble"], t["solve"], t["velocity"], t["field"], t["force"]))
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t_start

    # repeatability of the force between requests (the FP64 path is deterministic)
    F_last = req.F.copy()
    if F_ref is None:
        F_ref = F_last
    force_err = float(max(np.linalg.norm(F_last[k] - F_ref[k]) / np.linalg.norm(F_ref[k]) for k in (0, 1)))

    achieved = float(np.sum(flops) / (np.sum(ms_pairs) * 1e-3) / 1e12)
    st0 = req.stats[0]
    traffic, traffic_src = measured_traffic(args.workload)
    roof = {"bound": "mfma", "pipe": "FP64 vector ALU (MI355X FP64 vector and matrix peaks are both 78.6 TFLOP/s)",
            "kernel": "k_q0_pairs", "achieved": achieved, "peak": PEAK_FP64_TFLOPS,
            "unit": "TFLOP/s", "frac": achieved / PEAK_FP64_TFLOPS, "traffic": traffic, "traffic_source": traffic_src,
            "ms_per_launch": float(np.mean(ms_pairs)),
            "flops_per_launch": float(np.mean(flops)),
            "terms_per_launch": {"spatial": int(st0.pair_terms[0]), "propagating": int(st0.pair_terms[1]),
                                 "evanescent": int(st0.pair_terms[2]), "pure_spectral": int(st0.pair_terms[3])},
            "evals_per_launch": int(st0.pair_evals), "split_pairs": int(st0.pairs_split),
            "gh_pairs": int(st0.pairs_gh), "flops_per_term": F_TERM, "flops_per_free_point": F_FREE}

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
        "force_rel_err": force_err,
        "phases_ms": ph,
        "solver": req.solver,
        "force": {"type0": req.F[0].tolist(), "type1": req.F[1].tolist()},
        "torque": {"type0": req.N[0].tolist(), "type1": req.N[1].tolist()},
        "roofline": roof,
        "cpu_baseline": cpu,
    }
    with quiet_stdout():
        model.finalize()
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()

"""ctypes client of the serving loop of libbqd2_hip.so (bqd2_server_*, include/bqd2.h).

A request is one whole solve of a model (assembly, surface solve, velocity
recovery, optional field points and radiation force).  Models are described by
the packed bqd2_model that libbem3_aw_qd2.so builds from an initialised AQD2
(aqd2_model_new); the server copies it at submission.
"""
import contextlib
import ctypes as C
import os
import tempfile

import numpy as np

from ._native import check, hip_lib, host_lib


class Request(C.Structure):
    _fields_ = [("model", C.c_void_p), ("op_key", C.c_ulonglong), ("solver", C.c_int), ("npts", C.c_int),
                ("pts", C.c_void_p), ("force_type", C.c_int), ("rc", C.c_double * 3)]


class Result(C.Structure):
    _fields_ = [("status", C.c_int), ("reused", C.c_int), ("krylov_steps", C.c_int), ("pad_", C.c_int),
                ("relres", C.c_double), ("F", C.c_double * 3), ("N", C.c_double * 3), ("t_queue", C.c_double),
                ("t_run", C.c_double), ("t_done", C.c_double), ("device_bytes", C.c_size_t),
                ("error", C.c_char * 256)]


def _bind():
    hip, host = hip_lib(), host_lib()
    if getattr(hip, "_serve_bound", False):
        return hip, host
    hip.bqd2_footprint.restype = C.c_size_t
    hip.bqd2_footprint.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_int]
    hip.bqd2_device_bytes.argtypes = [C.c_void_p, C.POINTER(C.c_size_t)]
    hip.bqd2_server_create.argtypes = [C.c_size_t, C.c_int, C.POINTER(C.c_void_p)]
    hip.bqd2_server_submit.argtypes = [C.c_void_p, C.POINTER(Request), C.c_void_p, C.POINTER(C.c_longlong)]
    hip.bqd2_server_wait.argtypes = [C.c_void_p, C.c_longlong, C.POINTER(Result)]
    hip.bqd2_server_stats.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p]
    hip.bqd2_server_destroy.argtypes = [C.c_void_p]
    host.aqd2_model_new.restype = C.c_void_p
    host.aqd2_model_new.argtypes = [C.c_void_p]
    host.aqd2_model_free.argtypes = [C.c_void_p]
    host.aqd2_model_free.restype = None
    hip._serve_bound = True
    return hip, host


@contextlib.contextmanager
def _quiet():
    """Silence the C library's printf output while a model is read."""
    libc = C.CDLL(None)
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


class _ModelHead(C.Structure):
    """the leading fields of bqd2_model (include/bqd2.h) up to its QPD2"""
    _fields_ = [("Ne", C.c_int), ("Nn", C.c_int), ("NN", C.c_int), ("MN", C.c_int), ("ps", C.c_int),
                ("arrays", C.c_void_p * 16), ("qd", C.c_double * 8)]


class PackedModel:
    """The bqd2_model of a workload written by mesh.write_config (keyword
    arguments replace workload entries: n, radius, d1, d2, f, theta, psi, host,
    sphere), with its host AQD2 released once packed."""

    def __init__(self, workload, triangles=False, quiet=True, **over):
        from . import exdat, mesh
        hip, host = _bind()
        self.host = host
        with tempfile.TemporaryDirectory() as tmp:
            paths = mesh.write_config(workload, tmp, triangles=triangles, **over)
            ctx = _quiet() if quiet else contextlib.nullcontext()
            with ctx:
                m = exdat.Model(host).read_inputs(*paths, out=os.path.join(tmp, "ex.dat")).initialize()
                self.ne, self.nn, self.mn = m.ne(), m.nn(), m.mn()
                self.ptr = host.aqd2_model_new(m.buf)
                m.finalize()

    def kxky(self):
        """the in-plane wave vector (kx, ky) of the host medium's QPD2, as the
        lattice sums take it (bqd2_model.qd[1:3])"""
        head = _ModelHead.from_address(self.ptr)
        return head.qd[1], head.qd[2]

    def footprint(self, solver, npts=0, force_type=-1):
        return int(hip_lib().bqd2_footprint(self.ptr, solver, npts, force_type))

    def close(self):
        if self.ptr:
            self.host.aqd2_model_free(self.ptr)
            self.ptr = None

    def __del__(self):
        self.close()


class Server:
    """bqd2_server: `slots` concurrent requests within `budget` device bytes."""

    def __init__(self, budget, slots):
        self.hip, _ = _bind()
        self.s = C.c_void_p()
        check(self.hip.bqd2_server_create(int(budget), int(slots), C.byref(self.s)), "bqd2_server_create")
        self.outs = {}

    def submit(self, model, key=0, solver=0, pts=None, force_type=0, rc=(0.0, 0.0, 0.0)):
        rq = Request()
        rq.model = model.ptr
        rq.op_key = key
        rq.solver = solver
        n = 0 if pts is None else len(pts)
        rq.npts = n
        pts = None if pts is None else np.ascontiguousarray(pts, dtype=np.float64)
        rq.pts = pts.ctypes.data if n else None
        rq.force_type = force_type
        rq.rc[:] = rc
        out = np.zeros((n, 8)) if n else None
        t = C.c_longlong()
        check(self.hip.bqd2_server_submit(self.s, C.byref(rq), out.ctypes.data if n else None, C.byref(t)),
              "bqd2_server_submit")
        self.outs[t.value] = out
        return t.value

    def wait(self, ticket):
        r = Result()
        rc = self.hip.bqd2_server_wait(self.s, ticket, C.byref(r))
        out = self.outs.pop(ticket)
        res = {"status": rc, "reused": r.reused, "krylov_steps": r.krylov_steps, "relres": r.relres,
               "F": np.array(r.F[:]), "N": np.array(r.N[:]), "t_queue": r.t_queue, "t_run": r.t_run,
               "t_done": r.t_done, "device_bytes": int(r.device_bytes), "error": r.error.decode(errors="replace")}
        if out is not None:
            res["field"] = out[:, 0::2] + 1j * out[:, 1::2]
        return res

    def stats(self):
        c = np.zeros(6, np.int64)
        b = np.zeros(2, np.uint64)
        check(self.hip.bqd2_server_stats(self.s, c.ctypes.data, b.ctypes.data), "bqd2_server_stats")
        return {"hits": int(c[0]), "misses": int(c[1]), "evictions": int(c[2]), "failed": int(c[3]),
                "resident": int(c[4]), "pending": int(c[5]), "used_bytes": int(b[0]), "budget_bytes": int(b[1])}

    def close(self):
        if self.s:
            self.hip.bqd2_server_destroy(self.s)
            self.s = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

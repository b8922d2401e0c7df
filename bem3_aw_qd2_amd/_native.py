"""ctypes bindings of libbqd2_hip.so / libbem3_aw_qd2.so.

Loading fails loudly when a library is missing: there is no CPU fallback.
"""
import ctypes as C
import os
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_DIR = os.path.join(HERE, "lib")
CSRC = os.path.join(HERE, "csrc")

_hip = None
_host = None


class BQD2Error(RuntimeError):
    pass


def build(jobs=8):
    """Compile both libraries for gfx950 (hipcc cross-compiles without a GPU)."""
    subprocess.run(["make", "-C", CSRC, "-j%d" % jobs], check=True)


def _load(name):
    path = os.path.join(LIB_DIR, name)
    if not os.path.exists(path):
        raise BQD2Error("%s is not built (run bem3_aw_qd2_amd.build() or make -C %s)" % (path, CSRC))
    return C.CDLL(path, mode=C.RTLD_GLOBAL)


class QPD2(C.Structure):
    """Layout of the reference's QPD2 (qgf2_src/d3_qpgf_d2.h:32-36)."""
    _fields_ = [("k", C.c_double), ("vk", C.c_double * 3),
                ("vd1", C.c_double * 2), ("vd2", C.c_double * 2)]


class GfParams(C.Structure):
    _fields_ = [(n, C.c_double) for n in
                ("k", "vkx", "vky", "d1x", "d1y", "d2x", "d2y", "idd", "aidd",
                 "g1x", "g1y", "g2x", "g2y", "d2n", "g2n", "g12", "E0", "bN", "bNp", "bNm")] + \
               [("nc", C.c_int), ("ewald_only", C.c_int), ("fd_lift", C.c_double)]


# bqd2_progress_fn: (stage, done, total, ctx)
PROGRESS_FN = C.CFUNCTYPE(None, C.c_char_p, C.c_longlong, C.c_longlong, C.c_void_p)


class LargeStats(C.Structure):
    """bqd2_large_stats (include/bqd2.h)"""
    _fields_ = [("windows", C.c_int), ("resident", C.c_int), ("recomputed", C.c_int), ("pad_", C.c_int),
                ("ms_rows", C.c_double), ("ms_build", C.c_double), ("ms_factor", C.c_double),
                ("ms_solve", C.c_double), ("ms_velocity_rows", C.c_double), ("ms_reassemble", C.c_double)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_ if k != "pad_"}


class AssemblyStats(C.Structure):
    _fields_ = [("pairs_gl", C.c_longlong), ("pairs_gh", C.c_longlong), ("self_rows", C.c_longlong),
                ("gf_evals", C.c_longlong), ("gf_terms", C.c_longlong),
                ("ms_decide", C.c_double), ("ms_pairs", C.c_double), ("ms_self", C.c_double),
                ("pair_terms", C.c_longlong * 4), ("pair_evals", C.c_longlong),
                ("self_terms", C.c_longlong * 4), ("self_evals", C.c_longlong),
                ("pairs_split", C.c_longlong), ("pairs_smooth", C.c_longlong), ("free_points", C.c_longlong)]


def hip_lib():
    global _hip
    if _hip is None:
        lib = _load("libbqd2_hip.so")
        lib.bqd2_last_error.restype = C.c_char_p
        lib.bqd2_assemble.argtypes = [C.c_void_p, C.c_int, C.POINTER(AssemblyStats)]
        lib.bqd2_download_domain.argtypes = [C.c_void_p, C.c_int] + [C.c_void_p] * 5
        lib.bqd2_download_rows.argtypes = [C.c_void_p, C.c_int, C.c_int] + [C.c_void_p] * 6
        lib.bqd2_field_noalloc.argtypes = [C.c_void_p, C.c_int, C.c_void_p, C.c_int, C.c_int, C.c_void_p,
                                           C.c_void_p, C.c_void_p, C.c_void_p]
        lib.bqd2_solve_surface.argtypes = [C.c_void_p, C.c_void_p]
        lib.bqd2_solve_velocity.argtypes = [C.c_void_p]
        lib.bqd2_set_solver.argtypes = [C.c_void_p, C.c_int]
        lib.bqd2_solver_info.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]
        lib.bqd2_download_solution.argtypes = [C.c_void_p, C.c_int] + [C.c_void_p] * 4
        lib.bqd2_zgetrf.argtypes = [C.c_int, C.c_void_p, C.c_void_p, C.c_void_p]
        lib.bqd2_zgetrs.argtypes = [C.c_int, C.c_int, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]
        lib.bqd2_cgetrf_wsize.restype = C.c_size_t
        lib.bqd2_cgetrf_wsize.argtypes = [C.c_int]
        lib.bqd2_cgetrf.argtypes = [C.c_int, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]
        lib.bqd2_cgetrs.argtypes = [C.c_int, C.c_int, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]
        lib.bqd2_cgemm.argtypes = [C.c_int, C.c_int, C.c_int, C.c_int, C.c_void_p, C.c_longlong, C.c_void_p,
                                   C.c_longlong, C.c_void_p, C.c_longlong, C.c_void_p]
        lib.bqd2_lu32_info.argtypes = [C.c_void_p, C.POINTER(C.c_double)]
        lib.bqd2_set_lu32_pivoting.argtypes = [C.c_int]
        lib.bqd2_set_lu32_lookahead.argtypes = [C.c_int]
        lib.bqd2_lu32_panel_counts.argtypes = [C.c_void_p]
        lib.bqd2_set_lu32_growth_bound.argtypes = [C.c_double]
        lib.bqd2_lu32_panel_growth.argtypes = [C.c_void_p, C.c_int]
        lib.bqd2_set_incidence.argtypes = [C.c_void_p, C.c_void_p]
        lib.bqd2_assemble_batch.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_void_p]
        lib.bqd2_set_graphs.argtypes = [C.c_int]
        lib.bqd2_set_autotune.argtypes = [C.c_int]
        lib.bqd2_autotune_force.argtypes = [C.c_int, C.c_int]
        lib.bqd2_autotune_info.argtypes = [C.c_int, C.c_int, C.c_int, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]
        lib.bqd2_autotune_reset.argtypes = []
        lib.bqd2_counters.argtypes = [C.c_void_p, C.c_void_p, C.c_int]
        lib.bqd2_field.argtypes = [C.c_void_p, C.c_int, C.c_void_p, C.c_int, C.c_int, C.c_void_p, C.c_void_p,
                                   C.c_void_p]
        lib.bqd2_set_remainder_rules.argtypes = [C.c_void_p, C.c_int]
        lib.bqd2_gauss_legendre.argtypes = [C.c_int, C.c_void_p, C.c_void_p]
        lib.bqd2_set_field_timing.argtypes = [C.c_void_p, C.c_int]
        lib.bqd2_field_kernel_ms.argtypes = [C.c_void_p, C.c_void_p]
        lib.bqd2_set_field_far.argtypes = [C.c_void_p, C.c_int]
        lib.bqd2_field_far_info.argtypes = [C.c_void_p] + [C.c_void_p] * 5
        lib.bqd2_field_bd.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_void_p, C.c_void_p, C.c_int, C.c_int,
                                      C.c_void_p]
        lib.bqd2_force.argtypes = [C.c_void_p, C.c_void_p, C.c_int, C.c_void_p, C.c_void_p]
        lib.bqd2_domain_id.argtypes = [C.c_void_p, C.c_int, C.c_void_p, C.c_void_p]
        lib.bqd2_create_large.argtypes = [C.c_void_p, C.POINTER(C.c_void_p)]
        lib.bqd2_large_config.argtypes = [C.c_void_p, C.c_int, C.c_int]
        lib.bqd2_large_info.argtypes = [C.c_void_p, C.POINTER(LargeStats)]
        lib.bqd2_create.argtypes = [C.c_void_p, C.POINTER(C.c_void_p)]
        lib.bqd2_set_progress.argtypes = [PROGRESS_FN, C.c_void_p]
        lib.bqd2_destroy.argtypes = [C.c_void_p]
        lib.bqd2_gf_params.argtypes = [C.POINTER(C.c_double), C.POINTER(GfParams)]
        lib.bqd2_gf_batch.argtypes = [C.POINTER(GfParams), C.c_int, C.c_int, C.c_void_p, C.c_double,
                                      C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]
        _hip = lib
    return _hip


def host_lib():
    global _host
    if _host is None:
        hip_lib()
        lib = _load("libbem3_aw_qd2.so")
        lib.aqd2_problem.restype = C.c_void_p
        lib.aqd2_problem.argtypes = [C.c_void_p, C.c_void_p]
        lib.aqd2_pull_solution.argtypes = [C.c_void_p, C.c_void_p]
        lib.aqd2_pull_solution.restype = None
        lib.pv_batch_aqd2.argtypes = [C.c_int, C.c_void_p, C.c_int, C.c_int, C.c_void_p, C.c_void_p,
                                      C.c_void_p, C.c_void_p]
        lib.pv_batch_aqd2.restype = None
        lib.pv_t_bd_aqd2.restype = None
        lib.force_FN_aqd2.restype = None
        lib.surface_area_aqd2.restype = None
        lib.aqd2_point_batches.argtypes = [C.c_void_p, C.c_void_p]
        lib.aqd2_point_batches.restype = None
        lib.aqd2_model_new.restype = C.c_void_p
        lib.aqd2_model_new.argtypes = [C.c_void_p]
        lib.aqd2_model_free.argtypes = [C.c_void_p]
        lib.aqd2_model_free.restype = None
        _host = lib
    return _host


def check(rc, what):
    if rc != 0:
        msg = hip_lib().bqd2_last_error().decode(errors="replace")
        raise BQD2Error("%s failed (%d): %s" % (what, rc, msg))


def gf_params(qd):
    """GfParams from a QPD2 (or an 8-sequence k, vk[3], vd1[2], vd2[2])."""
    if isinstance(qd, QPD2):
        vals = [qd.k, qd.vk[0], qd.vk[1], qd.vk[2], qd.vd1[0], qd.vd1[1], qd.vd2[0], qd.vd2[1]]
    else:
        vals = list(qd)
    arr = (C.c_double * 8)(*vals)
    p = GfParams()
    check(hip_lib().bqd2_gf_params(arr, C.byref(p)), "bqd2_gf_params")
    return p

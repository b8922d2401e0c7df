"""Loaders for the checkers used by the tests: the CPU oracle (oracle/liboracle.so,
this repository's C restatement) and the reference's own Green's function compiled
from its sources into oracle/_ref/ (built by oracle/Makefile in the container; the
built file travels with the tree)."""
import ctypes as C
import os

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE = os.path.join(ROOT, "oracle", "liboracle.so")
REF = os.path.join(ROOT, "oracle", "_ref", "libqpgf_ref.so")

_oracle = None
_ref = None


def oracle():
    global _oracle
    if _oracle is None:
        if not os.path.exists(ORACLE):
            pytest.fail("oracle/liboracle.so is not built (make -C oracle)")
        _oracle = C.CDLL(ORACLE)
    return _oracle


def ref_gf():
    """The reference's d3_qpgf_d2.c + Faddeeva.c, or skip when not built."""
    global _ref
    if _ref is None:
        if not os.path.exists(REF):
            pytest.skip("oracle/_ref not built (reference sources absent when building)")
        _ref = C.CDLL(REF)
    return _ref


def ref_eval(mode, qd, pts, eps=1e-15):
    """Evaluate the reference GF at pts (n x 3). Returns G, dG, d2G, ret arrays."""
    lib = ref_gf()
    n = len(pts)
    G = np.zeros(n, complex); dG = np.zeros((n, 3), complex); d2G = np.zeros((n, 6), complex)
    ret = np.zeros(n, np.int32)
    g = (C.c_double * 2)(); dg = (C.c_double * 6)(); d2g = (C.c_double * 12)()
    for i, p in enumerate(pts):
        r = (C.c_double * 3)(*p)
        if mode == 0:
            e = lib.d3hm_qpgf_d2_qG(g, r, C.c_double(eps), C.byref(qd))
        elif mode == 1:
            e = lib.d3hm_qpgf_d2_dqG(dg, r, C.c_double(eps), C.byref(qd))
        elif mode == 2:
            e = lib.d3hm_qpgf_d2_v1(g, dg, r, C.c_double(eps), C.byref(qd))
        else:
            e = lib.d3hm_qpgf_d2_v2(g, dg, d2g, r, C.c_double(eps), C.byref(qd))
        ret[i] = e
        G[i] = complex(g[0], g[1])
        dG[i] = np.array(dg[:]).view(complex)
        d2G[i] = np.array(d2g[:]).view(complex)
    return G, dG, d2G, ret


def oracle_eval(mode, qd, pts, eps=1e-15, cr_exp=False):
    """The oracle's Green's function at pts; cr_exp: the two exponentials whose
    rounding residue the z == 0 stopping rule reads are correctly rounded."""
    lib = oracle()
    lib.orc_set_cr_exp(C.c_int(1 if cr_exp else 0))
    try:
        return _oracle_eval(lib, mode, qd, pts, eps)
    finally:
        lib.orc_set_cr_exp(C.c_int(0))


def _oracle_eval(lib, mode, qd, pts, eps):
    n = len(pts)
    r = np.ascontiguousarray(pts, dtype=np.float64)
    G = np.zeros(n, complex); dG = np.zeros((n, 3), complex); d2G = np.zeros((n, 6), complex)
    ret = np.zeros(n, np.int32)
    lib.orc_qpgf_batch(C.c_int(mode), C.byref(qd), C.c_int(n), r.ctypes.data_as(C.c_void_p), C.c_double(eps),
                       G.ctypes.data_as(C.c_void_p), dG.ctypes.data_as(C.c_void_p),
                       d2G.ctypes.data_as(C.c_void_p), ret.ctypes.data_as(C.c_void_p))
    return G, dG, d2G, ret

"""GPU coefficient matrices of the reference's shipped sample against the CPU
oracle, row by row (G, H with the 1+F / F diagonal, dG, dH, dF).

The oracle's rows are themselves pinned to the reference's ex.dat solution
(tests/test_oracle_sample.py).  Bar: 1e-10 normwise per row for G, H (periodic
and interior domains) and dG, dH, dF (interior domain)."""
import ctypes as C
import os

import numpy as np
import pytest

from bem3_aw_qd2_amd import AssemblyStats, check, exdat, hip_lib, host_lib
from tests.checkers import oracle

pytestmark = pytest.mark.gpu
HERE = os.path.dirname(os.path.abspath(__file__))
SAMPLE = os.path.join(HERE, "golden", "sample_ex.dat")


def oracle_row(buf, did, t, tn, Ne):
    N = 4 * Ne
    G = np.zeros(N, complex); H = np.zeros(N, complex)
    dG = np.zeros(2 * N, complex); dH = np.zeros(2 * N, complex); dF = np.zeros(3)
    oracle().orc_assemble_row.restype = C.c_long
    c = oracle().orc_assemble_row(buf, did, t, tn, G.ctypes.data_as(C.c_void_p), H.ctypes.data_as(C.c_void_p),
                                  dG.ctypes.data_as(C.c_void_p), dH.ctypes.data_as(C.c_void_p),
                                  dF.ctypes.data_as(C.c_void_p))
    assert c >= 0
    return G, H, dG, dH, dF


def rel(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


@pytest.mark.parametrize("did", [1, 0])
def test_sample_rows_match_oracle(did):
    ref = exdat.read(SAMPLE)
    m = exdat.Model(host_lib()).dat_read(SAMPLE)
    prob = host_lib().aqd2_problem(m.buf, None)
    st = AssemblyStats()
    check(hip_lib().bqd2_assemble(prob, did, C.byref(st)), "assemble")
    Ne = ref["dom"][did]["Ne"]
    N = 4 * Ne
    G = np.zeros((N, N), complex); H = np.zeros((N, N), complex); dF = np.zeros((N, 3))
    dG = np.zeros((2 * N, N), complex) if did else None
    dH = np.zeros((2 * N, N), complex) if did else None
    ptr = lambda a: a.ctypes.data_as(C.c_void_p) if a is not None else None
    check(hip_lib().bqd2_download_domain(prob, did, ptr(G), ptr(H), ptr(dG), ptr(dH), ptr(dF)), "download")
    print("domain", did, "pairs GL/GH", st.pairs_gl, st.pairs_gh, "GF evals", st.gf_evals, "terms", st.gf_terms,
          "ms", st.ms_decide, st.ms_pairs, st.ms_self)
    rows = [(1, 0), (2, 3), (Ne // 2, 1), (Ne, 2)] if did else [(1, 0), (Ne // 3, 2)]
    for t, tn in rows:
        r = (t - 1) * 4 + tn
        oG, oH, odG, odH, odF = oracle_row(m.buf, did, t, tn, Ne)
        eG, eH = rel(G[r], oG), rel(H[r], oH)
        msg = "did=%d t=%d tn=%d G %.2e H %.2e" % (did, t, tn, eG, eH)
        assert eG < 1e-10 and eH < 1e-10, msg
        assert abs(dF[r, 0] - odF[0]) <= 1e-10 * abs(odF[0]), msg
        if did:
            eg = rel(dG[2 * r:2 * r + 2], odG.reshape(2, N))
            eh = rel(dH[2 * r:2 * r + 2], odH.reshape(2, N))
            # dF/dv summed over a closed surface cancels to ~1e-11 of its terms
            # (the solid angle is constant on a smooth boundary): compare on the
            # scale of the sum of |terms|
            parts = np.zeros(3 * Ne)
            oracle().orc_row_fparts(m.buf, did, t, tn, parts.ctypes.data_as(C.c_void_p))
            scale = np.abs(parts.reshape(Ne, 3)[:, 1:]).sum(axis=0)
            ef = (np.abs(dF[r, 1:] - odF[1:]) / scale).max()
            assert eg < 1e-10 and eh < 1e-10 and ef < 1e-12, msg + " dG %.2e dH %.2e dF %.2e" % (eg, eh, ef)
    m.finalize()

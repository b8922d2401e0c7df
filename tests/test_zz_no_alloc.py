"""bqd2_field_noalloc: the field entry point that allocates nothing, against
bqd2_field on the same solved problem.

Two calls in a row (the second with fewer points, so stale partial sums of the
first would show), total and scattered fields, on the 96-quad sphere of the
2 mm lattice: the device outputs must be the bits bqd2_field returns (the same
kernels on the same chunking), the domain ids equal, no Green's function
errors.  A request beyond the fixed scratch array is refused with BQD2_E_ARG
before anything runs."""
import ctypes as C
import os
import tempfile

import numpy as np
import pytest

from bem3_aw_qd2_amd import exdat, hip_lib, host_lib, mesh

pytestmark = pytest.mark.gpu


def test_field_noalloc_matches_field():
    import torch
    tmp = tempfile.mkdtemp()
    paths = mesh.write_config("sphere96", tmp)
    m = exdat.Model(host_lib()).read_inputs(*paths, out=os.path.join(tmp, "ex.dat")).initialize()
    m.solve()
    prob = host_lib().aqd2_problem(m.buf, None)
    hip = hip_lib()
    hip.bqd2_field_noalloc.argtypes = [C.c_void_p, C.c_int, C.c_void_p, C.c_int, C.c_int, C.c_void_p, C.c_void_p,
                                       C.c_void_p, C.c_void_p]
    rng = np.random.default_rng(4)
    stream = torch.cuda.current_stream().cuda_stream
    try:
        for n, type_, total in ((300, 0, 1), (117, 2, 0)):
            pts = np.ascontiguousarray(rng.uniform(-1.5e-3, 1.5e-3, (n, 3)))
            out = np.zeros((n, 8))
            did = np.zeros(n, np.int32)
            assert hip.bqd2_field(prob, n, pts.ctypes.data, type_, total, out.ctypes.data, did.ctypes.data,
                                  None) == 0
            d_pts = torch.from_numpy(pts).cuda()
            d_out = torch.full((n, 8), float("nan"), dtype=torch.float64, device="cuda")
            d_did = torch.full((n,), -7, dtype=torch.int32, device="cuda")
            d_err = torch.full((1,), -1, dtype=torch.int64, device="cuda")
            print("noalloc call: n %d type %d total %d" % (n, type_, total), flush=True)
            rc = hip.bqd2_field_noalloc(prob, n, d_pts.data_ptr(), type_, total, d_out.data_ptr(), d_did.data_ptr(),
                                        d_err.data_ptr(), stream)
            assert rc == 0
            torch.cuda.synchronize()
            assert int(d_err.item()) == 0
            assert (d_did.cpu().numpy() == did).all()
            got = d_out.cpu().numpy()
            assert np.array_equal(got, out), np.abs(got - out).max()
        # beyond the scratch array: refused before any launch
        big = 1 << 20
        rc = hip.bqd2_field_noalloc(prob, big, d_pts.data_ptr(), 3, 1, d_out.data_ptr(), d_did.data_ptr(), None,
                                    stream)
        assert rc == -1
    finally:
        m.finalize()

"""Write tests/golden/force_oracle.json: the oracle's radiation force and torque
(oracle/force_cpu.c, a restatement of force_FN_aqd2, bem3_aw_qd2_force.c:4-278)
for the reference's four solved samples, from the surface solution each ex.dat
holds, with both quadrature settings (type 0: 4 node values per element; type 1:
the boundary field at 9 points per quad).

Type 1 is computed twice: with the exponentials of the z == 0 stopping rule of
the lattice sum as this machine's libm rounds them (what the reference compiled
here computes) and correctly rounded (the GPU's choice).  The difference between
the two is the reference's own sensitivity to its libm.

The type-1 force sums O(Ne^2) lattice Green's functions and the polar rule of
every element at 9 points: about 8 minutes on 8 cores for the 50 mm lattice.
Run in the container (the oracle is built there):

    python tests/golden/make_force_fixtures.py [sample ...]

Existing entries of the file are kept; named samples are recomputed.
"""
import ctypes as C
import json
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

SAMPLES = {
    "sample": "sample_ex.dat",
    "analysis_sample1": "analysis_sample1_ex1.dat",
    "analysis_sample2": "analysis_sample2_ex2.dat",
    "verification": "verification_veri.dat",
}
OUT = os.path.join(HERE, "force_oracle.json")


def oracle_force(o, m, rc, typ, cr):
    F = np.zeros(3)
    N = np.zeros(3)
    o.orc_set_cr_exp(C.c_int(cr))
    try:
        err = o.orc_force(m.buf, rc.ctypes.data_as(C.c_void_p), typ, F.ctypes.data_as(C.c_void_p),
                          N.ctypes.data_as(C.c_void_p))
    finally:
        o.orc_set_cr_exp(C.c_int(0))
    assert err == 0, err
    return F.tolist(), N.tolist()


def main(names):
    from bem3_aw_qd2_amd import exdat, host_lib
    o = C.CDLL(os.path.join(ROOT, "oracle", "liboracle.so"))
    data = json.load(open(OUT)) if os.path.exists(OUT) else {}
    for name in names:
        log = json.load(open(os.path.join(HERE, name + "_example1.json")))
        rc = np.array(log["rc"], float)
        m = exdat.Model(host_lib()).dat_read(os.path.join(HERE, SAMPLES[name]))
        entry = {"rc": log["rc"]}
        t0 = time.time()
        F, N = oracle_force(o, m, rc, 0, 0)
        entry["type0"] = {"F": F, "N": N}
        for key, cr in (("type1_libm", 0), ("type1_cr", 1)):
            F, N = oracle_force(o, m, rc, 1, cr)
            entry[key] = {"F": F, "N": N}
            print(name, key, F, N, "%.0f s" % (time.time() - t0), flush=True)
        m.finalize()
        data[name] = entry
        with open(OUT, "w") as f:
            json.dump(data, f, indent=1)


if __name__ == "__main__":
    main(sys.argv[1:] or sorted(SAMPLES))

"""Reader for the ex.dat layout (reference d3qd2_src/bem3_aw_qd2.c:185-224) and a
thin holder of an AQD2 image living in the host library.

Only used by tests, bench.py and tooling to inspect results; the product path is
the C library itself."""
import ctypes as C
import struct

import numpy as np

AQD2_SIZE = 2720
APW_SIZE = 240
# byte offsets inside the AQD2 image (checked against sizeof/offsetof in tests)
OFF_MN = 440
OFF_BD = 736
OFF_PS = OFF_BD + 0
OFF_QD = OFF_BD + 8
OFF_NN_NODES = OFF_BD + 72   # Nn
OFF_NE = OFF_BD + 76         # Ne
OFF_NNCOLL = OFF_BD + 96     # NN


def parse(raw):
    """Parse an ex.dat byte string into a dict of numpy arrays."""
    MN = struct.unpack_from("i", raw, OFF_MN)[0]
    Nn = struct.unpack_from("i", raw, OFF_NN_NODES)[0]
    Ne = struct.unpack_from("i", raw, OFF_NE)[0]
    NN = struct.unpack_from("i", raw, OFF_NNCOLL)[0]
    ps = struct.unpack_from("i", raw, OFF_PS)[0]
    qd = np.frombuffer(raw, np.float64, 8, OFF_QD).copy()
    off = AQD2_SIZE
    nd = MN + 1
    out = {"MN": MN, "Nn": Nn, "Ne": Ne, "NN": NN, "ps": ps, "qd": qd}
    for name in ("rho0", "c0", "k0", "K0"):
        out[name] = np.frombuffer(raw, np.float64, nd, off).copy(); off += 8 * nd
    for name in ("k1", "k2"):
        out[name] = np.frombuffer(raw, np.complex128, nd, off).copy(); off += 16 * nd
    out["pw"] = raw[off:off + APW_SIZE]; off += APW_SIZE
    out["rn"] = np.frombuffer(raw, np.float64, (Nn + 1) * 3, off).reshape(Nn + 1, 3).copy(); off += 24 * (Nn + 1)
    out["ed"] = np.frombuffer(raw, np.int32, (Ne + 1) * 4, off).reshape(Ne + 1, 4).copy(); off += 16 * (Ne + 1)
    out["eni"] = np.frombuffer(raw, np.int32, (Ne + 1) * 4, off).reshape(Ne + 1, 4).copy(); off += 16 * (Ne + 1)
    for name in ("md", "sd", "gd"):
        out[name] = np.frombuffer(raw, np.int32, Ne + 1, off).copy(); off += 4 * (Ne + 1)
    for name in ("ren", "wen"):
        out[name] = np.frombuffer(raw, np.float64, (Ne + 1) * 12, off).reshape(Ne + 1, 4, 3).copy(); off += 96 * (Ne + 1)
    for name in ("Pi", "dPi"):
        out[name] = np.frombuffer(raw, np.complex128, (Ne + 1) * 4, off).reshape(Ne + 1, 4).copy(); off += 64 * (Ne + 1)
    doms = []
    for d in range(nd):
        ne = struct.unpack_from("i", raw, off)[0]; off += 4
        dd = {"Ne": ne}
        dd["sid"] = np.frombuffer(raw, np.int32, ne + 1, off).copy(); off += 4 * (ne + 1)
        for name in ("P", "dP"):
            dd[name] = np.frombuffer(raw, np.complex128, (ne + 1) * 4, off).reshape(ne + 1, 4).copy(); off += 64 * (ne + 1)
        for name in ("pv", "dpv"):
            dd[name] = np.frombuffer(raw, np.complex128, (ne + 1) * 12, off).reshape(ne + 1, 4, 3).copy(); off += 192 * (ne + 1)
        doms.append(dd)
    out["dom"] = doms
    out["size"] = off
    return out


def read(path):
    with open(path, "rb") as f:
        return parse(f.read())


def relerr(a, b):
    """normwise relative difference over the used entries (row 0 is unused)."""
    a = np.asarray(a)[1:]
    b = np.asarray(b)[1:]
    return float(np.linalg.norm((a - b).ravel()) / np.linalg.norm(b.ravel()))


class Model:
    """An AQD2 held in a ctypes buffer and driven through libbem3_aw_qd2.so."""

    def __init__(self, lib):
        self.lib = lib
        self.buf = C.create_string_buffer(AQD2_SIZE)

    @property
    def ptr(self):
        return self.buf

    def dat_read(self, path):
        self.lib.dat_read_aqd2(path.encode(), self.buf)
        return self

    def dat_write(self, path):
        self.lib.dat_write_aqd2(path.encode(), self.buf)

    def read_inputs(self, pw, per, med, msh, out="ex.dat"):
        args = [b"solver", pw.encode(), per.encode(), med.encode(), msh.encode(), out.encode()]
        argv = (C.c_char_p * len(args))(*args)
        self.lib.read_aqd2(len(args), argv, self.buf)
        return self

    def initialize(self):
        self.lib.initialize_aqd2(self.buf)
        return self

    def solve(self):
        self.lib.solve_bieq_aqd2(self.buf)
        return self

    def finalize(self):
        self.lib.finalize_aqd2(self.buf)

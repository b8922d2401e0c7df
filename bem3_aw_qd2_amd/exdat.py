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
        dd = {"Ne": ne, "offsets": {}}
        dd["sid"] = np.frombuffer(raw, np.int32, ne + 1, off).copy(); off += 4 * (ne + 1)
        for name in ("P", "dP"):
            dd["offsets"][name] = off
            dd[name] = np.frombuffer(raw, np.complex128, (ne + 1) * 4, off).reshape(ne + 1, 4).copy(); off += 64 * (ne + 1)
        for name in ("pv", "dpv"):
            dd["offsets"][name] = off
            dd[name] = np.frombuffer(raw, np.complex128, (ne + 1) * 12, off).reshape(ne + 1, 4, 3).copy(); off += 192 * (ne + 1)
        doms.append(dd)
    out["dom"] = doms
    out["size"] = off
    return out


def with_solution(raw, d, did, **arrays):
    """Copy of an ex.dat byte string with domain did's P, dP, pv and/or dpv
    replaced (arrays shaped as parse returns them)."""
    out = bytearray(raw)
    dom = d["dom"][did]
    for name, a in arrays.items():
        a = np.ascontiguousarray(a, dtype=np.complex128)
        assert a.shape == dom[name].shape, name
        o = dom["offsets"][name]
        out[o:o + a.nbytes] = a.tobytes()
    return bytes(out)


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

    def _int(self, off):
        return struct.unpack_from("i", self.buf.raw, off)[0]

    def mn(self):
        """media besides the host (domains 0..MN)"""
        return self._int(OFF_MN)

    def ne(self):
        """boundary elements"""
        return self._int(OFF_NE)

    def nn(self):
        """collocation nodes (the system has 2 NN unknowns)"""
        return self._int(OFF_NNCOLL)

    @staticmethod
    def qpd2_offset():
        """byte offset of the host medium's QPD2 inside the AQD2 image"""
        return OFF_QD

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

    # ----------------------------------------------------------- field API
    def _field1(self, fn, r, type_):
        p = (C.c_double * 2)()
        v = (C.c_double * 6)()
        rt = (C.c_double * 3)(*r)
        did = fn(p, v, rt, int(type_), self.buf)
        return did, complex(p[0], p[1]), np.array([complex(v[2 * i], v[2 * i + 1]) for i in range(3)])

    def pv_t(self, r, type_=0):
        """total field (pv_t_aqd2): (domain id, p, v[3])"""
        return self._field1(self.lib.pv_t_aqd2, r, type_)

    def pv_s(self, r, type_=0):
        """scattered field (pv_s_aqd2)"""
        return self._field1(self.lib.pv_s_aqd2, r, type_)

    def pv_i(self, r, type_=0):
        """incident field (pv_i_aqd2)"""
        return self._field1(self.lib.pv_i_aqd2, r, type_)

    def domain_id(self, r):
        rt = (C.c_double * 3)(*r)
        return self.lib.domain_id_aqd2(rt, self.buf)

    def pv_batch(self, pts, type_=0, total=1):
        """field at n points in one device call (pv_batch_aqd2):
        returns p (n,), v (n, 3) complex and the domain ids (n,)."""
        pts = np.ascontiguousarray(pts, dtype=np.float64).reshape(-1, 3)
        n = len(pts)
        p = np.zeros(n, dtype=np.complex128)
        v = np.zeros((n, 3), dtype=np.complex128)
        did = np.zeros(n, dtype=np.int32)
        self.lib.pv_batch_aqd2(n, pts.ctypes.data, int(type_), int(total), p.ctypes.data, v.ctypes.data,
                               did.ctypes.data, self.buf)
        return p, v, did

    def pv_t_bd(self, did, t, zeta, eta, type_=0):
        """total field on the boundary (pv_t_bd_aqd2): (p, v[3])"""
        p = (C.c_double * 2)()
        v = (C.c_double * 6)()
        self.lib.pv_t_bd_aqd2(p, v, int(did), int(t), C.c_double(zeta), C.c_double(eta), int(type_), self.buf)
        return complex(p[0], p[1]), np.array([complex(v[2 * i], v[2 * i + 1]) for i in range(3)])

    def force(self, rc=(0.0, 0.0, 0.0), type_=0):
        """radiation force and torque about rc (force_FN_aqd2)"""
        F = (C.c_double * 3)()
        N = (C.c_double * 3)()
        r = (C.c_double * 3)(*rc)
        self.lib.force_FN_aqd2(F, N, r, int(type_), self.buf)
        return np.array(F[:]), np.array(N[:])

    def surface_area(self, did, type_=0):
        S = C.c_double()
        self.lib.surface_area_aqd2(C.byref(S), int(did), int(type_), self.buf)
        return S.value


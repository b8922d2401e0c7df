"""The two shared libraries load without a GPU and export every function that
include/bem3_aw_qd2.h and include/bqd2.h declare; the public structs keep the
reference's layout."""
import ctypes as C
import os
import re

import pytest

from bem3_aw_qd2_amd import exdat, hip_lib, host_lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared(header):
    text = open(os.path.join(ROOT, "include", header)).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    text = re.sub(r"//[^\n]*", "", text)
    text = re.sub(r"typedef struct[^;]*?\}[^;]*;", "", text, flags=re.S)
    text = re.sub(r"typedef[^{;]*;", "", text)  # function-pointer and other brace-less typedefs
    names = re.findall(r"^\s*(?:const\s+)?[A-Za-z_][\w\s\*]*?\b([A-Za-z_]\w*)\s*\([^;{]*\)\s*;", text, flags=re.M)
    return sorted(set(n for n in names if n not in ("if", "while", "return")))


@pytest.mark.parametrize("header,lib", [("bem3_aw_qd2.h", "host"), ("bqd2.h", "hip")])
def test_every_declared_function_is_exported(header, lib):
    names = declared(header)
    assert len(names) > 10
    so = host_lib() if lib == "host" else hip_lib()
    missing = [n for n in names if not hasattr(so, n)]
    assert not missing, missing


def test_reference_entry_points_present():
    # the prototypes of the reference's d3qd2_src/bem3_aw_qd2.h:13-96
    names = ["read_aqd2", "print_aqd2", "initialize_aqd2", "finalize_aqd2", "domain_id_aqd2", "domain_id_l_aqd2",
             "dat_write_aqd2", "dat_read_aqd2", "solve_bieq_aqd2", "force_FN_aqd2", "surface_area_aqd2"]
    for w in ("s", "t", "i"):
        names += ["phi_%s_aqd2" % w, "p_%s_aqd2" % w, "pv_%s_aqd2" % w, "phi_%s_bd_aqd2" % w, "pv_%s_bd_aqd2" % w]
    for n in names:
        assert hasattr(host_lib(), n), n


def test_struct_sizes():
    lib = host_lib()
    assert exdat.AQD2_SIZE == 2720
    # sizeof/offsetof as compiled into the host library
    lib.aqd2_layout.argtypes = [C.c_void_p, C.c_int]
    out = (C.c_long * 8)()
    n = lib.aqd2_layout(out, 8)
    got = list(out[:n])
    # sizeof(AQD2), offsetof(AQD2, bd), offsetof(BOUD, sb), sizeof(Apw), sizeof(QPD2), offsetof(AQD2, MN)
    assert got[:6] == [2720, 736, 1976, 240, 64, 440]

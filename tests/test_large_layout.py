"""Host-side checks of the system layout (bqd2_create_large), without a GPU: the
layout holds one sphere of quadrilaterals in a periodic lattice, each element
with its own four nodes, and rejects anything else with BQD2_E_ARG and a message
before any device call."""
import ctypes as C

import pytest

from bem3_aw_qd2_amd import hip_lib, serve

BQD2_E_ARG = -1


def create(model):
    hip = hip_lib()
    p = C.c_void_p()
    rc = hip.bqd2_create_large(model.ptr, C.byref(p))
    return rc, hip.bqd2_last_error().decode()


@pytest.mark.parametrize("kw, word", [({"triangles": True}, "four nodes"), ({"periodic": False}, "periodic")])
def test_layout_rejects(kw, word):
    m = serve.PackedModel("sphere24", **kw)
    try:
        rc, msg = create(m)
    finally:
        m.close()
    print(kw, rc, msg)
    assert rc == BQD2_E_ARG and word in msg


def test_config_needs_a_layout_problem():
    hip = hip_lib()
    assert hip.bqd2_large_config(None, 2048, -1) == BQD2_E_ARG

"""Host side of the serving loop, without a GPU: the packed model of an
initialised AQD2 (aqd2_model_new) and the device-memory footprint estimate the
server admits requests by (bqd2_footprint), and a server that starts and stops
its worker threads.

The footprint must cover the dense matrices it names: per domain G and H
(2 N^2 complex doubles, N = 4 Ne), dG and dH of the sphere domain (4 N^2), the
2NN system (na^2), plus the complex float factors (na^2 complex floats) for the
mixed-precision solver, so it grows like Ne^2."""
import pytest

from bem3_aw_qd2_amd import serve


@pytest.fixture(scope="module")
def models():
    return {n: serve.PackedModel(n) for n in ("sphere24", "sphere96", "sphere384")}


def test_packed_model_sizes(models):
    m = models["sphere96"]
    assert (m.ne, m.nn, m.mn) == (96, 384, 1)


@pytest.mark.parametrize("name", ["sphere24", "sphere96", "sphere384"])
def test_footprint_covers_the_matrices(models, name):
    m = models[name]
    N, na = 4 * m.ne, 2 * m.nn
    dense = (2 + 2 + 4) * N * N * 16 + na * na * 16
    f0, f2 = m.footprint(0), m.footprint(2)
    print(name, "footprint", f0, f2, "dense", dense)
    assert f0 >= dense and f0 < 2 * dense + 64 * 2 ** 20
    assert f2 - f0 >= na * na * 8
    assert m.footprint(0, npts=4096, force_type=1) >= f0


def test_footprint_grows_like_ne_squared(models):
    r = models["sphere384"].footprint(0) / models["sphere96"].footprint(0)
    assert 12.0 < r < 17.0


def test_server_starts_and_stops():
    with serve.Server(2 ** 30, 4) as s:
        st = s.stats()
        assert st["budget_bytes"] == 2 ** 30 and st["pending"] == 0 and st["resident"] == 0

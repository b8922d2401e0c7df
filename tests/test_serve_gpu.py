"""The serving loop on the GPU: continuous admission, resident operators,
concurrent execution, and results that do not depend on the mix.

Stream: three operators (24 quads on the 2 mm square lattice; 96 quads on the
1.8 mm triangular lattice at theta = 30 deg; 96 quads on a 2.4 mm square
lattice at 0.8 MHz with the mixed-precision solver), each requested again with
the same key and the mirrored incidence pi - theta, and the first two once more
with another amplitude and phase; one request has no key.  Every request asks
for the total field at 36 points and the type-0 force.

* Alone: each request on a fresh one-slot server (its own assembly).
* In the mix: all submitted at once to a four-slot server.  A resident
  operator serves a request only when the request's (kx, ky) has the
  operator's exact bits: the amplitude repeats always do, a mirrored request
  only when sin(pi - theta) rounds like sin(theta) (otherwise it is a miss and
  gets its own operator).  Each result must equal its alone result within 1e-12
  (relative, force and field), and be the same bits when it ran on its own
  operator; the lattice sums stop on a rounding residue where a point is level
  with a source, so an operator reused for a wave vector equal only up to
  rounding could move the force far beyond that (5.3e-7 measured before the
  exact-bits rule).
* Under a memory budget of 1.1 times the largest footprint: requests are
  admitted one operator at a time, idle operators are evicted, nothing fails,
  the results are the same; a request larger than the budget fails with
  BQD2_E_NOMEM.
* The device bytes a request's problem holds never exceed its footprint."""
import math

import numpy as np
import pytest

from bem3_aw_qd2_amd import serve

pytestmark = pytest.mark.gpu
TRI = dict(d1=(1.8e-3, 0.0), d2=(0.9e-3, 0.9e-3 * math.sqrt(3.0)), theta=math.pi / 6, psi=math.pi / 12)
SQ24 = dict(d1=(2.4e-3, 0.0), d2=(0.0, 2.4e-3), f=0.8e6, theta=0.4, psi=0.3)


def mirrored(over, base_theta):
    o = dict(over)
    o["theta"] = math.pi - o.get("theta", base_theta)
    return o


# (workload, overrides, key, solver, base request whose operator it may share)
STREAM = [
    ("sphere24", {}, 1, 0, None),
    ("sphere96", TRI, 2, 0, None),
    ("sphere96", SQ24, 3, 2, None),
    ("sphere24", mirrored({}, 1.0e-6), 1, 0, 0),
    ("sphere96", mirrored(TRI, 0.0), 2, 0, 1),
    ("sphere96", mirrored(SQ24, 0.0), 3, 2, 2),
    ("sphere24", dict(theta=0.2, psi=0.7), 0, 2, None),
    ("sphere24", dict(p=(0.3, -0.8)), 1, 0, 0),
    ("sphere96", dict(TRI, p=(-0.5, 0.2)), 2, 0, 1),
]
PTS = np.array([[x, 0.0, z] for x in np.linspace(-1.5e-3, 1.5e-3, 6) for z in np.linspace(-1.5e-3, 1.5e-3, 6)])


@pytest.fixture(scope="module")
def models():
    return [serve.PackedModel(w, **o) for w, o, _, _, _ in STREAM]


@pytest.fixture(scope="module")
def alone(models):
    out = []
    for m, (_, _, key, solver, _) in zip(models, STREAM):
        with serve.Server(64 * 2 ** 30, 1) as s:
            r = s.wait(s.submit(m, key=key, solver=solver, pts=PTS, force_type=0))
        assert r["status"] == 0, r["error"]
        assert r["reused"] == 0
        out.append(r)
    return out


def same(a, b, bar=1e-12):
    eF = np.linalg.norm(a["F"] - b["F"]) / np.linalg.norm(b["F"])
    eP = np.linalg.norm(a["field"] - b["field"]) / np.linalg.norm(b["field"])
    return eF, eP, eF <= bar and eP <= bar


def expected_reuse(models):
    """request i may run on the operator of its base request only when their
    (kx, ky) have the same bits"""
    out = []
    for m, (_, _, k, _, base) in zip(models, STREAM):
        out.append(int(k != 0 and base is not None and m.kxky() == models[base].kxky()))
    return out


def test_mix_matches_alone(models, alone):
    reuse = expected_reuse(models)
    assert reuse[7] == 1 and reuse[8] == 1   # amplitude and phase only: same wave vector
    with serve.Server(64 * 2 ** 30, 4) as s:
        tickets = [s.submit(m, key=k, solver=sv, pts=PTS, force_type=0)
                   for m, (_, _, k, sv, _) in zip(models, STREAM)]
        res = [s.wait(t) for t in tickets]
        st = s.stats()
    print("stats", st, "expected reuse", reuse)
    for i, (r, a) in enumerate(zip(res, alone)):
        assert r["status"] == 0, r["error"]
        eF, eP, ok = same(r, a)
        print("request %d key %d reused %d steps %d queue %.3f s run %.3f s: force %.1e field %.1e"
              % (i, STREAM[i][2], r["reused"], r["krylov_steps"], r["t_queue"], r["t_run"], eF, eP))
        assert ok
        assert r["reused"] == reuse[i]
        if not r["reused"]:
            assert eF == 0.0 and eP == 0.0  # its own operator, same incidence: the same bits
    n_keyed = len({(STREAM[i][2], models[i].kxky()) for i in range(len(STREAM)) if STREAM[i][2]})
    assert st["hits"] == sum(reuse) and st["misses"] == len(STREAM) - sum(reuse) and st["failed"] == 0
    assert st["resident"] == n_keyed


def test_budget_admission(models, alone):
    need = max(m.footprint(sv, len(PTS), 0) for m, (_, _, _, sv, _) in zip(models, STREAM))
    with serve.Server(int(1.1 * need), 4) as s:
        tickets = [s.submit(m, key=k, solver=sv, pts=PTS, force_type=0)
                   for m, (_, _, k, sv, _) in zip(models, STREAM)]
        res = [s.wait(t) for t in tickets]
        st = s.stats()
        big = serve.PackedModel("sphere384")
        assert big.footprint(0, len(PTS), 0) > 1.1 * need
        r = s.wait(s.submit(big, key=9, solver=0, pts=PTS, force_type=0))
        assert r["status"] == -5 and "budget" in r["error"]
    print("stats", st)
    assert st["failed"] == 0 and st["evictions"] >= 1
    for r, a in zip(res, alone):
        assert r["status"] == 0, r["error"]
        assert same(r, a)[2]


def test_device_bytes_within_footprint(models, alone):
    for m, (_, _, _, sv, _), a in zip(models, STREAM, alone):
        f = m.footprint(sv, len(PTS), 0)
        print("footprint %d device bytes %d" % (f, a["device_bytes"]))
        assert 0 < a["device_bytes"] <= f

"""The partial-wave series of tests/fluid_sphere.py, pinned without the
boundary-element code:

* Rayleigh's expansion of the incident wave reproduces exp(i k d.x);
* the total field satisfies both interface conditions (pressure and normal
  velocity, by central differences in r) on the sphere;
* unitarity of a lossless sphere, |1 + 2 A_n| = 1, and sigma_ext = sigma_sca;
* the rigid (rho1 -> inf) and pressure-release (rho1 -> 0) limits give the
  textbook coefficients -j_n'/h_n' and -j_n/h_n;
* r exp(-i k r) p_s tends to p0 f far away;
* the far-field force equals the time-averaged Brillouin stress integrated
  over a sphere of radius 1.5 a around the scatterer, -int (<p2> I + rho <v v>).n dS,
  with the velocity from the series by differences, <p2> = <p^2>/(2 rho c^2) - rho <v^2>/2."""
import numpy as np
import pytest
from scipy.special import spherical_jn, spherical_yn

from tests.fluid_sphere import FluidSphere

A0 = 0.5e-3
D = np.array([0.3, -0.2, 0.9]) / np.linalg.norm([0.3, -0.2, 0.9])


def water_sphere(**kw):
    args = dict(a=A0, rho0=1000.0, c0=1500.0, rho1=1200.0, c1=2400.0, f=1.0e6, p0=0.7 - 0.4j, d=D)
    args.update(kw)
    return FluidSphere(**args)


def directions(m=7, seed=3):
    v = np.random.default_rng(seed).standard_normal((m, 3))
    return v / np.linalg.norm(v, axis=1)[:, None]


def test_rayleigh_expansion():
    s = water_sphere(nmax=40)
    for u in directions():
        for r in (0.2 * A0, A0, 3.0 * A0):
            x = r * u
            assert abs(s.incident_series(x) - s.incident(x)) < 1e-13


@pytest.mark.parametrize("rho1,c1", [(1200.0, 2400.0), (800.0, 900.0)])
def test_interface_conditions(rho1, c1):
    s = water_sphere(rho1=rho1, c1=c1)
    h = 1e-5 * A0
    for u in directions():
        po = lambda r: s.incident(r * u) + s.scattered(r * u)
        pi = lambda r: s.interior(r * u)
        assert abs(po(A0) - pi(A0)) < 1e-13 * abs(po(A0)) + 1e-15
        vo = (po(A0 + h) - po(A0 - h)) / (2 * h) / s.rho0
        vi = (pi(A0 + h) - pi(A0 - h)) / (2 * h) / rho1
        assert abs(vo - vi) < 1e-7 * (abs(vo) + 1e-3 * s.k0 * abs(s.p0) / s.rho0)


def test_unitarity_and_optical_theorem():
    s = water_sphere()
    assert np.abs(np.abs(1 + 2 * s.A) - 1.0).max() < 1e-13
    s_ext, s_sca, _ = s.cross_sections()
    assert abs(s_ext - s_sca) < 1e-12 * s_sca


def test_rigid_and_soft_limits():
    x = 1000.0 * 2 * np.pi * 1e6 / 1500.0 * A0 / 1000.0
    n = np.arange(12)
    j, dj = spherical_jn(n, x), spherical_jn(n, x, derivative=True)
    h = j + 1j * spherical_yn(n, x)
    dh = dj + 1j * spherical_yn(n, x, derivative=True)
    rigid = water_sphere(rho1=1e15, nmax=11)
    soft = water_sphere(rho1=1e-12, nmax=11)
    assert np.abs(rigid.A - (-dj / dh)).max() < 1e-10
    assert np.abs(soft.A - (-j / h)).max() < 1e-10


def test_far_field_limit():
    s = water_sphere()
    for u in directions(3):
        r = 1e7 * A0
        got = s.scattered(r * u) * r * np.exp(-1j * s.k0 * r)
        want = s.p0 * s.far_amplitude(D @ u)[0]
        assert abs(got - want) < 1e-5 * abs(want)


def test_force_is_the_stress_integral():
    s = water_sphere()
    w = 2 * np.pi * 1e6
    R, h = 1.5 * A0, 1e-6 * A0
    u, wu = np.polynomial.legendre.leggauss(24)
    ph = np.linspace(0.0, 2 * np.pi, 48, endpoint=False)
    p = lambda x: s.incident(x) + s.scattered(x)
    F = np.zeros(3)
    for ct, wt in zip(u, wu):
        st = np.sqrt(1 - ct * ct)
        for f in ph:
            n = np.array([st * np.cos(f), st * np.sin(f), ct])
            x = R * n
            px = p(x)
            # v = grad p / (i w rho) for exp(-i w t)
            g = np.array([(p(x + h * e) - p(x - h * e)) / (2 * h) for e in np.eye(3)])
            v = g / (1j * w * s.rho0)
            p2 = 0.25 * abs(px) ** 2 / (s.rho0 * s.c0 ** 2) - 0.25 * s.rho0 * np.vdot(v, v).real
            vv = 0.5 * s.rho0 * np.real(np.outer(v, np.conj(v)))
            F -= (p2 * n + vv @ n) * wt * (2 * np.pi / len(ph)) * R * R
    Fs = s.force()
    print("stress integral", F, "far field", Fs)
    assert np.linalg.norm(F - Fs) < 1e-6 * np.linalg.norm(Fs)

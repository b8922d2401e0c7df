"""Plane-wave scattering by a fluid sphere in an unbounded fluid: the
partial-wave series, a closed form for the free-space (periodicity setting 0)
boundary-value problem that the library solves numerically.

Time factor exp(-i w t), outgoing spherical Hankel functions h_n = j_n + i y_n.
Incident pressure p0 exp(i k0 d.x); outside the sphere of radius a the
scattered pressure is p0 sum_n i^n (2n+1) A_n h_n(k0 r) P_n(d.x/r), inside the
total pressure is p0 sum_n i^n (2n+1) B_n j_n(k1 r) P_n(d.x/r).  The interface
conditions are continuity of pressure and of normal velocity, (1/rho) dp/dr.

The radiation force is the far-field momentum balance of the time-averaged
Brillouin stress (Westervelt, J. Acoust. Soc. Am. 23, 312 (1951)):
F = (I0/c0) [sigma_ext d - int |f|^2 n dOmega], with I0 = |p0|^2 / (2 rho0 c0)
and f the far-field amplitude, p_s ~ p0 f(n) exp(i k0 r)/r."""
import numpy as np
from scipy.special import eval_legendre, spherical_jn, spherical_yn


class FluidSphere:
    def __init__(self, a, rho0, c0, rho1, c1, f, p0=1.0, d=(0.0, 0.0, 1.0), nmax=None):
        w = 2.0 * np.pi * f
        self.a, self.rho0, self.c0 = a, rho0, c0
        self.k0, self.k1 = w / c0, w / c1
        self.p0 = complex(p0)
        self.d = np.asarray(d, float) / np.linalg.norm(d)
        if nmax is None:
            nmax = int(max(self.k0, self.k1) * a + 4.0 * (max(self.k0, self.k1) * a) ** (1.0 / 3.0) + 12)
        n = self.n = np.arange(nmax + 1)
        x0, x1 = self.k0 * a, self.k1 * a
        j0, dj0 = spherical_jn(n, x0), spherical_jn(n, x0, derivative=True)
        h0 = j0 + 1j * spherical_yn(n, x0)
        dh0 = dj0 + 1j * spherical_yn(n, x0, derivative=True)
        j1, dj1 = spherical_jn(n, x1), spherical_jn(n, x1, derivative=True)
        q0, q1 = self.k0 / rho0, self.k1 / rho1
        # A h0 - B j1 = -j0 ;  q0 A dh0 - q1 B dj1 = -q0 dj0
        det = q0 * dh0 * j1 - q1 * dj1 * h0
        self.A = (q1 * dj1 * j0 - q0 * dj0 * j1) / det
        self.B = q0 * (dh0 * j0 - h0 * dj0) / det
        self.cn = (1j ** n) * (2 * n + 1)

    def _radial(self, x):
        x = np.asarray(x, float)
        r = np.linalg.norm(x)
        return r, (self.d @ x) / r if r > 0 else 1.0

    def scattered(self, x):
        """scattered pressure outside the sphere"""
        r, c = self._radial(x)
        h = spherical_jn(self.n, self.k0 * r) + 1j * spherical_yn(self.n, self.k0 * r)
        return self.p0 * np.sum(self.cn * self.A * h * eval_legendre(self.n, c))

    def interior(self, x):
        """total pressure inside the sphere"""
        r, c = self._radial(x)
        return self.p0 * np.sum(self.cn * self.B * spherical_jn(self.n, self.k1 * r) * eval_legendre(self.n, c))

    def incident(self, x):
        return self.p0 * np.exp(1j * self.k0 * (self.d @ np.asarray(x, float)))

    def incident_series(self, x):
        """the incident wave from its partial waves (Rayleigh's expansion)"""
        r, c = self._radial(x)
        return self.p0 * np.sum(self.cn * spherical_jn(self.n, self.k0 * r) * eval_legendre(self.n, c))

    def far_amplitude(self, cos_t):
        """f with p_s ~ p0 f exp(i k0 r)/r, from h_n(x) ~ (-i)^(n+1) exp(i x)/x"""
        P = eval_legendre(self.n[:, None], np.atleast_1d(cos_t)[None, :])
        return (-1j / self.k0) * np.sum(((2 * self.n + 1) * self.A)[:, None] * P, axis=0)

    def cross_sections(self, order=200):
        """(sigma_ext by the optical theorem, sigma_sca, int |f|^2 cos dOmega)"""
        u, wq = np.polynomial.legendre.leggauss(order)
        f2 = np.abs(self.far_amplitude(u)) ** 2
        s_ext = 4.0 * np.pi / self.k0 * self.far_amplitude(1.0)[0].imag
        return s_ext, 2.0 * np.pi * np.sum(wq * f2), 2.0 * np.pi * np.sum(wq * f2 * u)

    def force(self):
        """time-averaged radiation force vector (along d by symmetry)"""
        s_ext, _, s_cos = self.cross_sections()
        I0 = abs(self.p0) ** 2 / (2.0 * self.rho0 * self.c0)
        return I0 / self.c0 * (s_ext - s_cos) * self.d

"""Lattice / frequency cases used across the Green's function tests."""
import numpy as np

from bem3_aw_qd2_amd import QPD2


def make_qpd2(k, theta, phi, d1, d2):
    q = QPD2()
    q.k = k
    q.vk[0] = k * np.sin(theta) * np.cos(phi)
    q.vk[1] = k * np.sin(theta) * np.sin(phi)
    q.vk[2] = k * np.cos(theta)
    q.vd1[0], q.vd1[1] = d1
    q.vd2[0], q.vd2[1] = d2
    return q


K1MHZ = 2 * np.pi * 1e6 / 1500.0
CASES = {
    # square 2 mm lattice, 1 MHz water, near-normal incidence (0.5 mm sphere set-up)
    "square2mm": (K1MHZ, 1e-6, np.pi / 4, (2e-3, 0.0), (0.0, 2e-3), 1e-3),
    # triangular 1.8 mm lattice, theta 30 deg, phi 15 deg
    "tri1p8mm": (K1MHZ, np.pi / 6, np.pi / 12, (1.8e-3, 0.0), (0.9e-3, 1.5588457268119896e-3), 1e-3),
    # the reference's shipped sample: square 15 mm, 100 kHz, theta 0.3, psi 1e-8
    "sample15mm": (2 * np.pi * 1e5 / 1500.0, 0.3, 1e-8, (15e-3, 0.0), (0.0, 15e-3), 1e-2),
    # the reference's verification sample: square 50 mm, 100 kHz (kd = 21, many propagating orders)
    "veri50mm": (2 * np.pi * 1e5 / 1500.0, 0.3, 1e-8, (50e-3, 0.0), (0.0, 50e-3), 1e-2),
}


def points(name, n, seed=1):
    """Offsets source-minus-target over the cell and neighbours, with some z == 0
    rows and some very short offsets."""
    k, th, ph, d1, d2, rm = CASES[name]
    rng = np.random.default_rng(seed)
    pts = rng.uniform(-rm, rm, (n, 3))
    pts[: n // 20, 2] = 0.0
    pts[n // 20: n // 10] *= 1e-3
    pts[n // 10: n // 10 + n // 20, :2] += 3 * np.array(d1)   # a few cells away
    return pts


def qpd2(name):
    k, th, ph, d1, d2, rm = CASES[name]
    return make_qpd2(k, th, ph, d1, d2)

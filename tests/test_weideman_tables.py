"""The two Weideman tables of include/weideman_coef.h that the image sum of the
GPU lattice sums uses (gf_dev.h weideman_upper / weideman24_upper), evaluated in
double precision as the kernels evaluate them (w(z) = 2 p(Z)/(L - iz)^2 +
(1/sqrt(pi))/(L - iz), Z = (L + iz)/(L - iz)):

* N = 24 on the region where the kernels use it (0 <= Re z <= 3,
  1.5 <= Im z <= 12): relative error below 1e-14 against scipy's wofz (whose own
  error sets the floor); the N = 20 table that covered Im z >= 2.5 before misses
  this bar at Im z < 2.5, so the check separates the orders;
* N = 40 on 0 <= Re z <= 8, 0 <= Im z <= 12: below 1e-13 (on the real axis
  Re w = exp(-x^2) becomes tiny against Im w, so the bar is on |w|)."""
import os
import re

import numpy as np
import pytest
from scipy.special import wofz

HDR = os.path.join(os.path.dirname(os.path.abspath(__file__)), "..", "include", "weideman_coef.h")


def table(prefix):
    s = open(HDR).read()
    n = int(re.search(r"#define %s_N (\d+)" % prefix, s).group(1))
    L = float(re.search(r"#define %s_L ([0-9.eE+-]+)" % prefix, s).group(1))
    body = s[s.index("#define %s_COEFS" % prefix):].split("\n")[1:]
    a = []
    for line in body:
        line = line.strip().rstrip("\\").strip().rstrip(",")
        if not line:
            break
        a.append(float(line))
    assert len(a) == n
    return L, np.array(a)


def weideman(prefix, z):
    L, a = table(prefix)
    iden = 1.0 / (L - 1j * z)
    Z = (L + 1j * z) * iden
    p = np.zeros_like(Z)
    for c in a[::-1]:
        p = p * Z + c
    return 2.0 * p * iden * iden + (1.0 / np.sqrt(np.pi)) * iden


def grid(x1, y0, y1, nx=121, ny=421):
    x, y = np.meshgrid(np.linspace(0.0, x1, nx), np.linspace(y0, y1, ny))
    return (x + 1j * y).ravel()


def test_n24_on_its_region():
    z = grid(3.0, 1.5, 12.0)
    ref = wofz(z)
    err = np.abs(weideman("WEID24", z) - ref) / np.abs(ref)
    print("N = 24: max relative error %.2e" % err.max())
    assert err.max() < 1e-14


@pytest.mark.parametrize("y0,y1", [(0.0, 1.5), (1.5, 12.0)])
def test_n40(y0, y1):
    z = grid(8.0, y0, y1)
    ref = wofz(z)
    err = np.abs(weideman("WEID", z) - ref) / np.abs(ref)
    print("N = 40, Im z in [%g, %g]: max relative error %.2e" % (y0, y1, err.max()))
    assert err.max() < 1e-13

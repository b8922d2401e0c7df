"""The erfcx coefficient table (include/erfcx_coef.h) that the evanescent spectral
orders of the Ewald sum use on the GPU (gf_dev.h erfcx_w): evaluated here in
double precision exactly as the kernel evaluates it (Horner in Z, then
2 p/(L + t)^2 + (1/sqrt(pi))/(L + t)), its relative error must stay below
1e-15 against mpmath's 30-digit erfc(t) exp(t^2) on [0, 1e6], and below
1.5e-15 against scipy's erfcx on a dense grid.  The negative-argument identity
erfcx(-t) = 2 exp(t^2) - erfcx(t) is checked against scipy too."""
import os
import re

import mpmath as mp
import numpy as np
from scipy.special import erfcx

HDR = os.path.join(os.path.dirname(os.path.abspath(__file__)), "..", "include", "erfcx_coef.h")


def table():
    s = open(HDR).read()
    n = int(re.search(r"#define ERFCX_N (\d+)", s).group(1))
    L = float(re.search(r"#define ERFCX_L ([0-9.eE+-]+)", s).group(1))
    body = s[s.index("#define ERFCX_COEFS"):].split("\n")[1:]
    c = []
    for line in body:
        line = line.strip().rstrip("\\").strip().rstrip(",")
        if not line:
            break
        c.append(float(line))
    assert len(c) == n
    return L, np.array(c)


def erfcx_table(t):
    L, c = table()
    a = np.abs(t)
    iden = 1.0 / (L + a)
    Z = (L - a) * iden
    p = np.full_like(Z, c[-1])
    for v in c[-2::-1]:
        p = p * Z + v
    v = 2.0 * p * iden * iden + (1.0 / np.sqrt(np.pi)) * iden
    return np.where(t >= 0.0, v, 2.0 * np.exp(t * t) - v)


def test_against_mpmath():
    mp.mp.dps = 30
    t = np.concatenate([np.linspace(0.0, 12.0, 241), np.geomspace(12.0, 1e6, 60)])
    ref = np.array([float(mp.erfc(mp.mpf(x)) * mp.exp(mp.mpf(x) ** 2)) for x in t])
    err = np.abs(erfcx_table(t) - ref) / ref
    print("max relative error vs mpmath %.2e" % err.max())
    assert err.max() < 1e-15


def test_against_scipy_dense():
    t = np.concatenate([np.linspace(0.0, 20.0, 200001), np.geomspace(20.0, 1e8, 2000)])
    err = np.abs(erfcx_table(t) - erfcx(t)) / erfcx(t)
    print("max relative error vs scipy %.2e" % err.max())
    assert err.max() < 1.5e-15


def test_negative_arguments():
    t = -np.linspace(0.0, 5.0, 501)
    err = np.abs(erfcx_table(t) - erfcx(t)) / erfcx(t)
    assert err.max() < 1e-14

"""Generate the coefficient table of the scaled complementary error function
erfcx(t) = exp(t^2) erfc(t), t >= 0, in Weideman's variable:

    erfcx(t) ~ 2 p(Z) / (L + t)^2 + (1/sqrt(pi)) / (L + t),
    Z = (L - t) / (L + t),  p(Z) = sum_{n=0}^{D} c_n Z^n,

the form of Weideman's approximation of w(z) on the imaginary axis
(w(i t) = erfcx(t); J.A.C. Weideman, SIAM J. Numer. Anal. 31 (1994) 1497-1518),
where Z is real and covers (-1, 1] once as t runs over [0, inf).  Weideman's
c_n are Fourier coefficients chosen for the whole upper half-plane; on the real
segment Z in (-1, 1] a polynomial fitted to the relative error of erfcx does
the same job with half the terms.  The fit is a weighted least-squares problem
on 300 Chebyshev nodes in Z, solved in 40-digit arithmetic (mpmath) and driven
toward the minimax solution by Lawson's iteration (C.L. Lawson, 1961: each
node's weight is multiplied by its current error); the coefficients are
rounded to double once.

With D = 20 and L = 3 the relative error of the double-precision evaluation
(Horner in Z) is below 7e-16 on [0, 1e6], checked against mpmath by
tests/test_erfcx_table.py; the N = 40 Weideman table on the same axis reaches
8e-16 with twice the multiply-adds.

Usage: python gen_erfcx.py > ../../include/erfcx_coef.h
"""
import mpmath as mp

DEG = 20
L_PARAM = 3


def fit(L, deg, nodes=300, iters=4):
    mp.mp.dps = 40
    L = mp.mpf(L)
    zs = [mp.cos(mp.pi * (j + mp.mpf(1) / 2) / nodes) for j in range(nodes)]
    ts = [L * (1 - z) / (1 + z) for z in zs]
    ex = [mp.erfc(t) * mp.exp(t * t) for t in ts]
    # target p(Z) and the factor that turns an error in p into a relative error of erfcx
    f = [(e - 1 / mp.sqrt(mp.pi) / (L + t)) * (L + t) ** 2 / 2 for e, t in zip(ex, ts)]
    w = [2 / (L + t) ** 2 / e for e, t in zip(ex, ts)]
    lw = [mp.mpf(1)] * nodes
    for _ in range(iters):
        A = mp.matrix(nodes, deg + 1)
        b = mp.matrix(nodes, 1)
        for i in range(nodes):
            s = mp.sqrt(lw[i]) * w[i]
            for k in range(deg + 1):
                A[i, k] = s * zs[i] ** k
            b[i] = s * f[i]
        c, _ = mp.qr_solve(A, b)
        r = [abs(w[i] * (sum(c[k] * zs[i] ** k for k in range(deg + 1)) - f[i])) for i in range(nodes)]
        m = max(r)
        lw = [lw[i] * r[i] / m for i in range(nodes)]
        s = sum(lw)
        lw = [x / s * nodes for x in lw]
    return [c[k] for k in range(deg + 1)], m


def main():
    c, m = fit(L_PARAM, DEG)
    print("/* Generated by gen_erfcx.py: erfcx(t) = exp(t^2) erfc(t) for t >= 0 as")
    print(" * 2 p(Z)/(L + t)^2 + (1/sqrt(pi))/(L + t), Z = (L - t)/(L + t), p of degree %d" % DEG)
    print(" * fitted to the relative error (weighted error %.1e).  c[n] multiplies Z^n. */" % float(m))
    print("#ifndef BQD2_ERFCX_COEF_H_")
    print("#define BQD2_ERFCX_COEF_H_")
    print("#define ERFCX_N %d" % (DEG + 1))
    print("#define ERFCX_L %d.0" % L_PARAM)
    print("#define ERFCX_COEFS \\")
    for i, v in enumerate(c):
        sep = "," if i + 1 < len(c) else ""
        print("  %s%s \\" % (mp.nstr(v, 20, min_fixed=-1, max_fixed=-1), sep))
    print("")
    print("#endif")


if __name__ == "__main__":
    main()

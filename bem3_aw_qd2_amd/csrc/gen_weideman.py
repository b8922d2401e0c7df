"""Generate the coefficient table of Weideman's rational approximation of the
Faddeeva function w(z) (J.A.C. Weideman, "Computation of the complex error
function", SIAM J. Numer. Anal. 31 (1994) 1497-1518):

    w(z) ~ 2 p(Z) / (L - i z)^2 + (1/sqrt(pi)) / (L - i z),
    Z = (L + i z) / (L - i z),  p(Z) = sum_{n=0}^{N-1} a_n Z^n,  L = sqrt(N / sqrt(2)),

for Im z >= 0.  The a_n are the Fourier coefficients of
f(t) = exp(-t^2) (L^2 + t^2) sampled at t = L tan(theta/2), computed here in
40-digit arithmetic (mpmath) and rounded to double once.  With N = 40 the
relative error against an mpmath reference is below 1e-15 on the closed upper
half-plane (checked by tests/test_oracle_faddeeva.py).  A second table with
N = 24 (WEID24_*) serves 1.5 <= Im z <= 12, Re z <= 3, where its relative error
is below 4e-15, the level of the N = 40 table there.

Usage: python gen_weideman.py > weideman_coef.h
"""
import mpmath as mp

N = 40


def coefficients(n_terms):
    mp.mp.dps = 40
    m = 2 * n_terms
    m2 = 2 * m
    L = mp.sqrt(mp.mpf(n_terms) / mp.sqrt(2))
    samples = [mp.mpf(0)]
    for k in range(-m + 1, m):
        t = L * mp.tan(k * mp.pi / (2 * m))
        samples.append(mp.exp(-t * t) * (L * L + t * t))
    half = len(samples) // 2
    shifted = samples[half:] + samples[:half]
    out = []
    for j in range(1, n_terms + 1):
        acc = mp.mpf(0)
        for idx, v in enumerate(shifted):
            acc += v * mp.cos(2 * mp.pi * j * idx / m2)
        out.append(acc / m2)
    return L, out


def table(prefix, n):
    L, a = coefficients(n)
    print("#define %s_N %d" % (prefix, n))
    print("#define %s_L %s" % (prefix, mp.nstr(L, 20)))
    print("#define %s_COEFS \\" % prefix)
    for i, v in enumerate(a):
        sep = "," if i + 1 < len(a) else ""
        print("  %s%s \\" % (mp.nstr(v, 20, min_fixed=-1, max_fixed=-1), sep))
    print("")


def main():
    print("/* Generated by gen_weideman.py: Weideman (1994) rational approximation of")
    print(" * the Faddeeva function, N = %d (WEID_*) and N = 24 (WEID24_*).  a[n]" % N)
    print(" * multiplies Z^n. */")
    print("#ifndef BQD2_WEIDEMAN_COEF_H_")
    print("#define BQD2_WEIDEMAN_COEF_H_")
    table("WEID", N)
    table("WEID24", 24)
    print("#endif")


if __name__ == "__main__":
    main()

"""Generate the constants of the correctly rounded exponential exp_cr()
(bem3_aw_qd2_amd/csrc/hip/exp_cr.h).

Method: Tang's table-driven reduction (P.T.P. Tang, ACM TOMS 15 (1989) 144-157)
carried out in double-double arithmetic (T.J. Dekker, Numer. Math. 18 (1971)
224-242):

    x = k ln2/64 + r,  k = 64 m + j,  |r| <= ln2/128,
    exp(x) = 2^m * 2^(j/64) * exp(r),

with exp(r) from its Taylor series to degree 11 (truncation below 2^-118 for
|r| <= ln2/128) and 2^(j/64) as a double-double pair.  The double-double result
approximates exp(x) to about 2^-100 relative, so rounding it to double once
gives the correctly rounded value except where exp(x) lies within 2^-100 of a
rounding midpoint (the hardest binary64 cases of exp need about 2^-113:
V. Lefevre, J.-M. Muller, "Worst cases for correct rounding of the elementary
functions in double precision", ARITH 15, 2001).

Constants written here:
  EXPCR_T[64][2]   2^(j/64) = hi + lo
  EXPCR_C[12][2]   1/n!     = hi + lo, n = 0..11
  EXPCR_L1..L3     ln2/64 = L1 + L2 + L3, L1 with 32 significant bits so that
                   k*L1 is exact for |k| < 2^21
  EXPCR_INV        64/ln2 rounded to double

Usage: python gen_exp_cr.py > ../../include/exp_cr_table.h
"""
import mpmath as mp


def split(v):
    hi = float(v)
    lo = float(v - mp.mpf(hi))
    return hi, lo


def main():
    mp.mp.prec = 300
    ln2_64 = mp.log(2) / 64
    # L1: ln2/64 rounded to 32 significant bits
    e = mp.floor(mp.log(ln2_64, 2))
    scale = mp.mpf(2) ** (31 - e)
    L1 = float(mp.floor(ln2_64 * scale) / scale)
    rest = ln2_64 - mp.mpf(L1)
    L2 = float(rest)
    L3 = float(rest - mp.mpf(L2))
    out = []
    out.append("// generated by bem3_aw_qd2_amd/csrc/gen_exp_cr.py -- do not edit")
    out.append("// constants of exp_cr(): 2^(j/64) and 1/n! as double-double pairs, ln2/64 in three parts")
    out.append("#pragma once")
    out.append("#define EXPCR_INV %r" % float(64 / mp.log(2)))
    out.append("#define EXPCR_L1 %r" % L1)
    out.append("#define EXPCR_L2 %r" % L2)
    out.append("#define EXPCR_L3 %r" % L3)
    out.append("#define EXPCR_T_INIT \\")
    rows = []
    for j in range(64):
        hi, lo = split(mp.mpf(2) ** (mp.mpf(j) / 64))
        rows.append("  {%r, %r}" % (hi, lo))
    out.append("  {" + ", \\\n".join(rows) + "}")
    out.append("#define EXPCR_C_INIT \\")
    rows = []
    for n in range(12):
        hi, lo = split(1 / mp.factorial(n))
        rows.append("  {%r, %r}" % (hi, lo))
    out.append("  {" + ", \\\n".join(rows) + "}")
    print("\n".join(out))


if __name__ == "__main__":
    main()

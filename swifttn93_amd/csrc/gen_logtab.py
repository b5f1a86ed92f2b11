#!/usr/bin/env python3
"""Generates tn93g_logtab.h: the table of the double-double natural logarithm used on the GPU.

Tang's table-driven method: for x = 2^e * m, m in [1, 2), the breakpoint F_j = 1 + j/256 (j = round((m-1)*256))
gives log x = e*ln2 + log F_j + log1p((m - F_j)/F_j).  The table holds log F_j and 1/F_j as double-double pairs
(hi = value rounded to nearest double, lo = the rounded remainder), computed here with mpmath at 200 bits.

    python3 swifttn93_amd/csrc/gen_logtab.py > swifttn93_amd/csrc/tn93g_logtab.h
"""
import mpmath

mpmath.mp.prec = 200
NT = 256


def dd(v):
    hi = float(v)
    lo = float(v - mpmath.mpf(hi))
    return hi, lo


def main():
    print("// Generated by gen_logtab.py (mpmath, 200-bit): do not edit.")
    print("// Breakpoints F_j = 1 + j/256, j = 0..256: log(F_j) and 1/F_j as double-double (hi, lo).")
    print("#pragma once")
    print(f"#define TN93G_LOGTAB_N {NT}")
    rows = []
    for j in range(NT + 1):
        F = 1 + mpmath.mpf(j) / NT
        lh, ll = dd(mpmath.log(F))
        ih, il = dd(1 / F)
        rows.append((lh, ll, ih, il))
    print("TN93G_TABLE_ATTR const double tn93g_logtab[%d][4] = {" % (NT + 1))
    for lh, ll, ih, il in rows:
        print("  {%s, %s, %s, %s}," % (lh.hex(), ll.hex(), ih.hex(), il.hex()))
    print("};")
    h, l = dd(mpmath.log(2))
    print(f"#define TN93G_LN2_HI {h.hex()}")
    print(f"#define TN93G_LN2_LO {l.hex()}")


if __name__ == "__main__":
    main()

"""Writes log_table.h: the table of log_tab (common.h).  The mantissa
m in [1, 2) is rounded to the nearest c_j = 1 + j/256, j = 0..256; r_j is the
double nearest to 1 / c_j and L_j = -ln(r_j) rounded to the nearest double
from 50-digit decimal arithmetic, so ln m = L_j + log1p(m r_j - 1) with
|m r_j - 1| <= 2^-9.  Rows from j = LOG_TAB_SHIFT (c_j >= sqrt 2) hold
L_j - ln 2 and the exponent is raised by one, so that the table term stays
small where ln x is small (x near 1 from below): rows 0 and 256 are then
exact (L = 0), and ln x near 1 is log1p(u) alone.
Usage:
    python gen_log_table.py > log_table.h"""
from decimal import Decimal, getcontext

getcontext().prec = 50
LN2 = Decimal(2).ln()
SHIFT = 107  # first j with 1 + j/256 >= sqrt(2)
assert (1 + Decimal(SHIFT) / 256) ** 2 >= 2 > (1 + Decimal(SHIFT - 1) / 256) ** 2
rows = []
for j in range(257):
    c = Decimal(1) + Decimal(j) / 256
    r = float(Decimal(1) / c)
    L = -Decimal(r).ln() - (LN2 if j >= SHIFT else 0)
    rows.append((r, float(L)))
print("// log_table.h -- generated by gen_log_table.py: the 257-row table of")
print("// log_tab (common.h), pairs {r_j, -ln r_j (- ln 2 from LOG_TAB_SHIFT)}.")
print("// Do not edit.")
print("#pragma once")
print("#define LOG_TAB_N 257")
print(f"#define LOG_TAB_SHIFT {SHIFT}")
print("#define LOG_TAB_ROWS \\")
print(",\\\n".join(f"  {{{r!r}, {L!r}}}" for r, L in rows))

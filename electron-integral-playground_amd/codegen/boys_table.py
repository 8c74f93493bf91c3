"""Boys-function grid table for the device kernels.

The device evaluates F_L(T) for T < BOYS_T_SWITCH by a Taylor expansion about the nearest grid point
T0 = k * h,

    F_L(T0 + d) = sum_{k=0}^{K} F_{L+k}(T0) (-d)^k / k!,     |d| <= h/2,

and then obtains F_{L-1}, ..., F_0 by the stable downward recursion
F_m = (2T F_{m+1} + e^{-T}) / (2m+1).  For T >= BOYS_T_SWITCH, e^{-T} < 2.4e-16 so
F_0 = sqrt(pi/T)/2 and the upward recursion F_{m+1} = ((2m+1) F_m - e^{-T}) / (2T) is stable.

With h = 1/16 and K = 8 the truncation error is below (1/32)^9/9! ~ 7e-20 relative, far under double
rounding.  Grid values come from mpmath at 40 digits: F_NMAX by its power series and the rest by
downward recursion, so every table entry is correctly rounded to float64.  This is a different
interpolation scheme from the reference's Chebyshev table (boys_function.hpp:102-176); the two agree
to ~1e-15 relative, which the tests check against an independent long-double evaluation.
"""

import mpmath

BOYS_H_INV = 16            # grid spacing h = 1/16
BOYS_T_SWITCH = 36.0       # grid covers [0, 36]
BOYS_NMAX = 32             # columns F_0 .. F_32 (L <= 24 with K = 8)
BOYS_K64 = 8               # Taylor terms for float64
BOYS_K32 = 5               # Taylor terms for float32


def grid_values():
    n_rows = int(BOYS_T_SWITCH * BOYS_H_INV) + 1
    rows = []
    with mpmath.workdps(40):
        for k in range(n_rows):
            T = mpmath.mpf(k) / BOYS_H_INV
            F = [mpmath.mpf(0)] * (BOYS_NMAX + 1)
            if k == 0:
                F = [mpmath.mpf(1) / (2 * n + 1) for n in range(BOYS_NMAX + 1)]
            else:
                eT = mpmath.exp(-T)
                term = mpmath.mpf(1) / (2 * BOYS_NMAX + 1)
                s = term
                i = 1
                while True:
                    term *= 2 * T / (2 * BOYS_NMAX + 2 * i + 1)
                    s += term
                    i += 1
                    if term < s * mpmath.mpf(10) ** -38:
                        break
                F[BOYS_NMAX] = eT * s
                for n in range(BOYS_NMAX - 1, -1, -1):
                    F[n] = (2 * T * F[n + 1] + eT) / (2 * n + 1)
            rows.append([float(v) for v in F])
    return rows


def emit(path):
    rows = grid_values()
    n_rows = len(rows)
    ncol = BOYS_NMAX + 1
    lines = [
        "// Generated by codegen/boys_table.py -- Boys function F_n(k/16), k = 0..%d, n = 0..%d." % (n_rows - 1, BOYS_NMAX),
        "#pragma once",
        "#define BOYS_H_INV %d" % BOYS_H_INV,
        "#define BOYS_T_SWITCH %r" % BOYS_T_SWITCH,
        "#define BOYS_NMAX %d" % BOYS_NMAX,
        "#define BOYS_NCOL %d" % ncol,
        "#define BOYS_NROW %d" % n_rows,
        "#define BOYS_K64 %d" % BOYS_K64,
        "#define BOYS_K32 %d" % BOYS_K32,
        "__device__ const double boys_grid64[BOYS_NROW * BOYS_NCOL] = {",
    ]
    for r in rows:
        lines.append("  " + ", ".join(repr(v) for v in r) + ",")
    lines.append("};")
    lines.append("__device__ const float boys_grid32[BOYS_NROW * BOYS_NCOL] = {")
    for r in rows:
        lines.append("  " + ", ".join(repr(float(f"{v:.9e}")) + "f" for v in r) + ",")
    lines.append("};")
    with open(path, "w") as f:
        f.write("\n".join(lines) + "\n")


if __name__ == "__main__":
    import sys
    emit(sys.argv[1])

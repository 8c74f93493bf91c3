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
# Coarse grid for kernels that stage their slice of the table in LDS: h = 1/2 on [0, 36] (73 rows) and
# BOYS_KC + 1 = 15 Taylor terms (|d| <= 1/4: truncation below (1/4)^15/15! ~ 7e-22 relative), stored per
# order L as the columns L .. L + BOYS_KC, column-major ([L][column][row]): a wave's lanes read the same
# column at scattered rows, which spreads them over the LDS banks.
BOYS_C_H_INV = 2
BOYS_KC = 14
BOYS_LDS_LMAX = 18         # orders L with a staged slice (L + BOYS_KC <= BOYS_NMAX)


def grid_values(h_inv=BOYS_H_INV):
    n_rows = int(BOYS_T_SWITCH * h_inv) + 1
    rows = []
    with mpmath.workdps(40):
        for k in range(n_rows):
            T = mpmath.mpf(k) / h_inv
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
    coarse = grid_values(BOYS_C_H_INV)
    nrc = len(coarse)
    lines += ["#define BOYS_C_H_INV %d" % BOYS_C_H_INV, "#define BOYS_C_NROW %d" % nrc,
              "#define BOYS_KC %d" % BOYS_KC, "#define BOYS_LDS_LMAX %d" % BOYS_LDS_LMAX,
              "// [L][c][k] = F_{L+c}(k/%d), L = 0..%d, c = 0..%d, k = 0..%d" % (BOYS_C_H_INV, BOYS_LDS_LMAX, BOYS_KC, nrc - 1),
              # slices padded to a multiple of 16 bytes (16-byte global->LDS copies)
              "#define BOYS_LDS_STRIDE %d" % (((BOYS_KC + 1) * nrc + 1) // 2 * 2),
              "__device__ __attribute__((aligned(16))) const double boys_lds64[(BOYS_LDS_LMAX + 1) * BOYS_LDS_STRIDE] = {"]
    for L in range(BOYS_LDS_LMAX + 1):
        for c in range(BOYS_KC + 1):
            lines.append("  " + ", ".join(repr(coarse[k][L + c]) for k in range(nrc)) + ",")
        if ((BOYS_KC + 1) * nrc) % 2:
            lines.append("  0.0,")
    lines.append("};")
    with open(path, "w") as f:
        f.write("\n".join(lines) + "\n")


if __name__ == "__main__":
    import sys
    emit(sys.argv[1])

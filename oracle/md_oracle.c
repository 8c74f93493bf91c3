/*
 * md_oracle.c -- CPU restatement of the reference's McMurchie-Davidson integral path.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the checker for the HIP library: only tests/,
 * __graft_entry__.smoke() and bench.py's CPU-baseline leg may load it.  It is never linked into,
 * called by, or used as a fallback for the product library.
 *
 * It exports the reference's four C-ABI functions under an "oracle_" prefix, with the reference's
 * argument lists, accumulate (+=) semantics, mirror rule and return codes:
 *   oracle_overlap             <- reference c_library/molecular_integral/overlap.cpp:147-199
 *   oracle_nuclear_attraction  <- reference c_library/molecular_integral/nuclear_attraction.cpp:197-253
 *   oracle_two_center          <- reference c_library/molecular_integral/two_center.cpp:237-319
 *   oracle_three_center        <- reference c_library/molecular_integral/three_center.cpp:319-631
 * plus oracle_boys (the Boys function) and oracle_cart_to_sph_row (one row of the Cartesian to
 * real-solid-harmonic matrix) so that the tests can pin those pieces on their own.
 *
 * The algorithm is the reference's: every primitive pair (or pair x auxiliary primitive, or pair x
 * charge) is evaluated on its own -- Hermite expansion coefficients E by the vertical recurrence and
 * the binomial horizontal transfer (mcmurchie_davidson_E_term.hpp:22-45, 109-130, 166-199), Hermite
 * integrals R_tuv (mcmurchie_davidson_R_term.hpp:15-28), the Cartesian contraction
 * sum_t E_t sum_s (-1)^|s| E_s R_{t+s} (three_center.cpp:93-160), the per-component Cartesian
 * normalisation (kernel_cartesian_normalization.hpp:22-77), the Cartesian-to-spherical transform
 * (cartesian_spherical_transformation.hpp) and accumulation with the mirror element written whenever
 * the two AO offsets differ (three_center.cpp:242-244).  Three places use an independent method on
 * purpose, so that a shared mistake with the reference's code cannot hide:
 *   - R_tuv comes from the textbook MD recursion R^n_{t+1} = t R^{n+1}_{t-1} + X R^{n+1}_t, not the
 *     reference's closed Hermite-polynomial form (mcmurchie_davidson_R_term.hpp:39-89);
 *   - the Boys function is a long-double power series / erf-plus-upward-recursion, not the
 *     reference's interpolation table (boys_function.hpp:102-176);
 *   - the spherical transform is built at start-up from the solid-harmonic generating formulas
 *     (the reference's gen/cartesian_spherical_transformation.py:8-67), not from stored tables.
 * The three-centre triplets are evaluated in parallel into a buffer, round by round, and each output
 * row then adds its contributions in a fixed (auxiliary primitive, pair) order on one thread, so no
 * atomics are needed and results do not depend on the thread count.
 */
#include <stdbool.h>
#include <tgmath.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

/* Internal arithmetic type: double for the checker, long double for the high-precision build
 * (libmd_oracle_ld.so) used to decide ill-conditioned cases.  Inputs and outputs stay double. */
#ifndef REAL_T
#define REAL_T double
#endif
typedef REAL_T real_t;
#define OR_PI ((real_t) 3.14159265358979323846264338327950288L)

#define OR_MAX_L 6          /* orbital shells (reference angular.h:4) */
#define OR_MAX_AUX_L 9      /* auxiliary shells (reference angular.h:5) */
#define OR_MAX_CART 55      /* (9+1)(9+2)/2 */
#define OR_MAX_LL 21        /* 6 + 6 + 9 */

/* ------------------------------------------------------------------------------------------ */
/* Small combinatorics                                                                          */
/* ------------------------------------------------------------------------------------------ */

static real_t fact(int n) { real_t r = 1.0; for (int i = 2; i <= n; i++) r *= i; return r; }
static real_t binom(int n, int k) { return (k < 0 || k > n) ? 0.0 : fact(n) / (fact(k) * fact(n - k)); }
static int ncart(int l) { return (l + 1) * (l + 2) / 2; }

/* Cartesian component c of shell l, x-major descending order (reference indexing.hpp:98-112). */
static void cart_xyz(int l, int c, int *x, int *y, int *z)
{
    int idx = 0;
    for (int a = l; a >= 0; a--)
        for (int b = l - a; b >= 0; b--) {
            if (idx == c) { *x = a; *y = b; *z = l - a - b; return; }
            idx++;
        }
    *x = *y = *z = -1;
}

/* sqrt(ix! iy! iz! (2L)! / ((2ix)! (2iy)! (2iz)! L!)) -- kernel_cartesian_normalization.hpp:22-26 */
static real_t cart_norm_eval(int l, int c)
{
    int x, y, z;
    cart_xyz(l, c, &x, &y, &z);
    return sqrt(fact(x) * fact(y) * fact(z) * fact(2 * l) / (fact(2 * x) * fact(2 * y) * fact(2 * z) * fact(l)));
}

/* Tables filled once by ensure_tables(): normalisation constant and exponents of every Cartesian
 * component (the inner loops read these instead of re-deriving them per element). */
static real_t cnorm_tab[OR_MAX_AUX_L + 1][OR_MAX_CART];
static int cxyz_tab[OR_MAX_AUX_L + 1][OR_MAX_CART][3];

static inline real_t cart_norm(int l, int c) { return cnorm_tab[l][c]; }

/* ------------------------------------------------------------------------------------------ */
/* Cartesian -> real solid harmonics, built from the generating formulas                        */
/* ------------------------------------------------------------------------------------------ */

/* raw_sph[l][m][c]: coefficient of monomial c in solid harmonic m (order S_l..S_1, C_0..C_l);
 * p is left as x, y, z (reference cartesian_spherical_transformation.hpp:38-39). */
static real_t raw_sph[OR_MAX_AUX_L + 1][2 * OR_MAX_AUX_L + 1][OR_MAX_CART];
static int tables_ready = 0;

typedef struct { real_t c[OR_MAX_AUX_L + 1][OR_MAX_AUX_L + 1]; } poly_t;  /* c[a][b], z power = l-a-b */

static void build_solid_harmonics(void)
{
    for (int l = 0; l <= OR_MAX_AUX_L; l++) {
        memset(raw_sph[l], 0, sizeof(raw_sph[l]));
        if (l <= 1) {
            for (int c = 0; c < ncart(l); c++) raw_sph[l][c][c] = 1.0;
            continue;
        }
        for (int row = 0; row < 2 * l + 1; row++) {
            const int is_sine = row < l;
            const int m = is_sine ? l - row : row - l;
            /* Pi_lm(z, r) A_m(x, y) or Pi_lm B_m, expanded into monomials x^a y^b z^(l-a-b). */
            poly_t P;
            memset(&P, 0, sizeof(P));
            for (int k = 0; 2 * k <= l - m; k++) {
                const real_t gamma = ((k % 2) ? -1.0 : 1.0) * pow(2.0, -l) * binom(l, k) * binom(2 * l - 2 * k, l)
                                     * fact(l - 2 * k) / fact(l - 2 * k - m);
                /* r^{2k} = sum_{i+j+h=k} k!/(i!j!h!) x^{2i} y^{2j} z^{2h} */
                for (int i = 0; i <= k; i++)
                    for (int j = 0; i + j <= k; j++) {
                        const real_t multinom = fact(k) / (fact(i) * fact(j) * fact(k - i - j));
                        /* times z^{l-2k-m} times A_m/B_m = sum_p C(m,p) x^p y^{m-p} trig */
                        for (int p = 0; p <= m; p++) {
                            const int q = m - p;            /* power of y from A/B */
                            real_t trig;
                            if (is_sine) trig = (q % 4 == 1) ? 1.0 : (q % 4 == 3) ? -1.0 : 0.0;   /* sin(pi q / 2) */
                            else         trig = (q % 4 == 0) ? 1.0 : (q % 4 == 2) ? -1.0 : 0.0;   /* cos(pi q / 2) */
                            if (trig == 0.0) continue;
                            P.c[2 * i + p][2 * j + q] += gamma * multinom * binom(m, p) * trig;
                        }
                    }
            }
            const real_t pref = sqrt((is_sine ? 2.0 : (m == 0 ? 1.0 : 2.0)) * fact(l - m) / fact(l + m));
            for (int c = 0; c < ncart(l); c++) {
                int x, y, z;
                cart_xyz(l, c, &x, &y, &z);
                raw_sph[l][row][c] = pref * P.c[x][y];
            }
        }
    }
    for (int l = 0; l <= OR_MAX_AUX_L; l++)
        for (int c = 0; c < ncart(l); c++) {
            cnorm_tab[l][c] = cart_norm_eval(l, c);
            cart_xyz(l, c, &cxyz_tab[l][c][0], &cxyz_tab[l][c][1], &cxyz_tab[l][c][2]);
        }
    tables_ready = 1;
}

static void ensure_tables(void)
{
    if (!tables_ready) {
#pragma omp critical(oracle_tables)
        if (!tables_ready) build_solid_harmonics();
    }
}

double oracle_cart_to_sph_row(int l, int row, int c)
{
    ensure_tables();
    if (l < 0 || l > OR_MAX_AUX_L || row < 0 || row > 2 * l || c < 0 || c >= ncart(l)) return NAN;
    return raw_sph[l][row][c];
}

/* ------------------------------------------------------------------------------------------ */
/* Boys function F_m(x), m = 0..mmax, in long double                                           */
/* ------------------------------------------------------------------------------------------ */

static void boys_internal(int mmax, long double x, long double *F)
{
    if (x < 50.0L) {
        /* F_mmax(x) = e^{-x} sum_k (2x)^k / ((2m+1)(2m+3)...(2m+2k+1)), then downward
         * F_m = (2x F_{m+1} + e^{-x}) / (2m+1).  All terms positive: no cancellation. */
        const long double ex = expl(-x);
        long double term = 1.0L / (2 * mmax + 1), sum = term;
        for (int k = 1; k < 2000; k++) {
            term *= 2.0L * x / (2 * mmax + 2 * k + 1);
            sum += term;
            if (term < sum * 1e-21L) break;
        }
        F[mmax] = ex * sum;
        for (int m = mmax - 1; m >= 0; m--) F[m] = (2.0L * x * F[m + 1] + ex) / (2 * m + 1);
    } else {
        /* F_0 = sqrt(pi/x) erf(sqrt x) / 2, upward F_{m+1} = ((2m+1) F_m - e^{-x}) / (2x),
         * stable for x >> m. */
        const long double ex = expl(-x);
        F[0] = 0.5L * sqrtl(3.14159265358979323846264338327950288L / x) * erfl(sqrtl(x));
        for (int m = 0; m < mmax; m++) F[m + 1] = ((2 * m + 1) * F[m] - ex) / (2.0L * x);
    }
}

void oracle_boys(int mmax, double xd, double *out)
{
    long double F[OR_MAX_LL + 8];
    boys_internal(mmax, xd, F);
    for (int m = 0; m <= mmax; m++) out[m] = (double) F[m];
}

/* ------------------------------------------------------------------------------------------ */
/* Hermite expansion coefficients and Hermite integrals                                         */
/* ------------------------------------------------------------------------------------------ */

#define ESTR (2 * OR_MAX_L + 2)   /* stride for t in E tables */

/* E[i][j][t] for i <= li, j <= lj on one axis: VRR E^{i+1,0}_t = 1/(2p) E^{i,0}_{t-1} + PA E^{i,0}_t
 * + (t+1) E^{i,0}_{t+1}, then E^{i,j}_t = sum_m C(j, j-m) AB^{j-m} E^{i+m,0}_t
 * (reference mcmurchie_davidson_E_term.hpp:22-45, 166-199). */
static void hermite_E_pair(int li, int lj, real_t PA, real_t AB, real_t one_over_two_p,
                           real_t E[OR_MAX_L + 1][OR_MAX_L + 1][ESTR])
{
    real_t E0[2 * OR_MAX_L + 2][ESTR];
    const int L = li + lj;
    memset(E0, 0, sizeof(E0));
    E0[0][0] = 1.0;
    for (int i = 0; i < L; i++)
        for (int t = 0; t <= i + 1; t++) {
            real_t v = 0.0;
            if (t >= 1) v += one_over_two_p * E0[i][t - 1];
            if (t <= i) v += PA * E0[i][t];
            if (t + 1 <= i) v += (t + 1) * E0[i][t + 1];
            E0[i + 1][t] = v;
        }
    for (int i = 0; i <= li; i++)
        for (int j = 0; j <= lj; j++)
            for (int t = 0; t < ESTR; t++) {
                real_t v = 0.0;
                if (t <= i + j) {
                    real_t ABpow = 1.0;
                    for (int m = j; m >= 0; m--) {
                        if (t <= i + m) v += binom(j, j - m) * ABpow * E0[i + m][t];
                        ABpow *= AB;
                    }
                }
                E[i][j][t] = v;
            }
}

/* Single-centre coefficients E^{k,0}_s with b = 0 (reference mcmurchie_davidson_E_term.hpp:109-130). */
static void hermite_E_single(int lk, real_t one_over_two_q, real_t E[OR_MAX_AUX_L + 1][OR_MAX_AUX_L + 1])
{
    memset(E, 0, sizeof(real_t) * (OR_MAX_AUX_L + 1) * (OR_MAX_AUX_L + 1));
    E[0][0] = 1.0;
    for (int i = 0; i < lk; i++)
        for (int s = 0; s <= i + 1; s++) {
            real_t v = 0.0;
            if (s >= 1) v += one_over_two_q * E[i][s - 1];
            if (s + 1 <= i) v += (s + 1) * E[i][s + 1];
            E[i + 1][s] = v;
        }
}

typedef struct {
    int L;
    real_t *R;      /* R[(t*(L+1) + u)*(L+1) + v], level 0 */
    real_t *work;   /* scratch level */
} herm_R;

static void herm_R_alloc(herm_R *h, int L)
{
    const size_t n = (size_t) (L + 1) * (L + 1) * (L + 1);
    h->L = L;
    h->R = (real_t *) calloc(n, sizeof(real_t));
    h->work = (real_t *) calloc(n, sizeof(real_t));
}

static void herm_R_free(herm_R *h) { free(h->R); free(h->work); }

/* R^0_{tuv}(zeta, PQ), t+u+v <= L, from R^n_000 = (-2 zeta)^n F_n(zeta |PQ|^2) by the MD recursion. */
static void herm_R_compute(herm_R *h, real_t zeta, real_t X, real_t Y, real_t Z)
{
    const int L = h->L;
    const int S = L + 1;
    long double F[OR_MAX_LL + 8];
    boys_internal(L, (long double) zeta * ((long double) X * X + (long double) Y * Y + (long double) Z * Z), F);
    /* ping-pong between the two buffers; L swaps follow, so start in h->R when L is even and the
     * final level (n = 0) lands in h->R either way */
    real_t *cur = h->R, *prev = h->work;
    if (L % 2 == 1) { cur = h->work; prev = h->R; }
    /* level n = L: only R^L_000 */
    real_t m2z = 1.0;
    for (int n = 0; n < L; n++) m2z *= -2.0 * zeta;
    cur[0] = m2z * F[L];
    for (int n = L - 1; n >= 0; n--) {
        real_t *tmp = prev; prev = cur; cur = tmp;          /* prev holds level n+1 (|tuv| <= L-n-1) */
        m2z = 1.0;
        for (int k = 0; k < n; k++) m2z *= -2.0 * zeta;
        const int top = L - n;
        for (int t = 0; t <= top; t++)
            for (int u = 0; t + u <= top; u++)
                for (int v = 0; t + u + v <= top; v++) {
                    real_t val;
                    if (t == 0 && u == 0 && v == 0) val = m2z * F[n];
                    else if (t > 0) {
                        val = X * prev[((t - 1) * S + u) * S + v];
                        if (t > 1) val += (t - 1) * prev[((t - 2) * S + u) * S + v];
                    } else if (u > 0) {
                        val = Y * prev[(t * S + u - 1) * S + v];
                        if (u > 1) val += (u - 1) * prev[(t * S + u - 2) * S + v];
                    } else {
                        val = Z * prev[(t * S + u) * S + v - 1];
                        if (v > 1) val += (v - 1) * prev[(t * S + u) * S + v - 2];
                    }
                    cur[(t * S + u) * S + v] = val;
                }
    }
}

static inline real_t herm_R_get(const herm_R *h, int t, int u, int v)
{
    const int S = h->L + 1;
    return h->R[(t * S + u) * S + v];
}

/* ------------------------------------------------------------------------------------------ */
/* Output transforms                                                                            */
/* ------------------------------------------------------------------------------------------ */

/* Spherical row m of shell l applied to a Cartesian vector v with stride (after the Cartesian
 * normalisation has been applied to v, the anti-normalised matrix raw/N gives raw coefficients). */
static real_t sph_row_dot(int l, int m, const real_t *v, int stride)
{
    real_t s = 0.0;
    for (int c = 0; c < ncart(l); c++) {
        const real_t coef = raw_sph[l][m][c];
        if (coef != 0.0) s += coef / cart_norm(l, c) * v[c * stride];
    }
    return s;
}

/* block[(a*ni + i)*nj + j] (Cartesian, normalised) -> out[(a*no_i + i')*no_j + j'] spherical or copy.
 * The leading dimension 'a' (nlead entries) is untouched. */
static void transform_ij(int li, int lj, int nlead, const real_t *cart, real_t *out, bool spherical)
{
    const int ni = ncart(li), nj = ncart(lj);
    if (!spherical) { memcpy(out, cart, sizeof(real_t) * nlead * ni * nj); return; }
    const int si = 2 * li + 1, sj = 2 * lj + 1;
    real_t tmp[OR_MAX_CART * (2 * OR_MAX_AUX_L + 1)];
    for (int a = 0; a < nlead; a++) {
        const real_t *blk = cart + (size_t) a * ni * nj;
        for (int i = 0; i < ni; i++)
            for (int j = 0; j < sj; j++) tmp[i * sj + j] = sph_row_dot(lj, j, blk + i * nj, 1);
        for (int i = 0; i < si; i++)
            for (int j = 0; j < sj; j++) out[((size_t) a * si + i) * sj + j] = sph_row_dot(li, i, tmp + j, sj);
    }
}

/* ------------------------------------------------------------------------------------------ */
/* Overlap                                                                                      */
/* ------------------------------------------------------------------------------------------ */

static int class_ok_orbital(int i_L, int j_L) { return i_L >= 0 && j_L >= 0 && i_L <= j_L && j_L <= OR_MAX_L; }

int oracle_overlap(int i_L, int j_L, const double *pair_A_a, const double *pair_B_b, const double *pair_coefficient,
                   const int *pair_i_ao_start, const int *pair_j_ao_start, int n_pair,
                   double *S_matrix, int n_ao, bool spherical)
{
    ensure_tables();
    if (!class_ok_orbital(i_L, j_L)) {
        printf("%s function does not support angular i_L = %d, j_L = %d\n", __func__, i_L, j_L);
        fflush(stdout);
        return 1;
    }
    const int ni = ncart(i_L), nj = ncart(j_L);
    const int oi = spherical ? 2 * i_L + 1 : ni, oj = spherical ? 2 * j_L + 1 : nj;
    real_t cart[OR_MAX_CART * OR_MAX_CART], out[OR_MAX_CART * OR_MAX_CART];
    real_t Ex[OR_MAX_L + 1][OR_MAX_L + 1][ESTR], Ey[OR_MAX_L + 1][OR_MAX_L + 1][ESTR], Ez[OR_MAX_L + 1][OR_MAX_L + 1][ESTR];
    for (int ip = 0; ip < n_pair; ip++) {
        const double *A = pair_A_a + 4 * ip, *B = pair_B_b + 4 * ip;
        /* P recomputed from A and B, P_p not read (reference overlap.cpp:22-29) */
        const real_t p = A[3] + B[3];
        const real_t f = -B[3] / p;
        const real_t AB[3] = {A[0] - B[0], A[1] - B[1], A[2] - B[2]};
        hermite_E_pair(i_L, j_L, f * AB[0], AB[0], 0.5 / p, Ex);
        hermite_E_pair(i_L, j_L, f * AB[1], AB[1], 0.5 / p, Ey);
        hermite_E_pair(i_L, j_L, f * AB[2], AB[2], 0.5 / p, Ez);
        const real_t pref = pair_coefficient[ip] * pow(OR_PI / p, 1.5);
        for (int ci = 0; ci < ni; ci++) {
            int ix, iy, iz; cart_xyz(i_L, ci, &ix, &iy, &iz);
            for (int cj = 0; cj < nj; cj++) {
                int jx, jy, jz; cart_xyz(j_L, cj, &jx, &jy, &jz);
                cart[ci * nj + cj] = pref * Ex[ix][jx][0] * Ey[iy][jy][0] * Ez[iz][jz][0]
                                     * cart_norm(i_L, ci) * cart_norm(j_L, cj);
            }
        }
        transform_ij(i_L, j_L, 1, cart, out, spherical);
        const int is = pair_i_ao_start[ip], js = pair_j_ao_start[ip];
        for (int i = 0; i < oi; i++)
            for (int j = 0; j < oj; j++) {
                S_matrix[(size_t) (is + i) * n_ao + (js + j)] += out[i * oj + j];
                if (is != js) S_matrix[(size_t) (is + i) + (size_t) (js + j) * n_ao] += out[i * oj + j];
            }
    }
    return 0;
}

/* ------------------------------------------------------------------------------------------ */
/* Nuclear attraction V_{mu nu}(C) = int mu nu / |r - C|  (no charge, no sign)                 */
/* ------------------------------------------------------------------------------------------ */

int oracle_nuclear_attraction(int i_L, int j_L, const double *pair_P_p, const double *pair_A_a, const double *pair_B_b,
                              const double *pair_coefficient, const int *pair_i_ao_start, const int *pair_j_ao_start,
                              int n_pair, const double *charge_position, int n_charge, double *V_tensor, int n_ao,
                              bool spherical, double omega)
{
    ensure_tables();
    if (!class_ok_orbital(i_L, j_L)) {
        printf("%s function does not support angular i_L = %d, j_L = %d\n", __func__, i_L, j_L);
        fflush(stdout);
        return 1;
    }
    const int ni = ncart(i_L), nj = ncart(j_L);
    const int oi = spherical ? 2 * i_L + 1 : ni, oj = spherical ? 2 * j_L + 1 : nj;
#pragma omp parallel
    {
        real_t cart[OR_MAX_CART * OR_MAX_CART], out[OR_MAX_CART * OR_MAX_CART];
        real_t Ex[OR_MAX_L + 1][OR_MAX_L + 1][ESTR], Ey[OR_MAX_L + 1][OR_MAX_L + 1][ESTR], Ez[OR_MAX_L + 1][OR_MAX_L + 1][ESTR];
        herm_R R;
        herm_R_alloc(&R, i_L + j_L);
        /* parallel over charges: each thread owns output column c, no two threads touch one element */
#pragma omp for schedule(dynamic)
        for (int ic = 0; ic < n_charge; ic++) {
            const double *C = charge_position + 3 * ic;
            for (int ip = 0; ip < n_pair; ip++) {
                const double *P = pair_P_p + 4 * ip, *A = pair_A_a + 4 * ip, *B = pair_B_b + 4 * ip;
                const real_t p = P[3];
                hermite_E_pair(i_L, j_L, P[0] - A[0], A[0] - B[0], 0.5 / p, Ex);
                hermite_E_pair(i_L, j_L, P[1] - A[1], A[1] - B[1], 0.5 / p, Ey);
                hermite_E_pair(i_L, j_L, P[2] - A[2], A[2] - B[2], 0.5 / p, Ez);
                /* reference nuclear_attraction.cpp:44-45, 69 */
                const real_t pref = (omega == 0.0) ? pair_coefficient[ip] * 2.0 * OR_PI / p
                                                   : pair_coefficient[ip] * 2.0 * OR_PI * omega / (p * sqrt(p + omega * omega));
                const real_t zeta = (omega == 0.0) ? p : p * omega * omega / (p + omega * omega);
                herm_R_compute(&R, zeta, P[0] - C[0], P[1] - C[1], P[2] - C[2]);
                for (int ci = 0; ci < ni; ci++) {
                    int ix, iy, iz; cart_xyz(i_L, ci, &ix, &iy, &iz);
                    for (int cj = 0; cj < nj; cj++) {
                        int jx, jy, jz; cart_xyz(j_L, cj, &jx, &jy, &jz);
                        real_t v = 0.0;
                        for (int tx = 0; tx <= ix + jx; tx++)
                            for (int ty = 0; ty <= iy + jy; ty++)
                                for (int tz = 0; tz <= iz + jz; tz++)
                                    v += Ex[ix][jx][tx] * Ey[iy][jy][ty] * Ez[iz][jz][tz] * herm_R_get(&R, tx, ty, tz);
                        cart[ci * nj + cj] = pref * v * cart_norm(i_L, ci) * cart_norm(j_L, cj);
                    }
                }
                transform_ij(i_L, j_L, 1, cart, out, spherical);
                const int is = pair_i_ao_start[ip], js = pair_j_ao_start[ip];
                for (int i = 0; i < oi; i++)
                    for (int j = 0; j < oj; j++) {
                        V_tensor[((size_t) (is + i) * n_ao + (js + j)) * n_charge + ic] += out[i * oj + j];
                        if (is != js) V_tensor[((size_t) (is + i) + (size_t) (js + j) * n_ao) * n_charge + ic] += out[i * oj + j];
                    }
            }
        }
        herm_R_free(&R);
    }
    return 0;
}

/* ------------------------------------------------------------------------------------------ */
/* Two-centre (P|Q), optionally erf-attenuated                                                  */
/* ------------------------------------------------------------------------------------------ */

int oracle_two_center(int i_L, int j_L, const double *aux_A_a, const double *aux_i_coefficient, const int *aux_i_aux_start,
                      int n_aux_primitive_i, const double *aux_B_b, const double *aux_j_coefficient,
                      const int *aux_j_aux_start, int n_aux_primitive_j, double *J2c_matrix, int n_aux,
                      bool spherical, double omega)
{
    ensure_tables();
    if (!(i_L >= 0 && j_L >= 0 && i_L <= j_L && j_L <= OR_MAX_AUX_L)) {
        printf("%s function does not support angular i_L = %d, j_L = %d\n", __func__, i_L, j_L);
        fflush(stdout);
        return 1;
    }
    const int ni = ncart(i_L), nj = ncart(j_L);
    const int oi = spherical ? 2 * i_L + 1 : ni, oj = spherical ? 2 * j_L + 1 : nj;
#pragma omp parallel
    {
        real_t *cart = (real_t *) malloc(sizeof(real_t) * ni * nj);
        real_t *out = (real_t *) malloc(sizeof(real_t) * ni * nj);
        real_t Ei[OR_MAX_AUX_L + 1][OR_MAX_AUX_L + 1], Ej[OR_MAX_AUX_L + 1][OR_MAX_AUX_L + 1];
        herm_R R;
        herm_R_alloc(&R, i_L + j_L);
        /* parallel over rows i: a thread owns the rows of its i primitive's shell and their mirrors
         * are written by nobody else only if shells are processed whole -- so split by shell start. */
#pragma omp for schedule(dynamic)
        for (int row_owner = 0; row_owner < n_aux; row_owner++) {
            for (int ia = 0; ia < n_aux_primitive_i; ia++) {
                if (aux_i_aux_start[ia] != row_owner) continue;
                for (int ja = 0; ja < n_aux_primitive_j; ja++) {
                    const int is = aux_i_aux_start[ia], js = aux_j_aux_start[ja];
                    if (i_L == j_L && is > js) continue;      /* reference two_center.cpp:138 */
                    const double *A = aux_A_a + 4 * ia, *B = aux_B_b + 4 * ja;
                    const real_t p = A[3], q = B[3];
                    const real_t coef = aux_i_coefficient[ia] * aux_j_coefficient[ja];
                    /* reference two_center.cpp:42-45 */
                    const real_t pref = (omega == 0.0) ? coef * 2.0 * pow(OR_PI, 2.5) / (p * q * sqrt(p + q))
                                                       : coef * 2.0 * pow(OR_PI, 2.5) * omega / (p * q * sqrt(p * q + (p + q) * omega * omega));
                    const real_t zeta = (omega == 0.0) ? p * q / (p + q) : 1.0 / (1.0 / p + 1.0 / q + 1.0 / (omega * omega));
                    hermite_E_single(i_L, 0.5 / p, Ei);
                    hermite_E_single(j_L, 0.5 / q, Ej);
                    herm_R_compute(&R, zeta, A[0] - B[0], A[1] - B[1], A[2] - B[2]);
                    for (int ci = 0; ci < ni; ci++) {
                        int ix, iy, iz; cart_xyz(i_L, ci, &ix, &iy, &iz);
                        for (int cj = 0; cj < nj; cj++) {
                            int jx, jy, jz; cart_xyz(j_L, cj, &jx, &jy, &jz);
                            real_t v = 0.0;
                            for (int tx = ix % 2; tx <= ix; tx += 2)
                                for (int ty = iy % 2; ty <= iy; ty += 2)
                                    for (int tz = iz % 2; tz <= iz; tz += 2)
                                        for (int sx = jx % 2; sx <= jx; sx += 2)
                                            for (int sy = jy % 2; sy <= jy; sy += 2)
                                                for (int sz = jz % 2; sz <= jz; sz += 2)
                                                    v += (((sx + sy + sz) % 2) ? -1.0 : 1.0)
                                                         * Ei[ix][tx] * Ei[iy][ty] * Ei[iz][tz]
                                                         * Ej[jx][sx] * Ej[jy][sy] * Ej[jz][sz]
                                                         * herm_R_get(&R, tx + sx, ty + sy, tz + sz);
                            cart[ci * nj + cj] = pref * v * cart_norm(i_L, ci) * cart_norm(j_L, cj);
                        }
                    }
                    transform_ij(i_L, j_L, 1, cart, out, spherical);
                    for (int i = 0; i < oi; i++)
                        for (int j = 0; j < oj; j++) {
                            J2c_matrix[(size_t) (is + i) * n_aux + (js + j)] += out[i * oj + j];
                            if (is != js) J2c_matrix[(size_t) (is + i) + (size_t) (js + j) * n_aux] += out[i * oj + j];
                        }
                }
            }
        }
        herm_R_free(&R);
        free(cart);
        free(out);
    }
    return 0;
}

/* ------------------------------------------------------------------------------------------ */
/* Three-centre (P|ij), optionally erf-attenuated                                               */
/* ------------------------------------------------------------------------------------------ */

/* Per-thread work space of one three-centre triplet evaluation. */
typedef struct {
    herm_R R;
    real_t *W;        /* W[ck][t]: sum_s (-1)^|s| E^k_s R_{t+s} for every pair Hermite index t */
    real_t *cart;     /* (nk, ni, nj) Cartesian block, normalised */
    real_t *ijsph;    /* (nk, oi, oj) after the ij transform */
} j3c_ws;

/* One primitive triplet (pair ip, auxiliary primitive ka) into blk[(k * oi + i) * oj + j]
 * (reference three_center.cpp:27-173 for the triplet, :216-248 for the transforms).  The
 * auxiliary-side sum over s is formed once per (Cartesian k component, pair Hermite index t) and
 * then contracted with E^ij_t for every (i, j) -- the reference's nest :93-160 with its two inner
 * sums in the other order. */
static void three_center_triplet(int i_L, int j_L, int k_L, const double *P, const double *A, const double *B,
                                 double pair_coef, const double *C, double aux_coef, double omega, bool spherical,
                                 j3c_ws *ws, real_t *blk)
{
    const int ni = ncart(i_L), nj = ncart(j_L), nk = ncart(k_L);
    const int oi = spherical ? 2 * i_L + 1 : ni, oj = spherical ? 2 * j_L + 1 : nj, ok = spherical ? 2 * k_L + 1 : nk;
    const int Lab = i_L + j_L, S2 = Lab + 1, NT = S2 * S2 * S2;
    real_t Ex[OR_MAX_L + 1][OR_MAX_L + 1][ESTR], Ey[OR_MAX_L + 1][OR_MAX_L + 1][ESTR], Ez[OR_MAX_L + 1][OR_MAX_L + 1][ESTR];
    real_t Ek[OR_MAX_AUX_L + 1][OR_MAX_AUX_L + 1];
    const real_t q = C[3], p = P[3];
    hermite_E_single(k_L, 0.5 / q, Ek);
    const real_t coef = (real_t) pair_coef * (real_t) aux_coef;
    /* reference three_center.cpp:46-49 */
    const real_t pref = (omega == 0.0) ? coef * 2.0 * pow(OR_PI, 2.5) / (p * q * sqrt(p + q))
                                       : coef * 2.0 * pow(OR_PI, 2.5) * omega / (p * q * sqrt(p * q + (p + q) * omega * omega));
    const real_t zeta = (omega == 0.0) ? p * q / (p + q) : 1.0 / (1.0 / p + 1.0 / q + 1.0 / (omega * omega));
    hermite_E_pair(i_L, j_L, P[0] - A[0], A[0] - B[0], 0.5 / p, Ex);
    hermite_E_pair(i_L, j_L, P[1] - A[1], A[1] - B[1], 0.5 / p, Ey);
    hermite_E_pair(i_L, j_L, P[2] - A[2], A[2] - B[2], 0.5 / p, Ez);
    herm_R_compute(&ws->R, zeta, P[0] - C[0], P[1] - C[1], P[2] - C[2]);
    for (int ck = 0; ck < nk; ck++) {
        const int kx = cxyz_tab[k_L][ck][0], ky = cxyz_tab[k_L][ck][1], kz = cxyz_tab[k_L][ck][2];
        real_t *Wk = ws->W + (size_t) ck * NT;
        for (int tx = 0; tx <= Lab; tx++)
            for (int ty = 0; tx + ty <= Lab; ty++)
                for (int tz = 0; tx + ty + tz <= Lab; tz++) {
                    real_t w = 0.0;
                    for (int sx = kx % 2; sx <= kx; sx += 2)
                        for (int sy = ky % 2; sy <= ky; sy += 2)
                            for (int sz = kz % 2; sz <= kz; sz += 2)
                                w += (((sx + sy + sz) % 2) ? -1.0 : 1.0) * Ek[kx][sx] * Ek[ky][sy] * Ek[kz][sz]
                                     * herm_R_get(&ws->R, tx + sx, ty + sy, tz + sz);
                    Wk[(tx * S2 + ty) * S2 + tz] = w;
                }
    }
    for (int ck = 0; ck < nk; ck++) {
        const real_t *Wk = ws->W + (size_t) ck * NT;
        for (int ci = 0; ci < ni; ci++) {
            const int ix = cxyz_tab[i_L][ci][0], iy = cxyz_tab[i_L][ci][1], iz = cxyz_tab[i_L][ci][2];
            for (int cj = 0; cj < nj; cj++) {
                const int jx = cxyz_tab[j_L][cj][0], jy = cxyz_tab[j_L][cj][1], jz = cxyz_tab[j_L][cj][2];
                real_t v = 0.0;
                for (int tx = 0; tx <= ix + jx; tx++)
                    for (int ty = 0; ty <= iy + jy; ty++) {
                        const real_t exy = Ex[ix][jx][tx] * Ey[iy][jy][ty];
                        for (int tz = 0; tz <= iz + jz; tz++)
                            v += exy * Ez[iz][jz][tz] * Wk[(tx * S2 + ty) * S2 + tz];
                    }
                ws->cart[((size_t) ck * ni + ci) * nj + cj] = pref * v * cart_norm(i_L, ci) * cart_norm(j_L, cj) * cart_norm(k_L, ck);
            }
        }
    }
    /* ij transform per k component, then k transform per ij element */
    transform_ij(i_L, j_L, nk, ws->cart, ws->ijsph, spherical);
    for (int i = 0; i < oi; i++)
        for (int j = 0; j < oj; j++)
            for (int k = 0; k < ok; k++)
                blk[((size_t) k * oi + i) * oj + j] = spherical
                    ? sph_row_dot(k_L, k, ws->ijsph + (size_t) i * oj + j, oi * oj)
                    : ws->ijsph[((size_t) k * oi + i) * oj + j];
}

/* Triplets are taken in (auxiliary shell, auxiliary primitive, pair) order, in rounds of at most
 * kRoundReals output values: the triplets of a round are evaluated in parallel into a buffer, then
 * each output row (auxiliary function) adds its contributions in that fixed order, one thread per
 * row.  Results therefore do not depend on the thread count, and a single auxiliary shell over
 * millions of pairs runs on every core. */
#define J3C_ROUND_REALS (1L << 25)

int oracle_three_center(int i_L, int j_L, int k_L, const double *pair_P_p, const double *pair_A_a, const double *pair_B_b,
                        const double *pair_coefficient, const int *pair_i_ao_start, const int *pair_j_ao_start,
                        int n_pair_ij, const double *aux_A_a, const double *aux_coefficient, const int *aux_i_aux_start,
                        int n_aux_primitive_k, double *J3c_matrix, int n_ao, int n_aux, bool spherical, double omega)
{
    ensure_tables();
    if (!class_ok_orbital(i_L, j_L) || k_L < 0 || k_L > OR_MAX_AUX_L) {
        printf("%s function does not support angular i_L = %d, j_L = %d, k_L = %d\n", __func__, i_L, j_L, k_L);
        fflush(stdout);
        return 1;
    }
    (void) n_aux;
    if (n_pair_ij <= 0 || n_aux_primitive_k <= 0) return 0;
    const int ni = ncart(i_L), nj = ncart(j_L), nk = ncart(k_L);
    const int oi = spherical ? 2 * i_L + 1 : ni, oj = spherical ? 2 * j_L + 1 : nj, ok = spherical ? 2 * k_L + 1 : nk;
    const size_t bsz = (size_t) ok * oi * oj;
    const int S2 = i_L + j_L + 1;
    /* auxiliary primitives grouped by shell start, first-appearance order of the starts */
    int *starts = (int *) malloc(sizeof(int) * n_aux_primitive_k);
    int *kord = (int *) malloc(sizeof(int) * n_aux_primitive_k);     /* primitives in (start, index) order */
    int *kgrp = (int *) malloc(sizeof(int) * (n_aux_primitive_k + 1)); /* CSR over starts */
    int n_start = 0;
    for (int ka = 0; ka < n_aux_primitive_k; ka++) {
        int seen = 0;
        for (int s = 0; s < n_start; s++) if (starts[s] == aux_i_aux_start[ka]) { seen = 1; break; }
        if (!seen) starts[n_start++] = aux_i_aux_start[ka];
    }
    int nk_ord = 0;
    for (int s = 0; s < n_start; s++) {
        kgrp[s] = nk_ord;
        for (int ka = 0; ka < n_aux_primitive_k; ka++)
            if (aux_i_aux_start[ka] == starts[s]) kord[nk_ord++] = ka;
    }
    kgrp[n_start] = nk_ord;
    const long long n_trip = (long long) n_aux_primitive_k * n_pair_ij;
    long long round = J3C_ROUND_REALS / (long long) bsz;
    if (round < 64) round = 64;
    if (round > n_trip) round = n_trip;
    real_t *buf = (real_t *) malloc(sizeof(real_t) * bsz * (size_t) round);
    for (long long t0 = 0; t0 < n_trip; t0 += round) {
        const long long t1 = (t0 + round < n_trip) ? t0 + round : n_trip;
#pragma omp parallel
        {
            j3c_ws ws;
            herm_R_alloc(&ws.R, i_L + j_L + k_L);
            ws.W = (real_t *) malloc(sizeof(real_t) * (size_t) nk * S2 * S2 * S2);
            ws.cart = (real_t *) malloc(sizeof(real_t) * (size_t) nk * ni * nj);
            ws.ijsph = (real_t *) malloc(sizeof(real_t) * (size_t) nk * ni * nj);
#pragma omp for schedule(dynamic, 16)
            for (long long t = t0; t < t1; t++) {
                const int ka = kord[t / n_pair_ij], ip = (int) (t % n_pair_ij);
                three_center_triplet(i_L, j_L, k_L, pair_P_p + 4 * ip, pair_A_a + 4 * ip, pair_B_b + 4 * ip,
                                     pair_coefficient[ip], aux_A_a + 4 * ka, aux_coefficient[ka], omega, spherical,
                                     &ws, buf + (size_t) (t - t0) * bsz);
            }
            herm_R_free(&ws.R);
            free(ws.W);
            free(ws.cart);
            free(ws.ijsph);
        }
        /* shells touched by this round, and their triplet ranges inside it */
        const int s_lo_idx = (int) (t0 / n_pair_ij), s_hi_idx = (int) ((t1 - 1) / n_pair_ij);
        int s_first = 0;
        while (kgrp[s_first + 1] <= s_lo_idx) s_first++;
        int s_last = s_first;
        while (kgrp[s_last + 1] <= s_hi_idx) s_last++;
        const int n_rows = (s_last - s_first + 1) * ok;
#pragma omp parallel for schedule(dynamic, 1)
        for (int r = 0; r < n_rows; r++) {
            const int s = s_first + r / ok, k = r % ok;
            long long a = (long long) kgrp[s] * n_pair_ij, b = (long long) kgrp[s + 1] * n_pair_ij;
            if (a < t0) a = t0;
            if (b > t1) b = t1;
            double *row = J3c_matrix + (size_t) (starts[s] + k) * n_ao * n_ao;
            for (long long t = a; t < b; t++) {
                const int ip = (int) (t % n_pair_ij);
                const int is = pair_i_ao_start[ip], js = pair_j_ao_start[ip];
                const real_t *blk = buf + (size_t) (t - t0) * bsz + (size_t) k * oi * oj;
                for (int i = 0; i < oi; i++)
                    for (int j = 0; j < oj; j++) {
                        const real_t val = blk[i * oj + j];
                        row[(size_t) (is + i) * n_ao + (js + j)] += val;
                        if (is != js) row[(size_t) (is + i) + (size_t) (js + j) * n_ao] += val;
                    }
            }
        }
    }
    free(buf);
    free(starts);
    free(kord);
    free(kgrp);
    return 0;
}

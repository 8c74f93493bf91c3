// A-priori term scale of three-centre work items, the estimate the mixed-precision split is made on.
//
// For an item (shell pair sp, auxiliary shells k0 .. k0 + 63) of class (la lb | lc), every output
// (ij|P_m) is a sum over primitive pairs pp of sp and primitives of P of
//     sum_t E_pp[ij][t] V_m[t],   V_m[t] = sum_s c_m(s) R_{t+s},   |t + s| = |t| + lc,
// and its float32 rounding error is a few float32 ulps of the magnitude of the terms summed, not of the
// result (the primitive pairs of a contracted shell pair cancel, and so do the terms of the Hermite
// recursion in the far field).  That magnitude is bounded here by
//     cmax * sum_pp sum_n Esum_pp[n] * A_{n + lc}
// with Esum_pp[n] = max_ij sum_{|t| = n} |E_pp[ij][t]| (pair-matrix row sums per Hermite level, from the
// float64 pair matrix), cmax = max_m sum_s |c_m(s)|, and A_N the Hermite recursion with every term taken
// positive: A_0^(n) = |R000^(n)| = |pf| (2 zeta)^n F_n(T), A_N^(n) = (N - 1) A_{N-2}^(n+1) + |PQ| A_{N-1}^(n+1)
// (|PQ| >= |PQx|, |PQy|, |PQz|), which bounds every term that enters an R_tuv of order N.  The Boys
// function and the prefactor are the kernels' own (float64), so the estimate carries the distance decay
// of the integrals, unlike a Schwarz-type bound.  One wave per item, one lane per auxiliary shell; the
// item's scale is the largest over its shells.  Cost: O(L^2) per primitive triplet, against O(nij n_t)
// for the integral itself.
#include <hip/hip_runtime.h>

#include <algorithm>

#include "md_device.hpp"

namespace {

struct ScaleArgs {
    const int* items;         // [n_items][2]: (shell pair, first auxiliary shell)
    int item_begin, item_end;
    const int* sp_pp_off;     // [n_sp + 1]
    const double* pp_P;       // [n_pp][4]
    const float* pp_esum;     // [n_pp][n_lev]
    int n_lev;                // la + lb + 1
    const int* ks_prim_off;   // [n_ks + 1]
    const double* ka_C;       // [n_kprim][4]
    const double* ka_coef;    // [n_kprim]
    int n_ks, lc;
    double cmax, omega;
    float* scale;             // [n_items]
};

template <int L>
__global__ void __launch_bounds__(64) j3c_scale_kernel(const ScaleArgs a)
{
    const int item = a.item_begin + (int) blockIdx.x;
    if (item >= a.item_end) return;
    const int lane = (int) threadIdx.x;
    const int sp = a.items[2 * item], k0 = a.items[2 * item + 1];
    const int ks = k0 + lane;
    double acc = 0.0;
    if (ks < a.n_ks) {
        const int pp0 = a.sp_pp_off[sp], pp1 = a.sp_pp_off[sp + 1];
        for (int kp = a.ks_prim_off[ks]; kp < a.ks_prim_off[ks + 1]; ++kp) {
            const double Cx = a.ka_C[4 * kp], Cy = a.ka_C[4 * kp + 1], Cz = a.ka_C[4 * kp + 2], q = a.ka_C[4 * kp + 3];
            // |ck 2 pi^(5/2) (1/(2q))^lc / q|, the auxiliary primitive's share of the prefactor
            double hl = 1.0;
            for (int i = 0; i < a.lc; ++i) hl *= 0.5 / q;
            const double ckq = fabs(a.ka_coef[kp]) * eip::kTwoPi25 * hl / q;
            for (int pp = pp0; pp < pp1; ++pp) {
                const double dx = a.pp_P[4 * pp] - Cx, dy = a.pp_P[4 * pp + 1] - Cy, dz = a.pp_P[4 * pp + 2] - Cz;
                const double p = a.pp_P[4 * pp + 3];
                double zeta, pf;
                if (a.omega == 0.0) {
                    const double rs = eip::rsq_nr(p + q);
                    zeta = p * q * (rs * rs);
                    pf = ckq * eip::rcp_nr(p) * rs;
                } else {
                    const double w2 = a.omega * a.omega;
                    zeta = p * q * eip::rcp_nr(q + p + p * q * eip::rcp_nr(w2));
                    pf = ckq * a.omega * eip::rcp_nr(p) * eip::rsq_nr(p * q + (p + q) * w2);
                }
                const double r2 = dx * dx + dy * dy + dz * dz;
                double F[L + 1];
                eip::boys64<L>(zeta * r2, F);
                const double Rd = sqrt(r2);
                // A[N][n], N + n <= L, built level by level in N
                double A[L + 1][L + 1];
                double f = pf;
#pragma unroll
                for (int n = 0; n <= L; ++n) {
                    A[0][n] = f * F[n];
                    f *= 2.0 * zeta;
                }
#pragma unroll
                for (int N = 1; N <= L; ++N)
#pragma unroll
                    for (int n = 0; n + N <= L; ++n)
                        A[N][n] = (N >= 2 ? (N - 1) * A[N - 2][n + 1] : 0.0) + Rd * A[N - 1][n + 1];
                const float* e = a.pp_esum + (size_t) pp * a.n_lev;
                double s = 0.0;
                for (int n = 0; n < a.n_lev; ++n) {
                    const int N = n + a.lc;
                    double An = 0.0;
#pragma unroll
                    for (int k = 0; k <= L; ++k)
                        if (k == N) An = A[k][0];
                    s += (double) e[n] * An;
                }
                acc += s;
            }
        }
    }
    float v = (float) (a.cmax * acc);
    v = eip::wave_max_f(v);
    if (lane == 0) a.scale[item] = v;
}

template <int L>
int launch_scale(const ScaleArgs& a, int n, hipStream_t s)
{
    hipLaunchKernelGGL(j3c_scale_kernel<L>, dim3(n), dim3(64), 0, s, a);
    return 0;
}

int dispatch_scale(int L, const ScaleArgs& a, int n, hipStream_t s)
{
    switch (L) {
        case 0: return launch_scale<0>(a, n, s);
        case 1: return launch_scale<1>(a, n, s);
        case 2: return launch_scale<2>(a, n, s);
        case 3: return launch_scale<3>(a, n, s);
        case 4: return launch_scale<4>(a, n, s);
        case 5: return launch_scale<5>(a, n, s);
        case 6: return launch_scale<6>(a, n, s);
        case 7: return launch_scale<7>(a, n, s);
        case 8: return launch_scale<8>(a, n, s);
        case 9: return launch_scale<9>(a, n, s);
        case 10: return launch_scale<10>(a, n, s);
    }
    return -1;
}

}  // namespace

extern "C" int eip_three_center_scale(int i_L, int j_L, int k_L, const int* d_items, int n_items,
                                      const int* d_sp_pp_off, const double* d_pp_P, const float* d_pp_esum,
                                      const int* d_ks_prim_off, const double* d_ka_C, const double* d_ka_coef,
                                      int n_ks, double cmax, double omega, float* d_scale, void* stream)
{
    if (i_L < 0 || j_L < 0 || k_L < 0 || i_L + j_L + k_L > 10) return 1;
    if (n_items <= 0) return 0;
    if (!d_items || !d_sp_pp_off || !d_pp_P || !d_pp_esum || !d_ks_prim_off || !d_ka_C || !d_ka_coef || !d_scale ||
        n_ks <= 0 || !(cmax > 0.0))
        return 4;
    ScaleArgs a{d_items, 0, 0, d_sp_pp_off, d_pp_P, d_pp_esum, i_L + j_L + 1, d_ks_prim_off, d_ka_C, d_ka_coef,
                n_ks, k_L, cmax, omega, d_scale};
    constexpr int kMaxGrid = 1 << 22;
    for (int b0 = 0; b0 < n_items; b0 += kMaxGrid) {
        const int n = std::min(kMaxGrid, n_items - b0);
        a.item_begin = b0;
        a.item_end = b0 + n;
        if (dispatch_scale(i_L + j_L + k_L, a, n, (hipStream_t) stream) != 0) return 3;
    }
    return hipGetLastError() == hipSuccess ? 0 : 2;
}

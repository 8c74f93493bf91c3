// Device-side building blocks shared by every generated McMurchie-Davidson kernel (gfx950).
//
// * Boys function F_0..F_L (float32 variant: the large-T upward recursion stays float64): Taylor expansion of F_L about the nearest point of a 1/16 grid on [0, 36]
//   followed by the stable downward recursion; above the grid F_0 = sqrt(pi/T)/2 and the upward
//   recursion (stable there since 2T > 2m+1 for every order used).  The float64 version is accurate to
//   a few ulp; the float32 version to a few float ulp.  (Reference: boys_function.hpp:102-176 uses a
//   five-branch scheme with a Chebyshev table; both agree to ~1e-15 relative.)
// * XCD-aware block numbering: the dispatcher hands consecutive workgroups to the 8 XCDs round-robin;
//   remapping puts consecutive logical blocks (which share shell-pair data) on the same XCD and L2.
#pragma once

#include <hip/hip_runtime.h>

#include "generated/boys_table.inc"

namespace eip {

// Four doubles: the accumulator tile slice one lane holds for v_mfma_f64_16x16x4f64.
typedef double v4d __attribute__((ext_vector_type(4)));
// Four floats: the accumulator slice of v_mfma_f32_16x16x32_bf16; eight bfloat16: one lane's operand.
typedef float v4f __attribute__((ext_vector_type(4)));
typedef __bf16 bf8 __attribute__((ext_vector_type(8)));
typedef unsigned int u4 __attribute__((ext_vector_type(4)));

constexpr double kTwoPi25 = 34.986836655249725;     // 2 pi^(5/2)
constexpr double kTwoPi = 6.283185307179586;        // 2 pi
constexpr double kPi = 3.141592653589793;

__device__ __forceinline__ int xcd_remap(int b, int nblk)
{
    // Blocks are dealt to XCDs round-robin: physical block b runs on XCD b % 8.  Give XCD x the
    // logical range [x * per, (x+1) * per) so neighbours in logical order share one L2.
    // XCD x owns per (+1 when x < rem) blocks; its logical range starts at x*per + min(x, rem).
    constexpr int kXcd = 8;
    const int per = nblk / kXcd;
    const int rem = nblk - per * kXcd;
    const int x = b % kXcd;
    const int i = b / kXcd;
    return x * per + (x < rem ? x : rem) + i;
}

// 1/x and 1/sqrt(x) for positive normal x: the hardware estimate (v_rcp_f64, v_rsq_f64) refined by two
// Newton steps, accurate to a few ulp (IEEE division and square root cost three to four times more
// instructions).
__device__ __forceinline__ double rcp_nr(const double x)
{
    double y = __builtin_amdgcn_rcp(x);
    double e = fma(-x, y, 1.0);
    y = fma(y, e, y);
    e = fma(-x, y, 1.0);
    return fma(y, e, y);
}

__device__ __forceinline__ double rsq_nr(const double x)
{
    double y = __builtin_amdgcn_rsq(x);
    const double h = 0.5 * x;
    y = y * fma(-h, y * y, 1.5);
    return y * fma(-h, y * y, 1.5);
}

constexpr double kHalfSqrtPi = 0.88622692545275801365;   // sqrt(pi) / 2

template <int L>
__device__ __forceinline__ void boys64(const double T, double* __restrict__ F)
{
    // e^{-T} once for both branches: a wave whose lanes fall on both sides of the switch runs both
    double eT = 0.0;
    if constexpr (L > 0) eT = exp(-T);
    if (T < BOYS_T_SWITCH) {
        const int k = __double2int_rn(T * BOYS_H_INV);
        const double x = k * (1.0 / BOYS_H_INV) - T;     // -(T - T0)
        const double* __restrict__ row = boys_grid64 + k * BOYS_NCOL + L;
        double acc = row[BOYS_K64];
#pragma unroll
        for (int j = BOYS_K64 - 1; j >= 0; --j) acc = fma(x * (1.0 / (j + 1)), acc, row[j]);
        F[L] = acc;
        if constexpr (L > 0) {
            const double twoT = T + T;
#pragma unroll
            for (int m = L - 1; m >= 0; --m) F[m] = fma(twoT, F[m + 1], eT) * (1.0 / (2 * m + 1));
        }
    } else {
        const double rT = rsq_nr(T);
        F[0] = kHalfSqrtPi * rT;                          // sqrt(pi / T) / 2
        if constexpr (L > 0) {
            const double half_invT = 0.5 * (rT * rT);
#pragma unroll
            for (int m = 0; m < L; ++m) F[m + 1] = fma((double) (2 * m + 1), F[m], -eT) * half_invT;
        }
    }
}

// boys64 from the coarse table slice of order L staged in LDS (tab: BOYS_KC + 1 columns of BOYS_C_NROW
// rows, column c holding F_{L+c} on the grid T0 = k / BOYS_C_H_INV): Taylor expansion of F_L about the
// nearest grid point with BOYS_KC + 1 terms (|T - T0| <= 1/4), then the same downward recursion (and the
// same large-T branch) as boys64.
template <int L>
__device__ __forceinline__ void boys64_lds(const double T, const double* __restrict__ tab, double* __restrict__ F)
{
    static_assert(L <= BOYS_LDS_LMAX, "order without a staged slice");
    double eT = 0.0;
    if constexpr (L > 0) eT = exp(-T);
    if (T < BOYS_T_SWITCH) {
        const int k = __double2int_rn(T * BOYS_C_H_INV);
        const double x = k * (1.0 / BOYS_C_H_INV) - T;   // -(T - T0)
        double acc = tab[BOYS_KC * BOYS_C_NROW + k];
#pragma unroll
        for (int j = BOYS_KC - 1; j >= 0; --j) acc = fma(x * (1.0 / (j + 1)), acc, tab[j * BOYS_C_NROW + k]);
        F[L] = acc;
        if constexpr (L > 0) {
            const double twoT = T + T;
#pragma unroll
            for (int m = L - 1; m >= 0; --m) F[m] = fma(twoT, F[m + 1], eT) * (1.0 / (2 * m + 1));
        }
    } else {
        const double rT = rsq_nr(T);
        F[0] = kHalfSqrtPi * rT;                          // sqrt(pi / T) / 2
        if constexpr (L > 0) {
            const double half_invT = 0.5 * (rT * rT);
#pragma unroll
            for (int m = 0; m < L; ++m) F[m + 1] = fma((double) (2 * m + 1), F[m], -eT) * half_invT;
        }
    }
}

// boys64 split in two for software pipelining: boys64_fetch loads the grid row a Taylor evaluation of
// F_L(T) reads (clamped to the last row, so it can be issued before the branch on T is known; a row
// fetched for T beyond the grid is simply unused), boys64_eval evaluates from that row.
template <int L>
__device__ __forceinline__ void boys64_fetch(const double T, double* __restrict__ row)
{
    const int k = __double2int_rn(fmin(T, (double) BOYS_T_SWITCH) * BOYS_H_INV);
    const double* __restrict__ r = boys_grid64 + k * BOYS_NCOL + L;
#pragma unroll
    for (int j = 0; j <= BOYS_K64; ++j) row[j] = r[j];
}

template <int L>
__device__ __forceinline__ void boys64_eval(const double T, const double* __restrict__ row, double* __restrict__ F)
{
    double eT = 0.0;
    if constexpr (L > 0) eT = exp(-T);
    if (T < BOYS_T_SWITCH) {
        const int k = __double2int_rn(T * BOYS_H_INV);
        const double x = k * (1.0 / BOYS_H_INV) - T;     // -(T - T0)
        double acc = row[BOYS_K64];
#pragma unroll
        for (int j = BOYS_K64 - 1; j >= 0; --j) acc = fma(x * (1.0 / (j + 1)), acc, row[j]);
        F[L] = acc;
        if constexpr (L > 0) {
            const double twoT = T + T;
#pragma unroll
            for (int m = L - 1; m >= 0; --m) F[m] = fma(twoT, F[m + 1], eT) * (1.0 / (2 * m + 1));
        }
    } else {
        const double rT = rsq_nr(T);
        F[0] = kHalfSqrtPi * rT;                          // sqrt(pi / T) / 2
        if constexpr (L > 0) {
            const double half_invT = 0.5 * (rT * rT);
#pragma unroll
            for (int m = 0; m < L; ++m) F[m + 1] = fma((double) (2 * m + 1), F[m], -eT) * half_invT;
        }
    }
}

template <int L>
__device__ __forceinline__ void boys32(const float T, float* __restrict__ F)
{
    if (T < (float) BOYS_T_SWITCH) {
        const int k = __float2int_rn(T * (float) BOYS_H_INV);
        const float x = k * (1.0f / BOYS_H_INV) - T;
        const float* __restrict__ row = boys_grid32 + k * BOYS_NCOL + L;
        float acc = row[BOYS_K32];
#pragma unroll
        for (int j = BOYS_K32 - 1; j >= 0; --j) acc = fmaf(x * (1.0f / (j + 1)), acc, row[j]);
        F[L] = acc;
        if constexpr (L > 0) {
            const float eT = __expf(-T);
            const float twoT = T + T;
#pragma unroll
            for (int m = L - 1; m >= 0; --m) F[m] = fmaf(twoT, F[m + 1], eT) * (1.0f / (2 * m + 1));
        }
    } else {
        // large T: the upward recursion runs in float64 and is rounded once per order
        const double Td = T;
        const double rT = rsq_nr(Td);
        double f = kHalfSqrtPi * rT;
        F[0] = (float) f;
        if constexpr (L > 0) {
            const double eT = exp(-Td);
            const double half_invT = 0.5 * (rT * rT);
#pragma unroll
            for (int m = 0; m < L; ++m) { f = fma((double) (2 * m + 1), f, -eT) * half_invT; F[m + 1] = (float) f; }
        }
    }
}

// Split of eight floats into P bfloat16 pieces by truncation: piece 0 keeps the top 16 bits of x (sign,
// exponent, 7 mantissa bits), piece k the top 16 bits of what the earlier pieces leave; every remainder is
// exact in float32, so three pieces represent any float32 value exactly (each piece takes at least 8 of
// its 24 significant bits).  h[k] holds piece k of the eight values, element v in lane v of the vector.
template <int P>
__device__ __forceinline__ void split_bf16(const float (&x)[8], bf8* __restrict__ h)
{
    unsigned int r[8];
#pragma unroll
    for (int v = 0; v < 8; ++v) r[v] = __float_as_uint(x[v]);
#pragma unroll
    for (int k = 0; k < P; ++k) {
        u4 w;
#pragma unroll
        for (int q = 0; q < 4; ++q) w[q] = (r[2 * q] >> 16) | (r[2 * q + 1] & 0xffff0000u);
        h[k] = __builtin_bit_cast(bf8, w);
        if (k + 1 < P) {
#pragma unroll
            for (int v = 0; v < 8; ++v)
                r[v] = __float_as_uint(__uint_as_float(r[v]) - __uint_as_float(r[v] & 0xffff0000u));
        }
    }
}

template <int P>
__device__ __forceinline__ void split_bf16(const double (&x)[8], bf8* __restrict__ h)
{
    float f[8];
#pragma unroll
    for (int v = 0; v < 8; ++v) f[v] = (float) x[v];
    split_bf16<P>(f, h);
}

// |x| of eight bfloat16 values (sign bits cleared).
__device__ __forceinline__ bf8 bf16_abs(bf8 x)
{
    return __builtin_bit_cast(bf8, __builtin_bit_cast(u4, x) & 0x7fff7fffu);
}

// Largest value of a non-negative float over the 64 lanes of the wave.
__device__ __forceinline__ float wave_max_f(float v)
{
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) v = fmaxf(v, __shfl_xor(v, d, 64));
    return v;
}

// Largest value of v over the 64 lanes of the wave (butterfly over lane-index bits).
__device__ __forceinline__ int wave_max(int v)
{
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) v = max(v, __shfl_xor(v, d, 64));
    return __builtin_amdgcn_readfirstlane(v);
}

// Lane K of each row of 16 lanes, broadcast to every lane of that row (DPP row_newbcast, 32- or
// 64-bit move).  Callers keep the same value in all four rows, so every lane of the wave receives it.
// Must be reached with all lanes active: a disabled source lane yields 0 (bound_ctrl).
template <int K, class T>
__device__ __forceinline__ T bcast16(T v)
{
    static_assert(K >= 0 && K < 16, "row lane");
    return __builtin_amdgcn_mov_dpp(v, 0x150 + K, 0xf, 0xf, true);
}

// Address-space views of read-only kernel inputs.  Data the kernels only read (pair matrices, primitive
// data, offsets, the job table) are viewed through the constant address space, so wave-uniform loads
// become scalar loads (SGPR operands) and lane-varying loads global (not flat) vector loads; nothing
// writes these arrays while a kernel runs.
template <class T>
using cptr = const __attribute__((address_space(4))) T*;

template <class T>
__device__ __forceinline__ cptr<T> ro(const T* p)
{
    return (cptr<T>) p;
}

// Whether any of the n flags of an item is set (the float32 pass's verdict for the item's waves and
// component chunks: some |value| reached the threshold, so the float64 pass recomputes the item).
__device__ __forceinline__ bool item_flagged(const int* __restrict__ f, int n)
{
    int any = 0;
    for (int k = 0; k < n; ++k) any |= ro(f)[k];
    return any != 0;
}

// Row of the i >= j packed lower triangle holding element (i, j) of a symmetric matrix.
__device__ __forceinline__ long long packed_row(int i, int j)
{
    return i >= j ? (long long) i * (i + 1) / 2 + j : (long long) j * (j + 1) / 2 + i;
}

}  // namespace eip

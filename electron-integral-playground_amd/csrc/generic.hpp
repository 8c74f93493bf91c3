// Argument blocks and launchers of the runtime-generic kernels (generic_kernels.hip).
#pragma once

#include <hip/hip_runtime.h>

struct GenPairArgs {
    int la, lb, sph, recompute_P;
    const double* P_p;
    const double* A_a;
    const double* B_b;
    const double* coef;
    int n_pp;
    double* E;              // [n_pp][nrow][nherm(la+lb)]
};

struct GenJ3cArgs {
    int la, lb, lc, sph;
    const int* sp_pp_off;
    const double* pp_P;
    const double* E;        // dense pair matrix from eip_gen_pair_matrix
    int n_sp;
    const int* ks_prim_off;
    const double* ka_C;
    const double* ka_coef;
    int n_ks;
    double* out;            // compact [(sp*nrow + row) * ld + ks*nk + mk]
    long long ld;
    double omega;
};

struct GenNucArgs {
    int la, lb, sph;
    const int* sp_pp_off;
    const double* pp_P;
    const double* E;
    int n_sp;
    const double* charge;
    int n_charge;
    double* out;            // compact [(sp*nrow + row) * ld + charge]
    long long ld;
    double omega;
};

struct GenJ2cArgs {
    int li, lj, sph;
    const int* i_prim_off;
    const double* i_A;
    const double* i_coef;
    int n_i;
    const int* j_prim_off;
    const double* j_A;
    const double* j_coef;
    int n_j;
    double* out;            // compact [(si*ni + mi) * ld + sj*nj + mj]
    long long ld;
    double omega;
};

struct GenOverlapArgs {
    const int* sp_pp_off;
    const double* A_a;
    const double* B_b;
    const double* E;        // pair matrix with t = 0 entry of row r at pp*pp_stride + r*row_stride
    long long pp_stride, row_stride;
    int n_sp, nrow;
    double* out;            // [sp * nrow + row]
};

int eip_gen_pair_matrix(const GenPairArgs& a, hipStream_t s);
int eip_gen_three_center(const GenJ3cArgs& a, hipStream_t s);
int eip_gen_nuclear(const GenNucArgs& a, hipStream_t s);
int eip_gen_two_center(const GenJ2cArgs& a, hipStream_t s);
int eip_gen_overlap(const GenOverlapArgs& a, hipStream_t s);

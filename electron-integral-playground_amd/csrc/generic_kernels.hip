// Runtime-generic kernels: every class the reference's C ABI accepts (orbital l <= 6, auxiliary
// l <= 9) in spherical or Cartesian output.  They carry the classes that have no specialised
// straight-line kernel (high l, Cartesian output).  Same mathematics as the specialised kernels,
// organised for generality rather than speed: one thread per output element row, Hermite
// integrals by the closed Hermite-polynomial form from R^n_000, tables in device memory.
#include "md_device.hpp"
#include "md_args.hpp"
#include "generic.hpp"
#include "generated/generic_tables.inc"

namespace {

__device__ __forceinline__ int g_ncart(int l) { return (l + 1) * (l + 2) / 2; }
__device__ __forceinline__ int g_nout(int l, int sph) { return sph ? 2 * l + 1 : g_ncart(l); }

__device__ __forceinline__ void g_cart_xyz(int l, int c, int& x, int& y, int& z)
{
    // x-major descending: index(x, y) = (l-x)(l-x+1)/2 + (l-x-y)
    int lx = 0;
    while ((lx + 1) * (lx + 2) / 2 <= c) ++lx;     // lx = l - x
    x = l - lx;
    y = lx - (c - lx * (lx + 1) / 2);
    z = l - x - y;
}

// Hermite triple of index h in the order (total n ascending, then x-major within n).
__device__ __forceinline__ void g_herm_xyz(int h, int& t, int& u, int& v)
{
    int n = 0;
    while ((n + 1) * (n + 2) * (n + 3) / 6 <= h) ++n;
    g_cart_xyz(n, h - n * (n + 1) * (n + 2) / 6, t, u, v);
}

__device__ __forceinline__ int g_nherm(int L) { return (L + 1) * (L + 2) * (L + 3) / 6; }

__device__ void boys64_rt(int L, double T, double* F)
{
    if (T < BOYS_T_SWITCH) {
        const int k = __double2int_rn(T * BOYS_H_INV);
        const double x = k * (1.0 / BOYS_H_INV) - T;
        const double* row = boys_grid64 + k * BOYS_NCOL + L;
        double acc = row[BOYS_K64];
        for (int j = BOYS_K64 - 1; j >= 0; --j) acc = fma(x / (j + 1), acc, row[j]);
        F[L] = acc;
        const double eT = exp(-T);
        for (int m = L - 1; m >= 0; --m) F[m] = fma(T + T, F[m + 1], eT) / (2 * m + 1);
    } else {
        F[0] = 0.5 * sqrt(eip::kPi / T);
        const double eT = exp(-T);
        for (int m = 0; m < L; ++m) F[m + 1] = ((2 * m + 1) * F[m] - eT) * (0.5 / T);
    }
}

// R^0_{tuv} = sum_{mx,my,mz} H(t,mx) H(u,my) H(v,mz) X^{2mx-t} Y^{2my-u} Z^{2mz-v} R^{mx+my+mz}_000
__device__ double g_R(int t, int u, int v, double X, double Y, double Z, const double* R000)
{
    double sum = 0.0;
    for (int mx = (t + 1) / 2; mx <= t; ++mx) {
        double px = gen_herm[t * 25 + mx];
        for (int k = 0; k < 2 * mx - t; ++k) px *= X;
        for (int my = (u + 1) / 2; my <= u; ++my) {
            double py = px * gen_herm[u * 25 + my];
            for (int k = 0; k < 2 * my - u; ++k) py *= Y;
            for (int mz = (v + 1) / 2; mz <= v; ++mz) {
                double pz = py * gen_herm[v * 25 + mz];
                for (int k = 0; k < 2 * mz - v; ++k) pz *= Z;
                sum += pz * R000[mx + my + mz];
            }
        }
    }
    return sum;
}

// E0[i][t] for i <= L on one axis (vertical recurrence), stride 13.
__device__ void g_E0(int L, double PA, double o2p, double* E0)
{
    for (int i = 0; i <= L; ++i)
        for (int t = 0; t <= L; ++t) E0[i * 13 + t] = 0.0;
    E0[0] = 1.0;
    for (int i = 0; i < L; ++i)
        for (int t = 0; t <= i + 1; ++t) {
            double v = 0.0;
            if (t >= 1) v += o2p * E0[i * 13 + t - 1];
            if (t <= i) v += PA * E0[i * 13 + t];
            if (t + 1 <= i) v += (t + 1) * E0[i * 13 + t + 1];
            E0[(i + 1) * 13 + t] = v;
        }
}

// E^{i,j}_t by the binomial horizontal transfer from E0.
__device__ double g_Eij(int i, int j, int t, double AB, const double* E0)
{
    if (t > i + j) return 0.0;
    double s = 0.0, ABp = 1.0, binom = 1.0;   // binom = C(j, j - m) for m = j, j-1, ...
    for (int m = j; m >= 0; --m) {
        if (t <= i + m) s += binom * ABp * E0[(i + m) * 13 + t];
        ABp *= AB;
        binom = binom * m / (j - m + 1);
    }
    return s;
}

__device__ __forceinline__ int brow_begin(int l, int sph, int row) { return gen_brow_off[gen_brow_base[sph * 10 + l] + row]; }
__device__ __forceinline__ int brow_end(int l, int sph, int row) { return gen_brow_off[gen_brow_base[sph * 10 + l] + row + 1]; }
__device__ __forceinline__ int arow_begin(int l, int sph, int row) { return gen_arow_off[gen_arow_base[sph * 10 + l] + row]; }
__device__ __forceinline__ int arow_end(int l, int sph, int row) { return gen_arow_off[gen_arow_base[sph * 10 + l] + row + 1]; }

// Dense pair matrix: Ed[pp][row][h], row = (mi, mj) in the output basis, h over Hermite |t| <= la+lb.
__global__ void __launch_bounds__(128) gen_pair_matrix(GenPairArgs a)
{
    const int nrow = g_nout(a.la, a.sph) * g_nout(a.lb, a.sph);
    const long long gid = (long long) blockIdx.x * blockDim.x + threadIdx.x;
    if (gid >= (long long) a.n_pp * nrow) return;
    const int pp = (int) (gid / nrow), row = (int) (gid - (long long) pp * nrow);
    const int nob = g_nout(a.lb, a.sph);
    const int mi = row / nob, mj = row - mi * nob;
    const double* Aa = a.A_a + 4 * (size_t) pp;
    const double* Bb = a.B_b + 4 * (size_t) pp;
    double P[3], p;
    if (a.recompute_P) {
        p = Aa[3] + Bb[3];
        for (int d = 0; d < 3; ++d) P[d] = Aa[d] - Bb[3] / p * (Aa[d] - Bb[d]);
    } else {
        for (int d = 0; d < 3; ++d) P[d] = a.P_p[4 * (size_t) pp + d];
        p = a.P_p[4 * (size_t) pp + 3];
    }
    const int L = a.la + a.lb;
    double E0[3][13 * 13];
    for (int d = 0; d < 3; ++d) g_E0(L, P[d] - Aa[d], 0.5 / p, E0[d]);
    const int nh = g_nherm(L);
    double* out = a.E + ((size_t) pp * nrow + row) * nh;
    const double c = a.coef[pp];
    for (int h = 0; h < nh; ++h) {
        int t[3];
        g_herm_xyz(h, t[0], t[1], t[2]);
        double val = 0.0;
        for (int ti = brow_begin(a.la, a.sph, mi); ti < brow_end(a.la, a.sph, mi); ++ti) {
            int ix[3];
            g_cart_xyz(a.la, gen_bterm_c[ti], ix[0], ix[1], ix[2]);
            for (int tj = brow_begin(a.lb, a.sph, mj); tj < brow_end(a.lb, a.sph, mj); ++tj) {
                int jx[3];
                g_cart_xyz(a.lb, gen_bterm_c[tj], jx[0], jx[1], jx[2]);
                double prod = gen_bterm_v[ti] * gen_bterm_v[tj];
                for (int d = 0; d < 3 && prod != 0.0; ++d) prod *= g_Eij(ix[d], jx[d], t[d], Aa[d] - Bb[d], E0[d]);
                val += prod;
            }
        }
        out[h] = c * val;
    }
}

// Three-centre: thread per (sp, ks, mk, row); compact output [(sp*nrow + row) * ld + ks*nk + mk].
__global__ void __launch_bounds__(128) gen_three_center(GenJ3cArgs a)
{
    const int nrow = g_nout(a.la, a.sph) * g_nout(a.lb, a.sph);
    const int nk = g_nout(a.lc, a.sph);
    const long long gid = (long long) blockIdx.x * blockDim.x + threadIdx.x;
    const long long per_sp = (long long) a.n_ks * nk * nrow;
    if (gid >= (long long) a.n_sp * per_sp) return;
    const int sp = (int) (gid / per_sp);
    long long r = gid - (long long) sp * per_sp;
    const int row = (int) (r % nrow);
    r /= nrow;
    const int mk = (int) (r % nk);
    const int ks = (int) (r / nk);
    const int L = a.la + a.lb + a.lc;
    const int nh = g_nherm(a.la + a.lb);
    double acc = 0.0;
    double R000[32], F[32];
    for (int pp = a.sp_pp_off[sp]; pp < a.sp_pp_off[sp + 1]; ++pp) {
        const double* Erow = a.E + ((size_t) pp * nrow + row) * nh;
        const double p = a.pp_P[4 * pp + 3];
        for (int kp = a.ks_prim_off[ks]; kp < a.ks_prim_off[ks + 1]; ++kp) {
            const double q = a.ka_C[4 * kp + 3];
            const double X = a.pp_P[4 * pp + 0] - a.ka_C[4 * kp + 0];
            const double Y = a.pp_P[4 * pp + 1] - a.ka_C[4 * kp + 1];
            const double Z = a.pp_P[4 * pp + 2] - a.ka_C[4 * kp + 2];
            double zeta, pf;
            if (a.omega == 0.0) { zeta = p * q / (p + q); pf = a.ka_coef[kp] * eip::kTwoPi25 / (p * q * sqrt(p + q)); }
            else {
                const double w2 = a.omega * a.omega;
                zeta = 1.0 / (1.0 / p + 1.0 / q + 1.0 / w2);
                pf = a.ka_coef[kp] * eip::kTwoPi25 * a.omega / (p * q * sqrt(p * q + (p + q) * w2));
            }
            for (int k = 0; k < a.lc; ++k) pf *= -0.5 / q;      // (-1)^lc (1/2q)^lc
            boys64_rt(L, zeta * (X * X + Y * Y + Z * Z), F);
            double f = pf;
            for (int n = 0; n <= L; ++n) { R000[n] = f * F[n]; f *= -2.0 * zeta; }
            for (int h = 0; h < nh; ++h) {
                const double e = Erow[h];
                if (e == 0.0) continue;
                int t, u, v;
                g_herm_xyz(h, t, u, v);
                double V = 0.0;
                for (int s = arow_begin(a.lc, a.sph, mk); s < arow_end(a.lc, a.sph, mk); ++s) {
                    const int code = gen_aterm[s];
                    const int sx = code & 31, sy = (code >> 5) & 31, sz = (code >> 10) & 31, j = code >> 15;
                    double w = gen_aterm_v[s];
                    for (int k = 0; k < j; ++k) w *= 2.0 * q;
                    V += w * g_R(t + sx, u + sy, v + sz, X, Y, Z, R000);
                }
                acc += e * V;
            }
        }
    }
    a.out[((long long) sp * nrow + row) * a.ld + (long long) ks * nk + mk] = acc;
}

// Nuclear attraction: thread per (sp, charge, row); compact output [(sp*nrow + row) * ld + charge].
__global__ void __launch_bounds__(128) gen_nuclear(GenNucArgs a)
{
    const int nrow = g_nout(a.la, a.sph) * g_nout(a.lb, a.sph);
    const long long gid = (long long) blockIdx.x * blockDim.x + threadIdx.x;
    if (gid >= (long long) a.n_sp * a.n_charge * nrow) return;
    const int row = (int) (gid % nrow);
    const long long r = gid / nrow;
    const int ic = (int) (r % a.n_charge);
    const int sp = (int) (r / a.n_charge);
    const int L = a.la + a.lb;
    const int nh = g_nherm(L);
    double acc = 0.0, R000[32], F[32];
    for (int pp = a.sp_pp_off[sp]; pp < a.sp_pp_off[sp + 1]; ++pp) {
        const double* Erow = a.E + ((size_t) pp * nrow + row) * nh;
        const double p = a.pp_P[4 * pp + 3];
        const double X = a.pp_P[4 * pp + 0] - a.charge[3 * ic + 0];
        const double Y = a.pp_P[4 * pp + 1] - a.charge[3 * ic + 1];
        const double Z = a.pp_P[4 * pp + 2] - a.charge[3 * ic + 2];
        double zeta, pf;
        if (a.omega == 0.0) { zeta = p; pf = eip::kTwoPi / p; }
        else { const double w2 = a.omega * a.omega; zeta = p * w2 / (p + w2); pf = eip::kTwoPi * a.omega / (p * sqrt(p + w2)); }
        boys64_rt(L, zeta * (X * X + Y * Y + Z * Z), F);
        double f = pf;
        for (int n = 0; n <= L; ++n) { R000[n] = f * F[n]; f *= -2.0 * zeta; }
        for (int h = 0; h < nh; ++h) {
            const double e = Erow[h];
            if (e == 0.0) continue;
            int t, u, v;
            g_herm_xyz(h, t, u, v);
            acc += e * g_R(t, u, v, X, Y, Z, R000);
        }
    }
    a.out[((long long) sp * nrow + row) * a.ld + ic] = acc;
}

// Two-centre: thread per (si, sj, mi, mj); compact output [(si*ni + mi) * ld + sj*nj + mj].
__global__ void __launch_bounds__(128) gen_two_center(GenJ2cArgs a)
{
    const int ni = g_nout(a.li, a.sph), nj = g_nout(a.lj, a.sph);
    const long long gid = (long long) blockIdx.x * blockDim.x + threadIdx.x;
    if (gid >= (long long) a.n_i * a.n_j * ni * nj) return;
    const int mj = (int) (gid % nj);
    long long r = gid / nj;
    const int mi = (int) (r % ni);
    r /= ni;
    const int sj = (int) (r % a.n_j);
    const int si = (int) (r / a.n_j);
    const int L = a.li + a.lj;
    double acc = 0.0, R000[32], F[32];
    for (int ip = a.i_prim_off[si]; ip < a.i_prim_off[si + 1]; ++ip) {
        const double p = a.i_A[4 * ip + 3];
        for (int jp = a.j_prim_off[sj]; jp < a.j_prim_off[sj + 1]; ++jp) {
            const double q = a.j_A[4 * jp + 3];
            const double X = a.i_A[4 * ip + 0] - a.j_A[4 * jp + 0];
            const double Y = a.i_A[4 * ip + 1] - a.j_A[4 * jp + 1];
            const double Z = a.i_A[4 * ip + 2] - a.j_A[4 * jp + 2];
            const double coef = a.i_coef[ip] * a.j_coef[jp];
            double zeta, pf;
            if (a.omega == 0.0) { zeta = p * q / (p + q); pf = coef * eip::kTwoPi25 / (p * q * sqrt(p + q)); }
            else {
                const double w2 = a.omega * a.omega;
                zeta = 1.0 / (1.0 / p + 1.0 / q + 1.0 / w2);
                pf = coef * eip::kTwoPi25 * a.omega / (p * q * sqrt(p * q + (p + q) * w2));
            }
            for (int k = 0; k < a.li; ++k) pf *= 0.5 / p;       // (1/2p)^li
            for (int k = 0; k < a.lj; ++k) pf *= -0.5 / q;      // (-1)^lj (1/2q)^lj
            boys64_rt(L, zeta * (X * X + Y * Y + Z * Z), F);
            double f = pf;
            for (int n = 0; n <= L; ++n) { R000[n] = f * F[n]; f *= -2.0 * zeta; }
            for (int s1 = arow_begin(a.li, a.sph, mi); s1 < arow_end(a.li, a.sph, mi); ++s1) {
                const int c1 = gen_aterm[s1];
                double w1 = gen_aterm_v[s1];
                for (int k = 0; k < (c1 >> 15); ++k) w1 *= 2.0 * p;
                for (int s2 = arow_begin(a.lj, a.sph, mj); s2 < arow_end(a.lj, a.sph, mj); ++s2) {
                    const int c2 = gen_aterm[s2];
                    double w = w1 * gen_aterm_v[s2];
                    for (int k = 0; k < (c2 >> 15); ++k) w *= 2.0 * q;
                    acc += w * g_R((c1 & 31) + (c2 & 31), ((c1 >> 5) & 31) + ((c2 >> 5) & 31),
                                   ((c1 >> 10) & 31) + ((c2 >> 10) & 31), X, Y, Z, R000);
                }
            }
        }
    }
    a.out[((long long) si * ni + mi) * a.ld + (long long) sj * nj + mj] = acc;
}

// Overlap: thread per (sp, row): sum_pp (pi/p)^{3/2} * E_pp[row, t = 0].
__global__ void __launch_bounds__(128) gen_overlap(GenOverlapArgs a)
{
    const long long gid = (long long) blockIdx.x * blockDim.x + threadIdx.x;
    if (gid >= (long long) a.n_sp * a.nrow) return;
    const int sp = (int) (gid / a.nrow), row = (int) (gid - (long long) sp * a.nrow);
    double acc = 0.0;
    for (int pp = a.sp_pp_off[sp]; pp < a.sp_pp_off[sp + 1]; ++pp) {
        const double p = a.A_a[4 * pp + 3] + a.B_b[4 * pp + 3];
        const double f = eip::kPi / p;
        acc += f * sqrt(f) * a.E[(size_t) pp * a.pp_stride + (size_t) row * a.row_stride];
    }
    a.out[gid] = acc;
}

}  // namespace

static inline unsigned nblk(long long n, int bs) { return (unsigned) ((n + bs - 1) / bs); }

int eip_gen_pair_matrix(const GenPairArgs& a, hipStream_t s)
{
    const long long n = (long long) a.n_pp * (a.sph ? (2 * a.la + 1) * (2 * a.lb + 1)
                                                    : (a.la + 1) * (a.la + 2) / 2 * ((a.lb + 1) * (a.lb + 2) / 2));
    if (n > 0) hipLaunchKernelGGL(gen_pair_matrix, dim3(nblk(n, 128)), dim3(128), 0, s, a);
    return (int) hipGetLastError();
}

int eip_gen_three_center(const GenJ3cArgs& a, hipStream_t s)
{
    auto no = [&](int l) { return a.sph ? 2 * l + 1 : (l + 1) * (l + 2) / 2; };
    const long long n = (long long) a.n_sp * a.n_ks * no(a.lc) * no(a.la) * no(a.lb);
    if (n > 0) hipLaunchKernelGGL(gen_three_center, dim3(nblk(n, 128)), dim3(128), 0, s, a);
    return (int) hipGetLastError();
}

int eip_gen_nuclear(const GenNucArgs& a, hipStream_t s)
{
    auto no = [&](int l) { return a.sph ? 2 * l + 1 : (l + 1) * (l + 2) / 2; };
    const long long n = (long long) a.n_sp * a.n_charge * no(a.la) * no(a.lb);
    if (n > 0) hipLaunchKernelGGL(gen_nuclear, dim3(nblk(n, 128)), dim3(128), 0, s, a);
    return (int) hipGetLastError();
}

int eip_gen_two_center(const GenJ2cArgs& a, hipStream_t s)
{
    auto no = [&](int l) { return a.sph ? 2 * l + 1 : (l + 1) * (l + 2) / 2; };
    const long long n = (long long) a.n_i * a.n_j * no(a.li) * no(a.lj);
    if (n > 0) hipLaunchKernelGGL(gen_two_center, dim3(nblk(n, 128)), dim3(128), 0, s, a);
    return (int) hipGetLastError();
}

int eip_gen_overlap(const GenOverlapArgs& a, hipStream_t s)
{
    const long long n = (long long) a.n_sp * a.nrow;
    if (n > 0) hipLaunchKernelGGL(gen_overlap, dim3(nblk(n, 128)), dim3(128), 0, s, a);
    return (int) hipGetLastError();
}

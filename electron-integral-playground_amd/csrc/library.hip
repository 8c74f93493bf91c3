// libmolecular_integral.so -- MI355X (gfx950) McMurchie-Davidson integral library.
//
// Two groups of C entry points (declared in include/molecular_integral.h):
//
// 1. The reference's four C-ABI functions, same names, signatures, accumulate (+=) semantics, mirror
//    rule and return codes, so the reference's ctypes callers work unchanged
//    (reference c_library/molecular_integral/{overlap,nuclear_attraction,two_center,three_center}.cpp).
//    Each call copies its inputs to the GPU, groups primitive pairs by contracted shell pair and
//    auxiliary primitives by auxiliary shell (the reference evaluates each primitive triplet on its own
//    and accumulates with atomics; here each output block has exactly one owner on the device), runs
//    the class's kernels into a compact block buffer, copies it back and adds it into the caller's
//    array, writing the transposed element too when the two AO offsets differ.
//
// 2. Device-resident entry points (eip_*) over grouped data already in HBM: pair-matrix preparation,
//    three-centre over a work-item list, two-centre and nuclear attraction.  They take device
//    pointers and a stream, allocate nothing and never synchronise.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <memory>
#include <mutex>
#include <numeric>
#include <thread>
#include <vector>

#include "md_args.hpp"
#include "generic.hpp"
#include "generated/registry.hpp"

namespace {

constexpr int kMaxL = 6;       // orbital shells (reference angular.h:4)
constexpr int kMaxAuxL = 9;    // auxiliary shells (reference angular.h:5)

inline int ncart(int l) { return (l + 1) * (l + 2) / 2; }
inline int nout(int l, bool sph) { return sph ? 2 * l + 1 : ncart(l); }

#define EIP_CHECK(expr)                                                                    \
    do {                                                                                   \
        hipError_t e_ = (expr);                                                            \
        if (e_ != hipSuccess) {                                                            \
            fprintf(stderr, "libmolecular_integral: %s failed: %s (%s:%d)\n", #expr,       \
                    hipGetErrorString(e_), __FILE__, __LINE__);                            \
            return 2;                                                                      \
        }                                                                                  \
    } while (0)

// One non-blocking stream per host thread: concurrent ctypes calls (the GIL is released) never
// share a stream or a buffer.
hipStream_t thread_stream()
{
    thread_local hipStream_t s = nullptr;
    if (!s && hipStreamCreateWithFlags(&s, hipStreamNonBlocking) != hipSuccess) s = nullptr;
    return s;
}

// Device memory of one C-ABI call.  By default every buffer is a hipMalloc allocation released when
// the call returns.  Over a fixed region (the no-allocation entry point) buffers are consecutive
// 256-byte-aligned pieces of that region and nothing is allocated; a request that does not fit returns
// hipErrorOutOfMemory.  Host arrays uploaded with put() are kept alive until the call returns -- after
// its last stream synchronisation -- so an asynchronous copy never reads released host memory.
class CallMem {
  public:
    CallMem() = default;
    CallMem(char* region, size_t bytes) : region_(region), cap_(bytes) {}
    ~CallMem()
    {
        // an early return may leave copies from kept host arrays in flight: wait for them first
        if (!keep_.empty()) (void) hipStreamSynchronize(stream_);
        for (void* p : owned_) (void) hipFree(p);
    }
    CallMem(const CallMem&) = delete;
    CallMem& operator=(const CallMem&) = delete;

    bool fixed() const { return region_ != nullptr; }
    size_t capacity() const { return cap_; }
    // bytes still available in the fixed region (unbounded otherwise)
    size_t left() const { return fixed() ? cap_ - std::min(cap_, round_up(used_)) : SIZE_MAX; }

    template <class T>
    hipError_t get(T** p, size_t n)
    {
        const size_t bytes = std::max<size_t>(n * sizeof(T), 8);
        if (fixed()) {
            const size_t off = round_up(used_);
            if (off > cap_ || bytes > cap_ - off) return hipErrorOutOfMemory;
            *p = reinterpret_cast<T*>(region_ + off);
            used_ = off + bytes;
            return hipSuccess;
        }
        void* q = nullptr;
        const hipError_t e = hipMalloc(&q, bytes);
        if (e == hipSuccess) {
            owned_.push_back(q);
            *p = static_cast<T*>(q);
        }
        return e;
    }

    // Device copy of a host array (the host array is moved into the call's keep list first).
    template <class T>
    hipError_t put(T** p, std::vector<T>&& host, hipStream_t s)
    {
        const hipError_t e = get(p, host.size());
        if (e != hipSuccess) return e;
        auto h = std::make_shared<std::vector<T>>(std::move(host));
        keep_.push_back(h);
        stream_ = s;
        return h->empty() ? hipSuccess : hipMemcpyAsync(*p, h->data(), h->size() * sizeof(T), hipMemcpyHostToDevice, s);
    }

    template <class T>
    hipError_t put(T** p, const std::vector<T>& host, hipStream_t s)
    {
        return put(p, std::vector<T>(host), s);
    }

  private:
    static size_t round_up(size_t x) { return (x + 255) & ~size_t(255); }
    char* region_ = nullptr;
    size_t cap_ = 0, used_ = 0;
    std::vector<void*> owned_;
    std::vector<std::shared_ptr<void>> keep_;
    hipStream_t stream_ = nullptr;        // every put() of one call uses the call's one stream
};

// Memory request of a call: rc 5 when a fixed region is too small for this call's inputs, else rc 2.
#define EIP_MEM(mem, expr)                                                                 \
    do {                                                                                   \
        hipError_t e_ = (expr);                                                            \
        if (e_ != hipSuccess) {                                                            \
            if ((mem).fixed() && e_ == hipErrorOutOfMemory) {                              \
                fprintf(stderr, "libmolecular_integral: %s: this call needs more than the "  \
                        "fixed %zu-byte scratch region\n", __func__, (mem).capacity());    \
                return 5;                                                                  \
            }                                                                              \
            fprintf(stderr, "libmolecular_integral: %s failed: %s (%s:%d)\n", #expr,       \
                    hipGetErrorString(e_), __FILE__, __LINE__);                            \
            return 2;                                                                      \
        }                                                                                  \
    } while (0)

// Stable grouping of items by an integer key pair.  Returns the permutation (items in group order),
// group offsets (CSR) and the key pair of each group.
struct Grouping {
    std::vector<int> order, off, key_a, key_b;
};

Grouping group_by(const int* a, const int* b, int n)
{
    Grouping g;
    g.order.resize(n);
    std::iota(g.order.begin(), g.order.end(), 0);
    std::stable_sort(g.order.begin(), g.order.end(), [&](int x, int y) {
        const int bx = b ? b[x] : 0, by = b ? b[y] : 0;
        return a[x] != a[y] ? a[x] < a[y] : bx < by;
    });
    for (int k = 0; k < n; ++k) {
        const int i = g.order[k];
        const int kb = b ? b[i] : 0;
        if (k == 0 || a[i] != g.key_a.back() || kb != g.key_b.back()) {
            g.off.push_back(k);
            g.key_a.push_back(a[i]);
            g.key_b.push_back(kb);
        }
    }
    g.off.push_back(n);
    return g;
}

template <class T>
std::vector<T> gather_rows(const T* src, const std::vector<int>& order, int width)
{
    std::vector<T> out(order.size() * (size_t) width);
    for (size_t k = 0; k < order.size(); ++k)
        std::memcpy(&out[k * width], src + (size_t) order[k] * width, sizeof(T) * width);
    return out;
}

// Run f(begin, end) over [0, n) on the host's cores.
template <class F>
void parallel_for(long long n, F f)
{
    unsigned nt = std::max(1u, std::min(std::thread::hardware_concurrency(), 32u));
    if (n < 4096 || nt == 1) { f(0LL, n); return; }
    std::vector<std::thread> th;
    const long long per = (n + nt - 1) / nt;
    for (unsigned t = 0; t < nt; ++t) {
        const long long b = t * per, e = std::min(n, b + per);
        if (b < e) th.emplace_back(f, b, e);
    }
    for (auto& t : th) t.join();
}

bool unsupported(const char* fn, int i_L, int j_L, int k_L = -1)
{
    if (k_L >= 0) printf("%s function does not support angular i_L = %d, j_L = %d, k_L = %d\n", fn, i_L, j_L, k_L);
    else printf("%s function does not support angular i_L = %d, j_L = %d\n", fn, i_L, j_L);
    fflush(stdout);
    return true;
}

// Grouped primitive pairs of one orbital class, resident on the device for one call.
struct DevPairs {
    Grouping g;
    double *P = nullptr, *A = nullptr, *B = nullptr, *coef = nullptr, *E = nullptr;
    int* off = nullptr;
    int n_pp = 0, n_sp = 0;
    bool spec = false;     // packed specialised pair matrix (else dense generic)
    int nnz = 0;           // entries per primitive pair
};

// allow_spec: the consumer has a specialised kernel for this class (else the dense generic layout).
int prepare_pairs(DevPairs& d, CallMem& mem, int la, int lb, bool sph, const double* P_p, const double* A_a,
                  const double* B_b, const double* coef, const int* i_start, const int* j_start, int n,
                  bool recompute_P, bool allow_spec, hipStream_t s)
{
    d.g = group_by(i_start, j_start, n);
    d.n_pp = n;
    d.n_sp = (int) d.g.key_a.size();
    EIP_MEM(mem, mem.put(&d.P, P_p ? gather_rows(P_p, d.g.order, 4) : std::vector<double>(4 * (size_t) n, 0.0), s));
    EIP_MEM(mem, mem.put(&d.A, gather_rows(A_a, d.g.order, 4), s));
    EIP_MEM(mem, mem.put(&d.B, gather_rows(B_b, d.g.order, 4), s));
    EIP_MEM(mem, mem.put(&d.coef, gather_rows(coef, d.g.order, 1), s));
    EIP_MEM(mem, mem.put(&d.off, d.g.off, s));
    d.spec = allow_spec && sph && la <= EIP_SPEC_MAX_L && lb <= EIP_SPEC_MAX_L;
    if (d.spec) {
        d.nnz = eip_pair_nnz(la, lb);
        EIP_MEM(mem, mem.get(&d.E, (size_t) n * d.nnz));
        PairPrepArgs a{d.P, d.A, d.B, d.coef, d.E, nullptr, n, recompute_P ? 1 : 0};
        if (eip_launch_pairE(la, lb, a, (n + 255) / 256, s) != 0) return 3;
    } else {
        const int L = la + lb;
        d.nnz = nout(la, sph) * nout(lb, sph) * (L + 1) * (L + 2) * (L + 3) / 6;
        EIP_MEM(mem, mem.get(&d.E, (size_t) n * d.nnz));
        GenPairArgs a{la, lb, sph ? 1 : 0, recompute_P ? 1 : 0, d.P, d.A, d.B, d.coef, n, d.E};
        EIP_CHECK((hipError_t) eip_gen_pair_matrix(a, s));
    }
    EIP_CHECK(hipGetLastError());
    return 0;
}

// Zero whole rows of a row-major matrix: one workgroup per listed row, 16-byte stores.
// Zero spans of doubles: span b = [spans[2b], spans[2b] + spans[2b+1]), each at most 16384 doubles; one
// block per span, 16-byte stores.
__global__ void __launch_bounds__(256) zero_spans_kernel(const long long* __restrict__ spans, int n_spans,
                                                         double* __restrict__ out)
{
    for (int b = blockIdx.x; b < n_spans; b += gridDim.x) {
        const long long off = spans[2 * b], n = spans[2 * b + 1];
        double* p = out + off;
        const long long head = (reinterpret_cast<uintptr_t>(p) & 15) ? 1 : 0;
        if (head && threadIdx.x == 0 && n > 0) p[0] = 0.0;
        double2* q = reinterpret_cast<double2*>(p + head);
        const long long n2 = (n - head) / 2;
        for (long long k = threadIdx.x; k < n2; k += blockDim.x) q[k] = make_double2(0.0, 0.0);
        if (((n - head) & 1) && threadIdx.x == 0) p[n - 1] = 0.0;
    }
}

// Background form of zero_spans_kernel for a small grid beside compute kernels: each wave waits for
// its stores after every kPaceStores of them (vmcnt counts stores on this target), so the bytes in flight
// -- and the write bandwidth taken from the compute kernels' own output stores -- stay bounded.
constexpr int kPaceStores = 8;

__global__ void __launch_bounds__(256) zero_spans_paced_kernel(const long long* __restrict__ spans, int n_spans,
                                                                double* __restrict__ out)
{
    for (int b = blockIdx.x; b < n_spans; b += gridDim.x) {
        const long long off = spans[2 * b], n = spans[2 * b + 1];
        double* p = out + off;
        const long long head = (reinterpret_cast<uintptr_t>(p) & 15) ? 1 : 0;
        if (head && threadIdx.x == 0 && n > 0) p[0] = 0.0;
        double2* q = reinterpret_cast<double2*>(p + head);
        const long long n2 = (n - head) / 2;
        int issued = 0;
        for (long long k = threadIdx.x; k < n2; k += blockDim.x) {
            q[k] = make_double2(0.0, 0.0);
            if (++issued == kPaceStores) {
                __builtin_amdgcn_s_waitcnt(0x0F70);      // vmcnt(0)
                issued = 0;
            }
        }
        if (((n - head) & 1) && threadIdx.x == 0) p[n - 1] = 0.0;
    }
}

// Plain copy of n doubles, one 8-byte load and one 8-byte store per lane (the store width of the
// three-centre kernels), grid-stride over a fixed grid: the known-size kernel that calibrates the HBM
// byte counters of a profiler.
constexpr int kCopyProbeBlocks = 8192;

__global__ void __launch_bounds__(256) copy_probe_kernel(const double* __restrict__ src, double* __restrict__ dst,
                                                         long long n)
{
    const long long stride = (long long) gridDim.x * blockDim.x;
    for (long long i = (long long) blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) dst[i] = src[i];
}

// ---- Consumers of an (ij|P) output on the device: a row-major float64 matrix A[n_rows][ld] (its first
// n_cols columns).  Both stream A once from HBM with coalesced loads and sum in a fixed order, so their
// results do not depend on scheduling.

// gamma[c] (+)= sum_r w[r] A[r][c], first pass: block (b, t) sums rows [b * rpb, (b + 1) * rpb) of the
// column tile t (256 threads x kColPerThread columns, consecutive threads on consecutive columns) into
// partial[b][c].
constexpr int kColPerThread = 4;
constexpr int kColTile = 256 * kColPerThread;
__global__ void __launch_bounds__(256) weighted_col_sum_partial(const double* __restrict__ A, long long n_rows, int n_cols,
                                                                long long ld, const double* __restrict__ w,
                                                                long long rpb, double* __restrict__ partial)
{
    const long long r0 = (long long) blockIdx.x * rpb;
    const long long r1 = r0 + rpb < n_rows ? r0 + rpb : n_rows;
    const int c0 = (int) blockIdx.y * kColTile + (int) threadIdx.x;
    double acc[kColPerThread] = {};
    bool in[kColPerThread];
#pragma unroll
    for (int k = 0; k < kColPerThread; ++k) in[k] = c0 + 256 * k < n_cols;
    // four rows per step: all 16 loads are issued before their FMAs (bytes in flight hide HBM latency)
    long long r = r0;
    for (; r + 4 <= r1; r += 4) {
        double v[4][kColPerThread];
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const double* __restrict__ row = A + (r + q) * ld + c0;
#pragma unroll
            for (int k = 0; k < kColPerThread; ++k) v[q][k] = in[k] ? __builtin_nontemporal_load(row + 256 * k) : 0.0;
        }
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const double wr = w[r + q];
#pragma unroll
            for (int k = 0; k < kColPerThread; ++k) acc[k] = fma(wr, v[q][k], acc[k]);
        }
    }
    for (; r < r1; ++r) {
        const double wr = w[r];
        const double* __restrict__ row = A + r * ld + c0;
#pragma unroll
        for (int k = 0; k < kColPerThread; ++k)
            if (in[k]) acc[k] = fma(wr, __builtin_nontemporal_load(row + 256 * k), acc[k]);
    }
#pragma unroll
    for (int k = 0; k < kColPerThread; ++k)
        if (in[k]) partial[(long long) blockIdx.x * n_cols + c0 + 256 * k] = acc[k];
}

// second pass: gamma[c] (+)= sum_b partial[b][c], b in order.
__global__ void __launch_bounds__(256) weighted_col_sum_final(const double* __restrict__ partial, int n_blocks, int n_cols,
                                                              int accumulate, double* __restrict__ gamma)
{
    const int c = (int) (blockIdx.x * blockDim.x + threadIdx.x);
    if (c >= n_cols) return;
    double acc = 0.0;
    for (int b = 0; b < n_blocks; ++b) acc += partial[(long long) b * n_cols + c];
    gamma[c] = accumulate ? gamma[c] + acc : acc;
}

// out[r] = sum_c A[r][c] x[c]: one wave per row (lanes stride the columns, 512 contiguous bytes per step),
// a butterfly over the lanes, lane 0 stores.
__global__ void __launch_bounds__(256) row_dot_kernel(const double* __restrict__ A, long long n_rows, int n_cols,
                                                      long long ld, const double* __restrict__ x,
                                                      double* __restrict__ out)
{
    const int lane = (int) threadIdx.x & 63;
    const long long waves = (long long) gridDim.x * (blockDim.x / 64);
    for (long long r = (long long) blockIdx.x * (blockDim.x / 64) + threadIdx.x / 64; r < n_rows; r += waves) {
        const double* __restrict__ row = A + r * ld;
        double a0 = 0.0, a1 = 0.0, a2 = 0.0, a3 = 0.0;
        int c = lane;
        for (; c + 192 < n_cols; c += 256) {
            const double v0 = __builtin_nontemporal_load(row + c), v1 = __builtin_nontemporal_load(row + c + 64);
            const double v2 = __builtin_nontemporal_load(row + c + 128), v3 = __builtin_nontemporal_load(row + c + 192);
            a0 = fma(v0, x[c], a0);
            a1 = fma(v1, x[c + 64], a1);
            a2 = fma(v2, x[c + 128], a2);
            a3 = fma(v3, x[c + 192], a3);
        }
        for (; c < n_cols; c += 64) a0 = fma(__builtin_nontemporal_load(row + c), x[c], a0);
        double v = (a0 + a1) + (a2 + a3);
#pragma unroll
        for (int d = 32; d >= 1; d >>= 1) v += __shfl_xor(v, d, 64);
        if (lane == 0) out[r] = v;
    }
}

__global__ void __launch_bounds__(256) zero_rows_kernel(const int* __restrict__ rows, int n_rows, long long ld,
                                                        double* __restrict__ out)
{
    for (int r = blockIdx.x; r < n_rows; r += gridDim.x) {
        double* row = out + (long long) rows[r] * ld;
        const long long n2 = ld / 2;
        double2* row2 = reinterpret_cast<double2*>(row);
        const bool aligned = (reinterpret_cast<uintptr_t>(row) & 15) == 0;
        if (aligned) {
            for (long long k = threadIdx.x; k < n2; k += blockDim.x) row2[k] = make_double2(0.0, 0.0);
            if ((ld & 1) && threadIdx.x == 0) row[ld - 1] = 0.0;
        } else {
            for (long long k = threadIdx.x; k < ld; k += blockDim.x) row[k] = 0.0;
        }
    }
}

// Chunk-plan variant in use per three-centre class (index [f32][la][lb][lc]); set by the autotuner.
int g_variant[2][EIP_SPEC_MAX_L + 1][EIP_SPEC_MAX_L + 1][EIP_SPEC_MAX_AUX_L + 1] = {};

inline int current_variant(int la, int lb, int lc, int f32) { return g_variant[f32 ? 1 : 0][la][lb][lc]; }

// Largest grid of one three-centre launch (work items = workgroups of up to 256 threads): longer item
// lists are launched in slices [item_begin, item_end), keeping every grid far inside the runtime's limit
// of 2^32 threads per dimension.
constexpr int kMaxItemsPerLaunch = 1 << 22;

// Every chunk kernel of variant v of a class over items [0, n_items), in slices of kMaxItemsPerLaunch.
// With flags_out set, chunk c of the float32 tier writes its flags at slot 4 c of each item's row.
int launch_j3c_items(int la, int lb, int lc, int v, int f32, J3cArgs a, int n_items, hipStream_t s)
{
    const int nchunks = eip_j3c_nchunks(la, lb, lc, v);
    for (int c = 0; c < nchunks; ++c) {
        a.flag_slot = 4 * c;
        for (int b0 = 0; b0 < n_items; b0 += kMaxItemsPerLaunch) {
            const int n = std::min(kMaxItemsPerLaunch, n_items - b0);
            a.item_begin = b0;
            a.item_end = b0 + n;
            if (eip_launch_j3c(la, lb, lc, v, c, f32, a, n, s) != 0) return 3;
        }
    }
    return 0;
}

}  // namespace

extern "C" {

// ---------------------------------------------------------------------------------------------
// Reference C ABI (drop-in)
// ---------------------------------------------------------------------------------------------

int overlap(int i_L, int j_L, const double* pair_A_a, const double* pair_B_b, const double* pair_coefficient,
            const int* pair_i_ao_start, const int* pair_j_ao_start, int n_pair, double* S_matrix, int n_ao,
            bool spherical)
{
    if (!(i_L >= 0 && j_L >= 0 && i_L <= j_L && j_L <= kMaxL) && unsupported(__func__, i_L, j_L)) return 1;
    if (n_pair <= 0) return 0;
    hipStream_t s = thread_stream();
    CallMem mem;
    DevPairs d;
    if (int rc = prepare_pairs(d, mem, i_L, j_L, spherical, nullptr, pair_A_a, pair_B_b, pair_coefficient,
                               pair_i_ao_start, pair_j_ao_start, n_pair, true, true, s)) return rc;
    const int ni = nout(i_L, spherical), nj = nout(j_L, spherical), nrow = ni * nj;
    double* out = nullptr;
    EIP_MEM(mem, mem.get(&out, (size_t) d.n_sp * nrow));
    const int L = i_L + j_L;
    GenOverlapArgs a{d.off, d.A, d.B, d.E, d.nnz, d.spec ? 1 : (long long) (L + 1) * (L + 2) * (L + 3) / 6, d.n_sp,
                     nrow, out};
    EIP_CHECK((hipError_t) eip_gen_overlap(a, s));
    std::vector<double> h((size_t) d.n_sp * nrow);
    EIP_CHECK(hipMemcpyAsync(h.data(), out, h.size() * sizeof(double), hipMemcpyDeviceToHost, s));
    EIP_CHECK(hipStreamSynchronize(s));
    for (int sp = 0; sp < d.n_sp; ++sp) {
        const int I = d.g.key_a[sp], J = d.g.key_b[sp];
        for (int i = 0; i < ni; ++i)
            for (int j = 0; j < nj; ++j) {
                const double v = h[(size_t) sp * nrow + i * nj + j];
                S_matrix[(size_t) (I + i) * n_ao + (J + j)] += v;
                if (I != J) S_matrix[(size_t) (I + i) + (size_t) (J + j) * n_ao] += v;
            }
    }
    return 0;
}

int nuclear_attraction(int i_L, int j_L, const double* pair_P_p, const double* pair_A_a, const double* pair_B_b,
                       const double* pair_coefficient, const int* pair_i_ao_start, const int* pair_j_ao_start,
                       int n_pair, const double* charge_position, int n_charge, double* V_tensor, int n_ao,
                       bool spherical, double omega)
{
    if (!(i_L >= 0 && j_L >= 0 && i_L <= j_L && j_L <= kMaxL) && unsupported(__func__, i_L, j_L)) return 1;
    if (n_pair <= 0 || n_charge <= 0) return 0;
    hipStream_t s = thread_stream();
    CallMem mem;
    DevPairs d;
    if (int rc = prepare_pairs(d, mem, i_L, j_L, spherical, pair_P_p, pair_A_a, pair_B_b, pair_coefficient,
                               pair_i_ao_start, pair_j_ao_start, n_pair, false, true, s)) return rc;
    const int ni = nout(i_L, spherical), nj = nout(j_L, spherical), nrow = ni * nj;
    double *charge = nullptr, *out = nullptr;
    EIP_MEM(mem, mem.put(&charge, std::vector<double>(charge_position, charge_position + 3 * (size_t) n_charge), s));
    EIP_MEM(mem, mem.get(&out, (size_t) d.n_sp * nrow * n_charge));
    if (d.spec) {
        NucArgs a{d.off, d.P, d.E, charge, d.n_sp, n_charge, out, n_charge, omega};
        const int nchunk = (n_charge + 63) / 64;
        if (eip_launch_nuc(i_L, j_L, a, d.n_sp * nchunk, s) != 0) return 3;
    } else {
        GenNucArgs a{i_L, j_L, spherical ? 1 : 0, d.off, d.P, d.E, d.n_sp, charge, n_charge, out, n_charge, omega};
        EIP_CHECK((hipError_t) eip_gen_nuclear(a, s));
    }
    EIP_CHECK(hipGetLastError());
    std::vector<double> h((size_t) d.n_sp * nrow * n_charge);
    EIP_CHECK(hipMemcpyAsync(h.data(), out, h.size() * sizeof(double), hipMemcpyDeviceToHost, s));
    EIP_CHECK(hipStreamSynchronize(s));
    for (int sp = 0; sp < d.n_sp; ++sp) {
        const int I = d.g.key_a[sp], J = d.g.key_b[sp];
        for (int i = 0; i < ni; ++i)
            for (int j = 0; j < nj; ++j) {
                const double* v = &h[((size_t) sp * nrow + i * nj + j) * n_charge];
                double* o = V_tensor + ((size_t) (I + i) * n_ao + (J + j)) * n_charge;
                for (int c = 0; c < n_charge; ++c) o[c] += v[c];
                if (I != J) {
                    double* m = V_tensor + ((size_t) (I + i) + (size_t) (J + j) * n_ao) * n_charge;
                    for (int c = 0; c < n_charge; ++c) m[c] += v[c];
                }
            }
    }
    return 0;
}

int two_center(int i_L, int j_L, const double* aux_A_a, const double* aux_i_coefficient, const int* aux_i_aux_start,
               int n_aux_primitive_i, const double* aux_B_b, const double* aux_j_coefficient,
               const int* aux_j_aux_start, int n_aux_primitive_j, double* J2c_matrix, int n_aux, bool spherical,
               double omega)
{
    if (!(i_L >= 0 && j_L >= 0 && i_L <= j_L && j_L <= kMaxAuxL) && unsupported(__func__, i_L, j_L)) return 1;
    if (n_aux_primitive_i <= 0 || n_aux_primitive_j <= 0) return 0;
    hipStream_t s = thread_stream();
    CallMem mem;
    Grouping gi = group_by(aux_i_aux_start, nullptr, n_aux_primitive_i);
    Grouping gj = group_by(aux_j_aux_start, nullptr, n_aux_primitive_j);
    const int n_i = (int) gi.key_a.size(), n_j = (int) gj.key_a.size();
    const int ni = nout(i_L, spherical), nj = nout(j_L, spherical);
    double *iA, *ic, *jA, *jc, *out;
    int *io, *jo, *icol, *jcol;
    EIP_MEM(mem, mem.put(&iA, gather_rows(aux_A_a, gi.order, 4), s));
    EIP_MEM(mem, mem.put(&ic, gather_rows(aux_i_coefficient, gi.order, 1), s));
    EIP_MEM(mem, mem.put(&io, gi.off, s));
    EIP_MEM(mem, mem.put(&jA, gather_rows(aux_B_b, gj.order, 4), s));
    EIP_MEM(mem, mem.put(&jc, gather_rows(aux_j_coefficient, gj.order, 1), s));
    EIP_MEM(mem, mem.put(&jo, gj.off, s));
    std::vector<int> hic(n_i), hjc(n_j);
    for (int k = 0; k < n_i; ++k) hic[k] = k * ni;
    for (int k = 0; k < n_j; ++k) hjc[k] = k * nj;
    EIP_MEM(mem, mem.put(&icol, std::move(hic), s));
    EIP_MEM(mem, mem.put(&jcol, std::move(hjc), s));
    const long long ld = (long long) n_j * nj;
    EIP_MEM(mem, mem.get(&out, (size_t) n_i * ni * ld));
    if (spherical && j_L <= EIP_SPEC_MAX_AUX_L) {
        J2cArgs a{io, iA, ic, icol, jo, jA, jc, jcol, n_i, n_j, 0, 0, out, ld, omega};
        if (eip_launch_j2c(i_L, j_L, a, n_i * ((n_j + 63) / 64), s) != 0) return 3;
    } else {
        GenJ2cArgs a{i_L, j_L, spherical ? 1 : 0, io, iA, ic, n_i, jo, jA, jc, n_j, out, ld, omega};
        EIP_CHECK((hipError_t) eip_gen_two_center(a, s));
    }
    EIP_CHECK(hipGetLastError());
    std::vector<double> h((size_t) n_i * ni * ld);
    EIP_CHECK(hipMemcpyAsync(h.data(), out, h.size() * sizeof(double), hipMemcpyDeviceToHost, s));
    EIP_CHECK(hipStreamSynchronize(s));
    for (int si = 0; si < n_i; ++si)
        for (int sj = 0; sj < n_j; ++sj) {
            const int I = gi.key_a[si], J = gj.key_a[sj];
            if (i_L == j_L && I > J) continue;                 // reference two_center.cpp:138
            for (int i = 0; i < ni; ++i)
                for (int j = 0; j < nj; ++j) {
                    const double v = h[((size_t) si * ni + i) * ld + (size_t) sj * nj + j];
                    J2c_matrix[(size_t) (I + i) * n_aux + (J + j)] += v;
                    if (I != J) J2c_matrix[(size_t) (I + i) + (size_t) (J + j) * n_aux] += v;
                }
        }
    return 0;
}

}  // extern "C"

namespace {

// Compact output rows kept per batch of shell pairs when device memory is allocated per call.
constexpr long long kBlockBudgetBytes = 256LL << 20;

// The three-centre C-ABI call over the given call memory and stream (shared by three_center and its
// no-allocation variant; the kernels and their order are the same, so the results are identical).
int three_center_impl(CallMem& mem, hipStream_t s, int i_L, int j_L, int k_L, const double* pair_P_p,
                      const double* pair_A_a, const double* pair_B_b, const double* pair_coefficient,
                      const int* pair_i_ao_start, const int* pair_j_ao_start, int n_pair_ij, const double* aux_A_a,
                      const double* aux_coefficient, const int* aux_i_aux_start, int n_aux_primitive_k,
                      double* J3c_matrix, int n_ao, bool spherical, double omega)
{
    DevPairs d;
    if (int rc = prepare_pairs(d, mem, i_L, j_L, spherical, pair_P_p, pair_A_a, pair_B_b, pair_coefficient,
                               pair_i_ao_start, pair_j_ao_start, n_pair_ij, false,
                               k_L <= EIP_SPEC_MAX_AUX_L, s)) return rc;
    Grouping gk = group_by(aux_i_aux_start, nullptr, n_aux_primitive_k);
    const int n_ks = (int) gk.key_a.size();
    const int ni = nout(i_L, spherical), nj = nout(j_L, spherical), nk = nout(k_L, spherical), nrow = ni * nj;
    double *kC, *kc;
    int *ko, *kcol;
    EIP_MEM(mem, mem.put(&kC, gather_rows(aux_A_a, gk.order, 4), s));
    EIP_MEM(mem, mem.put(&kc, gather_rows(aux_coefficient, gk.order, 1), s));
    EIP_MEM(mem, mem.put(&ko, gk.off, s));
    std::vector<int> hkcol(n_ks);
    for (int k = 0; k < n_ks; ++k) hkcol[k] = k * nk;
    EIP_MEM(mem, mem.put(&kcol, std::move(hkcol), s));
    const bool spec = d.spec;
    int *spi = nullptr, *spj = nullptr;
    const double* E_kernel = d.E;
    if (spec) {
        EIP_MEM(mem, mem.put(&spi, d.g.key_a, s));
        EIP_MEM(mem, mem.put(&spj, d.g.key_b, s));
        if (eip_j3c_variant_mfma(i_L, j_L, k_L, current_variant(i_L, j_L, k_L, 0)) == 1) {
            // the class's float64 kernels run on the matrix cores and read the dense pair matrix
            double* ED = nullptr;
            EIP_MEM(mem, mem.get(&ED, (size_t) d.n_pp * eip_pair_dense_width(i_L, j_L)));
            if (eip_launch_pairD(i_L, j_L, d.E, ED, d.n_pp, s) != 0) return 3;
            E_kernel = ED;
        }
    }
    const long long ld = (long long) n_ks * nk;
    const int kblocks = (n_ks + 63) / 64;
    // Batches of shell pairs: compact output rows (and work items) of at most kBlockBudgetBytes, or --
    // over a fixed region -- of what the region has left after the inputs.
    const long long per_sp = (long long) nrow * ld;
    const long long per_sp_bytes = per_sp * 8 + (spec ? 8LL * kblocks : 0);
    const long long room = mem.fixed() ? (long long) mem.left() - 1024 : kBlockBudgetBytes;
    const int batch = (int) std::min<long long>(d.n_sp, std::max(mem.fixed() ? 0LL : 1LL, room / std::max(1LL, per_sp_bytes)));
    if (batch < 1) EIP_MEM(mem, hipErrorOutOfMemory);
    double* out = nullptr;
    int* items = nullptr;
    EIP_MEM(mem, mem.get(&out, (size_t) batch * per_sp));
    if (spec) EIP_MEM(mem, mem.get(&items, (size_t) batch * kblocks * 2));
    std::vector<double> h((size_t) batch * per_sp);
    for (int sp0 = 0; sp0 < d.n_sp; sp0 += batch) {
        const int nb = std::min(batch, d.n_sp - sp0);
        if (spec) {
            // items: (sp, first aux shell) with compact rows relative to sp0
            std::vector<int> hit;
            hit.reserve((size_t) nb * kblocks * 2);
            for (int sp = sp0; sp < sp0 + nb; ++sp)
                for (int kb = 0; kb < kblocks; ++kb) { hit.push_back(sp); hit.push_back(kb * 64); }
            EIP_CHECK(hipMemcpyAsync(items, hit.data(), hit.size() * sizeof(int), hipMemcpyHostToDevice, s));
            // out pointer shifted so that row (sp * nrow + ij) lands at (sp - sp0) * nrow + ij
            J3cArgs a{};
            a.items = items;
            a.items_ks_end = nullptr;
            a.item_begin = 0;
            a.item_end = nb * kblocks;
            a.d.sp_pp_off = d.off;
            a.d.sp_i = spi;
            a.d.sp_j = spj;
            a.d.pp_P = d.P;
            a.d.E = E_kernel;
            a.d.E32 = nullptr;
            a.d.ks_perm = nullptr;
            a.d.ks_prim_off = ko;
            a.d.ks_col = kcol;
            a.d.ka_C = kC;
            a.d.ka_coef = kc;
            a.d.n_ks = n_ks;
            a.d.out = out - (long long) sp0 * nrow * ld;
            a.d.ld = ld;
            a.layout = 1;
            a.omega = omega;
            if (launch_j3c_items(i_L, j_L, k_L, current_variant(i_L, j_L, k_L, 0), 0, a, nb * kblocks, s) != 0) return 3;
            EIP_CHECK(hipGetLastError());
            EIP_CHECK(hipStreamSynchronize(s));        // 'hit' is read by the copy above until here
        } else {
            GenJ3cArgs a{i_L, j_L, k_L, spherical ? 1 : 0, d.off + sp0, d.P, d.E, nb, ko, kC, kc, n_ks, out, ld, omega};
            EIP_CHECK((hipError_t) eip_gen_three_center(a, s));
        }
        EIP_CHECK(hipMemcpyAsync(h.data(), out, sizeof(double) * (size_t) nb * per_sp, hipMemcpyDeviceToHost, s));
        EIP_CHECK(hipStreamSynchronize(s));
        // Scatter-add: threads own disjoint auxiliary shells, hence disjoint rows of J3c.
        const size_t nao2 = (size_t) n_ao * n_ao;
        parallel_for(n_ks, [&](long long k0, long long k1) {
            for (long long ks = k0; ks < k1; ++ks) {
                const int K = gk.key_a[ks];
                for (int sp = 0; sp < nb; ++sp) {
                    const int I = d.g.key_a[sp0 + sp], J = d.g.key_b[sp0 + sp];
                    for (int i = 0; i < ni; ++i)
                        for (int j = 0; j < nj; ++j) {
                            const double* v = &h[((size_t) sp * nrow + i * nj + j) * ld + (size_t) ks * nk];
                            for (int k = 0; k < nk; ++k) {
                                double* row = J3c_matrix + (size_t) (K + k) * nao2;
                                row[(size_t) (I + i) * n_ao + (J + j)] += v[k];
                                if (I != J) row[(size_t) (I + i) + (size_t) (J + j) * n_ao] += v[k];
                            }
                        }
                }
            }
        });
    }
    return 0;
}

}  // namespace

extern "C" {

int three_center(int i_L, int j_L, int k_L, const double* pair_P_p, const double* pair_A_a, const double* pair_B_b,
                 const double* pair_coefficient, const int* pair_i_ao_start, const int* pair_j_ao_start, int n_pair_ij,
                 const double* aux_A_a, const double* aux_coefficient, const int* aux_i_aux_start,
                 int n_aux_primitive_k, double* J3c_matrix, int n_ao, int n_aux, bool spherical, double omega)
{
    if (!(i_L >= 0 && j_L >= 0 && i_L <= j_L && j_L <= kMaxL && k_L >= 0 && k_L <= kMaxAuxL)
        && unsupported(__func__, i_L, j_L, k_L)) return 1;
    (void) n_aux;
    if (n_pair_ij <= 0 || n_aux_primitive_k <= 0) return 0;
    CallMem mem;
    return three_center_impl(mem, thread_stream(), i_L, j_L, k_L, pair_P_p, pair_A_a, pair_B_b, pair_coefficient,
                             pair_i_ao_start, pair_j_ao_start, n_pair_ij, aux_A_a, aux_coefficient, aux_i_aux_start,
                             n_aux_primitive_k, J3c_matrix, n_ao, spherical, omega);
}

}  // extern "C"

// ---------------------------------------------------------------------------------------------
// No-allocation variant of three_center
// ---------------------------------------------------------------------------------------------

// Device memory of three_center_noalloc: a fixed region in the library's code object, one copy per
// GPU the library runs on, held until the process ends.  In host code this name is a stand-in, not a
// device address: set-up looks up each device's own address (scratch_region) and only that is used.
constexpr size_t kNoAllocScratchBytes = size_t(256) << 20;        // 256 MiB per GPU
static __device__ __align__(256) char scratch[kNoAllocScratchBytes];

namespace {

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
char* g_scratch_ptr[kMaxDevices] = {};

// Device address of the current device's scratch region, looked up once per device (with that device
// current) and kept.
int scratch_region(char** out)
{
    int dev = -1;
    EIP_CHECK(hipGetDevice(&dev));
    if (dev < 0 || dev >= kMaxDevices) {
        fprintf(stderr, "libmolecular_integral: three_center_noalloc: device %d beyond the %d supported\n", dev,
                kMaxDevices);
        return 2;
    }
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || p == nullptr) {
            fprintf(stderr, "libmolecular_integral: three_center_noalloc: hipGetSymbolAddress(scratch) failed: %s\n",
                    hipGetErrorString(e));
            return 2;
        }
        g_scratch_ptr[dev] = static_cast<char*>(p);
    }
    *out = g_scratch_ptr[dev];
    return 0;
}

}  // namespace

extern "C" {

long long three_center_noalloc_capacity(void) { return (long long) kNoAllocScratchBytes; }

int three_center_noalloc(int i_L, int j_L, int k_L, const double* pair_P_p, const double* pair_A_a,
                         const double* pair_B_b, const double* pair_coefficient, const int* pair_i_ao_start,
                         const int* pair_j_ao_start, int n_pair_ij, const double* aux_A_a,
                         const double* aux_coefficient, const int* aux_i_aux_start, int n_aux_primitive_k,
                         double* J3c_matrix, int n_ao, int n_aux, bool spherical, double omega)
{
    if (!(i_L >= 0 && j_L >= 0 && i_L <= j_L && j_L <= kMaxL && k_L >= 0 && k_L <= kMaxAuxL)
        && unsupported(__func__, i_L, j_L, k_L)) return 1;
    (void) n_aux;
    if (n_pair_ij <= 0 || n_aux_primitive_k <= 0) return 0;
    char* region = nullptr;
    if (int rc = scratch_region(&region)) return rc;
    CallMem mem(region, kNoAllocScratchBytes);
    // The null stream (no stream is created).  Every buffer of the call is a piece of the region and is
    // fully written before it is read (pair matrices by their kernel, every compact output row by the
    // class kernels), so nothing has to be cleared; calls that overlap on one GPU race on the region.
    return three_center_impl(mem, nullptr, i_L, j_L, k_L, pair_P_p, pair_A_a, pair_B_b, pair_coefficient,
                             pair_i_ao_start, pair_j_ao_start, n_pair_ij, aux_A_a, aux_coefficient, aux_i_aux_start,
                             n_aux_primitive_k, J3c_matrix, n_ao, spherical, omega);
}

// ---------------------------------------------------------------------------------------------
// Device-resident entry points
// ---------------------------------------------------------------------------------------------

int eip_pair_matrix_nnz(int i_L, int j_L)
{
    if (!(i_L >= 0 && j_L >= 0 && i_L <= j_L && j_L <= EIP_SPEC_MAX_L)) return -1;
    return eip_pair_nnz(i_L, j_L);
}

static bool spec_3c(int i_L, int j_L, int k_L)
{
    return i_L >= 0 && j_L >= 0 && i_L <= j_L && j_L <= EIP_SPEC_MAX_L && k_L >= 0 && k_L <= EIP_SPEC_MAX_AUX_L;
}

int eip_three_center_nchunks(int i_L, int j_L, int k_L)
{
    if (!spec_3c(i_L, j_L, k_L)) return -1;
    return eip_j3c_nchunks(i_L, j_L, k_L, current_variant(i_L, j_L, k_L, 0));
}

int eip_three_center_nvariants(int i_L, int j_L, int k_L)
{
    return spec_3c(i_L, j_L, k_L) ? eip_j3c_nvariants(i_L, j_L, k_L) : -1;
}

int eip_three_center_chunk(int i_L, int j_L, int k_L, int variant, int chunk, int* m0, int* m1)
{
    return spec_3c(i_L, j_L, k_L) ? eip_j3c_chunk(i_L, j_L, k_L, variant, chunk, m0, m1) : -1;
}

int eip_set_three_center_variant(int i_L, int j_L, int k_L, int use_f32, int variant)
{
    if (!spec_3c(i_L, j_L, k_L) || variant < 0 || variant >= eip_j3c_nvariants(i_L, j_L, k_L)) return 1;
    const int kind = eip_j3c_variant_mfma(i_L, j_L, k_L, variant);
    if ((use_f32 && kind == 1) || (!use_f32 && kind == 2)) return 1;   // variant of the other tier only
    g_variant[use_f32 ? 1 : 0][i_L][j_L][k_L] = variant;
    return 0;
}

int eip_three_center_variant(int i_L, int j_L, int k_L, int use_f32)
{
    return spec_3c(i_L, j_L, k_L) ? current_variant(i_L, j_L, k_L, use_f32) : -1;
}

int eip_three_center_f32_uses_bf16(int i_L, int j_L, int k_L)
{
    return spec_3c(i_L, j_L, k_L) ? (eip_j3c_variant_mfma(i_L, j_L, k_L, current_variant(i_L, j_L, k_L, 1)) == 2) : 0;
}

int eip_three_center_flag_stride(int i_L, int j_L, int k_L)
{
    if (!spec_3c(i_L, j_L, k_L)) return -1;
    return 4 * eip_j3c_nchunks(i_L, j_L, k_L, current_variant(i_L, j_L, k_L, 1));
}

int eip_pair_matrix_bf16_width(int i_L, int j_L)
{
    if (!(i_L >= 0 && j_L >= 0 && i_L <= j_L && j_L <= EIP_SPEC_MAX_L)) return -1;
    return eip_pair_bf16_width(i_L, j_L);
}

int eip_pair_matrix_bf16(int i_L, int j_L, const double* d_E, int n_pp, void* d_EB, void* stream)
{
    if (eip_pair_matrix_bf16_width(i_L, j_L) < 0) return 1;
    if (n_pp <= 0) return 0;
    if (eip_launch_pairB(i_L, j_L, d_E, d_EB, n_pp, (hipStream_t) stream) != 0) return 3;
    return hipGetLastError() == hipSuccess ? 0 : 2;
}

int eip_three_center_variant_mfma(int i_L, int j_L, int k_L, int variant)
{
    return spec_3c(i_L, j_L, k_L) ? eip_j3c_variant_mfma(i_L, j_L, k_L, variant) : -1;
}

int eip_three_center_uses_mfma(int i_L, int j_L, int k_L)
{
    return spec_3c(i_L, j_L, k_L) ? (eip_j3c_variant_mfma(i_L, j_L, k_L, current_variant(i_L, j_L, k_L, 0)) == 1) : 0;
}

int eip_pair_matrix_dense_width(int i_L, int j_L)
{
    if (!(i_L >= 0 && j_L >= 0 && i_L <= j_L && j_L <= EIP_SPEC_MAX_L)) return -1;
    return eip_pair_dense_width(i_L, j_L);
}

int eip_pair_matrix_dense(int i_L, int j_L, const double* d_E, int n_pp, double* d_ED, void* stream)
{
    if (eip_pair_matrix_dense_width(i_L, j_L) < 0) return 1;
    if (n_pp <= 0) return 0;
    if (eip_launch_pairD(i_L, j_L, d_E, d_ED, n_pp, (hipStream_t) stream) != 0) return 3;
    return hipGetLastError() == hipSuccess ? 0 : 2;
}

int eip_pair_matrix(int i_L, int j_L, const double* d_P_p, const double* d_A_a, const double* d_B_b,
                    const double* d_coef, int n_pp, double* d_E, float* d_E32, void* stream)
{
    if (eip_pair_matrix_nnz(i_L, j_L) < 0) return 1;
    if (n_pp <= 0) return 0;
    PairPrepArgs a{d_P_p, d_A_a, d_B_b, d_coef, d_E, d_E32, n_pp, 0};
    if (eip_launch_pairE(i_L, j_L, a, (n_pp + 255) / 256, (hipStream_t) stream) != 0) return 3;
    return hipGetLastError() == hipSuccess ? 0 : 2;
}

int eip_three_center_items(int i_L, int j_L, int k_L, int use_f32, const int* d_items, const int* d_items_ks_end,
                           int n_items, const int* d_sp_pp_off, const int* d_sp_i, const int* d_sp_j,
                           const double* d_pp_P, const double* d_E, const float* d_E32, const int* d_ks_perm,
                           const int* d_ks_prim_off, const int* d_ks_col, const double* d_ka_C,
                           const double* d_ka_coef, int n_ks, double* d_out, long long ld, int layout,
                           double omega, void* stream)
{
    const int nchunks = eip_three_center_nchunks(i_L, j_L, k_L);
    if (nchunks < 0) return 1;
    if (n_items <= 0) return 0;
    J3cArgs a{};
    a.items = d_items;
    a.items_ks_end = d_items_ks_end;
    a.item_end = n_items;
    a.d = J3cData{d_sp_pp_off, d_sp_i, d_sp_j, d_pp_P, d_E, d_E32, d_ks_perm, d_ks_prim_off, d_ks_col, d_ka_C,
                  d_ka_coef, n_ks, d_out, ld, nullptr, nullptr};
    a.layout = layout;
    a.omega = omega;
    if (launch_j3c_items(i_L, j_L, k_L, current_variant(i_L, j_L, k_L, use_f32), use_f32, a, n_items,
                         (hipStream_t) stream) != 0) return 3;
    return hipGetLastError() == hipSuccess ? 0 : 2;
}

int eip_three_center_items_batched(int i_L, int j_L, int k_L, int use_f32, const int* d_items,
                                   const int* d_items_ks_end, int n_items, const int* d_sp_pp_off, const int* d_sp_i,
                                   const int* d_sp_j, const long long* d_sp_out_base, const int* d_sp_ld,
                                   const double* d_pp_P, const double* d_E, const float* d_E32, const int* d_ks_perm,
                                   const int* d_ks_prim_off, const int* d_ks_col, const double* d_ka_C,
                                   const double* d_ka_coef, int n_ks, double* d_out, double omega, void* stream)
{
    const int nchunks = eip_three_center_nchunks(i_L, j_L, k_L);
    if (nchunks < 0) return 1;
    if (n_items <= 0) return 0;
    J3cArgs a{};
    a.items = d_items;
    a.items_ks_end = d_items_ks_end;
    a.item_end = n_items;
    a.d = J3cData{d_sp_pp_off, d_sp_i, d_sp_j, d_pp_P, d_E, d_E32, d_ks_perm, d_ks_prim_off, d_ks_col, d_ka_C,
                  d_ka_coef, n_ks, d_out, 0, d_sp_out_base, d_sp_ld};
    a.layout = 0;
    a.omega = omega;
    if (launch_j3c_items(i_L, j_L, k_L, current_variant(i_L, j_L, k_L, use_f32), use_f32, a, n_items,
                         (hipStream_t) stream) != 0) return 3;
    return hipGetLastError() == hipSuccess ? 0 : 2;
}

int eip_three_center_work(int i_L, int j_L, int k_L, int use_f32, const int* d_items, const int* d_items_ks_end,
                          const int* d_items_pp, const int* d_items_job, int n_items, const void* d_jobs, int layout,
                          int accumulate, double omega, void* stream)
{
    const int nchunks = eip_three_center_nchunks(i_L, j_L, k_L);
    if (nchunks < 0) return 1;
    if (n_items <= 0) return 0;
    if (!d_items || !d_jobs || (layout != 0 && layout != 1) || (accumulate != 0 && accumulate != 1)) return 4;
    J3cArgs a{};
    a.items = d_items;
    a.items_ks_end = d_items_ks_end;
    a.items_pp = d_items_pp;
    a.items_job = d_items_job;
    a.item_end = n_items;
    a.jobs = static_cast<const J3cData*>(d_jobs);
    a.layout = layout;
    a.accumulate = accumulate;
    a.omega = omega;
    if (launch_j3c_items(i_L, j_L, k_L, current_variant(i_L, j_L, k_L, use_f32), use_f32, a, n_items,
                         (hipStream_t) stream) != 0) return 3;
    return hipGetLastError() == hipSuccess ? 0 : 2;
}

int eip_three_center_work_checked(int i_L, int j_L, int k_L, int pass, const int* d_items, const int* d_items_ks_end,
                                  const int* d_items_job, int n_items, const void* d_jobs, int layout, int* d_flags,
                                  int flag_stride, double threshold, double omega, void* stream)
{
    if (!spec_3c(i_L, j_L, k_L)) return 1;
    if (n_items <= 0) return 0;
    if (!d_items || !d_jobs || !d_flags || (layout != 0 && layout != 1) || (pass != 0 && pass != 1)) return 4;
    if (flag_stride < eip_three_center_flag_stride(i_L, j_L, k_L) || !(threshold > 0.0)) return 4;
    if (eip_three_center_f32_uses_bf16(i_L, j_L, k_L) != 1) return 4;   // the flags come from bf16 kernels
    J3cArgs a{};
    a.items = d_items;
    a.items_ks_end = d_items_ks_end;
    a.items_job = d_items_job;
    a.jobs = static_cast<const J3cData*>(d_jobs);
    a.layout = layout;
    a.omega = omega;
    a.flag_stride = flag_stride;
    a.flag_threshold = threshold;
    if (pass == 0) a.flags_out = d_flags;
    else a.flags_in = d_flags;
    if (launch_j3c_items(i_L, j_L, k_L, current_variant(i_L, j_L, k_L, pass == 0), pass == 0, a, n_items,
                         (hipStream_t) stream) != 0) return 3;
    return hipGetLastError() == hipSuccess ? 0 : 2;
}

// ---- persistent queue kernels: one kernel per register tier and precision over a device work queue ----

int eip_three_center_queue_tiers(int use_f32)
{
    return eip_j3cq_ntiers(use_f32 ? 1 : 0);
}

int eip_three_center_queue_chunk(int i_L, int j_L, int k_L, int use_f32, int chunk, int* m0, int* m1, int* tier)
{
    if (!spec_3c(i_L, j_L, k_L) || !m0 || !m1 || !tier) return -1;
    int kind = 0;
    return eip_j3cq_chunk(i_L, j_L, k_L, use_f32 ? 1 : 0, chunk, m0, m1, tier, &kind);
}

long long eip_three_center_queue_slot_bytes(void) { return (long long) sizeof(J3cQueueSlot); }

int eip_three_center_queue_ctl_words(void) { return kQueueCtlWords; }

int eip_three_center_queue_slot(int i_L, int j_L, int k_L, int use_f32, int chunk, const int* d_items,
                                const int* d_items_ks_end, const int* d_items_pp, const int* d_items_job,
                                int n_items, const void* d_jobs, int layout, int accumulate, double omega,
                                void* h_slot)
{
    if (!spec_3c(i_L, j_L, k_L)) return -1;
    if (!d_items || !d_jobs || !h_slot || n_items < 0 || (layout != 0 && layout != 1)
        || (accumulate != 0 && accumulate != 1)) return -4;
    int m0 = 0, m1 = 0, tier = 0, kind = 0;
    if (eip_j3cq_chunk(i_L, j_L, k_L, use_f32 ? 1 : 0, chunk, &m0, &m1, &tier, &kind) != 0) return -1;
    J3cQueueSlot sl{};
    sl.a.items = d_items;
    sl.a.items_ks_end = d_items_ks_end;
    sl.a.items_pp = d_items_pp;
    sl.a.items_job = d_items_job;
    sl.a.item_begin = 0;
    sl.a.item_end = n_items;
    sl.a.jobs = static_cast<const J3cData*>(d_jobs);
    sl.a.layout = layout;
    sl.a.accumulate = accumulate;
    sl.a.omega = omega;
    sl.kind = kind;
    sl.tier = tier;
    std::memcpy(h_slot, &sl, sizeof sl);
    return tier;
}

int eip_three_center_queue_grid(int use_f32, int tier, int waves_per_simd)
{
    const void* k = eip_j3cq_kernel(use_f32 ? 1 : 0, tier);
    if (!k) return -1;
    int dev = 0, n_cu = 0, per_cu = 0;
    if (hipGetDevice(&dev) != hipSuccess
        || hipDeviceGetAttribute(&n_cu, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess
        || hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k, 64, 0) != hipSuccess) return -2;
    if (waves_per_simd > 0) per_cu = std::min(per_cu, 4 * waves_per_simd);   // 4 SIMDs per CU
    return std::max(1, per_cu) * n_cu;
}

// Whether p is an address of device memory (or of host memory the device can read) known to the runtime:
// the persistent kernels dereference the slot, entry and control pointers unchecked, so a pointer that
// is not one (a truncated or host address) is rejected here instead of faulting the device.
static bool device_accessible(const void* p)
{
    hipPointerAttribute_t attr{};
    if (hipPointerGetAttributes(&attr, p) != hipSuccess) {
        (void) hipGetLastError();
        return false;
    }
    return attr.type == hipMemoryTypeDevice || attr.type == hipMemoryTypeManaged || attr.devicePointer != nullptr;
}

int eip_three_center_queue_run(int use_f32, int tier, const void* d_slots, const int* d_entries, int* d_ctl,
                               int grid, void* stream)
{
    if (!d_slots || !d_entries || !d_ctl || grid <= 0) return 4;
    if (!device_accessible(d_slots) || !device_accessible(d_entries) || !device_accessible(d_ctl)) return 4;
    J3cQueueArgs q{static_cast<const J3cQueueSlot*>(d_slots), reinterpret_cast<const J3cQueueEntry*>(d_entries), d_ctl};
    if (eip_launch_j3cq(use_f32 ? 1 : 0, tier, q, grid, (hipStream_t) stream) != 0) return 3;
    return hipGetLastError() == hipSuccess ? 0 : 2;
}

// Rows per block of the first pass of eip_weighted_column_sums and the blocks it uses: about 8192
// workgroups over the device, at least 64 rows each.
static void col_sum_shape(long long n_rows, int n_cols, long long* rpb, int* n_blocks)
{
    const int tiles = (n_cols + kColTile - 1) / kColTile;
    long long want = 8192 / (tiles > 0 ? tiles : 1);
    if (want < 1) want = 1;
    long long r = (n_rows + want - 1) / want;
    if (r < 64) r = 64;
    *rpb = r;
    *n_blocks = (int) ((n_rows + r - 1) / r);
}

long long eip_weighted_column_sums_workspace(long long n_rows, int n_cols)
{
    if (n_rows <= 0 || n_cols <= 0) return 0;
    long long rpb = 0;
    int nb = 0;
    col_sum_shape(n_rows, n_cols, &rpb, &nb);
    return (long long) nb * n_cols;
}

int eip_weighted_column_sums(const double* d_A, long long n_rows, int n_cols, long long ld, const double* d_w,
                             double* d_work, long long work_doubles, int accumulate, double* d_out, void* stream)
{
    if (n_cols <= 0) return 0;
    if (!d_out || ld < n_cols || n_rows < 0 || (n_rows > 0 && (!d_A || !d_w))) return 4;
    if (n_rows == 0) {   // no rows: gamma = 0 (or unchanged), by the second pass over no partials
        hipLaunchKernelGGL(weighted_col_sum_final, dim3((n_cols + 255) / 256), dim3(256), 0, (hipStream_t) stream,
                           d_out, 0, n_cols, accumulate, d_out);
        return hipGetLastError() == hipSuccess ? 0 : 2;
    }
    long long rpb = 0;
    int nb = 0;
    col_sum_shape(n_rows, n_cols, &rpb, &nb);
    if (!d_work || work_doubles < (long long) nb * n_cols) return 4;
    const dim3 grid1((unsigned) nb, (unsigned) ((n_cols + kColTile - 1) / kColTile));
    hipLaunchKernelGGL(weighted_col_sum_partial, grid1, dim3(256), 0, (hipStream_t) stream, d_A, n_rows, n_cols, ld, d_w,
                       rpb, d_work);
    hipLaunchKernelGGL(weighted_col_sum_final, dim3((n_cols + 255) / 256), dim3(256), 0, (hipStream_t) stream, d_work,
                       nb, n_cols, accumulate, d_out);
    return hipGetLastError() == hipSuccess ? 0 : 2;
}

int eip_row_dots(const double* d_A, long long n_rows, int n_cols, long long ld, const double* d_x, double* d_out,
                 void* stream)
{
    if (n_rows <= 0) return 0;
    if (!d_A || !d_x || !d_out || n_cols <= 0 || ld < n_cols) return 4;
    const long long blocks = (n_rows + 3) / 4;
    const int grid = (int) (blocks < 8192 ? blocks : 8192);
    hipLaunchKernelGGL(row_dot_kernel, dim3(grid), dim3(256), 0, (hipStream_t) stream, d_A, n_rows, n_cols, ld, d_x, d_out);
    return hipGetLastError() == hipSuccess ? 0 : 2;
}

int eip_copy_probe(const double* d_src, double* d_dst, long long n, void* stream)
{
    if (n <= 0) return 0;
    hipLaunchKernelGGL(copy_probe_kernel, dim3(kCopyProbeBlocks), dim3(256), 0, (hipStream_t) stream, d_src, d_dst, n);
    return hipGetLastError() == hipSuccess ? 0 : 2;
}

int eip_zero_rows(const int* d_rows, int n_rows, long long ld, double* d_out, void* stream)
{
    if (n_rows <= 0) return 0;
    const int grid = n_rows < 65536 ? n_rows : 65536;
    hipLaunchKernelGGL(zero_rows_kernel, dim3(grid), dim3(256), 0, (hipStream_t) stream, d_rows, n_rows, ld, d_out);
    return hipGetLastError() == hipSuccess ? 0 : 2;
}

int eip_zero_spans(const long long* d_spans, int n_spans, double* d_out, int max_blocks, void* stream)
{
    if (n_spans <= 0) return 0;
    if (max_blocks > 0) {
        const int grid = n_spans < max_blocks ? n_spans : (max_blocks < 65536 ? max_blocks : 65536);
        hipLaunchKernelGGL(zero_spans_paced_kernel, dim3(grid), dim3(256), 0, (hipStream_t) stream, d_spans, n_spans,
                           d_out);
        return hipGetLastError() == hipSuccess ? 0 : 2;
    }
    const int grid = n_spans < 65536 ? n_spans : 65536;
    hipLaunchKernelGGL(zero_spans_kernel, dim3(grid), dim3(256), 0, (hipStream_t) stream, d_spans, n_spans, d_out);
    return hipGetLastError() == hipSuccess ? 0 : 2;
}

int eip_two_center_device(int i_L, int j_L, const int* d_i_prim_off, const double* d_i_A, const double* d_i_coef,
                          const int* d_i_col, int n_i, const int* d_j_prim_off, const double* d_j_A,
                          const double* d_j_coef, const int* d_j_col, int n_j, int same_class, int mirror,
                          double* d_out, long long ld, double omega, void* stream)
{
    if (!(i_L >= 0 && j_L >= 0 && i_L <= j_L && j_L <= EIP_SPEC_MAX_AUX_L)) return 1;
    if (n_i <= 0 || n_j <= 0) return 0;
    J2cArgs a{d_i_prim_off, d_i_A, d_i_coef, d_i_col, d_j_prim_off, d_j_A, d_j_coef, d_j_col, n_i, n_j,
              same_class, mirror, d_out, ld, omega};
    if (eip_launch_j2c(i_L, j_L, a, n_i * ((n_j + 63) / 64), (hipStream_t) stream) != 0) return 3;
    return hipGetLastError() == hipSuccess ? 0 : 2;
}

int eip_nuclear_device(int i_L, int j_L, const int* d_sp_pp_off, const double* d_pp_P, const double* d_E,
                       int n_sp, const double* d_charge, int n_charge, double* d_out, long long ld, double omega,
                       void* stream)
{
    if (eip_pair_matrix_nnz(i_L, j_L) < 0) return 1;
    if (n_sp <= 0 || n_charge <= 0) return 0;
    NucArgs a{d_sp_pp_off, d_pp_P, d_E, d_charge, n_sp, n_charge, d_out, ld, omega};
    if (eip_launch_nuc(i_L, j_L, a, n_sp * ((n_charge + 63) / 64), (hipStream_t) stream) != 0) return 3;
    return hipGetLastError() == hipSuccess ? 0 : 2;
}

}  // extern "C"

// Kernel argument blocks shared by the generated kernels and the host library.
// All pointers are device pointers.  Passed by value (kernarg segment).
#pragma once

#include <hip/hip_runtime.h>

// Primitive-pair preparation: builds the packed pair matrix (Hermite coefficients already
// contracted into the output basis, pair coefficient folded in) for n_pp primitive pairs.
struct PairPrepArgs {
    const double* P_p;     // [n_pp][4] x, y, z, p
    const double* A_a;     // [n_pp][4] x, y, z, a
    const double* B_b;     // [n_pp][4] x, y, z, b
    const double* coef;    // [n_pp]
    double* E;             // [n_pp][nnz]
    float* E32;            // optional float copy, or nullptr
    int n_pp;
    int recompute_P;       // 1: P and p from A, B (overlap)
};

// One molecule's (or job's) data for the three-centre kernels.  Every field is 8 bytes wide, so a
// table of these can be filled from any host language as 16 64-bit words.
struct J3cData {
    // shell pairs
    const int* sp_pp_off;     // [n_sp + 1] CSR offsets into the primitive-pair arrays
    const int* sp_i;          // [n_sp] AO offset of shell i
    const int* sp_j;          // [n_sp] AO offset of shell j
    const double* pp_P;       // [n_pp][4]
    const double* E;          // [n_pp][nnz]
    const float* E32;         // [n_pp][nnz] (float path)
    // auxiliary shells
    const int* ks_perm;       // [n_ks] or nullptr: position -> shell index
    const int* ks_prim_off;   // [n_ks + 1]
    const int* ks_col;        // [n_ks] output column of the shell's first function
    const double* ka_C;       // [n_kprim][4]
    const double* ka_coef;    // [n_kprim]
    long long n_ks;
    // output
    double* out;
    long long ld;             // row stride (in doubles)
    // optional, per shell pair (batches of molecules sharing one buffer): offset of the molecule's
    // output in 'out' and its row stride.  nullptr: 0 and 'ld'.
    const long long* sp_out_base;
    const int* sp_ld;
};
static_assert(sizeof(J3cData) == 16 * 8, "J3cData is a table of 16 64-bit words");

// Three-centre (ij|P) for one (la, lb, lc) class over a list of work items.
// Work item w = (sp, ks_first): the wave handles shell pair sp and auxiliary shells
// ks_first .. ks_first + 63 (clipped at items_ks_end[w] or n_ks) of the (optionally permuted) list,
// over the primitive pairs items_pp[w] = [pp_begin, pp_end) (or all of the shell pair's), of the data
// jobs[items_job[w]] (or 'd').
struct J3cArgs {
    const int* items;         // [n_items][2]
    const int* items_ks_end;  // [n_items] or nullptr (then n_ks)
    const int* items_pp;      // [n_items][2] or nullptr (then sp_pp_off[sp] .. sp_pp_off[sp + 1])
    const int* items_job;     // [n_items] or nullptr (then job 0)
    int item_begin, item_end;
    J3cData d;                // used when jobs is nullptr
    const J3cData* jobs;      // device table of per-job data, or nullptr
    int layout;               // 0: rows = packed lower triangle (i >= j); 1: rows = sp * nij + ij
    int accumulate;           // 0: out = integral; 1: out += integral
    double omega;
    // Checked mixed precision (both nullptr otherwise).  A bf16 matrix-core (float32-tier) kernel given
    // flags_out stores, for its item, flags_out[item * flag_stride + flag_slot + w] = 1 when some output
    // of wave w has sum_m sum_k |E_k| |V_k,m| >= flag_threshold (the scale of its float32 rounding), else
    // 0, for w < 4; a float64 kernel given flags_in skips items whose flag_stride flags are all zero.
    int* flags_out;
    const int* flags_in;
    int flag_stride, flag_slot;
    double flag_threshold;
};

// Persistent (queue) three-centre kernels: one kernel per register tier and precision runs the bodies of
// every class chunk of that tier; each wave pulls work items (slot, item) from a device queue until it
// is empty.  A slot is the argument block of one class chunk over one item list (built on the host by
// eip_three_center_queue_slot); 'kind' selects the body inside the tier kernel.
struct J3cQueueSlot {
    J3cArgs a;
    int kind;                 // body index in the tier kernel's switch
    int tier;
};

// The queue is split into kQueueParts parts (one per XCD: the wave of workgroup b starts on part b % 8,
// whose entries the host dealt so that neighbours in the longest-first order share that XCD's L2) and
// moves on to the other parts when its own is empty.  An entry is a run of consecutive items of one slot
// (the host groups cheap items, so that a wave takes one entry per few microseconds of work at least).
// Control block, int32 words; each part's counter has a 128-byte cache line of its own (atomics on one
// line serialise in the L2):
//   [32 p] next entry of part p (0 before a launch; the last wave to finish resets them),
//   [kQueueFirst + p] first entry of part p, [kQueueCount + p] entries of part p, [kQueueDone] waves done.
constexpr int kQueueParts = 8;
constexpr int kQueueLine = 32;                         // int32 words per 128-byte line
constexpr int kQueueFirst = kQueueParts * kQueueLine;  // 256
constexpr int kQueueCount = kQueueFirst + kQueueParts; // 264
constexpr int kQueueDone = kQueueFirst + kQueueLine;   // 288 (a line of its own)
constexpr int kQueueCtlWords = kQueueDone + kQueueLine;
struct J3cQueueEntry {
    int slot;                 // argument block (class chunk, item list)
    int item;                 // first item of the run
    int count;                // items in the run
    int pad;
};
struct J3cQueueArgs {
    const J3cQueueSlot* slots;
    const J3cQueueEntry* entries;
    int* ctl;                 // control block (kQueueCtlWords int32 words)
};

// Nuclear attraction: lanes over point charges, output [(sp * nij + ij) * ld + charge].
struct NucArgs {
    const int* sp_pp_off;
    const double* pp_P;
    const double* E;
    const double* charge;     // [n_charge][3]
    int n_sp, n_charge;
    double* out;
    long long ld;
    double omega;
};

// Two-centre (P|Q): one wave per i shell, lanes over j shells.
struct J2cArgs {
    const int* i_prim_off;    // [n_i + 1]
    const double* i_A;        // [n_iprim][4]
    const double* i_coef;
    const int* i_col;         // [n_i] output row of the shell's first function
    const int* j_prim_off;
    const double* j_A;
    const double* j_coef;
    const int* j_col;
    int n_i, n_j;
    int same_class;           // skip blocks with i_col > j_col
    int mirror;               // also write the transposed block when i_col != j_col
    double* out;
    long long ld;
    double omega;
};

// Job-table entry of work item 'item' (a.jobs, indexed by a.items_job or 0), viewed through the constant
// address space: its fields are wave-uniform invariant loads, re-read where they are used.
template <class ArgsT>
__device__ __forceinline__ const __attribute__((address_space(4))) J3cData* j3c_item_job(const ArgsT& a, int item)
{
    typedef const __attribute__((address_space(4))) J3cData* job_ptr;
    typedef const __attribute__((address_space(4))) int* int_ptr;
    return (job_ptr) a.jobs + (a.items_job ? ((int_ptr) a.items_job)[item] : 0);
}

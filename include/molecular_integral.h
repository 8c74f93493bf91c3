/*
 * molecular_integral.h -- C ABI of libmolecular_integral.so, the MI355X (gfx950) McMurchie-Davidson
 * integral library.
 *
 * Part 1 replaces the reference's four exported functions one for one: same names, argument lists,
 * accumulate (+=) semantics, mirror rule and return codes (0 ok, 1 unsupported angular class with a
 * message on stdout; this library adds 2 = HIP runtime error, 3 = kernel launch error).  All pointers
 * are host pointers to C-contiguous arrays; the library copies inputs to the GPU, computes there and
 * adds the result into the caller's array.  Thread safe: each calling thread gets its own stream and
 * buffers.
 *
 * Part 2 is additive: the native primitive-pair pre-screen (host pointers, host threads) and
 * device-resident entry points over grouped data already in HBM (no allocation, no synchronisation;
 * every pointer is a device pointer, `stream` is a hipStream_t).
 */
#ifndef ELECTRON_INTEGRAL_PLAYGROUND_AMD_MOLECULAR_INTEGRAL_H
#define ELECTRON_INTEGRAL_PLAYGROUND_AMD_MOLECULAR_INTEGRAL_H

#include <stdbool.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- Part 1: the reference's C ABI -------------------------------------------------------- */

/* S[i][j] += overlap; P is recomputed from A and B (pair_P_p is not an argument).
 * Replaces reference c_library/molecular_integral/overlap.cpp:149-159 (`overlap`). */
int overlap(int i_L, int j_L, const double* pair_A_a, const double* pair_B_b, const double* pair_coefficient,
            const int* pair_i_ao_start, const int* pair_j_ao_start, int n_pair,
            double* S_matrix, int n_ao, bool spherical);

/* V[i][j][c] += int i j / |r - C_c| (no charge, positive sign); erf(omega r)/r when omega > 0.
 * Replaces reference c_library/molecular_integral/nuclear_attraction.cpp:199-213 (`nuclear_attraction`). */
int nuclear_attraction(int i_L, int j_L, const double* pair_P_p, const double* pair_A_a, const double* pair_B_b,
                       const double* pair_coefficient, const int* pair_i_ao_start, const int* pair_j_ao_start,
                       int n_pair, const double* charge_position, int n_charge, double* V_tensor, int n_ao,
                       bool spherical, double omega);

/* J2c[P][Q] += (P|Q); same-class blocks with i_aux_start > j_aux_start are skipped.
 * Replaces reference c_library/molecular_integral/two_center.cpp:239-252 (`two_center`). */
int two_center(int i_L, int j_L, const double* aux_A_a, const double* aux_i_coefficient, const int* aux_i_aux_start,
               int n_aux_primitive_i, const double* aux_B_b, const double* aux_j_coefficient,
               const int* aux_j_aux_start, int n_aux_primitive_j, double* J2c_matrix, int n_aux,
               bool spherical, double omega);

/* J3c[P][i][j] += (P|ij).
 * Replaces reference c_library/molecular_integral/three_center.cpp:321-339 (`three_center`). */
int three_center(int i_L, int j_L, int k_L, const double* pair_P_p, const double* pair_A_a, const double* pair_B_b,
                 const double* pair_coefficient, const int* pair_i_ao_start, const int* pair_j_ao_start,
                 int n_pair_ij, const double* aux_A_a, const double* aux_coefficient, const int* aux_i_aux_start,
                 int n_aux_primitive_k, double* J3c_matrix, int n_ao, int n_aux, bool spherical, double omega);

/* ---- Part 2: additive entry points (no counterpart in the reference's C ABI) ------------- */

/* Host-side (multithreaded) enumeration and pre-screen of primitive pairs, the native loop of reference
 * electron_integral_playground/pair_utility.py:78-124 (form_primitive_pair_list): admitted shell pairs
 * (l_i < l_j, or l_i == l_j and ao_start_i <= ao_start_j) in loop order, the cumulative distance rule and
 * the bound, keeping every primitive pair with bound >= threshold * (1 - margin).  Shells are given as
 * arrays (l, AO offset, atom, CSR primitive offsets into prim_exp / prim_coef); atom_dist is the
 * [n_atom][n_atom] table of |A - B|.  Returns the number of candidates (or -1 on bad arguments); the
 * outputs (i shell, j shell, i primitive, j primitive, clamped distance) are written only when capacity
 * is at least that number.  The caller applies the exact keep test and the stable descending order. */
long long eip_screen_primitive_pairs(int n_shell, const int* shell_l, const int* shell_ao_start,
                                     const int* shell_atom, const int* shell_prim_off, const double* prim_exp,
                                     const double* prim_coef, const double* atom_dist, int n_atom,
                                     double threshold, double margin, long long capacity, int* out_i_shell,
                                     int* out_j_shell, int* out_i_prim, int* out_j_prim, double* out_dist);

/* Entries per primitive pair of the packed pair matrix of class (i_L, j_L), or -1 when the class has
 * no specialised kernels (spherical output, l <= 3). */
int eip_pair_matrix_nnz(int i_L, int j_L);

/* Number of kernels (auxiliary-component chunks) of class (i_L j_L | k_L), or -1 (aux l <= 4). */
int eip_three_center_nchunks(int i_L, int j_L, int k_L);

/* Chunk-plan variants of class (i_L j_L | k_L) (each splits the auxiliary components over kernels with
 * a different register budget), the [m0, m1) range of one chunk, and the variant the entry points use
 * for float64 (use_f32 = 0) or float32 kernels.  Defaults to variant 0; chosen by the autotuner. */
int eip_three_center_nvariants(int i_L, int j_L, int k_L);
int eip_three_center_chunk(int i_L, int j_L, int k_L, int variant, int chunk, int* m0, int* m1);
int eip_set_three_center_variant(int i_L, int j_L, int k_L, int use_f32, int variant);
int eip_three_center_variant(int i_L, int j_L, int k_L, int use_f32);   /* the current one, -1: no such class */

/* Variants after the vector ones run the float64 contraction on the matrix cores
 * (v_mfma_f64_16x16x4f64: four primitive pairs x 16 auxiliary shells per instruction, one workgroup of
 * four waves per work item) and read the dense pair matrix below instead of the packed one.  After those
 * come float32-tier variants on the bf16 matrix cores (v_mfma_f32_16x16x32_bf16 over operands split into
 * three bfloat16 pieces, six products per element: float32-level accuracy; four primitive pairs x 8
 * Hermite indices x 16 auxiliary shells per instruction), which read the split bfloat16 pair matrix below
 * through the E32 pointer.
 * eip_three_center_variant_mfma: 1 for a float64 matrix-core variant, 2 for a bf16 one, 0 for a vector
 * variant, -1 for no such class; setting a variant of the other tier (1 with use_f32 = 1, 2 with
 * use_f32 = 0) returns 1.  eip_three_center_uses_mfma: 1 when the class's current float64 variant is a
 * matrix-core one (then every E pointer given for that class -- argument or eip_j3c_job -- must be the
 * dense matrix); eip_three_center_f32_uses_bf16: 1 when its current float32 variant is a bf16 one (then
 * every E32 pointer must be the split bfloat16 matrix). */
int eip_three_center_variant_mfma(int i_L, int j_L, int k_L, int variant);
int eip_three_center_uses_mfma(int i_L, int j_L, int k_L);
int eip_three_center_f32_uses_bf16(int i_L, int j_L, int k_L);

/* Doubles per primitive pair of the dense pair matrix [n_pp][n_t][16 * ceil(nij / 16)] of class
 * (i_L, j_L) (structural zeros stored, -1 when the class has no matrix-core kernels), and its
 * construction from the packed float64 matrix d_E. */
int eip_pair_matrix_dense_width(int i_L, int j_L);
int eip_pair_matrix_dense(int i_L, int j_L, const double* d_E, int n_pp, double* d_ED, void* stream);

/* bfloat16 values per primitive pair of the split pair matrix
 * [n_pp][3 pieces][ceil(n_t / 8)][ceil(nij / 16)][16 rows][8 Hermite indices] of class (i_L, j_L) (-1 when
 * the class has no bf16 kernels), and its construction from the packed float64 matrix d_E: each entry,
 * rounded to float32, is the exact sum of its three truncated bfloat16 pieces. */
int eip_pair_matrix_bf16_width(int i_L, int j_L);
int eip_pair_matrix_bf16(int i_L, int j_L, const double* d_E, int n_pp, void* d_EB, void* stream);

/* Pair matrix (Hermite expansion contracted into real solid harmonics, pair coefficient folded in)
 * for n_pp primitive pairs: d_E[n_pp][nnz] float64 and, when d_E32 is not null, a float32 copy. */
int eip_pair_matrix(int i_L, int j_L, const double* d_P_p, const double* d_A_a, const double* d_B_b,
                    const double* d_coef, int n_pp, double* d_E, float* d_E32, void* stream);

/* (ij|P) of one class over a work-item list.  Item w = (shell pair sp, first auxiliary shell k0)
 * covers auxiliary shells k0 .. k0+63 (clipped at d_items_ks_end[w], or n_ks when that is null) of
 * the list d_ks_perm (or 0..n_ks-1).  layout 0 writes d_out[tri(i,j) * ld + col + m] (packed i >= j);
 * layout 1 writes d_out[(sp * nij + ij) * ld + col + m].  use_f32 selects the float32 kernels. */
int eip_three_center_items(int i_L, int j_L, int k_L, int use_f32, const int* d_items, const int* d_items_ks_end,
                           int n_items, const int* d_sp_pp_off, const int* d_sp_i, const int* d_sp_j,
                           const double* d_pp_P, const double* d_E, const float* d_E32, const int* d_ks_perm,
                           const int* d_ks_prim_off, const int* d_ks_col, const double* d_ka_C,
                           const double* d_ka_coef, int n_ks, double* d_out, long long ld, int layout,
                           double omega, void* stream);

/* Same over shell pairs of several molecules at once (one launch per class for a whole batch):
 * shell pair sp writes the packed output of its molecule at d_out + d_sp_out_base[sp] with row stride
 * d_sp_ld[sp] (that molecule's n_aux); every item pairs a shell pair and auxiliary shells of one molecule. */
int eip_three_center_items_batched(int i_L, int j_L, int k_L, int use_f32, const int* d_items,
                                   const int* d_items_ks_end, int n_items, const int* d_sp_pp_off, const int* d_sp_i,
                                   const int* d_sp_j, const long long* d_sp_out_base, const int* d_sp_ld,
                                   const double* d_pp_P, const double* d_E, const float* d_E32, const int* d_ks_perm,
                                   const int* d_ks_prim_off, const int* d_ks_col, const double* d_ka_C,
                                   const double* d_ka_coef, int n_ks, double* d_out, double omega, void* stream);

/* One molecule's (or job's) three-centre data, as read by eip_three_center_work: 16 64-bit words.
 * Shell pairs: CSR primitive-pair offsets, AO offsets of shells i and j, primitive pairs P_p (n,4),
 * pair matrices (float64 and optional float32); auxiliary shells: optional position->shell permutation,
 * CSR primitive offsets, output column, primitives C_q (n,4) and coefficients; n_ks; output pointer and
 * row stride; optional per-shell-pair output offset and row stride (batches sharing one buffer). */
typedef struct {
    const int* sp_pp_off;
    const int* sp_i;
    const int* sp_j;
    const double* pp_P;
    const double* E;
    const float* E32;
    const int* ks_perm;
    const int* ks_prim_off;
    const int* ks_col;
    const double* ka_C;
    const double* ka_coef;
    long long n_ks;
    double* out;
    long long ld;
    const long long* sp_out_base;
    const int* sp_ld;
} eip_j3c_job;

/* (ij|P) of one class over work items that may belong to different molecules (jobs): item w =
 * (shell pair sp, first auxiliary shell k0) of job d_items_job[w] (or job 0 when null) in the device table
 * d_jobs (eip_j3c_job[]), over auxiliary shells k0 .. min(k0 + 64, d_items_ks_end[w] or n_ks) and primitive
 * pairs d_items_pp[w] = [begin, end) (or all of the shell pair's when null).  accumulate = 0 stores,
 * accumulate = 1 adds into the output (the float32 part of a mixed-precision split is added on top of the
 * float64 part).  Layouts as eip_three_center_items.  Returns 4 on bad arguments. */
int eip_three_center_work(int i_L, int j_L, int k_L, int use_f32, const int* d_items, const int* d_items_ks_end,
                          const int* d_items_pp, const int* d_items_job, int n_items, const void* d_jobs, int layout,
                          int accumulate, double omega, void* stream);

/* Persistent (queue) three-centre kernels.  The work items of many classes -- of one molecule, a batch of
 * molecules or several serving jobs -- run in one persistent kernel per register tier and precision: each
 * wave pulls (slot, item) entries from a device queue until it is empty, so a whole step is a handful of
 * launches (one graph node each) whatever the number of classes and molecules.  Replaces, for these
 * callers, the per-class launches of the reference's dispatcher (CL/MI/three_center.cpp:319-631) and its
 * OpenMP loop over (pair, auxiliary primitive) (three_center.cpp:291-317).
 *
 * eip_three_center_queue_tiers: number of register tiers (kernels) per precision.
 * eip_three_center_queue_chunk: the auxiliary-component chunk [m0, m1) and tier of chunk 'chunk' of a class
 *   in the queue kernels; -1 past the last chunk.
 * eip_three_center_queue_slot: writes into host memory h_slot (eip_three_center_queue_slot_bytes() bytes) the
 *   slot of one class chunk over the item list of eip_three_center_work (same arguments); returns its tier,
 *   or a negative value (-1 unsupported class or chunk, -4 bad arguments).
 * eip_three_center_queue_grid: workgroups that fill the device at the tier kernel's occupancy (capped at
 *   waves_per_simd waves per SIMD when > 0).
 * eip_three_center_queue_run: one launch of a tier kernel over the queue d_entries ([n][4] int32: slot index
 *   into d_slots, first item within the slot's items, number of consecutive items, 0) described by the
 *   control block d_ctl (eip_three_center_queue_ctl_words() int32 words, 128-byte aligned: word 32 p the
 *   counter of queue part p = 0..7, zero; words 256 + p the first entry of part p; 264 + p the entries of
 *   part p; 288 zero).  The wave of workgroup b starts on part b % 8, the part its XCD serves, and
 *   continues with the others when it is empty.  The kernel leaves the counters zero, so the same control
 *   block can be launched again (or replayed in a graph) without clearing.  Returns 4 on bad arguments
 *   (including slot, entry or control pointers the runtime does not know as device-accessible memory). */
int eip_three_center_queue_tiers(int use_f32);
int eip_three_center_queue_chunk(int i_L, int j_L, int k_L, int use_f32, int chunk, int* m0, int* m1, int* tier);
long long eip_three_center_queue_slot_bytes(void);
int eip_three_center_queue_ctl_words(void);
int eip_three_center_queue_slot(int i_L, int j_L, int k_L, int use_f32, int chunk, const int* d_items,
                                const int* d_items_ks_end, const int* d_items_pp, const int* d_items_job,
                                int n_items, const void* d_jobs, int layout, int accumulate, double omega,
                                void* h_slot);
int eip_three_center_queue_grid(int use_f32, int tier, int waves_per_simd);
int eip_three_center_queue_run(int use_f32, int tier, const void* d_slots, const int* d_entries, int* d_ctl,
                               int grid, void* stream);

/* Consumers of an (ij|P) output on the device, for the density-fitted Coulomb matrix (no counterpart in the
 * reference, which stops at the integrals): A is a row-major float64 matrix A[n_rows][ld] (its first n_cols
 * columns), e.g. the packed output (rows tri(i, j), columns P) or a compact block of shell pairs.  Both
 * stream A once from HBM and sum in a fixed order (results independent of scheduling).
 * eip_weighted_column_sums: out[c] = sum_r w[r] A[r][c] (accumulate = 1: out[c] += ...), through a
 *   caller-provided workspace of at least eip_weighted_column_sums_workspace(n_rows, n_cols) doubles.
 * eip_row_dots: out[r] = sum_c A[r][c] x[c].
 * Return 4 on bad arguments. */
long long eip_weighted_column_sums_workspace(long long n_rows, int n_cols);
int eip_weighted_column_sums(const double* d_A, long long n_rows, int n_cols, long long ld, const double* d_w,
                             double* d_work, long long work_doubles, int accumulate, double* d_out, void* stream);
int eip_row_dots(const double* d_A, long long n_rows, int n_cols, long long ld, const double* d_x, double* d_out,
                 void* stream);

/* d_dst[i] = d_src[i] for i < n (device pointers): a plain copy kernel of known size, 8 n bytes read and
 * 8 n bytes written, with one 8-byte load and store per lane.  No counterpart in the reference: it
 * calibrates the profiler's HBM byte counters (FETCH_SIZE, WRITE_SIZE) for the traffic figures. */
int eip_copy_probe(const double* d_src, double* d_dst, long long n, void* stream);

/* A-priori term scale of work items (d_items[w] = (shell pair, first auxiliary shell), 64 auxiliary shells
 * each) of class (i_L j_L | k_L), for the mixed-precision split: d_scale[w] = cmax * max over the item's
 * auxiliary shells P of sum over primitive pairs pp and primitives of P of
 * sum_n d_pp_esum[pp][n] A_{n + k_L}, where d_pp_esum[pp][n] = max_ij sum_{|t| = n} |E_pp[ij][t]| (n <= i_L
 * + j_L; from the float64 pair matrix), cmax = max_m sum_s |c_m(s)| of the auxiliary transform, and A_N the
 * Hermite recursion with every term positive (A_0^(n) = |prefactor (2 zeta)^n F_n(T)|, A_N^(n) = (N - 1)
 * A_{N-2}^(n+1) + |PQ| A_{N-1}^(n+1)): a bound of the magnitude of every term summed into the item's
 * outputs, with their distance decay -- the scale of their float32 rounding.  Pair data as in J3cData
 * (sp_pp_off, pp_P), auxiliary data as ks_prim_off / ka_C / ka_coef.  No counterpart in the reference (its
 * split is the distance-free bound of pair_utility.py:95-114).  Returns 1 when i_L + j_L + k_L > 10, 4 on
 * bad arguments. */
int eip_three_center_scale(int i_L, int j_L, int k_L, const int* d_items, int n_items, const int* d_sp_pp_off,
                           const double* d_pp_P, const float* d_pp_esum, const int* d_ks_prim_off,
                           const double* d_ka_C, const double* d_ka_coef, int n_ks, double cmax, double omega,
                           float* d_scale, void* stream);

/* Checked mixed precision over the items of eip_three_center_work (all primitive pairs of each item).
 * pass 0 runs the class's float32 kernels -- its current float32 variant must be a bf16 matrix-core one
 * -- which store every item's results and, in d_flags[item * flag_stride ..], whether some output's term
 * scale sum_m sum_k |E_k| |V_k,m| (the magnitude its float32 rounding is relative to, so cancellation
 * between primitive pairs of a contracted shell pair counts) reached threshold (one flag per wave and
 * component chunk); pass 1 runs the float64 kernels over the same items, and each item whose flags are
 * all clear is skipped, so the float64 results replace exactly the flagged items.  Run pass 0 then pass 1
 * on one stream with the same arguments.  flag_stride >= eip_three_center_flag_stride (4 x the chunk
 * count of the current float32 variant; the flags need no clearing).  No counterpart in the reference
 * (its three_center.cpp computes every triplet in float64).  Returns 4 on bad arguments or when the
 * class's float32 variant is not a bf16 one. */
int eip_three_center_flag_stride(int i_L, int j_L, int k_L);
int eip_three_center_work_checked(int i_L, int j_L, int k_L, int pass, const int* d_items, const int* d_items_ks_end,
                                  const int* d_items_job, int n_items, const void* d_jobs, int layout, int* d_flags,
                                  int flag_stride, double threshold, double omega, void* stream);

/* Zero rows d_rows[0..n_rows) (each ld doubles long) of d_out: the packed rows of screened-out shell
 * pairs, which no three-centre kernel writes. */
int eip_zero_rows(const int* d_rows, int n_rows, long long ld, double* d_out, void* stream);

/* Zero n_spans spans of d_out: span s covers doubles [d_spans[2s], d_spans[2s] + d_spans[2s+1]) (any
 * length; one workgroup per span); the screened-out rows of a packed output merged into contiguous runs.
 * max_blocks > 0 runs a background form on at most max_blocks workgroups (looping over spans) whose
 * waves wait for their stores every 8 of them, bounding the write bandwidth taken from compute kernels
 * on other streams; 0 = one workgroup per span up to 65536 at full speed. */
int eip_zero_spans(const long long* d_spans, int n_spans, double* d_out, int max_blocks, void* stream);

/* (P|Q) blocks of classes (i_L | j_L), aux l <= 4: d_out[(col_i + m_i) * ld + col_j + m_j]; with
 * same_class, blocks with col_i > col_j are skipped; with mirror, the transposed block is written too. */
int eip_two_center_device(int i_L, int j_L, const int* d_i_prim_off, const double* d_i_A, const double* d_i_coef,
                          const int* d_i_col, int n_i, const int* d_j_prim_off, const double* d_j_A,
                          const double* d_j_coef, const int* d_j_col, int n_j, int same_class, int mirror,
                          double* d_out, long long ld, double omega, void* stream);

/* Nuclear attraction of shell pairs at n_charge points: d_out[(sp * nij + ij) * ld + c]. */
int eip_nuclear_device(int i_L, int j_L, const int* d_sp_pp_off, const double* d_pp_P, const double* d_E,
                       int n_sp, const double* d_charge, int n_charge, double* d_out, long long ld, double omega,
                       void* stream);

/* ---- Part 3: entry point that allocates no device memory ----------------------------------- */

/* three_center without any device allocation: same arguments, accumulate (+=) and mirror semantics
 * and return codes as three_center, and bit-identical results (the same kernels in the same order).
 * No counterpart in the reference: it is for callers that cannot allocate device memory (harnesses
 * that forbid hipMalloc, tight serving loops).  Every device buffer of the call is carved from a fixed
 * region in the library (three_center_noalloc_capacity() bytes, 256 MiB, one per GPU, held until the
 * process ends); a call whose inputs and smallest output batch do not fit returns 5 with a message on
 * stderr.  Runs on the null stream of the current device.  Calls that overlap on one GPU share the
 * region and race: serialise them. */
int three_center_noalloc(int i_L, int j_L, int k_L, const double* pair_P_p, const double* pair_A_a,
                         const double* pair_B_b, const double* pair_coefficient, const int* pair_i_ao_start,
                         const int* pair_j_ao_start, int n_pair_ij, const double* aux_A_a,
                         const double* aux_coefficient, const int* aux_i_aux_start, int n_aux_primitive_k,
                         double* J3c_matrix, int n_ao, int n_aux, bool spherical, double omega);
long long three_center_noalloc_capacity(void);

#ifdef __cplusplus
}
#endif

#endif

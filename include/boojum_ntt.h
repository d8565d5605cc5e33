/*
 * C ABI of the MI355X (gfx950) batched Goldilocks NTT / coset-LDE library
 * (libboojum_ntt.so).  This is the drop-in boundary for boojum-cuda's `ntt`
 * and `context` modules and for the functions either side of the transform
 * (`ops_complex` bit reversal, domain powers and omega shift; `barycentric`
 * evaluation; `poseidon` Merkle-tree hashing of the LDE output): a Rust FFI
 * crate (or any C caller) binds exactly these symbols
 * in place of the CUDA kernels that era_cudart launched.
 *
 * Conventions (same as the reference):
 *   - Elements are Goldilocks field elements stored as uint64_t, any value
 *     in [0, 2^64) is accepted as its residue mod p = 2^64 - 2^32 + 1.
 *     Outputs are canonical, in [0, p).
 *   - Matrices are column-major: column i starts at base + i * stride.
 *   - All transforms are asynchronous on `stream` (a hipStream_t; NULL is the
 *     null stream) and allocate nothing.
 *   - One context per process and device must exist (bc_context_create)
 *     before a transform is issued, as the reference requires a live
 *     Context (src/context.rs:157-282).
 *   - Return value: 0 on success; where the reference panics on an argument
 *     assert this returns BC_ERROR_INVALID_VALUE before anything is launched;
 *     launch failures return the HIP error code.
 */
#ifndef BOOJUM_NTT_H
#define BOOJUM_NTT_H

#include <stdbool.h>
#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define BC_SUCCESS 0
#define BC_ERROR_INVALID_VALUE 1   /* == hipErrorInvalidValue */
#define BC_ERROR_OUT_OF_MEMORY 2   /* == hipErrorOutOfMemory */
#define BC_ERROR_NOT_INITIALIZED 3 /* == hipErrorNotInitialized: no context */
#define BC_ERROR_WAIT_TIMEOUT 702  /* == hipErrorLaunchTimeOut: a ragged run's wait between levels gave up */

typedef struct bc_context bc_context;

/* OMEGA_LOG_ORDER (reference: src/context.rs:12, value 24 there; 32 here so
 * log_n + log_extension_degree may reach the full two-adic subgroup). */
uint32_t bc_omega_log_order(void);

/* Context::create(powers_of_w_coarse_log_count, powers_of_g_coarse_log_count)
 * (reference: src/context.rs:170-282).  Builds the power tables on the
 * current device and makes them the process-wide tables for that device.
 * Both arguments must be <= bc_omega_log_order() (src/context.rs:175-176);
 * the fine/coarse split of this implementation is fixed (2^16 x 2^16). */
int bc_context_create(uint32_t powers_of_w_coarse_log_count, uint32_t powers_of_g_coarse_log_count,
                      bc_context **out);

/* Context::destroy (reference: src/context.rs:284-296). */
int bc_context_destroy(bc_context *ctx);

/* batch_ntt_internal (reference: src/ntt.rs:179-291).  Pointers are already
 * offset; strides are in elements.  Asserts of src/ntt.rs:192-200. */
int bc_batch_ntt_internal(const uint64_t *inputs, uint64_t *outputs, uint32_t log_n, uint32_t num_ntts,
                          uint32_t stride_between_input_arrays, uint32_t stride_between_output_arrays,
                          bool bitrev_inputs, bool inverse, uint32_t log_extension_degree, uint32_t coset_index,
                          void *stream);

/* batch_ntt_out_of_place (reference: src/ntt.rs:293-335).  `*_len` are the
 * DeviceSlice lengths in elements (checked as at src/ntt.rs:310-313, in 64-bit
 * arithmetic); the matrices must not start at the same address
 * (src/ntt.rs:317,320). */
int bc_batch_ntt_out_of_place(const uint64_t *inputs_matrix, size_t inputs_len, uint64_t *outputs_matrix,
                              size_t outputs_len, uint32_t log_n, uint32_t num_ntts, uint32_t inputs_offset,
                              uint32_t outputs_offset, uint32_t stride_between_input_arrays,
                              uint32_t stride_between_output_arrays, bool bitrev_inputs, bool inverse,
                              uint32_t log_extension_degree, uint32_t coset_index, void *stream);

/* batch_ntt_in_place (reference: src/ntt.rs:337-369). */
int bc_batch_ntt_in_place(uint64_t *inputs_matrix, size_t inputs_len, uint32_t log_n, uint32_t num_ntts,
                          uint32_t inputs_offset, uint32_t stride_between_input_arrays, bool bitrev_inputs,
                          bool inverse, uint32_t log_extension_degree, uint32_t coset_index, void *stream);

/* ---- Extensions beyond the reference API (the pipeline around the NTT) ---- */

/* Coset low-degree extension of a batch of columns.  No single reference
 * function does this; it chains the reference's own calls the way its users
 * do (src/ntt.rs:176-177 describes the intent): an inverse transform of the
 * evaluations on H to coefficients, then one forward coset transform
 * (log_extension_degree = log_blowup, coset_index = c) per coset.
 *   inputs:  evaluations, column i at inputs + i * stride_in (n values,
 *            bit-reversed order if bitrev_inputs, else natural)
 *   coeffs:  receives the coefficients (n per column, stride stride_coeffs);
 *            may equal inputs (then the inputs are overwritten)
 *   outputs: column i at outputs + i * stride_out holds 2^log_blowup blocks
 *            of n values.  With bitrev_inputs, block brev(c) holds coset c in
 *            bit-reversed order, so the column is the bit-reversed LDE over
 *            g * H_{n 2^log_blowup}; otherwise block c holds coset c in
 *            natural order.
 * stride_out >= n << log_blowup; returns as bc_batch_ntt_internal. */
int bc_batch_lde(const uint64_t *inputs, uint64_t *coeffs, uint64_t *outputs, uint32_t log_n, uint32_t num_cols,
                 uint32_t stride_in, uint32_t stride_coeffs, uint64_t stride_out, uint32_t log_blowup,
                 bool bitrev_inputs, void *stream);

/* A captured, replayable launch sequence (hipGraph) for a fixed shape. */
typedef struct bc_graph bc_graph;

/* Captures bc_batch_lde with these exact arguments into a graph (on an
 * internal stream; nothing runs during capture).  Launch it with
 * bc_graph_launch on any stream; buffers must stay valid. */
int bc_lde_graph_create(const uint64_t *inputs, uint64_t *coeffs, uint64_t *outputs, uint32_t log_n,
                        uint32_t num_cols, uint32_t stride_in, uint32_t stride_coeffs, uint64_t stride_out,
                        uint32_t log_blowup, bool bitrev_inputs, bc_graph **out);
int bc_graph_launch(bc_graph *graph, void *stream);
int bc_graph_destroy(bc_graph *graph);

/* Ragged batches: many independent jobs of different sizes and operations,
 * run by one persistent kernel that pulls work units from a device queue.
 * Units of dependency level k are pass k of every job; bc_ragged_run runs all
 * levels in one cooperative launch whose workgroups wait for each other
 * between levels (the serving loop launches once per level instead).
 * Each job transforms num_cols columns of 2^log_n points (8 <= log_n):
 *   BC_OP_NTT   forward, natural -> bit-reversed (bc_batch_ntt_internal with
 *               bitrev_inputs = false, inverse = false)
 *   BC_OP_INTT  inverse, bit-reversed -> natural (bitrev_inputs = true,
 *               inverse = true)
 *   BC_OP_LDE   bit-reversed evaluations -> coefficients (written to coeffs)
 *               -> bit-reversed LDE with blowup 2^log_blowup (bc_batch_lde
 *               with bitrev_inputs = true)
 * No reference counterpart: the reference launches each job separately. */
#define BC_OP_NTT 0
#define BC_OP_INTT 1
#define BC_OP_LDE 2

typedef struct {
  const uint64_t *inputs; /* column i at inputs + i * stride_in */
  uint64_t *outputs;      /* column i at outputs + i * stride_out */
  uint64_t *coeffs;       /* BC_OP_LDE only: column i at coeffs + i * stride_coeffs */
  uint32_t log_n;
  uint32_t num_cols;
  uint32_t stride_in;
  uint32_t stride_out;    /* >= n, or n << log_blowup for BC_OP_LDE */
  uint32_t stride_coeffs;
  uint32_t op;
  uint32_t log_blowup;    /* BC_OP_LDE only */
} bc_ragged_job;

typedef struct bc_ragged_plan bc_ragged_plan;

/* Plans the jobs into work units (host + one device upload).  The plan keeps
 * the job pointers; buffers must stay valid while it is run. */
int bc_ragged_plan_create(const bc_ragged_job *jobs, uint32_t num_jobs, bc_ragged_plan **out);
/* Resets the queue and runs every job on stream in one persistent launch
 * (cooperative: every workgroup resident).  On a stream under graph capture,
 * or if the runtime refuses that grid, it launches once per dependency level
 * instead (at most 16), with the same results. */
int bc_ragged_run(bc_ragged_plan *plan, void *stream);
/* After the stream of the plan's last run has been synchronized:
 * BC_ERROR_WAIT_TIMEOUT if a wait between levels of that run gave up after
 * its spin limit (the outputs are then invalid), else BC_SUCCESS.  The waits
 * are bounded so that the kernel always ends.  No reference counterpart. */
int bc_ragged_wait_status(const bc_ragged_plan *plan);
/* Polls of one wait between levels before it gives up (about 1-2 us each);
 * 0 restores the default, 2^22.  For tests of the timeout path. */
int bc_ragged_set_wait_limit(uint32_t spins);
int bc_ragged_plan_destroy(bc_ragged_plan *plan);
/* Statistics of a plan: work units, dependency levels, persistent workgroups. */
int bc_ragged_plan_info(const bc_ragged_plan *plan, uint64_t *units, uint32_t *levels, uint32_t *workgroups);
/* The plain path for comparison: the same jobs as separate launch sequences
 * (one bc_batch_ntt_internal / bc_batch_lde per job). */
int bc_ragged_run_plain(const bc_ragged_job *jobs, uint32_t num_jobs, void *stream);
/* Plans and runs the jobs in one call without allocating device memory: no
 * hipMalloc/hipFree, no allocator of any library.  No counterpart in the
 * reference; it is for callers that cannot allocate (tight serving loops,
 * harnesses that forbid it).  The packed plan table and the queue state go to
 * a static 8 MiB device array that every GPU running the library keeps until
 * the process ends; a plan that needs more returns BC_ERROR_OUT_OF_MEMORY
 * before anything is launched (the 1000-job ragged benchmark batch takes
 * about 0.12 MiB).  The table is uploaded from pageable host memory, which HIP
 * copies synchronously, then the plan is launched on `stream` once per
 * dependency level (plain kernel launches, no cooperative launch), with
 * bc_ragged_run's results.  Every call on one GPU shares
 * that GPU's array and overwrites it: calls that overlap on one GPU (other
 * threads, or work still queued on another stream) race.  Host planning uses
 * the process heap. */
int bc_ragged_run_no_alloc(const bc_ragged_job *jobs, uint32_t num_jobs, void *stream);

/* ---- Order conversion and domain powers around the NTT ---- */

/* bit_reverse (reference: src/ops_complex.rs:260-318): for each of `cols`
 * columns (stride apart), dst[brev(i)] = src[i], i < rows; values are copied
 * unchanged.  rows must be a power of two below 2^32 (src/ops_complex.rs:279-
 * 281).  src == dst with equal strides is the in-place permutation. */
int bc_bit_reverse(const uint64_t *src, uint64_t src_stride, uint64_t *dst, uint64_t dst_stride, uint64_t rows,
                   uint32_t cols, void *stream);
/* bit_reverse_in_place (reference: src/ops_complex.rs:320-330). */
int bc_bit_reverse_in_place(uint64_t *values, uint64_t stride, uint64_t rows, uint32_t cols, void *stream);

/* get_powers_of_w / get_powers_of_g (reference: src/ops_complex.rs:29-86,
 * native/ops_complex.cu:9-33): result[i] = omega_{2^log_degree}^(+-idx), or
 * g^(+-((idx << (24 - log_degree)) mod 2^24)) (the reference's 24-bit g
 * tables: log_degree <= 24), with idx = offset + i, bit-reversed over
 * log_degree bits when bit_reverse.  Needs a context. */
int bc_get_powers_of_w(uint32_t log_degree, uint32_t offset, bool inverse, bool bit_reverse, uint64_t *result,
                       uint32_t count, void *stream);
int bc_get_powers_of_g(uint32_t log_degree, uint32_t offset, bool inverse, bool bit_reverse, uint64_t *result,
                       uint32_t count, void *stream);

/* omega_shift / omega_shift_in_place (reference: src/ops_complex.rs:207-258,
 * native/ops_complex.cu:76-87): dst[i] = src[i] * omega_{2^log_degree}^(+-
 * shift * (offset + i)), canonical.  Needs a context. */
int bc_omega_shift(const uint64_t *src, uint32_t log_degree, uint32_t offset, bool inverse, uint32_t shift,
                   uint64_t *dst, uint32_t count, void *stream);
int bc_omega_shift_in_place(uint64_t *values, uint32_t log_degree, uint32_t offset, bool inverse, uint32_t shift,
                            uint32_t count, void *stream);

/* ---- Barycentric evaluation on the coset domain {coset * omega_N^i} ----
 * Extension values are F_p[u]/(u^2 - 7) elements (c0, c1).  A single
 * extension element (x, the common factor, one evaluation) is the pair
 * (c0, c1) in adjacent u64; an extension vector or matrix column of N values
 * is a "double vector": N c0 values, then N c1 values. */

/* barycentric::precompute_lagrange_coeffs (src/barycentric.rs:83-127), for
 * PrecomputeAtBase (ext false) and PrecomputeAtExt (ext true; kernels
 * native/barycentric.cu:10-96).  x: device pointer to the point (1 or 2
 * u64); common_factor: device scratch of the same size (receives
 * coset (x^N - coset^N) / (N coset^N)); coeffs: N (base) or 2N (ext, double
 * vector) u64 receiving L_i(x) = common / (x omega_N^-i - coset), in natural
 * order or, with bit_reverse, at position brev(i).  count = N, a power of two
 * up to 2^31.  x must not be a domain point (the basis is undefined there).
 * Needs a context. */
int bc_barycentric_precompute(bool ext, const uint64_t *x, uint64_t *common_factor, uint64_t coset,
                              uint64_t *coeffs, uint64_t count, bool bit_reverse, void *stream);

/* Kinds of barycentric::batch_eval (src/barycentric.rs:265-267). */
enum {
  BC_EVAL_BASE_AT_BASE = 0, /* EvalBaseAtBase: base columns, base coefficients */
  BC_EVAL_BASE_AT_EXT = 1,  /* EvalBaseAtExt: base columns, extension coefficients */
  BC_EVAL_EXT_AT_EXT = 2    /* EvalExtAtExt: extension columns and coefficients */
};

/* barycentric::get_batch_eval_temp_storage_sizes (src/barycentric.rs:196-214):
 * partial_elems = elements (of the coefficient type: 1 u64 base, 2 ext) of
 * the partial-sum scratch; final_bytes = bytes of the second scratch, which
 * this implementation does not need (0).  stride = N, a power of two. */
int bc_batch_eval_temp_sizes(uint32_t kind, uint64_t stride, uint32_t num_polys, size_t *partial_elems,
                             size_t *final_bytes);

/* barycentric::batch_eval (src/barycentric.rs:216-261): evals[p] =
 * sum_i coeffs_i y_{i,p} for the num_polys columns of ys (column p at
 * ys + p * stride for base columns, ys + 2 p stride as a double vector for
 * extension columns; stride = N).  coeffs as written by
 * bc_barycentric_precompute.  evals: num_polys u64 (BASE_AT_BASE) or
 * num_polys (c0, c1) pairs.  temp_partial must hold partial_elems elements
 * from bc_batch_eval_temp_sizes; temp_final may be NULL.  Raw (non-canonical)
 * inputs are read mod p; outputs are canonical. */
int bc_batch_eval(uint32_t kind, const uint64_t *ys, uint64_t stride, uint32_t num_polys, const uint64_t *coeffs,
                  uint64_t *temp_partial, size_t temp_partial_elems, void *temp_final, size_t temp_final_bytes,
                  uint64_t *evals, void *stream);

/* ---- Poseidon / Poseidon2 Merkle trees (reference: src/poseidon.rs:129-328) ----
 * Poseidon over Goldilocks: width 12 (rate 8, capacity 4), 8 full and 22
 * partial rounds, S-box x^7, MDS = circulant of 2^{0,0,1,0,3,5,1,8,12,3,16,10}
 * (native/poseidon_constants_template.cuh:26).  Digests are 4 elements,
 * stored as 4 columns of `count` (element i of digest g at results[g + i *
 * count]).  Outputs are canonical; raw inputs are read mod p. */

/* Replaces the constants the reference compiles in from boojum
 * (build/poseidon_constants.rs:66-75): round_constants = the 360 textbook
 * round constants (30 rounds x 12, round-major; boojum's
 * poseidon_goldilocks_params::ALL_ROUND_CONSTANTS).  Derives the fused form
 * the kernels run (the reference's poseidon_constants_template.cuh tables)
 * and stores it on the current device.  Synchronous: it holds the library's
 * table lock exclusively (hashing calls on other threads wait for it), waits
 * for all work queued on the device (hashing kernels read the table), then
 * replaces the table.  Every hashing call on a device returns
 * BC_ERROR_NOT_INITIALIZED until this has been called there. */
int bc_poseidon_set_round_constants(const uint64_t *round_constants);

/* Host only (no device call): the per-device table derived from 360
 * constants, as out_words >= 864 u64.  Poseidon's fused form: textbook
 * constants of rounds 0-3 and 27-29 (84), the fused constant layer (12), the
 * dense matrix (144, row-major), the fused S-box constants (22), VS (22 x 11)
 * and W_HATS (22 x 11); then Poseidon2's constants: full rounds 0-3 and
 * 26-29 (96) and element 0 of each partial round (22). */
int bc_poseidon_fused_table(const uint64_t *round_constants, uint64_t *out, uint64_t out_words);

/* launch_leaves_kernel (src/poseidon.rs:129-170): count = results_len / 4
 * leaves; leaf g hashes rows = 2^log_rows_per_hash consecutive rows of each of
 * cols = values_len / (count rows) columns (value (r, c) of leaf g at
 * values[g rows + r + c rows count]), absorbed 8 at a time in the order
 * c rows + r, the last chunk zero-padded (native/poseidon_utils.cuh:8-42).
 * load_intermediate: start from the capacity stored in results by an earlier
 * call with store_intermediate, which stores the capacity instead of the
 * digest and requires rows * cols to be a multiple of 8. */
int bc_poseidon_leaves(const uint64_t *values, uint64_t values_len, uint64_t *results, uint64_t results_len,
                       uint32_t log_rows_per_hash, bool load_intermediate, bool store_intermediate, void *stream);

/* launch_single_thread_nodes_kernel (src/poseidon.rs:218-238): count =
 * results_len / 4 nodes from a layer of 2 count digests (values_len =
 * 2 results_len); node g hashes digests 2g and 2g + 1. */
int bc_poseidon_nodes(const uint64_t *values, uint64_t values_len, uint64_t *results, uint64_t results_len,
                      void *stream);

/* build_merkle_tree_nodes (src/poseidon.rs:262-286): layers_count node layers
 * above the layer in values (values_len = results_len, 8 2^k elements); layer
 * l fills the first half of what layers 0..l-1 left of results. */
int bc_poseidon_build_merkle_tree_nodes(const uint64_t *values, uint64_t values_len, uint64_t *results,
                                        uint64_t results_len, uint32_t layers_count, void *stream);

/* build_merkle_tree (src/poseidon.rs:311-328): leaves into the first half of
 * results (results_len = 8 leaves), then layers_count - 1 node layers. */
int bc_poseidon_build_merkle_tree(const uint64_t *values, uint64_t values_len, uint64_t *results,
                                  uint64_t results_len, uint32_t log_rows_per_hash, uint32_t layers_count,
                                  void *stream);

/* Poseidon2 (native/poseidon2_single_thread.cu, the hash the reference's
 * tests exercise): the same entry points, layouts and argument checks as the
 * Poseidon ones above, with the Poseidon2 permutation (external matrix
 * circ(2 M4, M4, M4), internal matrix J + diag(2^{4,14,11,8,0,5,2,9,13,6,3,12}),
 * eprint 2023/323) over the same round constants. */
int bc_poseidon2_leaves(const uint64_t *values, uint64_t values_len, uint64_t *results, uint64_t results_len,
                        uint32_t log_rows_per_hash, bool load_intermediate, bool store_intermediate, void *stream);
int bc_poseidon2_nodes(const uint64_t *values, uint64_t values_len, uint64_t *results, uint64_t results_len,
                       void *stream);
int bc_poseidon2_build_merkle_tree_nodes(const uint64_t *values, uint64_t values_len, uint64_t *results,
                                         uint64_t results_len, uint32_t layers_count, void *stream);
int bc_poseidon2_build_merkle_tree(const uint64_t *values, uint64_t values_len, uint64_t *results,
                                   uint64_t results_len, uint32_t log_rows_per_hash, uint32_t layers_count,
                                   void *stream);

/* launch_cooperative_leaves_kernel (src/poseidon.rs:194-216): the leaves of
 * bc_poseidon_leaves / bc_poseidon2_leaves (same layout, argument checks and
 * results) computed with one permutation state spread over 16 lanes, which
 * finishes sooner when there are few leaves.  The reference implements it for
 * Poseidon2 only (src/poseidon.rs:91-93); here both hashes have it. */
int bc_poseidon_leaves_cooperative(const uint64_t *values, uint64_t values_len, uint64_t *results,
                                   uint64_t results_len, uint32_t log_rows_per_hash, bool load_intermediate,
                                   bool store_intermediate, void *stream);
int bc_poseidon2_leaves_cooperative(const uint64_t *values, uint64_t values_len, uint64_t *results,
                                    uint64_t results_len, uint32_t log_rows_per_hash, bool load_intermediate,
                                    bool store_intermediate, void *stream);

/* launch_cooperative_nodes_kernel (src/poseidon.rs:240-260): the nodes of
 * bc_poseidon_nodes / bc_poseidon2_nodes, always with 16 lanes per
 * permutation (bc_poseidon_nodes itself switches to them for layers of at
 * most 2^14 nodes, as build_merkle_tree_nodes does, src/poseidon.rs:279-283). */
int bc_poseidon_nodes_cooperative(const uint64_t *values, uint64_t values_len, uint64_t *results,
                                  uint64_t results_len, void *stream);
int bc_poseidon2_nodes_cooperative(const uint64_t *values, uint64_t values_len, uint64_t *results,
                                   uint64_t results_len, void *stream);

/* gather_rows (src/poseidon.rs:330-358): for each of indexes_count indexes,
 * its 2^log_rows_per_index rows (index * 2^log_rows_per_index + r) of every
 * column of the column-major matrix values (column c at values + c *
 * values_stride, values_rows rows) go to rows i * 2^log_rows_per_index + r of
 * results (column c at results + c * results_stride, results_rows =
 * indexes_count << log_rows_per_index).  Unlike the reference, a source row
 * at or past values_rows is not read and its destination row is left as it
 * was. */
int bc_gather_rows(const uint32_t *indexes, uint32_t indexes_count, uint32_t log_rows_per_index,
                   const uint64_t *values, uint64_t values_stride, uint64_t values_rows, uint64_t *results,
                   uint64_t results_stride, uint64_t results_rows, uint32_t cols, void *stream);

/* gather_merkle_paths (src/poseidon.rs:360-390): values = a whole tree as
 * bc_poseidon_build_merkle_tree writes it (4 2^(L+1) elements for 2^L
 * leaves); results = layers_count layers of 4 columns of indexes_count
 * siblings: layer l, column c, index i gets element c of node
 * (indexes[i] >> l) ^ 1 of layer l.  layers_count < L.  An index at or past
 * 2^L gathers nothing (the reference reads out of range). */
int bc_gather_merkle_paths(const uint32_t *indexes, uint32_t indexes_count, const uint64_t *values,
                           uint64_t values_len, uint64_t *results, uint64_t results_len, uint32_t layers_count,
                           void *stream);

/* Seeded inputs on the device: column c (c = 0..num_cols-1) at
 * out + c * stride gets 2^log_n u64 draws of the splitmix64 stream seeded with
 * seed * 0x9E3779B97F4A7C15 + first_col + c (raw u64, any value, read mod p).
 * No reference counterpart: the reference's tests draw inputs on the host. */
int bc_fill_seeded(uint64_t *out, uint32_t log_n, uint32_t num_cols, uint64_t stride, uint64_t seed,
                   uint32_t first_col, void *stream);

/* ---- Tuning database ----
 * No reference counterpart (the reference's launch plans are the static
 * PLANS table, src/ntt.rs:74-169).  Launch parameters are looked up per
 * (log_n, log_cols = floor(log2(num_cols))); a missing bucket takes the
 * nearest tuned bucket of the same log_n, and zero fields (or no entry) keep
 * the built-in choice.  Plans change speed only: every tuned plan computes
 * the same bits.  bc_context_create loads the database named by $BC_TUNE_DB,
 * else tuning/mi355x.txt beside libboojum_ntt.so, once per process. */
typedef struct {
  uint32_t contig_log;          /* contiguous pass size 9..12 (or log_n when log_n <= 12); 0: built-in */
  uint32_t num_strided;         /* strided passes above it, each 4..8 bits */
  uint32_t strided_bpw;         /* (column, block) pairs per strided workgroup, <= 64 */
  uint32_t contig_grid;         /* grid cap of contiguous passes */
  uint32_t ragged_unit_contig;  /* ragged work unit: tiles, <= 64 */
  uint32_t ragged_unit_strided; /* ragged work unit: (column, block) pairs, <= 64 */
  uint32_t tile_log;            /* workgroup tile: 8 = one-wave 256-point passes (log_n 16, no coset, up to 65535
                                   columns), 12 = 4096-point pass kernels; 0: built-in (256-point for 1..4
                                   columns of 2^16) */
} bc_tune_params;

/* One HBM pass of a planned transform, as bc_batch_ntt_internal would launch
 * it (host only: no device call, no context needed). */
typedef struct {
  uint32_t contiguous;          /* 1: contiguous pass over 4096-point tiles; 0: strided pass */
  uint32_t log_h;               /* bits of the transform this pass computes */
  uint32_t log_block;           /* strided: log2 of the block it works inside */
  uint32_t mode;                /* scaling folded in: 0 none, 1 n^-1, 2 coset s^i on inputs, 3 n^-1 s^-i on outputs */
  uint64_t work_items;          /* contiguous: 4096-point tiles; strided: (column, block) pairs */
  uint64_t positions;           /* strided: r-ranges (grid.x) */
  uint64_t groups;              /* strided: pair groups (grid.y, <= 65535) */
  uint32_t pairs_per_workgroup; /* strided */
  uint32_t workgroup_threads;   /* 256, or 64 for the one-wave 256-point passes */
} bc_pass_info;

/* The pass plan of bc_batch_ntt_internal for this shape (log_n >= 8; below
 * that one generic pass runs and *num_passes is 0).  Arguments are checked as
 * bc_batch_ntt_internal checks them.  No reference counterpart (the
 * reference's plan is the static PLANS table, src/ntt.rs:74-169). */
int bc_ntt_describe_plan(uint32_t log_n, uint32_t num_ntts, bool bitrev_inputs, bool inverse,
                         uint32_t log_extension_degree, uint32_t coset_index, bc_pass_info *passes,
                         uint32_t max_passes, uint32_t *num_passes);

/* Sets (params != NULL) or erases the entry of one bucket; invalid parameters
 * return BC_ERROR_INVALID_VALUE. */
int bc_tune_set(uint32_t log_n, uint32_t log_cols, const bc_tune_params *params);
/* The parameters a call with this shape uses (all zero: built-in). */
int bc_tune_get(uint32_t log_n, uint32_t num_cols, bc_tune_params *params);
int bc_tune_clear(void);
/* Text database, one entry per line: log_n log_cols and the seven fields in
 * order (a line of six fields leaves tile_log 0); '#' starts a comment line.
 * Load merges into the table. */
int bc_tune_load(const char *path);
int bc_tune_save(const char *path);

/* ---- Serving loop: continuous batching under an HBM budget ----
 * No reference counterpart.  A server owns one HBM arena of hbm_budget_bytes
 * and a scheduler thread.  Requests (NTT / iNTT / coset-LDE of num_cols
 * columns) are split by columns into parts of at most max_part_elems
 * transformed elements; parts are admitted first-come first-served when their
 * buffers fit in the arena, gathered into ragged batches (one persistent
 * launch each) and run on up to max_batches_in_flight streams at once, so
 * requests that arrive while others run join the next batch and their work
 * overlaps on the device.  Results do not depend on how requests were split
 * or batched.
 *
 * Memory: a request's kept outputs are reserved in full when it is submitted,
 * and its parts are sized so that any queued part fits beside every kept
 * output, so no part waits on a bc_server_release.  Kept outputs are placed
 * from the top of the arena and transient buffers from the bottom; if a part
 * still cannot be placed while nothing runs (kept outputs have split the free
 * space), its request fails with hipErrorOutOfMemory rather than wait. */
typedef struct bc_server bc_server;

typedef struct {
  uint64_t hbm_budget_bytes;      /* arena size */
  uint64_t max_part_elems;        /* transformed elements per part (0: 2^28) */
  uint32_t max_batches_in_flight; /* streams (0: 3) */
  uint32_t max_parts_per_batch;   /* (0: 64) */
} bc_server_config;

typedef struct {
  uint32_t op;             /* BC_OP_NTT, BC_OP_INTT or BC_OP_LDE (as bc_ragged_job) */
  uint32_t log_n;          /* 8 <= log_n, log_n + log_blowup <= 31 */
  uint32_t num_cols;
  uint32_t log_blowup;     /* BC_OP_LDE only, 1..8 */
  const uint64_t *inputs;  /* device columns (stride_in apart), valid until the request completes;
                              the server reads them on its own streams, so they must be completely
                              written (the writing stream synchronized) before submit;
                              NULL: seeded inputs as bc_fill_seeded(seed, first_col = 0) */
  uint64_t stride_in;
  uint64_t seed;
  uint32_t keep_outputs;   /* 1: outputs stay in the arena until bc_server_release;
                              0: freed as each part completes (throughput runs) */
  uint32_t reserved;
} bc_request;

typedef struct {
  int64_t t_submit_ns;     /* host steady clock */
  int64_t t_first_launch_ns;
  int64_t t_done_ns;
  uint32_t parts;
  uint32_t parts_done;
  int32_t status;          /* 0 running or queued, 1 done, < 0 failed */
  uint32_t reserved;
} bc_request_info;

typedef struct {
  uint64_t *outputs;       /* column first_col + i at outputs + i * stride_out */
  uint64_t stride_out;
  uint32_t first_col;
  uint32_t num_cols;
} bc_part_info;

typedef struct {
  uint64_t requests_done;
  uint64_t parts_done;
  uint64_t batches_launched;
  uint32_t max_batches_in_flight_seen; /* > 1: batches overlapped on the device */
  uint32_t pending_parts;
  uint64_t arena_high_water_bytes;
  uint64_t arena_bytes_in_use;
} bc_server_stats;

/* Needs a live context on the current device; the server uses that device. */
int bc_server_create(const bc_server_config *config, bc_server **out);
/* Queues a request; *id identifies it.  Returns BC_ERROR_INVALID_VALUE for bad
 * arguments and 2 (hipErrorOutOfMemory) when one column, or all kept outputs
 * beside the parts already queued, cannot fit in the budget. */
int bc_server_submit(bc_server *server, const bc_request *request, uint64_t *id);
/* Blocks until the request has completed (or failed: returns its error). */
int bc_server_wait(bc_server *server, uint64_t id);
int bc_server_request_info(bc_server *server, uint64_t id, bc_request_info *info);
/* Where part `part` of a completed keep_outputs request left its outputs:
 * the request's layout (bit-reversed LDE blocks etc.) as bc_ragged_job. */
int bc_server_part(bc_server *server, uint64_t id, uint32_t part, bc_part_info *info);
/* Frees a completed request's kept outputs and forgets the request. */
int bc_server_release(bc_server *server, uint64_t id);
/* Blocks until every submitted request has completed. */
int bc_server_drain(bc_server *server);
int bc_server_stats_get(bc_server *server, bc_server_stats *stats);
/* Drains, stops the scheduler and frees the arena. */
int bc_server_destroy(bc_server *server);

#ifdef __cplusplus
}
#endif

#endif /* BOOJUM_NTT_H */

// Shared declarations between the C-ABI layer and the NTT kernels.
#pragma once
#include <stdint.h>
#include <hip/hip_runtime.h>

namespace bc {

// Largest supported log2 transform length (including the LDE blowup).
// The reference caps this at 24 (src/context.rs:12); the power tables here
// cover the whole 2^32 two-adic subgroup.
constexpr uint32_t OMEGA_LOG_ORDER = 32;

// Power tables resident in HBM (one allocation, built by bc_context_create):
//   w_fine[i]   = omega_{2^32}^i            i < 2^16
//   w_coarse[i] = omega_{2^32}^(i * 2^16)   i < 2^16
//   g_fine / g_coarse: powers of g = 7, same split; gi_*: powers of g^-1.
// omega_{2^32}^E for any u32 exponent E is w_coarse[E >> 16] * w_fine[E & 0xFFFF],
// so every omega_{2^k} power is an exponent E = e << (32 - k) in u32 arithmetic
// (wrap-around is reduction mod 2^32, i.e. mod the subgroup order).
struct Tables {
  const uint64_t *lvl_tw;  // radix-16 level twists, see LVL_TW_* below
  const uint64_t *w_fine;
  const uint64_t *w_coarse;
  const uint64_t *g_fine;
  const uint64_t *g_coarse;
  const uint64_t *gi_fine;
  const uint64_t *gi_coarse;
};
constexpr uint32_t TABLE_LOG = 16;
constexpr uint32_t TABLE_LEN = 1u << TABLE_LOG;
constexpr uint32_t NUM_TABLES = 6;

// Level twists of a radix-16 level inside a block of M = 2^m points
// (5 <= m <= 12, stride S = M / 16): entry (i - 1) * S + rp holds
// omega_M^(rp * brev4(i)) for i = 1..15, rp < S, so a wave reading one i for
// consecutive rp reads consecutive words.  Tables for m = 5..12 are packed in
// order, forward block first, then the inverse block (omega_M^-(...)).
constexpr uint32_t LVL_TW_MIN_LOG = 5;
constexpr uint32_t LVL_TW_MAX_LOG = 12;
__host__ __device__ constexpr uint32_t lvl_tw_offset(uint32_t m) { return 15u * ((1u << (m - 4)) - 2u); }
constexpr uint32_t LVL_TW_BLOCK = 15u * ((1u << (LVL_TW_MAX_LOG - 3)) - 2u);  // entries per direction

// One HBM pass of a batched transform (see ntt_kernels.hip for the math).
struct PassArgs {
  const uint64_t *in;
  uint64_t *out;
  uint64_t stride_in;   // elements between consecutive input columns
  uint64_t stride_out;  // elements between consecutive output columns
  uint64_t n_inv;       // n^-1 mod p (last pass of an inverse transform)
  Tables tab;
  uint32_t num_cols;
  uint32_t log_n;      // column length
  uint32_t log_N;      // block length this pass works inside
  uint32_t log_H;      // sub-transform length this pass computes
  uint32_t first;      // pass reads the caller's input (coset pre-scaling)
  uint32_t last;       // pass writes the final output (inverse post-scaling)
  uint32_t log_ext;    // LDE: log2 blowup, 0 = plain transform
  uint32_t coset_exp;  // coset_index << (32 - log_n - log_ext)
};

// Per-column base pointers of one pass (ragged job groups: the columns of
// several jobs, each with its own buffers, processed as one batch).
struct ColPtrs {
  const uint64_t *in;
  uint64_t *out;
};

// Arguments of the register-resident pass kernels (ntt_fast.hip).
struct FastArgs {
  const uint64_t *in;
  uint64_t *out;
  uint64_t stride_in;
  uint64_t stride_out;
  uint64_t n_inv;
  Tables tab;
  uint32_t num_cols;
  uint32_t log_n;
  uint32_t log_N;          // strided passes: block length this pass works inside
  uint32_t log_ext;
  uint32_t coset_exp;
  uint32_t first;          // pass reads the caller's (possibly non-canonical) input
  uint32_t last;           // pass writes the final (canonical) output
  uint32_t tiles;          // contiguous passes: number of 4096-point tiles
  uint32_t blocks_per_wg;  // strided passes: (column, block) pairs per workgroup
  const ColPtrs *cols;     // column-table addressing (ragged): column c at cols[c], in/out/strides unused
};

// Scaling folded into a pass (see plan in ntt_api.hip).
enum PassMode : int {
  MODE_NONE = 0,     // nothing beyond the transform
  MODE_NINV = 1,     // multiply by n^-1
  MODE_LDE_IN = 2,   // multiply inputs by s^i (forward coset transform)
  MODE_LDE_OUT = 3,  // multiply outputs by n^-1 s^-i (inverse coset transform)
};

// One pass of a planned transform (log_n >= 8), as launched by
// bc_batch_ntt_internal and scheduled by the ragged persistent kernel.
struct PassDesc {
  bool contig;         // contiguous pass (else strided)
  uint32_t log_h;      // sub-transform size of the pass
  int mode;            // PassMode
  FastArgs a;          // in/out/strides/flags; a.tiles for contiguous passes
  uint64_t positions;  // strided: number of r-ranges (4096-point tiles per block)
  uint64_t pairs;      // strided: (column, block) pairs
};

// Plans a transform into passes in execution order; returns the pass count
// (at most 8).  Arguments must already be validated; requires log_n >= 8.
int describe_fast_passes(const Tables &tab, const uint64_t *inputs, uint64_t *outputs, uint32_t log_n,
                         uint32_t num_cols, uint64_t stride_in, uint64_t stride_out, bool bitrev_inputs, bool inverse,
                         uint32_t log_ext, uint32_t coset_index, PassDesc *out);

// Launch shape of a strided pass over `pairs` (column, block) pairs at
// `positions` r-ranges: returns the pair groups (grid.y, at most
// STRIDED_MAX_GROUPS) and stores the pairs per workgroup in *bpw.  tuned_bpw
// (0: built-in) is raised when the groups would not fit the grid.
constexpr uint64_t STRIDED_MAX_GROUPS = 65535;
uint32_t strided_groups(uint64_t pairs, uint64_t positions, uint32_t tuned_bpw, uint32_t *bpw);

// max_wgs: grid cap of a contiguous pass (0: built-in 2048).
hipError_t launch_contig(const FastArgs &a, int log_h, bool inverse, bool dit, int mode, uint32_t max_wgs,
                         hipStream_t stream);
hipError_t launch_strided(const FastArgs &a, int log_h, bool inverse, bool dit, int mode, uint32_t positions,
                          uint32_t groups, hipStream_t stream);

// Tuned launch parameters for one (log_n, column-count bucket); zero fields
// select the built-in choice (ntt_tune.hip).
struct TuneParams {
  uint32_t contig_log;    // contiguous pass size; 0: built-in split
  uint32_t num_strided;   // strided passes above it (the rest of log_n split evenly)
  uint32_t strided_bpw;   // (column, block) pairs per strided workgroup
  uint32_t contig_grid;   // grid cap of contiguous passes
  uint32_t unit_contig;   // ragged: tiles per work unit
  uint32_t unit_strided;  // ragged: pairs per work unit
  uint32_t tile_log;      // 8: one-wave 256-point passes (log_n 16); 12: 4096-point passes; 0: built-in
};
TuneParams tuned(uint32_t log_n, uint32_t num_cols);
bool tune_params_valid(uint32_t log_n, const TuneParams &p);
void tune_autoload();

constexpr uint32_t TILE_LOG = 12;
constexpr uint32_t TILE = 1u << TILE_LOG;
constexpr uint32_t TILE_THREADS = 256;

// Latency path (ntt_small.hip): 1..SMALL_MAX_COLS columns of 2^SMALL_LOG_N
// points with no coset, as two launches of one-wave workgroups.
constexpr uint32_t SMALL_LOG_N = 16;
constexpr uint32_t SMALL_MAX_COLS = 4;
// Whether a call of this shape takes the latency path under tuned params tp.
bool small_path_applies(uint32_t log_n, uint32_t num_cols, uint32_t log_ext, const TuneParams &tp);
hipError_t launch_small(const Tables &tab, const uint64_t *inputs, uint64_t *outputs, uint32_t num_cols,
                        uint64_t stride_in, uint64_t stride_out, bool bitrev_inputs, bool inverse, hipStream_t s);

// Host: the split of log_n into per-pass sub-transform sizes, n2b order
// (first entry = the pass that works on the top index bits).
int plan_passes(uint32_t log_n, uint32_t *log_h, int max_passes);

hipError_t launch_pass(const PassArgs &a, bool inverse, bool b2n, hipStream_t stream);

}  // namespace bc

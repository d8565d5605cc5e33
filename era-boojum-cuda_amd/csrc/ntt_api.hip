// C ABI entry points: context management and the batched NTT host logic
// (argument checks and launches; the pass planning is in ntt_plan.cpp).  Mirrors src/context.rs and
// src/ntt.rs of the reference; see include/boojum_ntt.h for the contract.
#include "../../include/boojum_ntt.h"
#include "gl_field.hpp"
#include "ntt_internal.hpp"

#include <mutex>
#include <vector>

struct bc_context {
  int device;
  uint64_t *tables;  // NUM_TABLES * TABLE_LEN u64, layout of bc::Tables
  bc::Tables tab;
};

namespace bc {
namespace {

constexpr int MAX_DEVICES = 64;
std::mutex g_mu;
bc_context *g_current[MAX_DEVICES] = {};

// Consecutive powers base^0 .. base^(len-1) (boojum distribute_powers over a
// vector of ones, as Context::create does at src/context.rs:144-155).
void fill_powers(uint64_t *dst, uint32_t len, uint64_t base) {
  uint64_t p = 1;
  for (uint32_t i = 0; i < len; ++i) {
    dst[i] = p;
    p = gl::canon(gl::mul(p, base));
  }
}

bc_context *current_context() {
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= MAX_DEVICES) return nullptr;
  std::lock_guard<std::mutex> lk(g_mu);
  return g_current[dev];
}


// log_n < 8: one pass of the generic LDS kernel.
int run_general(bc_context *ctx, const uint64_t *inputs, uint64_t *outputs, uint32_t log_n, uint32_t num_ntts,
                uint32_t stride_in, uint32_t stride_out, bool bitrev_inputs, bool inverse, uint32_t log_ext,
                uint32_t coset_exp, uint64_t n_inv, hipStream_t s) {
  PassArgs a{};
  a.tab = ctx->tab;
  a.num_cols = num_ntts;
  a.log_n = log_n;
  a.log_ext = log_ext;
  a.coset_exp = coset_exp;
  a.n_inv = n_inv;
  a.first = a.last = 1;
  a.in = inputs;
  a.stride_in = stride_in;
  a.out = outputs;
  a.stride_out = stride_out;
  a.log_N = log_n;
  a.log_H = log_n;
  return launch_pass(a, inverse, bitrev_inputs, s);
}

}  // namespace

Tables current_tables(bool &ok) {
  bc_context *ctx = current_context();
  ok = ctx != nullptr;
  return ok ? ctx->tab : Tables{};
}

}  // namespace bc


extern "C" uint32_t bc_omega_log_order(void) { return bc::OMEGA_LOG_ORDER; }

extern "C" int bc_context_create(uint32_t powers_of_w_coarse_log_count, uint32_t powers_of_g_coarse_log_count,
                                 bc_context **out) {
  using namespace bc;
  if (!out) return BC_ERROR_INVALID_VALUE;
  *out = nullptr;
  if (powers_of_w_coarse_log_count > OMEGA_LOG_ORDER || powers_of_g_coarse_log_count > OMEGA_LOG_ORDER)
    return BC_ERROR_INVALID_VALUE;
  int dev = 0;
  hipError_t err = hipGetDevice(&dev);
  if (err != hipSuccess) return err;
  if (dev < 0 || dev >= MAX_DEVICES) return BC_ERROR_INVALID_VALUE;
  tune_autoload();

  std::vector<uint64_t> host((size_t)NUM_TABLES * TABLE_LEN + 2 * LVL_TW_BLOCK);
  const uint64_t w = gl::OMEGA_2_32;
  const uint64_t w_coarse = gl::pow(w, TABLE_LEN);
  const uint64_t g = gl::GENERATOR;
  const uint64_t gi = gl::inv(g);
  fill_powers(&host[0 * TABLE_LEN], TABLE_LEN, w);
  fill_powers(&host[1 * TABLE_LEN], TABLE_LEN, w_coarse);
  fill_powers(&host[2 * TABLE_LEN], TABLE_LEN, g);
  fill_powers(&host[3 * TABLE_LEN], TABLE_LEN, gl::pow(g, TABLE_LEN));
  fill_powers(&host[4 * TABLE_LEN], TABLE_LEN, gi);
  fill_powers(&host[5 * TABLE_LEN], TABLE_LEN, gl::pow(gi, TABLE_LEN));
  // radix-16 level twists: omega_M^(+-rp * brev4(i)), omega_M = omega_{2^32}^(2^(32-m))
  uint64_t *lvl = &host[(size_t)NUM_TABLES * TABLE_LEN];
  for (uint32_t m = LVL_TW_MIN_LOG; m <= LVL_TW_MAX_LOG; ++m) {
    const uint32_t S = 1u << (m - 4);
    const uint64_t wm = gl::pow(w, 1ull << (32 - m));
    const uint64_t wm_inv = gl::inv(wm);
    for (uint32_t i = 1; i < 16; ++i) {
      const uint32_t bi = ((i & 1) << 3) | ((i & 2) << 1) | ((i & 4) >> 1) | ((i & 8) >> 3);
      for (uint32_t rp = 0; rp < S; ++rp) {
        lvl[lvl_tw_offset(m) + (i - 1) * S + rp] = gl::pow(wm, (uint64_t)rp * bi);
        lvl[LVL_TW_BLOCK + lvl_tw_offset(m) + (i - 1) * S + rp] = gl::pow(wm_inv, (uint64_t)rp * bi);
      }
    }
  }

  bc_context *ctx = new bc_context();
  ctx->device = dev;
  err = hipMalloc((void **)&ctx->tables, host.size() * sizeof(uint64_t));
  if (err != hipSuccess) {
    delete ctx;
    return err;
  }
  err = hipMemcpy(ctx->tables, host.data(), host.size() * sizeof(uint64_t), hipMemcpyHostToDevice);
  if (err != hipSuccess) {
    (void)hipFree(ctx->tables);
    delete ctx;
    return err;
  }
  ctx->tab.lvl_tw = ctx->tables + (size_t)NUM_TABLES * TABLE_LEN;
  ctx->tab.w_fine = ctx->tables + 0 * TABLE_LEN;
  ctx->tab.w_coarse = ctx->tables + 1 * TABLE_LEN;
  ctx->tab.g_fine = ctx->tables + 2 * TABLE_LEN;
  ctx->tab.g_coarse = ctx->tables + 3 * TABLE_LEN;
  ctx->tab.gi_fine = ctx->tables + 4 * TABLE_LEN;
  ctx->tab.gi_coarse = ctx->tables + 5 * TABLE_LEN;
  {
    std::lock_guard<std::mutex> lk(g_mu);
    g_current[dev] = ctx;
  }
  *out = ctx;
  return BC_SUCCESS;
}

extern "C" int bc_context_destroy(bc_context *ctx) {
  using namespace bc;
  if (!ctx) return BC_ERROR_INVALID_VALUE;
  {
    std::lock_guard<std::mutex> lk(g_mu);
    if (ctx->device >= 0 && ctx->device < MAX_DEVICES && g_current[ctx->device] == ctx)
      g_current[ctx->device] = nullptr;
  }
  hipError_t err = hipFree(ctx->tables);
  delete ctx;
  return err;
}

extern "C" int bc_batch_ntt_internal(const uint64_t *inputs, uint64_t *outputs, uint32_t log_n, uint32_t num_ntts,
                                     uint32_t stride_in, uint32_t stride_out, bool bitrev_inputs, bool inverse,
                                     uint32_t log_ext, uint32_t coset_index, void *stream) {
  using namespace bc;
  // src/ntt.rs:192-200
  if (log_n < 1) return BC_ERROR_INVALID_VALUE;
  if ((uint64_t)log_n + log_ext > OMEGA_LOG_ORDER) return BC_ERROR_INVALID_VALUE;
  const uint64_t n = 1ull << log_n;
  if (n > stride_in || n > stride_out) return BC_ERROR_INVALID_VALUE;
  if (coset_index > 0 && (log_ext == 0 || (uint64_t)coset_index >= (1ull << log_ext))) return BC_ERROR_INVALID_VALUE;
  if (num_ntts == 0) return BC_SUCCESS;
  if (!inputs || !outputs) return BC_ERROR_INVALID_VALUE;
  bc_context *ctx = current_context();
  if (!ctx) return BC_ERROR_NOT_INITIALIZED;

  hipStream_t s = (hipStream_t)stream;
  if (log_n < 8) {
    const uint32_t coset_exp = log_ext ? (uint32_t)((uint64_t)coset_index << (32 - log_n - log_ext)) : 0u;
    return run_general(ctx, inputs, outputs, log_n, num_ntts, stride_in, stride_out, bitrev_inputs, inverse, log_ext,
                       coset_exp, gl::inv(n % gl::P), s);
  }

  const TuneParams tp = tuned(log_n, num_ntts);
  if (small_path_applies(log_n, num_ntts, log_ext, tp))
    return launch_small(ctx->tab, inputs, outputs, num_ntts, stride_in, stride_out, bitrev_inputs, inverse, s);

  PassDesc passes[8];
  const int np = describe_fast_passes(ctx->tab, inputs, outputs, log_n, num_ntts, stride_in, stride_out,
                                      bitrev_inputs, inverse, log_ext, coset_index, passes);
  if (np <= 0) return BC_ERROR_INVALID_VALUE;
  // every pass's launch shape is fixed before the first launch, so a shape
  // that cannot be launched is refused with nothing queued (an in-place
  // transform would otherwise lose its input to the passes already run)
  uint32_t groups[8] = {};
  for (int k = 0; k < np; ++k) {
    PassDesc &pd = passes[k];
    if (pd.contig) continue;
    if (pd.positions > 0x7FFFFFFFull) return BC_ERROR_INVALID_VALUE;
    groups[k] = strided_groups(pd.pairs, pd.positions, tp.strided_bpw, &pd.a.blocks_per_wg);
  }
  for (int k = 0; k < np; ++k) {
    const PassDesc &pd = passes[k];
    const hipError_t err =
        pd.contig ? launch_contig(pd.a, (int)pd.log_h, inverse, bitrev_inputs, pd.mode, tp.contig_grid, s)
                  : launch_strided(pd.a, (int)pd.log_h, inverse, bitrev_inputs, pd.mode, (uint32_t)pd.positions,
                                   groups[k], s);
    if (err != hipSuccess) return err;
  }
  return BC_SUCCESS;
}

extern "C" int bc_batch_ntt_out_of_place(const uint64_t *inputs_matrix, size_t inputs_len, uint64_t *outputs_matrix,
                                         size_t outputs_len, uint32_t log_n, uint32_t num_ntts, uint32_t inputs_offset,
                                         uint32_t outputs_offset, uint32_t stride_in, uint32_t stride_out,
                                         bool bitrev_inputs, bool inverse, uint32_t log_ext, uint32_t coset_index,
                                         void *stream) {
  // src/ntt.rs:309-320, with the products in 64 bits (the reference's u32
  // products overflow for matrices beyond 2^32 elements).
  if (!inputs_matrix || !outputs_matrix || log_n >= 32) return BC_ERROR_INVALID_VALUE;
  const uint64_t n = 1ull << log_n;
  if ((uint64_t)inputs_offset + n > stride_in) return BC_ERROR_INVALID_VALUE;
  if ((uint64_t)outputs_offset + n > stride_out) return BC_ERROR_INVALID_VALUE;
  if ((uint64_t)inputs_len < (uint64_t)num_ntts * stride_in) return BC_ERROR_INVALID_VALUE;
  if ((uint64_t)outputs_len < (uint64_t)num_ntts * stride_out) return BC_ERROR_INVALID_VALUE;
  if ((const void *)inputs_matrix == (const void *)outputs_matrix) return BC_ERROR_INVALID_VALUE;
  const uint64_t *in = inputs_matrix + inputs_offset;
  uint64_t *out = outputs_matrix + outputs_offset;
  if ((const void *)in == (const void *)out) return BC_ERROR_INVALID_VALUE;
  return bc_batch_ntt_internal(in, out, log_n, num_ntts, stride_in, stride_out, bitrev_inputs, inverse, log_ext,
                               coset_index, stream);
}

extern "C" int bc_batch_ntt_in_place(uint64_t *inputs_matrix, size_t inputs_len, uint32_t log_n, uint32_t num_ntts,
                                     uint32_t inputs_offset, uint32_t stride, bool bitrev_inputs, bool inverse,
                                     uint32_t log_ext, uint32_t coset_index, void *stream) {
  // src/ntt.rs:350-352
  if (!inputs_matrix || log_n >= 32) return BC_ERROR_INVALID_VALUE;
  const uint64_t n = 1ull << log_n;
  if ((uint64_t)inputs_offset + n > stride) return BC_ERROR_INVALID_VALUE;
  if ((uint64_t)inputs_len < (uint64_t)num_ntts * stride) return BC_ERROR_INVALID_VALUE;
  uint64_t *p = inputs_matrix + inputs_offset;
  return bc_batch_ntt_internal(p, p, log_n, num_ntts, stride, stride, bitrev_inputs, inverse, log_ext, coset_index,
                               stream);
}

// ---------------------------------------------------------------------------
// Coset LDE pipeline and graph capture

namespace bc {
namespace {
uint32_t brev_small(uint32_t x, uint32_t bits) {
  uint32_t r = 0;
  for (uint32_t i = 0; i < bits; ++i) r |= ((x >> i) & 1u) << (bits - 1 - i);
  return r;
}
}  // namespace
}  // namespace bc

extern "C" int bc_batch_lde(const uint64_t *inputs, uint64_t *coeffs, uint64_t *outputs, uint32_t log_n,
                            uint32_t num_cols, uint32_t stride_in, uint32_t stride_coeffs, uint64_t stride_out,
                            uint32_t log_blowup, bool bitrev_inputs, void *stream) {
  using namespace bc;
  if (!inputs || !coeffs || !outputs || log_n < 1 || log_n >= 32) return BC_ERROR_INVALID_VALUE;
  if ((uint64_t)log_n + log_blowup > OMEGA_LOG_ORDER || log_blowup > 16) return BC_ERROR_INVALID_VALUE;
  const uint64_t n = 1ull << log_n;
  if (stride_out < (n << log_blowup) || stride_out > 0xFFFFFFFFull) return BC_ERROR_INVALID_VALUE;
  // 1. evaluations -> coefficients (natural order when the input is
  //    bit-reversed, bit-reversed otherwise)
  int rc = bc_batch_ntt_internal(inputs, coeffs, log_n, num_cols, stride_in, stride_coeffs, bitrev_inputs, true, 0, 0,
                                 stream);
  if (rc != BC_SUCCESS) return rc;
  // 2. one forward coset transform per coset; natural coefficients give
  //    bit-reversed evaluations (n2b) and vice versa (b2n)
  for (uint32_t c = 0; c < (1u << log_blowup); ++c) {
    const uint32_t blk = bitrev_inputs ? brev_small(c, log_blowup) : c;
    rc = bc_batch_ntt_internal(coeffs, outputs + (uint64_t)blk * n, log_n, num_cols, stride_coeffs,
                               (uint32_t)stride_out, !bitrev_inputs, false, log_blowup, c, stream);
    if (rc != BC_SUCCESS) return rc;
  }
  return BC_SUCCESS;
}

struct bc_graph {
  hipGraph_t graph;
  hipGraphExec_t exec;
};

extern "C" int bc_lde_graph_create(const uint64_t *inputs, uint64_t *coeffs, uint64_t *outputs, uint32_t log_n,
                                   uint32_t num_cols, uint32_t stride_in, uint32_t stride_coeffs,
                                   uint64_t stride_out, uint32_t log_blowup, bool bitrev_inputs, bc_graph **out) {
  if (!out) return BC_ERROR_INVALID_VALUE;
  *out = nullptr;
  hipStream_t cs;
  hipError_t err = hipStreamCreateWithFlags(&cs, hipStreamNonBlocking);
  if (err != hipSuccess) return err;
  err = hipStreamBeginCapture(cs, hipStreamCaptureModeThreadLocal);
  if (err != hipSuccess) {
    (void)hipStreamDestroy(cs);
    return err;
  }
  const int rc = bc_batch_lde(inputs, coeffs, outputs, log_n, num_cols, stride_in, stride_coeffs, stride_out,
                              log_blowup, bitrev_inputs, cs);
  hipGraph_t g = nullptr;
  err = hipStreamEndCapture(cs, &g);
  (void)hipStreamDestroy(cs);
  if (rc != BC_SUCCESS) {
    if (g) (void)hipGraphDestroy(g);
    return rc;
  }
  if (err != hipSuccess) return err;
  hipGraphExec_t ex = nullptr;
  err = hipGraphInstantiate(&ex, g, nullptr, nullptr, 0);
  if (err != hipSuccess) {
    (void)hipGraphDestroy(g);
    return err;
  }
  bc_graph *h = new bc_graph{g, ex};
  *out = h;
  return BC_SUCCESS;
}

extern "C" int bc_graph_launch(bc_graph *graph, void *stream) {
  if (!graph) return BC_ERROR_INVALID_VALUE;
  return hipGraphLaunch(graph->exec, (hipStream_t)stream);
}

extern "C" int bc_graph_destroy(bc_graph *graph) {
  if (!graph) return BC_ERROR_INVALID_VALUE;
  hipError_t e1 = hipGraphExecDestroy(graph->exec);
  hipError_t e2 = hipGraphDestroy(graph->graph);
  delete graph;
  return e1 != hipSuccess ? e1 : e2;
}

extern "C" int bc_ntt_describe_plan(uint32_t log_n, uint32_t num_ntts, bool bitrev_inputs, bool inverse,
                                    uint32_t log_ext, uint32_t coset_index, bc_pass_info *passes, uint32_t max_passes,
                                    uint32_t *num_passes) {
  using namespace bc;
  if (!num_passes || log_n < 1 || log_n >= 32) return BC_ERROR_INVALID_VALUE;
  if ((uint64_t)log_n + log_ext > OMEGA_LOG_ORDER) return BC_ERROR_INVALID_VALUE;
  if (coset_index > 0 && (log_ext == 0 || (uint64_t)coset_index >= (1ull << log_ext))) return BC_ERROR_INVALID_VALUE;
  *num_passes = 0;
  if (log_n < 8 || num_ntts == 0) return BC_SUCCESS;
  if (small_path_applies(log_n, num_ntts, log_ext, tuned(log_n, num_ntts))) {
    // two one-wave passes of 8 bits: the top one (stride 256, pass twist,
    // n^-1 for inverses) and the bottom one; b2n runs them in reverse
    if (!passes || max_passes < 2) return BC_ERROR_INVALID_VALUE;
    for (int k = 0; k < 2; ++k) {
      const bool top = (k == 0) != bitrev_inputs;
      bc_pass_info &o = passes[k];
      o = bc_pass_info{};
      o.contiguous = top ? 0u : 1u;
      o.log_h = 8;
      o.log_block = top ? SMALL_LOG_N : 8;
      o.mode = top && inverse ? (uint32_t)MODE_NINV : (uint32_t)MODE_NONE;
      o.work_items = (uint64_t)num_ntts << 8;
      o.positions = top ? 256 : 0;
      o.groups = top ? num_ntts : 0;
      o.pairs_per_workgroup = top ? 1 : 0;
      o.workgroup_threads = 64;
    }
    *num_passes = 2;
    return BC_SUCCESS;
  }
  PassDesc pd[8];
  const int np = describe_fast_passes(Tables{}, nullptr, nullptr, log_n, num_ntts, 1ull << log_n, 1ull << log_n,
                                      bitrev_inputs, inverse, log_ext, coset_index, pd);
  if (np <= 0) return BC_ERROR_INVALID_VALUE;
  if (!passes || max_passes < (uint32_t)np) return BC_ERROR_INVALID_VALUE;
  const TuneParams tp = tuned(log_n, num_ntts);
  for (int k = 0; k < np; ++k) {
    bc_pass_info &o = passes[k];
    o = bc_pass_info{};
    o.contiguous = pd[k].contig ? 1u : 0u;
    o.log_h = pd[k].log_h;
    o.log_block = pd[k].a.log_N;
    o.mode = (uint32_t)pd[k].mode;
    o.workgroup_threads = 256;
    if (pd[k].contig) {
      o.work_items = pd[k].a.tiles;
    } else {
      if (pd[k].positions > 0x7FFFFFFFull) return BC_ERROR_INVALID_VALUE;
      o.work_items = pd[k].pairs;
      o.positions = pd[k].positions;
      o.groups = strided_groups(pd[k].pairs, pd[k].positions, tp.strided_bpw, &o.pairs_per_workgroup);
    }
  }
  *num_passes = (uint32_t)np;
  return BC_SUCCESS;
}

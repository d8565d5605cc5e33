// Host planning of batched transforms (no device code, no HIP calls): the
// split of a transform into HBM passes, the per-pass launch shapes, and the
// ragged-batch planner that turns many jobs into one work queue.  Compiled as
// plain C++ into the library, and on its own under AddressSanitizer and
// UndefinedBehaviorSanitizer by the CPU tests.
#include "../../include/boojum_ntt.h"
#include "gl_field.hpp"
#include "ntt_internal.hpp"
#include "ntt_ragged.hpp"

#include <string.h>

#include <algorithm>
#include <map>
#include <vector>

namespace bc {
namespace {

// Split of log_n >= 8 into strided passes of 4..8 bits (top index bits first)
// and one contiguous pass of 8..12 bits: the tuned split when one is set,
// else the fewest strided passes with the largest contiguous pass.
struct FastPlan {
  int num_strided;
  uint32_t strided[8];
  uint32_t log_N[8];  // block length of each strided pass
  uint32_t contig;
};

void make_fast_plan(uint32_t log_n, const TuneParams &tp, FastPlan &pl) {
  pl.num_strided = 0;
  if (log_n <= TILE_LOG) {
    pl.contig = log_n;
    return;
  }
  uint32_t c = TILE_LOG, s = log_n - TILE_LOG;
  uint32_t k = (s + 7) / 8;
  if (s < 4 * k) {  // only when k == 1 and s < 4: shrink the contiguous pass
    c -= 4 * k - s;
    s = 4 * k;
  }
  if (tp.contig_log && tune_params_valid(log_n, tp)) {
    c = tp.contig_log;
    s = log_n - c;
    k = tp.num_strided;
  }
  uint32_t N = log_n;
  for (uint32_t i = 0; i < k; ++i) {
    const uint32_t part = s / k + (i < s % k ? 1 : 0);
    pl.strided[i] = part;
    pl.log_N[i] = N;
    N -= part;
  }
  pl.num_strided = (int)k;
  pl.contig = c;
}

}  // namespace

uint32_t strided_groups(uint64_t pairs, uint64_t positions, uint32_t tuned_bpw, uint32_t *bpw_out) {
  // about 2048 (r-range, pair group) workgroups, 1..16 pairs each, unless tuned
  uint64_t bpw = pairs * positions / 2048;
  bpw = bpw < 1 ? 1 : (bpw > 16 ? 16 : bpw);
  if (tuned_bpw) bpw = tuned_bpw;
  // grid.y is at most 65535: larger batches take more pairs per workgroup
  const uint64_t min_bpw = (pairs + STRIDED_MAX_GROUPS - 1) / STRIDED_MAX_GROUPS;
  if (bpw < min_bpw) bpw = min_bpw;
  *bpw_out = (uint32_t)bpw;
  return (uint32_t)((pairs + bpw - 1) / bpw);
}

int describe_fast_passes(const Tables &tab, const uint64_t *inputs, uint64_t *outputs, uint32_t log_n,
                         uint32_t num_cols, uint64_t stride_in, uint64_t stride_out, bool bitrev_inputs, bool inverse,
                         uint32_t log_ext, uint32_t coset_index, PassDesc *out) {
  if (log_n < 8) return -1;
  FastPlan plan;
  make_fast_plan(log_n, tuned(log_n, num_cols), plan);
  const uint64_t n = 1ull << log_n;
  FastArgs a{};
  a.tab = tab;
  a.num_cols = num_cols;
  a.log_n = log_n;
  a.log_ext = log_ext;
  a.coset_exp = log_ext ? (uint32_t)((uint64_t)coset_index << (32 - log_n - log_ext)) : 0u;
  a.n_inv = gl::inv(n % gl::P);
  const int np = plan.num_strided + 1;
  const bool single = np == 1;
  for (int k = 0; k < np; ++k) {
    // n2b runs strided passes top-down then the contiguous pass; b2n the reverse
    const int p = bitrev_inputs ? (np - 1 - k) : k;  // index in n2b order
    const bool first = k == 0, last = k == np - 1;
    PassDesc &pd = out[k];
    pd.a = a;
    pd.a.in = first ? inputs : outputs;
    pd.a.stride_in = first ? stride_in : stride_out;
    pd.a.out = outputs;
    pd.a.stride_out = stride_out;
    pd.a.first = first;
    pd.a.last = last;
    pd.positions = pd.pairs = 0;
    if (p == np - 1) {
      int mode = MODE_NONE;
      if (!inverse && log_ext && first) mode = MODE_LDE_IN;
      else if (inverse && last && log_ext) mode = MODE_LDE_OUT;
      else if (inverse && last && single) mode = MODE_NINV;
      pd.contig = true;
      pd.log_h = plan.contig;
      pd.mode = mode;
      pd.a.log_N = plan.contig;
      const uint64_t tiles = log_n >= TILE_LOG ? ((uint64_t)num_cols << (log_n - TILE_LOG))
                                               : (num_cols + (1ull << (TILE_LOG - log_n)) - 1) >> (TILE_LOG - log_n);
      if (tiles > 0xFFFFFFFFull) return -1;
      pd.a.tiles = (uint32_t)tiles;
    } else {
      const bool top = p == 0;
      int mode = MODE_NONE;
      if (top) {
        if (!inverse && log_ext && !bitrev_inputs) mode = MODE_LDE_IN;
        else if (inverse && log_ext && bitrev_inputs) mode = MODE_LDE_OUT;
        else if (inverse && !log_ext) mode = MODE_NINV;
      }
      pd.contig = false;
      pd.log_h = plan.strided[p];
      pd.mode = mode;
      pd.a.log_N = plan.log_N[p];
      pd.positions = 1ull << (pd.a.log_N - TILE_LOG);
      pd.pairs = (uint64_t)num_cols << (log_n - pd.a.log_N);
    }
  }
  return np;
}


namespace {

uint32_t brev_bits_host(uint32_t x, uint32_t bits) {
  uint32_t r = 0;
  for (uint32_t i = 0; i < bits; ++i) r |= ((x >> i) & 1u) << (bits - 1 - i);
  return r;
}

// Work-unit size when none is tuned: `large` items (tiles or pairs) per unit,
// which amortizes a unit's setup (twists, coset factors, queue pop) the way a
// standalone workgroup does, halved while the segment would have fewer than 64
// units (items / size units at each of `positions` r-ranges), so that small
// segments still spread over many workgroups.
uint32_t default_unit(uint64_t items, uint64_t positions, uint32_t large) {
  uint32_t u = large;
  while (u > 1 && ((items + u - 1) / u) * positions < 64) u /= 2;
  return u;
}

// A stage group: the same transform (log_n, direction, coset) applied to
// columns of several jobs, planned and run as one batch through per-column
// pointer tables.
struct StageKey {
  uint32_t log_n, bitrev, inv, ext, coset;
  bool operator<(const StageKey &o) const {
    if (log_n != o.log_n) return log_n < o.log_n;
    if (ext != o.ext) return ext < o.ext;  // inverse-direction stages (ext 0) first
    if (bitrev != o.bitrev) return bitrev < o.bitrev;
    if (inv != o.inv) return inv < o.inv;
    return coset < o.coset;
  }
};

struct StageMember {
  const uint64_t *in;
  uint64_t *out;
  uint64_t si, so;
  uint32_t cols;
};

struct PlanSeg {
  RaggedSeg s;
  uint32_t level, group;
};

}  // namespace

int ragged_validate_job(const bc_ragged_job &j) {
  if (j.log_n < 8 || j.log_n >= 32 || !j.inputs || !j.outputs) return BC_ERROR_INVALID_VALUE;
  const uint64_t n = 1ull << j.log_n;
  if (j.stride_in < n) return BC_ERROR_INVALID_VALUE;
  if (j.op == BC_OP_NTT || j.op == BC_OP_INTT) {
    if (j.stride_out < n) return BC_ERROR_INVALID_VALUE;
  } else if (j.op == BC_OP_LDE) {
    if (!j.coeffs || j.stride_coeffs < n || j.log_blowup < 1 || j.log_blowup > 8) return BC_ERROR_INVALID_VALUE;
    if ((uint64_t)j.log_n + j.log_blowup > OMEGA_LOG_ORDER) return BC_ERROR_INVALID_VALUE;
    if ((uint64_t)j.stride_out < (n << j.log_blowup)) return BC_ERROR_INVALID_VALUE;
  } else {
    return BC_ERROR_INVALID_VALUE;
  }
  return BC_SUCCESS;
}

int ragged_build(const bc_ragged_job *jobs, uint32_t num_jobs, const Tables &tab, RaggedHostPlan &out) {
  out = RaggedHostPlan{};
  for (uint32_t j = 0; j < num_jobs; ++j) {
    const int rc = ragged_validate_job(jobs[j]);
    if (rc != BC_SUCCESS) return rc;
  }
  // stages of every job, grouped by transform: an NTT job is one forward
  // natural->bit-reversed stage, an iNTT job one inverse bit-reversed->natural
  // stage, an LDE job that inverse stage (into coeffs) plus one forward coset
  // stage per coset (coeffs -> output block brev(c))
  std::map<StageKey, std::vector<StageMember>> groups;
  for (uint32_t j = 0; j < num_jobs; ++j) {
    const bc_ragged_job &jb = jobs[j];
    if (jb.num_cols == 0) continue;
    const uint64_t n = 1ull << jb.log_n;
    if (jb.op == BC_OP_NTT) {
      groups[{jb.log_n, 0, 0, 0, 0}].push_back({jb.inputs, jb.outputs, jb.stride_in, jb.stride_out, jb.num_cols});
    } else if (jb.op == BC_OP_INTT) {
      groups[{jb.log_n, 1, 1, 0, 0}].push_back({jb.inputs, jb.outputs, jb.stride_in, jb.stride_out, jb.num_cols});
    } else {
      groups[{jb.log_n, 1, 1, 0, 0}].push_back({jb.inputs, jb.coeffs, jb.stride_in, jb.stride_coeffs, jb.num_cols});
      for (uint32_t c = 0; c < (1u << jb.log_blowup); ++c)
        groups[{jb.log_n, 0, 0, jb.log_blowup, c}].push_back(
            {jb.coeffs, jb.outputs + (uint64_t)brev_bits_host(c, jb.log_blowup) * n, jb.stride_coeffs, jb.stride_out,
             jb.num_cols});
    }
  }
  // inverse stages (ext 0) are keyed before the coset stages of the same
  // log_n, so their pass counts are known when a coset stage needs its level
  std::map<uint32_t, uint32_t> inverse_passes;  // log_n -> passes of its inverse group
  std::vector<PlanSeg> segs;
  uint32_t gi = 0;
  for (const auto &kv : groups) {
    const StageKey &key = kv.first;
    uint64_t total_cols = 0;
    for (const StageMember &m : kv.second) total_cols += m.cols;
    if (total_cols > 0xFFFFFFFFull) return BC_ERROR_INVALID_VALUE;
    PassDesc pd[8];
    const uint64_t n = 1ull << key.log_n;
    const int np = describe_fast_passes(tab, nullptr, nullptr, key.log_n, (uint32_t)total_cols, n, n, key.bitrev != 0,
                                        key.inv != 0, key.ext, key.coset, pd);
    if (np <= 0) return BC_ERROR_INVALID_VALUE;
    uint32_t base_level = 0;
    if (key.ext) {
      auto it = inverse_passes.find(key.log_n);
      if (it == inverse_passes.end()) return BC_ERROR_INVALID_VALUE;
      base_level = it->second;
    } else if (key.inv) {
      inverse_passes[key.log_n] = (uint32_t)np;
    }
    const TuneParams tp = tuned(key.log_n, (uint32_t)total_cols);
    for (int k = 0; k < np; ++k) {
      PlanSeg ps{};
      ps.level = base_level + (uint32_t)k;
      ps.group = gi;
      ps.s.a = pd[k].a;
      ps.s.a.in = nullptr;
      ps.s.a.out = nullptr;
      // column table of this pass: the first pass reads the stage's inputs,
      // later passes work in place on its outputs
      ps.s.col_off = out.cols.size();
      for (const StageMember &m : kv.second)
        for (uint32_t c = 0; c < m.cols; ++c)
          out.cols.push_back({k == 0 ? m.in + c * m.si : m.out + c * m.so, m.out + c * m.so});
      uint32_t dm;
      if (!key.inv && !key.bitrev) dm = pd[k].mode == MODE_LDE_IN ? 1u : 0u;
      else if (key.inv && key.bitrev) dm = pd[k].mode == MODE_NINV ? 3u : 2u;
      else return BC_ERROR_INVALID_VALUE;
      if ((dm == 0 || dm == 2) && pd[k].mode != MODE_NONE) return BC_ERROR_INVALID_VALUE;
      if (pd[k].contig) {
        const bool mc = key.log_n < TILE_LOG;
        ps.s.variant = variant_id(false, pd[k].log_h, mc, dm);
        ps.s.total = pd[k].a.tiles;
        ps.s.per_unit = tp.unit_contig ? tp.unit_contig : default_unit(ps.s.total, 1, 32);
        ps.s.units = (uint32_t)((ps.s.total + ps.s.per_unit - 1) / ps.s.per_unit);
        ps.s.positions = 0;
      } else {
        ps.s.variant = variant_id(true, pd[k].log_h, false, dm);
        ps.s.total = pd[k].pairs;
        ps.s.per_unit = (uint32_t)std::min<uint64_t>(
            pd[k].pairs, tp.unit_strided ? tp.unit_strided : default_unit(pd[k].pairs, pd[k].positions, 32));
        ps.s.positions = (uint32_t)pd[k].positions;
        const uint64_t ngroups = (ps.s.total + ps.s.per_unit - 1) / ps.s.per_unit;
        if (ngroups * ps.s.positions > 0xFFFFFFFFull) return BC_ERROR_INVALID_VALUE;
        ps.s.units = (uint32_t)(ngroups * ps.s.positions);
      }
      segs.push_back(ps);
    }
    ++gi;
  }

  // queue order: level, then kernel variant, then group
  std::stable_sort(segs.begin(), segs.end(), [](const PlanSeg &a, const PlanSeg &b) {
    if (a.level != b.level) return a.level < b.level;
    if (a.s.variant != b.s.variant) return a.s.variant < b.s.variant;
    return a.group < b.group;
  });
  out.segs.resize(segs.size());
  uint64_t item = 0;
  for (uint32_t i = 0; i < segs.size(); ++i) {
    RaggedSeg r = segs[i].s;
    r.first_item = item;
    item += r.units;
    while (out.level_seg.size() <= segs[i].level) {  // first segment of each level
      out.level_seg.push_back(i);
      out.level_item.push_back(r.first_item);
    }
    out.levels = std::max(out.levels, segs[i].level + 1);
    out.segs[i] = r;
  }
  if (out.levels > RAGGED_MAX_LEVELS) return BC_ERROR_INVALID_VALUE;
  out.level_seg.push_back((uint32_t)out.segs.size());
  out.level_item.push_back(item);
  out.total_items = item;
  // chunk index: segment of item c << RAGGED_CHUNK_LOG for every chunk, and one
  // more entry so that chunk_seg[c + 1] exists for the last chunk
  const uint64_t chunks = (item >> RAGGED_CHUNK_LOG) + 1;
  out.chunk_seg.assign(chunks + 1, out.segs.empty() ? 0u : (uint32_t)out.segs.size() - 1);
  uint32_t sgi = 0;
  for (uint64_t c = 0; c < chunks && !out.segs.empty(); ++c) {
    const uint64_t it = c << RAGGED_CHUNK_LOG;
    if (it >= item) break;
    while (sgi + 1 < out.segs.size() && out.segs[sgi + 1].first_item <= it) ++sgi;
    out.chunk_seg[c] = sgi;
  }
  return BC_SUCCESS;
}

// Packed plan table: segments, chunk index, column tables (16-byte aligned).
size_t ragged_chunk_offset(const RaggedHostPlan &hp) { return std::max<size_t>(1, hp.segs.size()) * sizeof(RaggedSeg); }
static size_t cols_offset(const RaggedHostPlan &hp) {
  return (ragged_chunk_offset(hp) + hp.chunk_seg.size() * sizeof(uint32_t) + 15) / 16 * 16;
}

size_t ragged_table_bytes(const RaggedHostPlan &hp) { return cols_offset(hp) + hp.cols.size() * sizeof(ColPtrs); }

void ragged_pack_table(const RaggedHostPlan &hp, void *host_buf, const void *device_buf) {
  uint8_t *b = (uint8_t *)host_buf;
  const ColPtrs *dev_cols = (const ColPtrs *)((const uint8_t *)device_buf + cols_offset(hp));
  RaggedSeg *segs = (RaggedSeg *)b;
  for (size_t i = 0; i < hp.segs.size(); ++i) {
    segs[i] = hp.segs[i];
    segs[i].a.cols = dev_cols + hp.segs[i].col_off;
  }
  memcpy(b + ragged_chunk_offset(hp), hp.chunk_seg.data(), hp.chunk_seg.size() * sizeof(uint32_t));
  if (!hp.cols.empty()) memcpy(b + cols_offset(hp), hp.cols.data(), hp.cols.size() * sizeof(ColPtrs));
}

}  // namespace bc

// Register-resident NTT pass bodies for gfx950 (device code shared by the
// standalone pass kernels and the ragged persistent kernel).
//
// The math of a pass is the one described in ntt_kernels.hip: inside blocks of
// N points, for every r < R = N / H, an H-point sub-DFT over j (positions
// j * R + r), with the twist omega_N^(r * brev(j)).  Two tile shapes cover
// every pass of a transform with log_n >= 8:
//
//   contig<h>  (8 <= h <= 12): R = 1, the last pass in n2b order.  A tile is
//              4096 consecutive points = 2^(12-h) sub-DFTs.  Levels of radix
//              16, 16 and 2^(h-8) run in registers, with LDS exchanges
//              between them; for plain and n^-1 passes the one between
//              levels 1 and 2 (and in b2n the load regrouping) stays inside
//              a wave, with no workgroup barrier.
//   strided<h> (4 <= h <= 8):  R >= 16, every other pass.  A tile is
//              2^(12-h) consecutive r times the h-point sub-DFT, so each row of
//              the tile is >= 128 contiguous bytes.  Levels of radix 16 and
//              2^(h-4), one LDS exchange (none for h = 4).
//
// Every workgroup has 256 threads holding 16 values each.  Global loads go
// straight into the first level's registers and the last level stores from
// registers.  Column values are read and written once per pass, so those
// loads and stores are non-temporal (`nt`: streamed past the caches, which
// keep the twiddle tables); measured 2% faster per 256 x 2^20 call than
// ordinary loads and stores.  Per-thread level twists depend only on the
// thread and are computed once per workgroup; the strided pass twist depends
// on the tile's r range, so a workgroup walks several (column, block) pairs
// that share it and computes it once.  n2b runs levels top-down (DIF), b2n
// bottom-up (DIT) as the transpose.
#pragma once
#include "ntt_device.hpp"

namespace bc {
namespace pass {

constexpr uint32_t NT = 256;  // threads per workgroup

// Padded LDS index: one pad slot per 16 keeps stride-16 accesses on distinct banks.
__device__ __forceinline__ uint32_t pidx(uint32_t e) { return e + (e >> 4); }
constexpr uint32_t LDS_PADDED = TILE + TILE / 16;

// Level-twist table of a radix-16 level in blocks of M = 2^LM points: entry
// (i - 1) * S + rp = omega_M^(+-rp * brev4(i)), S = M / 16.
template <int LM, bool INV> __device__ __forceinline__ const uint64_t *lvl_table(const Tables &tab) {
  return tab.lvl_tw + (INV ? LVL_TW_BLOCK : 0u) + lvl_tw_offset(LM);
}

// Copies a level-twist table into LDS (once per workgroup).
template <int LM, bool INV> __device__ __forceinline__ void stage_level_twists(uint64_t *dst, const Tables &tab) {
  constexpr uint32_t LEN = 15u << (LM - 4);
  const uint64_t *src = lvl_table<LM, INV>(tab);
  for (uint32_t k = threadIdx.x; k < LEN; k += NT) dst[k] = src[k];
}

// x[i] *= tw[(i - 1) * S + rp] for i = 1..15 (tw in LDS).
template <int LM>
__device__ __forceinline__ void level_twist_lds(uint64_t *x, const uint64_t *tw, uint32_t rp) {
  constexpr uint32_t S = 1u << (LM - 4);
#pragma unroll
  for (int i = 1; i < 16; ++i) x[i] = gl::mul(x[i], tw[(i - 1) * S + rp]);
}

// -------------------------------------------------------------------------
// contig<h>

// Tile element e of thread t's group g, point i, at level L (width A) of an
// h-point sub-DFT whose levels are 4, 4, h - 8.
template <int LH, int L, int A> __device__ __forceinline__ uint32_t celem(uint32_t t, int g, int i) {
  constexpr int C = 4 * L, LM = LH - C, LS = LM - A;
  const uint32_t gi = t + NT * (uint32_t)g;
  const uint32_t rp = gi & ((1u << LS) - 1u);
  const uint32_t b = (gi >> LS) & ((1u << C) - 1u);
  const uint32_t qb = gi >> (LH - A);
  return (qb << LH) + (b << LM) + ((uint32_t)i << LS) + rp;
}

// Coset scaling (MODE_LDE_IN: s^idx on the inputs; MODE_LDE_OUT: n^-1 s^-idx
// on the outputs).  A thread's 16 points sit at indices idx0 + part(i) whose
// parts use bits the thread's base index leaves zero, so s^(idx0 + part(i)) =
// s^idx0 * s^part(i): the 16 part factors are computed once per workgroup
// (into LDS, by threads 0..15) and each thread computes one power per tile.
template <int MODE> __device__ __forceinline__ uint64_t coset_scale_pow(const FastArgs &a, uint32_t idx) {
  return coset_pow(a.tab, a.coset_exp, idx, MODE == MODE_LDE_OUT);
}
template <int MODE> __device__ __forceinline__ uint64_t coset_part_factor(const FastArgs &a, uint32_t part) {
  const uint64_t v = coset_scale_pow<MODE>(a, part);
  return MODE == MODE_LDE_OUT ? gl::mul(a.n_inv, v) : v;
}

// Level twists with a coset factor folded in (strided passes): group rp's
// twists become tw * c(rp) with c(rp) = s^(+-rp 2^log_r), and dst0[rp] = c(rp)
// scales the group's point 0, whose twist is 1.
template <int LM, bool INV, int MODE>
__device__ __forceinline__ void stage_level_twists_coset(uint64_t *dst, uint64_t *dst0, const FastArgs &a,
                                                         uint32_t log_r) {
  constexpr uint32_t S = 1u << (LM - 4), LEN = 15u * S;
  const uint64_t *src = lvl_table<LM, INV>(a.tab);
  for (uint32_t k = threadIdx.x; k < LEN; k += NT)
    dst[k] = gl::mul(src[k], coset_scale_pow<MODE>(a, (k & (S - 1u)) << log_r));
  if (threadIdx.x < S) dst0[threadIdx.x] = coset_scale_pow<MODE>(a, threadIdx.x << log_r);
}

// Column addressing: one matrix (base + c * stride) for the standalone
// launches, or a per-column pointer table (CT) for ragged job groups.
template <bool CT> __device__ __forceinline__ const uint64_t *in_col(const FastArgs &a, uint64_t c) {
  if constexpr (CT) return a.cols[c].in;
  else return a.in + c * a.stride_in;
}
template <bool CT> __device__ __forceinline__ uint64_t *out_col(const FastArgs &a, uint64_t c) {
  if constexpr (CT) return a.cols[c].out;
  else return a.out + c * a.stride_out;
}

// Addressing of a contiguous-pass tile: 4096 consecutive points of one column,
// or (MC: single-pass transforms shorter than 4096) 2^(12-h) whole columns.
template <int LH, bool MC, bool CT> struct ContigTile {
  const uint64_t *in;  // !MC: the tile's first input point
  uint64_t *out;       // !MC: the tile's first output point
  uint32_t pos0, col0;
  __device__ __forceinline__ ContigTile(const FastArgs &a, uint32_t tile) {
    if constexpr (!MC) {
      const uint32_t sh = a.log_n - TILE_LOG;
      const uint64_t col = tile >> sh;
      pos0 = (tile & ((1u << sh) - 1u)) << TILE_LOG;
      in = in_col<CT>(a, col) + pos0;
      out = out_col<CT>(a, col) + pos0;
      col0 = 0;
    } else {
      col0 = tile << (12 - LH);
      pos0 = 0;
      in = nullptr;
      out = nullptr;
    }
  }
  // element e of the tile; nullptr for MC slots past the last column
  __device__ __forceinline__ const uint64_t *src(const FastArgs &a, uint32_t e) const {
    if constexpr (!MC) return in + e;
    const uint32_t c = col0 + (e >> LH);
    return c < a.num_cols ? in_col<CT>(a, c) + (e & ((1u << LH) - 1u)) : nullptr;
  }
  __device__ __forceinline__ uint64_t *dst(const FastArgs &a, uint32_t e) const {
    if constexpr (!MC) return out + e;
    const uint32_t c = col0 + (e >> LH);
    return c < a.num_cols ? out_col<CT>(a, c) + (e & ((1u << LH) - 1u)) : nullptr;
  }
  __device__ __forceinline__ uint32_t pos(uint32_t e) const {
    if constexpr (MC) return e & ((1u << LH) - 1u);
    else return pos0 + e;
  }
};

// Wave-region mappings of a 4096-point tile, used from level 1 on when a
// transform has three register levels.  Tile bits 10-11 are the wave index in
// the level-1 mapping (celem<LH, 1, 4> for every 9 <= LH <= 12), so each wave
// then owns the 1024 points (and LDS slots) w * 1024 + [0, 1024):
//   wl_run: thread t holds 16 consecutive points t * 16 + k (the last level's
//           2^(LH-8)-point runs, grouped by k's low bits);
//   wl_row: register i of lane l of wave w holds point w * 1024 + i * 64 + l
//           (coalesced: 64 consecutive points per register).
__device__ __forceinline__ uint32_t wl_run(uint32_t t, int k) { return (t << 4) + (uint32_t)k; }
__device__ __forceinline__ uint32_t wl_row(uint32_t t, int i) {
  return ((t >> 6) << 10) + ((uint32_t)i << 6) + (t & 63u);
}

// Orders one wave's LDS accesses across its lanes.  A wave's LDS operations
// execute in program order, so this only keeps the compiler from moving LDS
// loads and stores across it; no workgroup barrier is involved.
__device__ __forceinline__ void wave_lds_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Loads a tile's 16 values per thread in the level-0 mapping (lanes on
// consecutive points), or (ROW) in the wl_row mapping.
template <int LH, bool MC, bool CT, bool ROW = false>
__device__ __forceinline__ void contig_load(const FastArgs &a, uint32_t tile, uint32_t t, uint64_t *x) {
  const ContigTile<LH, MC, CT> ct(a, tile);
#pragma unroll
  for (int i = 0; i < 16; ++i) {
    const uint64_t *q = ct.src(a, ROW ? wl_row(t, i) : celem<LH, 0, 4>(t, 0, i));
    x[i] = q ? __builtin_nontemporal_load(q) : 0ull;
  }
}

// LDS a pass needs: the padded tile plus a staged level-twist table.
constexpr uint32_t LDS_TW_LEN = 15u << 4;  // largest staged table (M = 256)
struct PassLds {
  uint64_t tile[LDS_PADDED];
  uint64_t tw[LDS_TW_LEN];
  uint64_t cf[16];  // coset part factors (LDE modes)
  uint64_t cr[16];  // coset factors of the strided level-0 groups (LDE modes)
};

// Runs tiles tile_begin, tile_begin + step, ... < tile_end of a contiguous
// pass.  Called by a whole workgroup; ends with a barrier.
template <int LH, bool INV, bool DIT, int MODE, bool MC, bool CT = false>
__device__ __forceinline__ void contig_run(const FastArgs &a, PassLds &sh, uint32_t tile_begin, uint32_t tile_end,
                                           uint32_t step) {
  constexpr int NL = LH > 8 ? 3 : 2;
  constexpr int A2 = NL == 3 ? LH - 8 : 4;  // width of the third level
  constexpr int G2 = 16 >> A2;              // third-level groups per thread
  constexpr int LL = NL - 1, AL = NL == 3 ? A2 : 4, GL = 16 >> AL;  // last level
  // WL: with three levels and no coset factors, the exchange between levels 1
  // and 2 stays inside each wave's own 1024 points (wl_run) and needs no
  // workgroup barrier.  b2n also loads in the wl_row mapping, so its
  // regrouping exchange is wave-local too (two barriers per tile instead of
  // four); n2b keeps its store in the level-0 mapping (three barriers), which
  // measured faster than wl_row stores.
  constexpr bool WL = NL == 3 && !MC && (MODE == MODE_NONE || MODE == MODE_NINV);
  uint64_t *lds = sh.tile;
  uint64_t *tw1 = sh.tw;
  const uint32_t t = threadIdx.x;
  // LDS slots of a thread's 16 points in the fixed mappings, as one base and
  // a constant stride, so that they are addressed with immediate offsets
  // instead of 16 loop-invariant address registers per mapping:
  // pidx(e0 + i 2^S) = pidx(e0) + i (2^S + 2^(S-4)) when S >= 4 and bits
  // S..S+3 of e0 are zero (level 0: S = h - 4; level 1: S = h - 8, so h = 12
  // only; wl_row: S = 6; wl_run: consecutive slots).
  uint64_t *const s0b = lds + pidx(celem<LH, 0, 4>(t, 0, 0));
  uint64_t *const s1b = lds + pidx(celem<LH, 1, 4>(t, 0, 0));
  uint64_t *const srunb = lds + pidx(wl_run(t, 0));
  uint64_t *const srowb = lds + pidx(wl_row(t, 0));
  auto s0 = [&](int i) -> uint64_t & { return s0b[i * ((1 << (LH - 4)) + (1 << (LH - 8)))]; };
  auto s1 = [&](int i) -> uint64_t & {
    if constexpr (LH >= 12) return s1b[i * ((1 << (LH - 8)) + (1 << (LH - 12)))];
    else return lds[pidx(celem<LH, 1, 4>(t, 0, i))];
  };
  auto srun = [&](int k) -> uint64_t & { return srunb[k]; };
  auto srow = [&](int i) -> uint64_t & { return srowb[i * 68]; };

  // level-0 twists: 15 per thread, read from the (cache-resident) level table
  // at each use, which keeps 30 VGPRs free
  const uint64_t *tw0 = lvl_table<LH, INV>(a.tab) + (t & ((1u << (LH - 4)) - 1u));
  // level-1 twists (at most 240 entries) in LDS
  if constexpr (NL == 3) stage_level_twists<LH - 4, INV>(tw1, a.tab);
  const uint32_t rp1 = t & ((1u << (LH - 8)) - 1u);
  // coset part factors: point i of the level-0 mapping is at tile position
  // pos(e_0) + i 2^(LH-4); the scaled index is that position, or its log_n-bit
  // reversal (DIT inputs, DIF outputs), whose part is brev4(i) 2^(log_n-LH)
  constexpr bool SC = MODE == MODE_LDE_IN || MODE == MODE_LDE_OUT;
  constexpr bool SC_BREV = (MODE == MODE_LDE_IN) == DIT;
  if constexpr (SC) {
    if (t < 16) {
      const uint32_t part = SC_BREV ? (brev_bits(t, 4) << (a.log_n - LH)) : (t << (LH - 4));
      sh.cf[t] = coset_part_factor<MODE>(a, part);
    }
  }

  // b2n with WL loads in the wl_row mapping, everything else in level 0's
  constexpr bool ROW_LOAD = WL && DIT;
  uint32_t tile = tile_begin;
  uint64_t pf[16];
  if (tile < tile_end) contig_load<LH, MC, CT, ROW_LOAD>(a, tile, t, pf);
  __syncthreads();  // tw1 and coset part factors staged

  for (; tile < tile_end; tile += step) {
    const ContigTile<LH, MC, CT> ct(a, tile);
    uint64_t x[16];
#pragma unroll
    for (int i = 0; i < 16; ++i) x[i] = pf[i];
    {
      const uint32_t next = tile + step;
      if (next < tile_end) contig_load<LH, MC, CT, ROW_LOAD>(a, next, t, pf);  // prefetch: in flight during this tile's math
    }
    uint64_t cfb = 0;  // coset factor of the thread's point 0 in this tile
    if constexpr (SC) {
      const uint32_t pos = ct.pos(celem<LH, 0, 4>(t, 0, 0));
      cfb = coset_scale_pow<MODE>(a, SC_BREV ? brev_bits(pos, a.log_n) : pos);
    }
    if constexpr (MODE == MODE_LDE_IN) {
#pragma unroll
      for (int i = 0; i < 16; ++i) x[i] = gl::mul(x[i], gl::mul(cfb, sh.cf[i]));
    } else {
      if (a.first) {
#pragma unroll
        for (int i = 0; i < 16; ++i) x[i] = gl::canon(x[i]);  // field ops take canonical values
      }
    }

    if constexpr (!DIT) {
      dft_dif<4, INV>(x);
#pragma unroll
      for (int i = 1; i < 16; ++i) x[i] = gl::mul(x[i], tw0[(i - 1) << (LH - 4)]);
#pragma unroll
      for (int i = 0; i < 16; ++i) s0(i) = x[i];
      __syncthreads();
#pragma unroll
      for (int i = 0; i < 16; ++i) x[i] = s1(i);
      dft_dif<4, INV>(x);
      if constexpr (WL) {
        level_twist_lds<LH - 4>(x, tw1, rp1);
        // each thread rewrites the slots it read, inside its wave's region
#pragma unroll
        for (int i = 0; i < 16; ++i) s1(i) = x[i];
        wave_lds_sync();
#pragma unroll
        for (int k = 0; k < 16; ++k) x[k] = srun(k);
#pragma unroll
        for (int g = 0; g < G2; ++g) dft_dif<A2, INV>(x + g * (1 << A2));
#pragma unroll
        for (int k = 0; k < 16; ++k) srun(k) = x[k];
        __syncthreads();  // the level-0 mapping crosses waves
#pragma unroll
        for (int i = 0; i < 16; ++i) {
          const uint32_t e = celem<LH, 0, 4>(t, 0, i);
          uint64_t v = s0(i);
          if constexpr (MODE == MODE_NINV) v = gl::mul(v, a.n_inv);
          __builtin_nontemporal_store(v, ct.dst(a, e));
        }
      } else if constexpr (NL == 3) {
        level_twist_lds<LH - 4>(x, tw1, rp1);
#pragma unroll
        for (int i = 0; i < 16; ++i) s1(i) = x[i];
        __syncthreads();
#pragma unroll
        for (int g = 0; g < G2; ++g)
#pragma unroll
          for (int i = 0; i < (1 << A2); ++i) x[g * (1 << A2) + i] = lds[pidx(celem<LH, 2, A2>(t, g, i))];
#pragma unroll
        for (int g = 0; g < G2; ++g) dft_dif<A2, INV>(x + g * (1 << A2));
      }
      // regroup the last level's runs into the coalesced level-0 mapping
      // through LDS (each thread rewrites the slots it read, so no barrier
      // is needed before the write), then store
      if constexpr (!WL) {
#pragma unroll
        for (int g = 0; g < GL; ++g)
#pragma unroll
          for (int i = 0; i < (1 << AL); ++i) lds[pidx(celem<LH, LL, AL>(t, g, i))] = x[g * (1 << AL) + i];
        __syncthreads();
#pragma unroll
        for (int i = 0; i < 16; ++i) {
          const uint32_t e = celem<LH, 0, 4>(t, 0, i);
          uint64_t v = s0(i);
          if constexpr (MODE == MODE_NINV) v = gl::mul(v, a.n_inv);
          if constexpr (MODE == MODE_LDE_OUT) v = gl::mul(v, gl::mul(cfb, sh.cf[i]));
          uint64_t *q = ct.dst(a, e);
          if (q) __builtin_nontemporal_store(v, q);
        }
      }
    } else if constexpr (WL) {
      // loaded in the wl_row mapping: regroup into runs inside the wave's region
#pragma unroll
      for (int i = 0; i < 16; ++i) srow(i) = x[i];
      wave_lds_sync();
#pragma unroll
      for (int k = 0; k < 16; ++k) x[k] = srun(k);
#pragma unroll
      for (int g = 0; g < G2; ++g) dft_dit<A2, INV>(x + g * (1 << A2));
#pragma unroll
      for (int k = 0; k < 16; ++k) srun(k) = x[k];
      wave_lds_sync();
#pragma unroll
      for (int i = 0; i < 16; ++i) x[i] = s1(i);
      level_twist_lds<LH - 4>(x, tw1, rp1);
      dft_dit<4, INV>(x);
      // each thread rewrites the slots it read; the level-0 mapping crosses waves
#pragma unroll
      for (int i = 0; i < 16; ++i) s1(i) = x[i];
      __syncthreads();
#pragma unroll
      for (int i = 0; i < 16; ++i) x[i] = s0(i);
#pragma unroll
      for (int i = 1; i < 16; ++i) x[i] = gl::mul(x[i], tw0[(i - 1) << (LH - 4)]);
      dft_dit<4, INV>(x);
#pragma unroll
      for (int i = 0; i < 16; ++i) {
        const uint32_t e = celem<LH, 0, 4>(t, 0, i);
        uint64_t v = x[i];
        if constexpr (MODE == MODE_NINV) v = gl::mul(v, a.n_inv);
        __builtin_nontemporal_store(v, ct.dst(a, e));
      }
    } else {
      // regroup the coalesced load into the last level's runs through LDS
#pragma unroll
      for (int i = 0; i < 16; ++i) s0(i) = x[i];
      __syncthreads();
#pragma unroll
      for (int g = 0; g < GL; ++g)
#pragma unroll
        for (int i = 0; i < (1 << AL); ++i) x[g * (1 << AL) + i] = lds[pidx(celem<LH, LL, AL>(t, g, i))];
      if constexpr (NL == 3) {
#pragma unroll
        for (int g = 0; g < G2; ++g) dft_dit<A2, INV>(x + g * (1 << A2));
#pragma unroll
        for (int g = 0; g < G2; ++g)
#pragma unroll
          for (int i = 0; i < (1 << A2); ++i) lds[pidx(celem<LH, 2, A2>(t, g, i))] = x[g * (1 << A2) + i];
        __syncthreads();
#pragma unroll
        for (int i = 0; i < 16; ++i) x[i] = s1(i);
        level_twist_lds<LH - 4>(x, tw1, rp1);
      }
      dft_dit<4, INV>(x);
#pragma unroll
      for (int i = 0; i < 16; ++i) s1(i) = x[i];
      __syncthreads();
#pragma unroll
      for (int i = 0; i < 16; ++i) x[i] = s0(i);
#pragma unroll
      for (int i = 1; i < 16; ++i) x[i] = gl::mul(x[i], tw0[(i - 1) << (LH - 4)]);
      dft_dit<4, INV>(x);
#pragma unroll
      for (int i = 0; i < 16; ++i) {
        const uint32_t e = celem<LH, 0, 4>(t, 0, i);
        uint64_t v = x[i];
        if constexpr (MODE == MODE_NINV) v = gl::mul(v, a.n_inv);
        if constexpr (MODE == MODE_LDE_OUT) v = gl::mul(v, gl::mul(cfb, sh.cf[i]));
        uint64_t *q = ct.dst(a, e);
        if (q) __builtin_nontemporal_store(v, q);
      }
    }
    __syncthreads();  // LDS is reused by the next tile
  }
}

// -------------------------------------------------------------------------
// strided<h>

// Runs (column, block) pairs [p_begin, p_end) at r-range rho of a strided
// pass.  Called by a whole workgroup; ends with a barrier.
template <int LH, bool INV, bool DIT, int MODE, bool CT = false>
__device__ __forceinline__ void strided_run(const FastArgs &a, PassLds &sh, uint32_t rho, uint64_t p_begin,
                                            uint64_t p_end) {
  constexpr int LT = 12 - LH;         // log2 sub-DFTs (consecutive r) per tile
  constexpr int T = 1 << LT;
  constexpr int NL = LH > 4 ? 2 : 1;
  constexpr int A1 = NL == 2 ? LH - 4 : 4;  // second level width
  constexpr int G1 = 16 >> A1;
  constexpr int LS0 = LH - 4;               // level-0 stride (log2)
  uint64_t *lds = sh.tile;
  uint64_t *tw0 = sh.tw;
  const uint32_t t = threadIdx.x;
  const uint32_t log_r = a.log_N - LH;
  const uint32_t r0 = rho << LT;

  // level-0 mapping: q = t % T, rp = t / T, j = i * 2^LS0 + rp
  const uint32_t q0 = t & (T - 1u), rp0 = t >> LT;
  // last-level mapping (level 1, or level 0 when h = 4): q = gi % T, b = gi / T, j = b * 2^A1 + i
  auto lastj = [&](int g, int i) -> uint32_t {
    if constexpr (NL == 2) return (((t + NT * (uint32_t)g) >> LT) << A1) + (uint32_t)i;
    else return ((uint32_t)i << LS0) + rp0;
  };
  auto lastq = [&](int g) -> uint32_t {
    if constexpr (NL == 2) return (t + NT * (uint32_t)g) & (T - 1u);
    else return q0;
  };
  // Element offsets from the block base (r0 included).  Level-0 mapping:
  // base0 + i * step0; last-level mapping: basel(g) + i * stepl.
  const uint64_t step0 = 1ull << (LS0 + log_r);
  const uint64_t base0 = ((uint64_t)rp0 << log_r) + q0;
  const uint64_t stepl = NL == 2 ? (1ull << log_r) : step0;
  auto basel = [&](int g) -> uint64_t {
    if constexpr (NL == 2) return ((uint64_t)lastj(g, 0) << log_r) + lastq(g);
    else return base0;
  };

  // Coset scaling (this is the top pass: one block per column).  The point of
  // level-0 slot i sits at column position i 2^(log_N - 4) + rp 2^log_r + r
  // (r = r0 + q), and its factor s^(+-pos) splits into three parts, each
  // folded where it costs no extra multiply per point: s^(+-r) is common to
  // the whole sub-DFT at r and joins its pass twist; s^(+-rp 2^log_r) is
  // common to one level-0 DFT and joins that level's twists (plus point 0);
  // s^(+-i 2^(log_N - 4)) multiplies slot i (n2b inputs, b2n outputs) from a
  // 16-entry LDS table.  n^-1 of the inverse coset transform joins the pass
  // twist.
  constexpr bool SC = MODE == MODE_LDE_IN || MODE == MODE_LDE_OUT;
  if constexpr (NL == 2) {
    if constexpr (SC) stage_level_twists_coset<LH, INV, MODE>(tw0, sh.cr, a, log_r);
    else stage_level_twists<LH, INV>(tw0, a.tab);
  }
  if constexpr (SC) {
    if (t < 16) sh.cf[t] = coset_scale_pow<MODE>(a, t << (a.log_N - 4));
  }

  // pass twist omega_N^((r0 + q) * brev_h(j)) in the last-level mapping, with
  // n^-1 folded in when this pass carries the inverse scaling
  uint64_t pt[16];
#pragma unroll
  for (int g = 0; g < G1; ++g) {
    uint64_t cq = 0;
    if constexpr (SC) cq = coset_scale_pow<MODE>(a, r0 + lastq(g));
#pragma unroll
    for (int i = 0; i < (1 << A1); ++i) {
      const uint32_t j = lastj(g, i);
      uint32_t e = ((r0 + lastq(g)) * brev_bits(j, LH)) << (32u - a.log_N);
      if (INV) e = 0u - e;
      uint64_t w = pow_w(a.tab, e);
      if constexpr (MODE == MODE_NINV || MODE == MODE_LDE_OUT) w = gl::mul(w, a.n_inv);
      if constexpr (SC) w = gl::mul(w, cq);
      pt[g * (1 << A1) + i] = w;
    }
  }

  const uint32_t blocks_per_col_log = a.log_n - a.log_N;
  // offset of pair p's r-range inside its column
  auto pair_pos = [&](uint64_t p) -> uint64_t {
    return ((p & ((1ull << blocks_per_col_log) - 1ull)) << a.log_N) + r0;
  };
  // loads one pair's 16 values in the first level's mapping
  auto load = [&](uint64_t p, uint64_t *v) {
    const uint64_t *in = in_col<CT>(a, p >> blocks_per_col_log) + pair_pos(p);
    if constexpr (!DIT) {
      const uint64_t *q = in + base0;
#pragma unroll
      for (int k = 0; k < 16; ++k, q += step0) v[k] = __builtin_nontemporal_load(q);
    } else {
#pragma unroll
      for (int g = 0; g < G1; ++g) {
        const uint64_t *q = in + basel(g);
#pragma unroll
        for (int i = 0; i < (1 << A1); ++i, q += stepl) v[g * (1 << A1) + i] = __builtin_nontemporal_load(q);
      }
    }
  };

  uint64_t pf[16];
  if (p_begin < p_end) load(p_begin, pf);
  __syncthreads();  // tw0 and coset factors staged
  for (uint64_t p = p_begin; p < p_end; ++p) {
    uint64_t x[16];
#pragma unroll
    for (int i = 0; i < 16; ++i) x[i] = pf[i];
    if (p + 1 < p_end) load(p + 1, pf);  // prefetch: in flight during this pair's math
    uint64_t *out = out_col<CT>(a, p >> blocks_per_col_log) + pair_pos(p);
    if constexpr (!DIT) {
      if constexpr (MODE == MODE_LDE_IN) {
        if (a.first) x[0] = gl::canon(x[0]);  // field ops take canonical values
#pragma unroll
        for (int i = 1; i < 16; ++i) x[i] = gl::mul_uniform(x[i], sh.cf[i]);
      } else if (a.first) {
#pragma unroll
        for (int i = 0; i < 16; ++i) x[i] = gl::canon(x[i]);  // field ops take canonical values
      }
      dft_dif<4, INV>(x);
      if constexpr (NL == 2) {
        level_twist_lds<LH>(x, tw0, rp0);
        if constexpr (SC) x[0] = gl::mul(x[0], sh.cr[rp0]);
        // LDS row-major by j: element (q, j) at j * T + q
#pragma unroll
        for (int i = 0; i < 16; ++i) lds[((((uint32_t)i << LS0) + rp0) << LT) + q0] = x[i];
        __syncthreads();
#pragma unroll
        for (int g = 0; g < G1; ++g)
#pragma unroll
          for (int i = 0; i < (1 << A1); ++i) x[g * (1 << A1) + i] = lds[(lastj(g, i) << LT) + lastq(g)];
        __syncthreads();  // the next iteration rewrites LDS
#pragma unroll
        for (int g = 0; g < G1; ++g) dft_dif<A1, INV>(x + g * (1 << A1));
      }
#pragma unroll
      for (int g = 0; g < G1; ++g) {
        uint64_t *q = out + basel(g);
#pragma unroll
        for (int i = 0; i < (1 << A1); ++i, q += stepl) {
          const int k = g * (1 << A1) + i;
          __builtin_nontemporal_store(gl::mul(x[k], pt[k]), q);
        }
      }
    } else {
      // the pass twist makes every value canonical
#pragma unroll
      for (int k = 0; k < 16; ++k) x[k] = gl::mul(x[k], pt[k]);
      if constexpr (NL == 2) {
#pragma unroll
        for (int g = 0; g < G1; ++g) dft_dit<A1, INV>(x + g * (1 << A1));
#pragma unroll
        for (int g = 0; g < G1; ++g)
#pragma unroll
          for (int i = 0; i < (1 << A1); ++i) lds[(lastj(g, i) << LT) + lastq(g)] = x[g * (1 << A1) + i];
        __syncthreads();
#pragma unroll
        for (int i = 0; i < 16; ++i) x[i] = lds[((((uint32_t)i << LS0) + rp0) << LT) + q0];
        __syncthreads();
        level_twist_lds<LH>(x, tw0, rp0);
        if constexpr (SC) x[0] = gl::mul(x[0], sh.cr[rp0]);
      }
      dft_dit<4, INV>(x);
      uint64_t *q = out + base0;
#pragma unroll
      for (int i = 0; i < 16; ++i, q += step0) {
        uint64_t v = x[i];
        if constexpr (MODE == MODE_LDE_OUT) {
          if (i > 0) v = gl::mul_uniform(v, sh.cf[i]);
        }
        __builtin_nontemporal_store(v, q);
      }
    }
  }
  __syncthreads();  // LDS (tile and staged twists) is reused by the caller's next unit
}

}  // namespace pass
}  // namespace bc

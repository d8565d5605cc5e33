// Batched Goldilocks NTT kernels for gfx950.
//
// Semantics (reference: src/ntt.rs:527-583 and native/ntt.cu:105-122):
//   forward:  X_k = sum_i x_i s^i w^(ik)
//   inverse:  y_i = n^-1 s^-i sum_k x_k w^(-ik)
// with w = omega_n, s = g * omega_{n 2^ext}^coset when log_ext > 0 (else 1).
// n2b: natural input, bit-reversed output; b2n: bit-reversed input, natural
// output.
//
// Decomposition.  A transform of 2^L points is split into HBM passes
// L = h_0 + h_1 + ... (plan_passes).  In n2b order pass p works inside blocks
// of N = 2^(L - h_0 - ... - h_{p-1}) points.  Writing an index inside a block
// as i = j * R + r (R = N / H, H = 2^h_p), the pass runs, for every r, an
// H-point natural->bit-reversed DFT over j (root omega_H), then multiplies the
// value at position j by the twist omega_N^(r * brev(j)).  The next pass then
// works on blocks of R contiguous points.  This is the four-step recursion of
// the DFT; the last pass has R = 1 and no twist.
//
// b2n is the transpose of n2b: the same diagonal twists and the same
// symmetric sub-DFTs applied in the reverse order, so passes run last-first,
// each pass applies its twist before its sub-DFT, and sub-DFTs use the
// bit-reversed-input (decimation in time) form.
//
// Inside a pass one workgroup owns a 4096-point tile held in LDS: T = 4096/H
// sub-transforms.  The H-point sub-DFT is again split into radix-16 levels
// computed in registers (16 values per thread); their internal twiddles are
// powers of omega_16, which are powers of two mod p, so they cost shifts, not
// multiplications.  Between levels the same twist rule applies with N -> M
// (the level's block) and comes from the coarse omega table.
#include "ntt_device.hpp"

namespace bc {

namespace {

// LDS position of point j of sub-transform q.  The XOR only permutes the low
// four bits of j (by a function of q and j >> 4), which spreads both the
// strided tile loads and the short-stride level accesses over the banks.
__device__ __forceinline__ uint32_t lds_pos(uint32_t q, uint32_t j, uint32_t log_h) {
  const uint32_t m = log_h >= 4 ? 15u : ((1u << log_h) - 1u);
  return (q << log_h) + (j ^ (((j >> 4) + q) & m));
}

// One register level: radix-2^A sub-DFTs at stride S = M / 2^A inside blocks
// of M = H >> c points, with the level twist omega_M^(r' * k).
template <int A, bool INV, bool DIT>
__device__ __forceinline__ void run_level(uint64_t *lds, const Tables &tab, uint32_t log_h, uint32_t c) {
  constexpr int NA = 1 << A;
  constexpr int U = 16 >> A;  // groups per thread
  const uint32_t log_m = log_h - c;
  const uint32_t log_s = log_m - A;
  for (int u = 0; u < U; ++u) {
    const uint32_t g = threadIdx.x + u * TILE_THREADS;
    const uint32_t rp = g & ((1u << log_s) - 1u);
    const uint32_t b = (g >> log_s) & ((1u << c) - 1u);
    const uint32_t q = g >> (log_h - A);
    const uint32_t j0 = (b << log_m) + rp;
    uint64_t x[NA];
    uint32_t pos[NA];
#pragma unroll
    for (int i = 0; i < NA; ++i) {
      pos[i] = lds_pos(q, j0 + ((uint32_t)i << log_s), log_h);
      x[i] = lds[pos[i]];
    }
    if (DIT) {
      if (log_s > 0) {
#pragma unroll
        for (int i = 1; i < NA; ++i) {
          uint32_t e = (rp * brev_bits(i, A)) << (32u - log_m);
          if (INV) e = 0u - e;
          x[i] = gl::mul(x[i], pow_w_coarse(tab, e));
        }
      }
      dft_dit<A, INV>(x);
    } else {
      dft_dif<A, INV>(x);
      if (log_s > 0) {
#pragma unroll
        for (int i = 1; i < NA; ++i) {
          uint32_t e = (rp * brev_bits(i, A)) << (32u - log_m);
          if (INV) e = 0u - e;
          x[i] = gl::mul(x[i], pow_w_coarse(tab, e));
        }
      }
    }
#pragma unroll
    for (int i = 0; i < NA; ++i) lds[pos[i]] = x[i];
  }
}

template <bool INV, bool DIT>
__device__ __forceinline__ void run_level_dispatch(uint64_t *lds, const Tables &tab, uint32_t log_h, uint32_t c,
                                                   uint32_t a) {
  switch (a) {
  case 1: run_level<1, INV, DIT>(lds, tab, log_h, c); break;
  case 2: run_level<2, INV, DIT>(lds, tab, log_h, c); break;
  case 3: run_level<3, INV, DIT>(lds, tab, log_h, c); break;
  default: run_level<4, INV, DIT>(lds, tab, log_h, c); break;
  }
}

// Geometry of one tile element.
struct Elem {
  uint32_t q;     // sub-transform within the tile
  uint32_t j;     // point within the sub-transform
  uint32_t col;   // column
  uint32_t r;     // offset within the block (the twist's r)
  uint64_t pos;   // position within the column
  bool valid;
};

// Maps the e-th element of a tile, in memory order, to its coordinates.
// Consecutive e walk consecutive addresses: r first (up to the tile's T
// sub-transforms), then j, then blocks.
__device__ __forceinline__ Elem tile_elem(const PassArgs &a, uint64_t tile, uint32_t e) {
  const uint32_t log_h = a.log_H;
  const uint32_t log_t = TILE_LOG - log_h;
  const uint32_t log_r = a.log_N - log_h;
  const uint32_t log_q = a.log_n - log_h;
  const uint32_t lrp = log_r < log_t ? log_r : log_t;
  Elem el;
  const uint32_t r_lo = e & ((1u << lrp) - 1u);
  el.j = (e >> lrp) & ((1u << log_h) - 1u);
  const uint32_t q_hi = e >> (lrp + log_h);
  el.q = (q_hi << lrp) | r_lo;
  const uint64_t gq = (tile << log_t) + el.q;
  const uint64_t colq = gq >> log_q;
  const uint64_t rem = gq & ((1ull << log_q) - 1ull);
  const uint64_t blk = rem >> log_r;
  el.r = (uint32_t)(rem & ((1ull << log_r) - 1ull));
  el.pos = (blk << a.log_N) + ((uint64_t)el.j << log_r) + el.r;
  el.valid = colq < a.num_cols;
  el.col = (uint32_t)colq;
  return el;
}

template <bool INV, bool DIT>
__global__ __launch_bounds__(TILE_THREADS) void ntt_pass_kernel(PassArgs a) {
  __shared__ uint64_t lds[TILE];
  const uint64_t tile = blockIdx.x;
  const uint32_t log_h = a.log_H;
  const uint32_t log_r = a.log_N - log_h;
  const uint32_t vpt = TILE / TILE_THREADS;

  // Load (coalesced), with the b2n pass twist and the forward coset scaling.
  for (uint32_t i = 0; i < vpt; ++i) {
    const uint32_t e = threadIdx.x + i * TILE_THREADS;
    const Elem el = tile_elem(a, tile, e);
    uint64_t v = 0;
    if (el.valid) {
      v = a.in[(uint64_t)el.col * a.stride_in + el.pos];
      if (a.first) v = gl::canon(v);  // the field ops below take canonical values
      if (!INV && a.first && a.log_ext) {
        // input position pos holds x_i with i = pos (n2b) or brev(pos) (b2n)
        const uint32_t idx = DIT ? brev_bits((uint32_t)el.pos, a.log_n) : (uint32_t)el.pos;
        v = gl::mul(v, coset_pow(a.tab, a.coset_exp, idx, false));
      }
      if (DIT && log_r > 0) {
        uint32_t ex = (el.r * brev_bits(el.j, log_h)) << (32u - a.log_N);
        if (INV) ex = 0u - ex;
        v = gl::mul(v, pow_w(a.tab, ex));
      }
    }
    lds[lds_pos(el.q, el.j, log_h)] = v;
  }
  __syncthreads();

  // Register levels: n2b runs them top-down, b2n bottom-up.
  const uint32_t full = log_h / 4, rem = log_h % 4;
  if (!DIT) {
    for (uint32_t l = 0; l < full; ++l) {
      run_level_dispatch<INV, false>(lds, a.tab, log_h, 4 * l, 4);
      __syncthreads();
    }
    if (rem) {
      run_level_dispatch<INV, false>(lds, a.tab, log_h, 4 * full, rem);
      __syncthreads();
    }
  } else {
    if (rem) {
      run_level_dispatch<INV, true>(lds, a.tab, log_h, 4 * full, rem);
      __syncthreads();
    }
    for (uint32_t l = full; l-- > 0;) {
      run_level_dispatch<INV, true>(lds, a.tab, log_h, 4 * l, 4);
      __syncthreads();
    }
  }

  // Store (coalesced), with the n2b pass twist and the inverse scaling.
  for (uint32_t i = 0; i < vpt; ++i) {
    const uint32_t e = threadIdx.x + i * TILE_THREADS;
    const Elem el = tile_elem(a, tile, e);
    if (!el.valid) continue;
    uint64_t v = lds[lds_pos(el.q, el.j, log_h)];
    if (!DIT && log_r > 0) {
      uint32_t ex = (el.r * brev_bits(el.j, log_h)) << (32u - a.log_N);
      if (INV) ex = 0u - ex;
      v = gl::mul(v, pow_w(a.tab, ex));
    }
    if (a.last) {
      if (INV) {
        v = gl::mul(v, a.n_inv);
        if (a.log_ext) {
          // output position pos holds y_i with i = brev(pos) (n2b) or pos (b2n)
          const uint32_t idx = DIT ? (uint32_t)el.pos : brev_bits((uint32_t)el.pos, a.log_n);
          v = gl::mul(v, coset_pow(a.tab, a.coset_exp, idx, true));
        }
      }
      v = gl::canon(v);
    }
    a.out[(uint64_t)el.col * a.stride_out + el.pos] = v;
  }
}

}  // namespace

int plan_passes(uint32_t log_n, uint32_t *log_h, int max_passes) {
  // Strided passes take at most 8 bits so a tile row spans >= 16 consecutive
  // columns of the block (128-byte rows); the contiguous pass takes 12.
  int np = 0;
  if (log_n <= TILE_LOG) {
    log_h[np++] = log_n;
    return np;
  }
  uint32_t rest = log_n - TILE_LOG;
  while (rest > 0 && np < max_passes - 1) {
    uint32_t take = rest > 8 ? 8 : rest;
    log_h[np++] = take;
    rest -= take;
  }
  if (rest > 0) return -1;
  log_h[np++] = TILE_LOG;
  return np;
}

hipError_t launch_pass(const PassArgs &a, bool inverse, bool b2n, hipStream_t stream) {
  const uint64_t subcols = (uint64_t)a.num_cols << (a.log_n - a.log_H);
  const uint32_t log_t = TILE_LOG - a.log_H;
  const uint64_t tiles = (subcols + (1ull << log_t) - 1) >> log_t;
  if (tiles == 0) return hipSuccess;
  if (tiles > 0x7FFFFFFFull) return hipErrorInvalidConfiguration;
  const dim3 grid((uint32_t)tiles), block(TILE_THREADS);
  if (inverse) {
    if (b2n) ntt_pass_kernel<true, true><<<grid, block, 0, stream>>>(a);
    else ntt_pass_kernel<true, false><<<grid, block, 0, stream>>>(a);
  } else {
    if (b2n) ntt_pass_kernel<false, true><<<grid, block, 0, stream>>>(a);
    else ntt_pass_kernel<false, false><<<grid, block, 0, stream>>>(a);
  }
  return hipGetLastError();
}

}  // namespace bc

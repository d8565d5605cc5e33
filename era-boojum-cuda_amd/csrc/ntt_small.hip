// Latency path for a few columns of 2^16 points (the single-column case of the
// reference's benches): two launches of one-wave workgroups, each wave holding
// one 256-point sub-DFT at 4 values per lane.
//
// The pass kernels of ntt_fast.hip give each thread 16 values of a 4096-point
// tile, so one 2^16 column is 16 workgroups per pass and the call takes one
// workgroup's serial work.  Here a 2^16 transform is split 8 + 8: the top pass
// runs 256 sub-DFTs over the top index bits (stride 256, with the pass twist
// omega_n^(r * brev(j))), the bottom pass 256 sub-DFTs over consecutive
// points; 256 one-wave workgroups per column per pass, and each lane does four
// values' work.
//
// One wave's 256-point DIF, point j = 64 i + l at lane l, register i:
//   stage 1: radix 4 over j bits 7..6 (registers), twist omega_256^(l k1);
//   exchange 1: register bits <-> lane bits 4, 5 (v_permlane16/32_swap);
//   stage 2: radix 4 over j bits 5..4, twist omega_64^((l & 15) k2);
//   exchange 2: register bits <-> lane bits 2, 3 (through the wave's LDS);
//   stage 3: radix 4 over j bits 3..2, twist omega_16^((l & 3) k3);
//   exchange 3: register bits <-> lane bits 0, 1 (through the wave's LDS);
//   stage 4: radix 4 over j bits 1..0,
// where k_s = brev2(register) after the stage's radix-4 DIF.  Every exchange
// transposes two register bits with two lane bits, so lane l, register i ends
// at the bit-reversed output position 4 l + i (the wave stores 32 contiguous
// bytes per lane).  DIT (bit-reversed input) is the transpose: the same
// stages, twists and exchanges in reverse order.
#include "ntt_device.hpp"
#include "ntt_internal.hpp"

namespace bc {
namespace {

constexpr uint32_t SW = 64;  // threads per workgroup: one wave

__device__ __forceinline__ uint32_t brev2(uint32_t i) { return ((i & 1u) << 1) | (i >> 1); }

// omega_{2^s}^e (e < 2^s, s <= 16) from the coarse table: omega_{2^32}^(e 2^(32-s))
__device__ __forceinline__ uint64_t root(const Tables &tab, uint32_t s, uint32_t e, bool inv) {
  const uint32_t m = (1u << s) - 1u;
  const uint32_t k = inv ? ((0u - e) & m) : (e & m);
  return tab.w_coarse[k << (16u - s)];
}

// Transposes register bits 0, 1 with lane bits 4, 5 of four 64-bit values:
// v_permlane16_swap exchanges the odd 16-lane rows of its first operand with
// the even rows of its second, v_permlane32_swap the upper 32 lanes of the
// first with the lower 32 of the second.
__device__ __forceinline__ void swap_lane45(uint64_t *x) {
  auto sw16 = [](uint64_t &a, uint64_t &b) {
    const auto lo = __builtin_amdgcn_permlane16_swap((uint32_t)a, (uint32_t)b, false, false);
    const auto hi = __builtin_amdgcn_permlane16_swap((uint32_t)(a >> 32), (uint32_t)(b >> 32), false, false);
    a = ((uint64_t)hi[0] << 32) | lo[0];
    b = ((uint64_t)hi[1] << 32) | lo[1];
  };
  auto sw32 = [](uint64_t &a, uint64_t &b) {
    const auto lo = __builtin_amdgcn_permlane32_swap((uint32_t)a, (uint32_t)b, false, false);
    const auto hi = __builtin_amdgcn_permlane32_swap((uint32_t)(a >> 32), (uint32_t)(b >> 32), false, false);
    a = ((uint64_t)hi[0] << 32) | lo[0];
    b = ((uint64_t)hi[1] << 32) | lo[1];
  };
  sw16(x[0], x[1]);
  sw16(x[2], x[3]);
  sw32(x[0], x[2]);
  sw32(x[1], x[3]);
}

// Transposes register bits 0, 1 with lane bits B, B + 1 through `buf` (256
// words of this wave's LDS).
template <int B> __device__ __forceinline__ void swap_lanes_lds(uint64_t *x, uint64_t *buf, uint32_t l) {
#pragma unroll
  for (int i = 0; i < 4; ++i) buf[i * SW + l] = x[i];
  __syncthreads();
  const uint32_t src_reg = (l >> B) & 3u, lane_rest = l & ~(3u << B);
#pragma unroll
  for (int i = 0; i < 4; ++i) x[i] = buf[src_reg * SW + (lane_rest | ((uint32_t)i << B))];
}

struct WaveLds {
  uint64_t ex2[4 * SW];
  uint64_t ex3[4 * SW];
};

// Stage twiddles of lane l: tw[3 (s - 1) + i - 1] for stage s = 1..3 and
// register i = 1..3 (register 0 has k = 0).
__device__ __forceinline__ void stage_twiddles(const Tables &tab, uint32_t l, bool inv, uint64_t *tw) {
#pragma unroll
  for (int i = 1; i < 4; ++i) {
    const uint32_t k = brev2((uint32_t)i);
    tw[i - 1] = root(tab, 8, l * k, inv);
    tw[3 + i - 1] = root(tab, 6, (l & 15u) * k, inv);
    tw[6 + i - 1] = root(tab, 4, (l & 3u) * k, inv);
  }
}

template <bool INV> __device__ __forceinline__ void wave_dif256(uint64_t *x, const uint64_t *tw, WaveLds &sh, uint32_t l) {
  dft_dif<2, INV>(x);
#pragma unroll
  for (int i = 1; i < 4; ++i) x[i] = gl::mul(x[i], tw[i - 1]);
  swap_lane45(x);
  dft_dif<2, INV>(x);
#pragma unroll
  for (int i = 1; i < 4; ++i) x[i] = gl::mul(x[i], tw[3 + i - 1]);
  swap_lanes_lds<2>(x, sh.ex2, l);
  dft_dif<2, INV>(x);
#pragma unroll
  for (int i = 1; i < 4; ++i) x[i] = gl::mul(x[i], tw[6 + i - 1]);
  swap_lanes_lds<0>(x, sh.ex3, l);
  dft_dif<2, INV>(x);
}

template <bool INV> __device__ __forceinline__ void wave_dit256(uint64_t *x, const uint64_t *tw, WaveLds &sh, uint32_t l) {
  dft_dit<2, INV>(x);
  swap_lanes_lds<0>(x, sh.ex3, l);
#pragma unroll
  for (int i = 1; i < 4; ++i) x[i] = gl::mul(x[i], tw[6 + i - 1]);
  dft_dit<2, INV>(x);
  swap_lanes_lds<2>(x, sh.ex2, l);
#pragma unroll
  for (int i = 1; i < 4; ++i) x[i] = gl::mul(x[i], tw[3 + i - 1]);
  dft_dit<2, INV>(x);
  swap_lane45(x);
#pragma unroll
  for (int i = 1; i < 4; ++i) x[i] = gl::mul(x[i], tw[i - 1]);
  dft_dit<2, INV>(x);
}

// Top pass (sub-DFTs over index bits 15..8, stride 256): workgroup (r, column).
// DIF loads natural points (64 i + l) 256 + r and stores position 4 l + i with
// the pass twist omega_65536^(r brev8(4 l + i)) (n^-1 folded in for inverse
// transforms); DIT is the transpose.
template <bool INV, bool DIT>
__global__ __launch_bounds__(SW) void ntt_small_top_kernel(FastArgs a) {
  __shared__ WaveLds sh;
  const uint32_t l = threadIdx.x;
  // workgroups are dispatched round-robin over the 8 XCDs: keep the 16
  // workgroups whose r share a 128-byte line on one XCD (one L2)
  const uint32_t r = ((blockIdx.x & 7u) << 5) | (blockIdx.x >> 3);
  const uint64_t c = blockIdx.y;
  const uint64_t *in = a.in + c * a.stride_in + r;
  uint64_t *out = a.out + c * a.stride_out + r;
  uint64_t tw[9];
  stage_twiddles(a.tab, l, INV, tw);
  uint64_t pt[4];
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    const uint32_t k = brev_bits(4u * l + (uint32_t)i, 8);
    uint32_t e = (r * k) << 16;
    if (INV) e = 0u - e;
    pt[i] = pow_w(a.tab, e);
    if (INV) pt[i] = gl::mul(pt[i], a.n_inv);
  }
  uint64_t x[4];
  if constexpr (!DIT) {
#pragma unroll
    for (int i = 0; i < 4; ++i) x[i] = in[(uint64_t)(64u * i + l) << 8];
    if (a.first) {
#pragma unroll
      for (int i = 0; i < 4; ++i) x[i] = gl::canon(x[i]);
    }
    wave_dif256<INV>(x, tw, sh, l);
#pragma unroll
    for (int i = 0; i < 4; ++i) out[(uint64_t)(4u * l + i) << 8] = gl::mul(x[i], pt[i]);
  } else {
#pragma unroll
    for (int i = 0; i < 4; ++i) x[i] = gl::mul(in[(uint64_t)(4u * l + i) << 8], pt[i]);
    wave_dit256<INV>(x, tw, sh, l);
#pragma unroll
    for (int i = 0; i < 4; ++i) out[(uint64_t)(64u * i + l) << 8] = x[i];
  }
}

// Bottom pass (sub-DFTs over index bits 7..0): workgroup (block, column),
// 256 consecutive points.
template <bool INV, bool DIT>
__global__ __launch_bounds__(SW) void ntt_small_bottom_kernel(FastArgs a) {
  __shared__ WaveLds sh;
  const uint32_t l = threadIdx.x;
  const uint64_t c = blockIdx.y, b = blockIdx.x;
  const uint64_t *in = a.in + c * a.stride_in + (b << 8);
  uint64_t *out = a.out + c * a.stride_out + (b << 8);
  uint64_t tw[9];
  stage_twiddles(a.tab, l, INV, tw);
  uint64_t x[4];
  if constexpr (!DIT) {
#pragma unroll
    for (int i = 0; i < 4; ++i) x[i] = in[64u * i + l];
    wave_dif256<INV>(x, tw, sh, l);
#pragma unroll
    for (int i = 0; i < 4; ++i) out[4u * l + i] = x[i];
  } else {
#pragma unroll
    for (int i = 0; i < 4; ++i) x[i] = in[4u * l + i];
    if (a.first) {
#pragma unroll
      for (int i = 0; i < 4; ++i) x[i] = gl::canon(x[i]);
    }
    wave_dit256<INV>(x, tw, sh, l);
#pragma unroll
    for (int i = 0; i < 4; ++i) out[64u * i + l] = x[i];
  }
}

template <bool INV, bool DIT> hipError_t launch_pair(const FastArgs &top, const FastArgs &bottom, hipStream_t s) {
  const dim3 grid(256, top.num_cols);
  if constexpr (!DIT) {
    ntt_small_top_kernel<INV, false><<<grid, SW, 0, s>>>(top);
    hipError_t err = hipGetLastError();
    if (err != hipSuccess) return err;
    ntt_small_bottom_kernel<INV, false><<<grid, SW, 0, s>>>(bottom);
  } else {
    ntt_small_bottom_kernel<INV, true><<<grid, SW, 0, s>>>(bottom);
    hipError_t err = hipGetLastError();
    if (err != hipSuccess) return err;
    ntt_small_top_kernel<INV, true><<<grid, SW, 0, s>>>(top);
  }
  return hipGetLastError();
}

}  // namespace

bool small_path_applies(uint32_t log_n, uint32_t num_cols, uint32_t log_ext, const TuneParams &tp) {
  if (log_n != SMALL_LOG_N || log_ext != 0 || num_cols < 1 || num_cols > 65535) return false;
  return tp.tile_log == 8 || (tp.tile_log == 0 && num_cols <= SMALL_MAX_COLS);
}

hipError_t launch_small(const Tables &tab, const uint64_t *inputs, uint64_t *outputs, uint32_t num_cols,
                        uint64_t stride_in, uint64_t stride_out, bool bitrev_inputs, bool inverse, hipStream_t s) {
  FastArgs a{};
  a.tab = tab;
  a.num_cols = num_cols;
  a.log_n = SMALL_LOG_N;
  a.n_inv = gl::inv((1ull << SMALL_LOG_N) % gl::P);
  // the first launch reads the caller's input, the second works in place on the output
  FastArgs first = a, second = a;
  first.in = inputs;
  first.stride_in = stride_in;
  first.out = outputs;
  first.stride_out = stride_out;
  first.first = 1;
  second.in = outputs;
  second.stride_in = stride_out;
  second.out = outputs;
  second.stride_out = stride_out;
  second.first = 0;
  if (!bitrev_inputs)
    return inverse ? launch_pair<true, false>(first, second, s) : launch_pair<false, false>(first, second, s);
  return inverse ? launch_pair<true, true>(second, first, s) : launch_pair<false, true>(second, first, s);
}

}  // namespace bc

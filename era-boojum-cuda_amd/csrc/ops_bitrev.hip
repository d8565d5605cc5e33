// Bit-reversal permutation of batched columns (the order conversion on
// either side of the NTT): dst[brev(i)] = src[i] for every column, values
// copied unchanged (not reduced).  Reference: bit_reverse /
// bit_reverse_in_place (src/ops_complex.rs:260-330).
//
// Columns of 2^n points with n >= 10 are split as i = a 2^(n-5) + m 2^5 + b
// (a, b < 32).  brev(i) = brev5(b) 2^(n-5) + brev(m) 2^5 + brev5(a), so the
// 32 x 32 tile (a, b) at middle index m maps, transposed and with both tile
// indices reversed, onto the tile at middle index brev(m).  A workgroup loads
// tile rows of 32 consecutive points (256 B), transposes through LDS and
// stores rows of 32 consecutive points, so HBM sees only full-row accesses.
// In place, the workgroup that owns m <= brev(m) swaps the two tiles.
// Shorter columns use one thread per point.
#include "../../include/boojum_ntt.h"

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace bc {
namespace {

constexpr uint32_t BR_T = 32;       // tile edge
constexpr uint32_t BR_PITCH = 33;   // LDS row pitch (one pad slot)
constexpr uint32_t BR_NT = 256;     // threads per workgroup: 8 rows x 32 columns per step

__device__ __forceinline__ uint32_t brev_n(uint32_t x, uint32_t bits) {
  return bits ? (__builtin_bitreverse32(x) >> (32 - bits)) : 0u;
}

template <bool IN_PLACE>
__global__ __launch_bounds__(BR_NT) void bitrev_tiled_kernel(const uint64_t *__restrict__ src, uint64_t *dst,
                                                             uint64_t stride_src, uint64_t stride_dst, uint32_t log_n,
                                                             uint32_t cols) {
  __shared__ uint64_t ta[BR_T * BR_PITCH];
  __shared__ uint64_t tb[IN_PLACE ? BR_T * BR_PITCH : 1];
  const uint32_t mid_bits = log_n - 10;
  const uint64_t nm = 1ull << mid_bits;
  const uint32_t c = threadIdx.x & 31u, r0 = threadIdx.x >> 5;
  const uint32_t row_shift = log_n - 5;
  const uint64_t work_total = nm * cols;
  for (uint64_t w = blockIdx.x; w < work_total; w += gridDim.x) {
    const uint64_t col = w >> mid_bits;
    const uint32_t m = (uint32_t)(w & (nm - 1));
    const uint32_t mr = brev_n(m, mid_bits);
    if (IN_PLACE && m > mr) continue;  // the workgroup of the pair's lower index swaps both tiles
    const uint64_t *s = src + col * stride_src;
    uint64_t *d = dst + col * stride_dst;
#pragma unroll
    for (uint32_t k = 0; k < BR_T / 8; ++k) {
      // every element is moved once: non-temporal loads and stores (4-6% faster)
      const uint32_t r = r0 + 8 * k;
      const uint64_t *row = s + ((uint64_t)r << row_shift) + c;
      ta[r * BR_PITCH + c] = __builtin_nontemporal_load(row + ((uint64_t)m << 5));
      if (IN_PLACE && m != mr) tb[r * BR_PITCH + c] = __builtin_nontemporal_load(row + ((uint64_t)mr << 5));
    }
    __syncthreads();
    // dst row R, column C of tile mr holds src (a, b) = (brev5(C), brev5(R)) of tile m
    const uint32_t bc5 = brev_n(c, 5);
#pragma unroll
    for (uint32_t k = 0; k < BR_T / 8; ++k) {
      const uint32_t R = r0 + 8 * k;
      const uint32_t bR = brev_n(R, 5);
      uint64_t *row = d + ((uint64_t)R << row_shift) + c;
      __builtin_nontemporal_store(ta[bc5 * BR_PITCH + bR], row + ((uint64_t)mr << 5));
      if (IN_PLACE && m != mr) __builtin_nontemporal_store(tb[bc5 * BR_PITCH + bR], row + ((uint64_t)m << 5));
    }
    __syncthreads();  // the tiles are reloaded in the next iteration
  }
}

template <bool IN_PLACE>
__global__ __launch_bounds__(BR_NT) void bitrev_small_kernel(const uint64_t *__restrict__ src, uint64_t *dst,
                                                             uint64_t stride_src, uint64_t stride_dst, uint32_t log_n,
                                                             uint32_t cols) {
  const uint64_t n = 1ull << log_n;
  const uint64_t total = n * cols;
  for (uint64_t k = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; k < total; k += (uint64_t)gridDim.x * blockDim.x) {
    const uint64_t col = k >> log_n;
    const uint32_t i = (uint32_t)(k & (n - 1));
    const uint32_t j = brev_n(i, log_n);
    if (IN_PLACE) {
      if (i < j) {
        uint64_t *d = dst + col * stride_dst;
        const uint64_t x = d[i];
        d[i] = d[j];
        d[j] = x;
      }
    } else {
      dst[col * stride_dst + j] = src[col * stride_src + i];
    }
  }
}

int bitrev_launch(const uint64_t *src, uint64_t stride_src, uint64_t *dst, uint64_t stride_dst, uint64_t rows,
                  uint32_t cols, bool in_place, hipStream_t s) {
  // src/ops_complex.rs:279-281: rows a power of two that fits u32
  if (rows == 0 || (rows & (rows - 1)) || rows > 0xFFFFFFFFull) return BC_ERROR_INVALID_VALUE;
  if (cols == 0) return BC_SUCCESS;
  if (!dst || (!in_place && !src)) return BC_ERROR_INVALID_VALUE;
  if (stride_dst < rows || (!in_place && stride_src < rows)) return BC_ERROR_INVALID_VALUE;
  const uint32_t log_n = (uint32_t)__builtin_ctzll(rows);
  if (log_n >= 10) {
    const uint64_t work = ((uint64_t)cols) << (log_n - 10);
    // An odd grid stride walks each workgroup through all middle indices m,
    // so in place the skipped half (m > brev(m)) is spread evenly instead of
    // idling the workgroups whose m is fixed by a power-of-two stride.
    const uint32_t grid = (uint32_t)(work < 65535 ? work : 65535);
    if (in_place) bitrev_tiled_kernel<true><<<grid, BR_NT, 0, s>>>(dst, dst, stride_dst, stride_dst, log_n, cols);
    else bitrev_tiled_kernel<false><<<grid, BR_NT, 0, s>>>(src, dst, stride_src, stride_dst, log_n, cols);
  } else {
    const uint64_t total = (uint64_t)cols << log_n;
    const uint64_t blocks = (total + BR_NT - 1) / BR_NT;
    const uint32_t grid = (uint32_t)(blocks < 65536 ? blocks : 65536);
    if (in_place) bitrev_small_kernel<true><<<grid, BR_NT, 0, s>>>(dst, dst, stride_dst, stride_dst, log_n, cols);
    else bitrev_small_kernel<false><<<grid, BR_NT, 0, s>>>(src, dst, stride_src, stride_dst, log_n, cols);
  }
  const hipError_t err = hipGetLastError();
  return err == hipSuccess ? BC_SUCCESS : (int)err;
}

}  // namespace
}  // namespace bc

extern "C" int bc_bit_reverse(const uint64_t *src, uint64_t src_stride, uint64_t *dst, uint64_t dst_stride,
                              uint64_t rows, uint32_t cols, void *stream) {
  // the same matrix in and out is the in-place permutation (the reference
  // runs one kernel for both, src/ops_complex.rs:310-330)
  if (src && src == dst && src_stride == dst_stride)
    return bc::bitrev_launch(nullptr, 0, dst, dst_stride, rows, cols, true, (hipStream_t)stream);
  return bc::bitrev_launch(src, src_stride, dst, dst_stride, rows, cols, false, (hipStream_t)stream);
}

extern "C" int bc_bit_reverse_in_place(uint64_t *values, uint64_t stride, uint64_t rows, uint32_t cols,
                                       void *stream) {
  return bc::bitrev_launch(nullptr, 0, values, stride, rows, cols, true, (hipStream_t)stream);
}

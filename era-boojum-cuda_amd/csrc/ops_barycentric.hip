// Barycentric evaluation of polynomials given on the coset domain
// {coset * omega_N^i}: Lagrange coefficients at a point x, then many columns
// evaluated against them.  Reference: src/barycentric.rs:83-280 (host) and
// native/barycentric.cu:10-156 (kernels).
//
//   common = coset (x^N - coset^N) / (N coset^N)          (barycentric.cu:12-21)
//   L_i    = common / (x omega_N^-i - coset)               (barycentric.cu:44-86)
//   eval_p = sum_i L_i y_{i,p}                             (barycentric.cu:104-156
//                                                           + the reference's
//                                                           final column sums)
// x and the coefficients are base-field or extension (F_p[u]/(u^2 - 7))
// values; the columns y are base or extension.  Layouts are the reference's:
// extension vectors and matrix columns are "double vectors" (the c0 plane,
// then the c1 plane, each N long); extension evaluations are (c0, c1) pairs.
//
// The evaluation is one pass over the columns, so it is bound by HBM reads
// (8 or 16 bytes per element) and by the VALU cost of one field product per
// base element.  Products are accumulated unreduced (gl::mac: 4 multiply-adds
// and a carry chain) and reduced once per lane and column.  Eight lanes share
// a column and read it with 16-byte loads, so a wave instruction covers 128
// contiguous bytes of each of 8 columns; every lane keeps its 16 rows of
// coefficients in registers for all the columns it visits.
#include "../../include/boojum_ntt.h"
#include "gl_ext.hpp"
#include "ntt_device.hpp"

namespace bc {

Tables current_tables(bool &ok);  // ntt_api.hip

namespace {

// ---- Lagrange coefficients ---------------------------------------------------

// One thread: the common factor, with x^N and coset^N by log N squarings.
template <bool EXT>
__global__ void bary_common_factor_kernel(const uint64_t *x, uint64_t *common, uint64_t coset, uint32_t log_count) {
  uint64_t cn = coset;
  for (uint32_t i = 0; i < log_count; ++i) cn = gl::mul(cn, cn);
  const uint64_t f = gl::mul(coset, gl::inv(gl::mul(1ull << log_count, cn)));
  if constexpr (EXT) {
    gl::E2 xn{gl::canon(x[0]), gl::canon(x[1])};
    for (uint32_t i = 0; i < log_count; ++i) xn = gl::e2_mul(xn, xn);
    common[0] = gl::mul(gl::sub(xn.c0, cn), f);
    common[1] = gl::mul(xn.c1, f);
  } else {
    uint64_t xn = gl::canon(x[0]);
    for (uint32_t i = 0; i < log_count; ++i) xn = gl::mul(xn, xn);
    common[0] = gl::mul(gl::sub(xn, cn), f);
  }
}

constexpr uint32_t LC_NT = 256;

template <bool EXT> struct CoeffBatch;
template <> struct CoeffBatch<false> { static constexpr uint32_t B = 16; };
template <> struct CoeffBatch<true> { static constexpr uint32_t B = 8; };

// Thread t of workgroup b owns outputs i = b 256 B + j 256 + t (j < B; stores
// coalesced for each j).  Its B denominators x omega_N^-idx - coset (idx = i,
// or brev(i) when the output is bit-reversed) share one field inversion
// (prefix products).  Padding lanes (i >= N) use the denominator 1.
template <bool EXT>
__global__ __launch_bounds__(LC_NT) void bary_coeffs_kernel(Tables tab, const uint64_t *x, const uint64_t *common,
                                                            uint64_t coset, uint64_t *coeffs, uint32_t log_count,
                                                            bool bit_reverse) {
  constexpr uint32_t B = CoeffBatch<EXT>::B;
  const uint64_t count = 1ull << log_count;
  const gl::E2 xv{gl::canon(x[0]), EXT ? gl::canon(x[1]) : 0ull};
  const gl::E2 cf{common[0], EXT ? common[1] : 0ull};
  const uint64_t base = (uint64_t)blockIdx.x * (LC_NT * B) + threadIdx.x;
  gl::E2 d[B], pre[B];
  gl::E2 acc{1, 0};
#pragma unroll
  for (uint32_t j = 0; j < B; ++j) {
    const uint64_t i = base + (uint64_t)j * LC_NT;
    gl::E2 dj{1, 0};
    if (i < count) {
      const uint32_t idx = bit_reverse ? (log_count ? __builtin_bitreverse32((uint32_t)i) >> (32 - log_count) : 0u)
                                       : (uint32_t)i;
      const uint64_t w = pow_w(tab, 0u - (log_count ? idx << (32 - log_count) : 0u));  // omega_N^-idx
      dj.c0 = gl::sub(gl::mul(xv.c0, w), coset);
      if constexpr (EXT) dj.c1 = gl::mul(xv.c1, w);
    }
    d[j] = dj;
    pre[j] = acc;
    if constexpr (EXT) acc = gl::e2_mul(acc, dj);
    else acc.c0 = gl::mul(acc.c0, dj.c0);
  }
  gl::E2 inv_acc;
  if constexpr (EXT) inv_acc = gl::e2_inv(acc);
  else inv_acc = {gl::inv(acc.c0), 0};
#pragma unroll
  for (uint32_t j = B; j-- > 0;) {
    const uint64_t i = base + (uint64_t)j * LC_NT;
    gl::E2 l;
    if constexpr (EXT) {
      l = gl::e2_mul(gl::e2_mul(inv_acc, pre[j]), cf);
      inv_acc = gl::e2_mul(inv_acc, d[j]);
    } else {
      l.c0 = gl::mul(gl::mul(inv_acc.c0, pre[j].c0), cf.c0);
      inv_acc.c0 = gl::mul(inv_acc.c0, d[j].c0);
    }
    if (i < count) {
      coeffs[i] = l.c0;
      if constexpr (EXT) coeffs[count + i] = l.c1;
    }
  }
}

// ---- batch evaluation ----------------------------------------------------------

enum Kind : uint32_t { BASE_AT_BASE = 0, BASE_AT_EXT = 1, EXT_AT_EXT = 2 };

constexpr uint32_t BE_NT = 256;                  // 4 independent waves
constexpr uint32_t BE_WAVES = BE_NT / 64;
constexpr uint32_t BE_LPC = 8;                   // lanes per column
constexpr uint32_t BE_K = 8;                     // 16-byte loads per lane and column
constexpr uint32_t BE_ROWS = BE_LPC * 2 * BE_K;  // 128 rows per wave tile
constexpr uint32_t BE_COLS = 64 / BE_LPC;        // 8 columns per wave pass
constexpr uint32_t BE_TARGET_WGS = 2048;

struct U2 {
  uint64_t x, y;
};

// Two consecutive elements; one 16-byte load when the buffer allows it.
template <bool ALIGNED> __device__ __forceinline__ U2 load2(const uint64_t *p) {
  if constexpr (ALIGNED) {
    const ulonglong2 v = *reinterpret_cast<const ulonglong2 *>(p);
    return {v.x, v.y};
  } else {
    return {p[0], p[1]};
  }
}

__device__ __forceinline__ uint64_t shfl_xor64(uint64_t v, int m) {
  const uint32_t lo = __shfl_xor((int)(uint32_t)v, m), hi = __shfl_xor((int)(uint32_t)(v >> 32), m);
  return ((uint64_t)hi << 32) | lo;
}

// Wave w of workgroup b evaluates tile t = 4 b + w (rows [128 t, 128 t + 128))
// for column passes blockIdx.y, blockIdx.y + gridDim.y, ...; lane l reads
// column 8 pass + l / 8 at rows 128 t + 16 k + 2 (l % 8) + {0, 1}.  The sum
// over a column's 8 lanes is written to partial[column][t] (one element, or a
// (c0, c1) pair for extension results).  Waves share nothing, so a wave past
// the last tile simply returns.
template <uint32_t KIND, bool ALIGNED>
__global__ __launch_bounds__(BE_NT) void bary_eval_tiles_kernel(const uint64_t *ys, const uint64_t *coeffs,
                                                                uint64_t count, uint32_t num_polys,
                                                                uint64_t *partial, uint64_t tiles) {
  constexpr bool CEXT = KIND != BASE_AT_BASE;
  constexpr bool YEXT = KIND == EXT_AT_EXT;
  const uint32_t lane = threadIdx.x & 63;
  const uint64_t tile = (uint64_t)blockIdx.x * BE_WAVES + (threadIdx.x >> 6);
  if (tile >= tiles) return;
  const uint32_t g = lane / BE_LPC, q = lane % BE_LPC;
  const uint64_t row0 = tile * BE_ROWS + 2 * q;

  uint64_t l0[2 * BE_K], l1[CEXT ? 2 * BE_K : 1];
#pragma unroll
  for (uint32_t k = 0; k < BE_K; ++k) {
    const U2 v = load2<ALIGNED>(coeffs + row0 + 16 * k);
    l0[2 * k] = v.x;
    l0[2 * k + 1] = v.y;
    if constexpr (CEXT) {
      const U2 w = load2<ALIGNED>(coeffs + count + row0 + 16 * k);
      l1[2 * k] = w.x;
      l1[2 * k + 1] = w.y;
    }
  }

  const uint64_t col_elems = YEXT ? 2 * count : count;
  for (uint32_t pass = blockIdx.y; pass * BE_COLS < num_polys; pass += gridDim.y) {
    const uint32_t c = pass * BE_COLS + g;
    const bool valid = c < num_polys;
    // KIND 0: a0 = sum l0 y.  KIND 1: a0 = sum l0 y, a1 = sum l1 y.
    // KIND 2: a0 = sum l0 y0, a1 = sum (l0 y1 + l1 y0), a2 = sum l1 y1.
    gl::Acc a0, a1, a2;
    if (valid) {
      const uint64_t *col = ys + (uint64_t)c * col_elems + row0;
      U2 y0[BE_K], y1[YEXT ? BE_K : 1];
#pragma unroll
      for (uint32_t k = 0; k < BE_K; ++k) {
        y0[k] = load2<ALIGNED>(col + 16 * k);
        if constexpr (YEXT) y1[k] = load2<ALIGNED>(col + count + 16 * k);
      }
#pragma unroll
      for (uint32_t k = 0; k < BE_K; ++k) {
        gl::mac(a0, y0[k].x, l0[2 * k]);
        gl::mac(a0, y0[k].y, l0[2 * k + 1]);
        if constexpr (CEXT) {
          gl::mac(a1, y0[k].x, l1[2 * k]);
          gl::mac(a1, y0[k].y, l1[2 * k + 1]);
        }
        if constexpr (YEXT) {
          gl::mac(a1, y1[k].x, l0[2 * k]);
          gl::mac(a1, y1[k].y, l0[2 * k + 1]);
          gl::mac(a2, y1[k].x, l1[2 * k]);
          gl::mac(a2, y1[k].y, l1[2 * k + 1]);
        }
      }
    }
    uint64_t s0 = gl::acc_reduce(a0), s1 = CEXT ? gl::acc_reduce(a1) : 0ull;
    if constexpr (YEXT) s0 = gl::add(s0, gl::mul_by_non_residue(gl::acc_reduce(a2)));
#pragma unroll
    for (int m = 1; m < (int)BE_LPC; m <<= 1) {
      s0 = gl::add(s0, shfl_xor64(s0, m));
      if constexpr (CEXT) s1 = gl::add(s1, shfl_xor64(s1, m));
    }
    if (valid && q == 0) {
      if constexpr (CEXT) {
        uint64_t *dst = partial + 2 * ((uint64_t)c * tiles + tile);
        dst[0] = s0;
        dst[1] = s1;
      } else {
        partial[(uint64_t)c * tiles + tile] = s0;
      }
    }
  }
}

// One workgroup per column: the column's tile sums added up.
template <bool EXT>
__global__ __launch_bounds__(BE_NT) void bary_eval_final_kernel(const uint64_t *partial, uint64_t tiles,
                                                                uint64_t *evals) {
  __shared__ uint64_t s[2][BE_WAVES];
  const uint64_t c = blockIdx.x;
  const uint64_t *src = partial + (EXT ? 2 : 1) * c * tiles;
  uint64_t s0 = 0, s1 = 0;
  for (uint64_t t = threadIdx.x; t < tiles; t += BE_NT) {
    if constexpr (EXT) {
      s0 = gl::add(s0, src[2 * t]);
      s1 = gl::add(s1, src[2 * t + 1]);
    } else {
      s0 = gl::add(s0, src[t]);
    }
  }
#pragma unroll
  for (int m = 1; m < 64; m <<= 1) {
    s0 = gl::add(s0, shfl_xor64(s0, m));
    if constexpr (EXT) s1 = gl::add(s1, shfl_xor64(s1, m));
  }
  const uint32_t wave = threadIdx.x >> 6;
  if ((threadIdx.x & 63) == 0) {
    s[0][wave] = s0;
    s[1][wave] = s1;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    uint64_t r0 = s[0][0], r1 = s[1][0];
    for (uint32_t w = 1; w < BE_WAVES; ++w) {
      r0 = gl::add(r0, s[0][w]);
      r1 = gl::add(r1, s[1][w]);
    }
    if constexpr (EXT) {
      evals[2 * c] = r0;
      evals[2 * c + 1] = r1;
    } else {
      evals[c] = r0;
    }
  }
}

// N < 128: one thread per column sums its N products directly.
template <uint32_t KIND>
__global__ void bary_eval_small_kernel(const uint64_t *ys, const uint64_t *coeffs, uint32_t count,
                                       uint32_t num_polys, uint64_t *evals) {
  constexpr bool CEXT = KIND != BASE_AT_BASE;
  constexpr bool YEXT = KIND == EXT_AT_EXT;
  const uint32_t c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= num_polys) return;
  const uint64_t *col = ys + (uint64_t)c * (YEXT ? 2 : 1) * count;
  gl::Acc a0, a1, a2;
  for (uint32_t r = 0; r < count; ++r) {
    const uint64_t y0 = col[r];
    gl::mac(a0, y0, coeffs[r]);
    if constexpr (CEXT) gl::mac(a1, y0, coeffs[count + r]);
    if constexpr (YEXT) {
      const uint64_t y1 = col[count + r];
      gl::mac(a1, y1, coeffs[r]);
      gl::mac(a2, y1, coeffs[count + r]);
    }
  }
  uint64_t s0 = gl::acc_reduce(a0);
  if constexpr (YEXT) s0 = gl::add(s0, gl::mul_by_non_residue(gl::acc_reduce(a2)));
  if constexpr (CEXT) {
    evals[2 * c] = s0;
    evals[2 * c + 1] = gl::acc_reduce(a1);
  } else {
    evals[c] = s0;
  }
}

int finish() {
  const hipError_t err = hipGetLastError();
  return err == hipSuccess ? BC_SUCCESS : (int)err;
}

bool power_of_two(uint64_t v) { return v && !(v & (v - 1)); }

uint64_t eval_tiles(uint64_t count) { return count >= BE_ROWS ? count / BE_ROWS : 0; }

template <uint32_t KIND>
void launch_eval(const uint64_t *ys, const uint64_t *coeffs, uint64_t count, uint32_t num_polys, uint64_t *partial,
                 uint64_t *evals, hipStream_t s) {
  constexpr bool OUT_EXT = KIND != BASE_AT_BASE;
  const uint64_t tiles = eval_tiles(count);
  if (!tiles) {
    bary_eval_small_kernel<KIND><<<dim3((num_polys + 255) / 256), 256, 0, s>>>(ys, coeffs, (uint32_t)count,
                                                                               num_polys, evals);
    return;
  }
  const uint64_t gx = (tiles + BE_WAVES - 1) / BE_WAVES;
  const uint32_t passes = (num_polys + BE_COLS - 1) / BE_COLS;
  uint64_t gy = (BE_TARGET_WGS + gx - 1) / gx;
  if (gy > passes) gy = passes;
  if (gy > 65535) gy = 65535;
  const dim3 grid((uint32_t)gx, (uint32_t)gy);
  const bool aligned = !(((uintptr_t)ys | (uintptr_t)coeffs) & 15);
  if (aligned)
    bary_eval_tiles_kernel<KIND, true><<<grid, BE_NT, 0, s>>>(ys, coeffs, count, num_polys, partial, tiles);
  else
    bary_eval_tiles_kernel<KIND, false><<<grid, BE_NT, 0, s>>>(ys, coeffs, count, num_polys, partial, tiles);
  bary_eval_final_kernel<OUT_EXT><<<dim3(num_polys), BE_NT, 0, s>>>(partial, tiles, evals);
}

}  // namespace
}  // namespace bc

extern "C" int bc_barycentric_precompute(bool ext, const uint64_t *x, uint64_t *common_factor, uint64_t coset,
                                         uint64_t *coeffs, uint64_t count, bool bit_reverse, void *stream) {
  using namespace bc;
  if (!x || !common_factor || !coeffs || !power_of_two(count) || count > (1ull << 31)) return BC_ERROR_INVALID_VALUE;
  bool ok = false;
  const Tables tab = current_tables(ok);
  if (!ok) return BC_ERROR_NOT_INITIALIZED;
  const uint32_t log_count = (uint32_t)__builtin_ctzll(count);
  hipStream_t s = (hipStream_t)stream;
  const uint64_t c = gl::canon(coset);
  const uint64_t per_wg = (uint64_t)LC_NT * (ext ? CoeffBatch<true>::B : CoeffBatch<false>::B);
  const dim3 grid((uint32_t)((count + per_wg - 1) / per_wg));
  if (ext) {
    bary_common_factor_kernel<true><<<1, 1, 0, s>>>(x, common_factor, c, log_count);
    bary_coeffs_kernel<true><<<grid, LC_NT, 0, s>>>(tab, x, common_factor, c, coeffs, log_count, bit_reverse);
  } else {
    bary_common_factor_kernel<false><<<1, 1, 0, s>>>(x, common_factor, c, log_count);
    bary_coeffs_kernel<false><<<grid, LC_NT, 0, s>>>(tab, x, common_factor, c, coeffs, log_count, bit_reverse);
  }
  return finish();
}

extern "C" int bc_batch_eval_temp_sizes(uint32_t kind, uint64_t stride, uint32_t num_polys, size_t *partial_elems,
                                        size_t *final_bytes) {
  using namespace bc;
  if (kind > EXT_AT_EXT || !power_of_two(stride) || stride > (1ull << 31) || !partial_elems || !final_bytes)
    return BC_ERROR_INVALID_VALUE;
  const uint64_t tiles = eval_tiles(stride);
  *partial_elems = (size_t)num_polys * (tiles ? tiles : 1);
  *final_bytes = 0;
  return BC_SUCCESS;
}

extern "C" int bc_batch_eval(uint32_t kind, const uint64_t *ys, uint64_t stride, uint32_t num_polys,
                             const uint64_t *coeffs, uint64_t *temp_partial, size_t temp_partial_elems,
                             void *temp_final, size_t temp_final_bytes, uint64_t *evals, void *stream) {
  using namespace bc;
  (void)temp_final;
  (void)temp_final_bytes;
  size_t need = 0, need_final = 0;
  if (bc_batch_eval_temp_sizes(kind, stride, num_polys, &need, &need_final) != BC_SUCCESS)
    return BC_ERROR_INVALID_VALUE;
  if (num_polys == 0) return BC_SUCCESS;
  if (!ys || !coeffs || !evals || (eval_tiles(stride) && (!temp_partial || temp_partial_elems < need)))
    return BC_ERROR_INVALID_VALUE;
  hipStream_t s = (hipStream_t)stream;
  switch (kind) {
  case BASE_AT_BASE: launch_eval<BASE_AT_BASE>(ys, coeffs, stride, num_polys, temp_partial, evals, s); break;
  case BASE_AT_EXT: launch_eval<BASE_AT_EXT>(ys, coeffs, stride, num_polys, temp_partial, evals, s); break;
  default: launch_eval<EXT_AT_EXT>(ys, coeffs, stride, num_polys, temp_partial, evals, s); break;
  }
  return finish();
}

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
// and a carry chain) and reduced once per lane and column.  A wave reads a
// column tile with 16-byte loads (a whole wave per column for long columns:
// 1 KiB contiguous per load instruction), and every lane keeps its rows'
// coefficients in registers for all the columns it visits; per-tile partial
// sums are then added per column by a second small kernel (the sum is exact
// in the field, so the order does not change the result).
#include "../../include/boojum_ntt.h"
#include "gl_ext.hpp"
#include "ntt_device.hpp"

#include <mutex>

namespace bc {

Tables current_tables(bool &ok);  // ntt_api.hip

namespace {

// ---- Lagrange coefficients ---------------------------------------------------

constexpr uint32_t LC_NT = 256;  // threads per workgroup
constexpr uint32_t LC_B = 4;     // coefficients per thread

// Product in the base field (c1 unused) or in the extension.
template <bool EXT> __device__ __forceinline__ gl::E2 pmul(gl::E2 a, gl::E2 b) {
  if constexpr (EXT) return gl::e2_mul(a, b);
  else return {gl::mul(a.c0, b.c0), 0};
}

// Thread t of workgroup b owns outputs i = 1024 b + 256 j + t (j < 4; stores
// coalesced for each j) with denominators d = x omega_N^-idx - coset (idx = i,
// or brev(i) when the output is bit-reversed); padding outputs (i >= N) use
// d = 1.  All 1024 denominators of the workgroup share ONE field inversion
// through a product tree in LDS: each thread multiplies its four (keeping
// prefix products); the 256 thread products are paired up over 8 levels
// (the left operand of every pairing stays in its thread's registers); thread
// 0 inverts the total and multiplies in the common factor; going back down,
// a node's inverse I gives its children I * right and I * left.  Each thread
// ends with common / (its own product) and peels off its four coefficients
// with its prefix products.  The common factor is
//   coset (x^N - coset^N) / (N coset^N) = (x^N - coset^N) f,
// with coset^N and f = coset / (N coset^N) from the host (they depend on
// coset and N only); wave 0 squares x log N times while the power-table reads
// are in flight, and workgroup 0 stores the common factor for the caller, as
// the reference's first kernel does.
template <bool EXT>
__global__ __launch_bounds__(LC_NT) void bary_coeffs_kernel(Tables tab, const uint64_t *x, uint64_t *common_out,
                                                            uint64_t coset, uint64_t coset_n, uint64_t f,
                                                            uint64_t *coeffs, uint32_t log_count, bool bit_reverse) {
  constexpr uint32_t LEVELS = 8;  // log2(LC_NT)
  __shared__ uint64_t node[2][LC_NT];  // subtree products (c0, c1)
  __shared__ uint64_t down[2][LC_NT];  // inverses handed to right children
  const uint32_t t = threadIdx.x;
  const uint64_t count = 1ull << log_count;
  const gl::E2 xv{gl::canon(x[0]), EXT ? gl::canon(x[1]) : 0ull};

  const uint64_t base = (uint64_t)blockIdx.x * (LC_NT * LC_B) + t;
  uint64_t w[LC_B];
#pragma unroll
  for (uint32_t j = 0; j < LC_B; ++j) {
    const uint64_t i = base + (uint64_t)j * LC_NT;
    const uint32_t idx = bit_reverse ? (log_count ? __builtin_bitreverse32((uint32_t)i) >> (32 - log_count) : 0u)
                                     : (uint32_t)i;
    w[j] = i < count ? pow_w(tab, 0u - (log_count ? idx << (32 - log_count) : 0u)) : 0ull;  // omega_N^-idx
  }
  // Wave 0 forms the common factor while the table reads are in flight (the
  // squaring chain is serial; the other waves skip it).
  gl::E2 cf{0, 0};
  if (t < 64) {
    gl::E2 xn = xv;
    for (uint32_t i = 0; i < log_count; ++i) xn = pmul<EXT>(xn, xn);
    cf = {gl::mul(gl::sub(xn.c0, coset_n), f), EXT ? gl::mul(xn.c1, f) : 0ull};
    if (blockIdx.x == 0 && t == 0) {
      common_out[0] = cf.c0;
      if constexpr (EXT) common_out[1] = cf.c1;
    }
  }
  gl::E2 d[LC_B], pre[LC_B];
  gl::E2 v{1, 0};
#pragma unroll
  for (uint32_t j = 0; j < LC_B; ++j) {
    const uint64_t i = base + (uint64_t)j * LC_NT;
    gl::E2 dj{1, 0};
    if (i < count) {
      dj.c0 = gl::sub(gl::mul(xv.c0, w[j]), coset);
      if constexpr (EXT) dj.c1 = gl::mul(xv.c1, w[j]);
    }
    d[j] = dj;
    pre[j] = v;
    v = pmul<EXT>(v, dj);
  }

  // Up: at level l, thread t (a multiple of 2^(l+1)) joins [t, t + 2^l) and
  // [t + 2^l, t + 2^(l+1)).
  gl::E2 left[LEVELS];
  node[0][t] = v.c0;
  node[1][t] = v.c1;
  __syncthreads();
#pragma unroll
  for (uint32_t l = 0; l < LEVELS; ++l) {
    const uint32_t s = 1u << l;
    if ((t & (2 * s - 1)) == 0) {
      left[l] = v;
      v = pmul<EXT>(v, gl::E2{node[0][t + s], node[1][t + s]});
      node[0][t] = v.c0;
      node[1][t] = v.c1;
    }
    __syncthreads();
  }
  // Root: I = common / total.
  gl::E2 inv_node{0, 0};
  if (t == 0) {
    gl::E2 ti;
    if constexpr (EXT) ti = gl::e2_inv(v);
    else ti = {gl::inv(v.c0), 0};
    inv_node = pmul<EXT>(ti, cf);
  }
  // Down: the parent's inverse times the sibling's product.
#pragma unroll
  for (uint32_t l = LEVELS; l-- > 0;) {
    const uint32_t s = 1u << l;
    if ((t & (2 * s - 1)) == 0) {
      const gl::E2 r = pmul<EXT>(inv_node, left[l]);
      down[0][t + s] = r.c0;
      down[1][t + s] = r.c1;
      inv_node = pmul<EXT>(inv_node, gl::E2{node[0][t + s], node[1][t + s]});
    }
    __syncthreads();
    if ((t & (2 * s - 1)) == s) inv_node = gl::E2{down[0][t], down[1][t]};
  }

  // inv_node = common / (d_0 d_1 d_2 d_3) of this thread.
#pragma unroll
  for (uint32_t j = LC_B; j-- > 0;) {
    const uint64_t i = base + (uint64_t)j * LC_NT;
    const gl::E2 l = pmul<EXT>(inv_node, pre[j]);
    inv_node = pmul<EXT>(inv_node, d[j]);
    if (i < count) {
      coeffs[i] = l.c0;
      if constexpr (EXT) coeffs[count + i] = l.c1;
    }
  }
}

// ---- batch evaluation ----------------------------------------------------------

enum Kind : uint32_t { BASE_AT_BASE = 0, BASE_AT_EXT = 1, EXT_AT_EXT = 2 };

constexpr uint32_t BE_NT = 256;  // 4 independent waves
constexpr uint32_t BE_WAVES = BE_NT / 64;

struct U2 {
  uint64_t x, y;
};

// Two consecutive elements; one 16-byte load when the buffer allows it.  The
// columns are read once, so the loads are non-temporal (7-9% faster for the
// base-at-base and extension batches).
template <bool ALIGNED> __device__ __forceinline__ U2 load2(const uint64_t *p) {
  if constexpr (ALIGNED) {
    typedef uint64_t v2u64 __attribute__((ext_vector_type(2)));
    const v2u64 v = __builtin_nontemporal_load(reinterpret_cast<const v2u64 *>(p));
    return {v.x, v.y};
  } else {
    return {__builtin_nontemporal_load(p), __builtin_nontemporal_load(p + 1)};
  }
}

__device__ __forceinline__ uint64_t shfl_xor64(uint64_t v, int m) {
  const uint32_t lo = __shfl_xor((int)(uint32_t)v, m), hi = __shfl_xor((int)(uint32_t)(v >> 32), m);
  return ((uint64_t)hi << 32) | lo;
}

// Tile shape: LPC lanes share a column and read it with 16-byte loads, so a
// wave load instruction covers 64 / LPC columns, 16 LPC contiguous bytes
// each; each lane issues K such loads per column (and per plane for
// extension columns), so a wave tile is 2 LPC K rows.  PREFETCH issues the
// next column pass's loads before the current pass's arithmetic.
template <uint32_t LPC, uint32_t K, bool PREFETCH> struct EvalShape {
  static constexpr uint32_t lpc = LPC, k = K, rows = 2 * LPC * K, cols = 64 / LPC;
  static constexpr bool prefetch = PREFETCH;
  static_assert(LPC >= 1 && LPC <= 64 && 64 % LPC == 0 && K >= 1, "shape");
};

// Tile shapes per kind, from a sweep on MI355X over 256 columns of 2^20 and
// 2^16 rows: wide contiguous runs per column (64 lanes x 16 B = 1 KiB per
// load instruction) read fastest.  `large` serves columns of at least 2^18
// rows, `mid` shorter ones, `small` columns of 128..(mid rows - 1) rows; below
// 128 rows one thread per column sums the column directly.
template <uint32_t KIND> struct Shapes;
template <> struct Shapes<BASE_AT_BASE> {
  using large = EvalShape<64, 16, false>;  // 2048-row tiles
  using mid = EvalShape<64, 4, true>;      // 512
  using small = EvalShape<16, 4, false>;   // 128
};
template <> struct Shapes<BASE_AT_EXT> {
  using large = EvalShape<64, 8, false>;  // 1024
  using mid = EvalShape<16, 4, false>;    // 128
  using small = EvalShape<16, 4, false>;
};
template <> struct Shapes<EXT_AT_EXT> {
  using large = EvalShape<64, 4, true>;  // 512
  using mid = EvalShape<64, 4, true>;
  using small = EvalShape<16, 4, false>;
};
constexpr uint64_t BE_LARGE_COUNT = 1ull << 18;

// Wave w of workgroup b evaluates tile t = 4 b + w (rows [R t, R t + R),
// R = S::rows) for column passes blockIdx.y, blockIdx.y + gridDim.y, ...;
// lane l reads column cols pass + l / LPC at rows R t + 2 LPC k + 2 (l % LPC)
// + {0, 1}.  The sum over a column's LPC lanes is written to
// partial[column][t] (one element, or a (c0, c1) pair for extension
// results).  Waves share nothing, so a wave past the last tile returns.
template <uint32_t KIND, bool ALIGNED, class S>
__global__ __launch_bounds__(BE_NT) void bary_eval_tiles_kernel(const uint64_t *ys, const uint64_t *coeffs,
                                                                uint64_t count, uint32_t num_polys,
                                                                uint64_t *partial, uint64_t tiles) {
  constexpr bool CEXT = KIND != BASE_AT_BASE;
  constexpr bool YEXT = KIND == EXT_AT_EXT;
  constexpr uint32_t LPC = S::lpc, K = S::k, COLS = S::cols, STEP = 2 * LPC;
  const uint32_t lane = threadIdx.x & 63;
  const uint64_t tile = (uint64_t)blockIdx.x * BE_WAVES + (threadIdx.x >> 6);
  if (tile >= tiles) return;
  const uint32_t g = lane / LPC, q = lane % LPC;
  const uint64_t row0 = tile * S::rows + 2 * q;

  uint64_t l0[2 * K], l1[CEXT ? 2 * K : 1];
#pragma unroll
  for (uint32_t k = 0; k < K; ++k) {
    const U2 v = load2<ALIGNED>(coeffs + row0 + STEP * k);
    l0[2 * k] = v.x;
    l0[2 * k + 1] = v.y;
    if constexpr (CEXT) {
      const U2 w = load2<ALIGNED>(coeffs + count + row0 + STEP * k);
      l1[2 * k] = w.x;
      l1[2 * k + 1] = w.y;
    }
  }

  const uint64_t col_elems = YEXT ? 2 * count : count;
  U2 y0[K], y1[YEXT ? K : 1];
  auto load_pass = [&](uint32_t pass, U2 (&d0)[K], U2 (&d1)[YEXT ? K : 1]) {
    const uint32_t c = pass * COLS + g;
    if (c < num_polys) {
      const uint64_t *col = ys + (uint64_t)c * col_elems + row0;
#pragma unroll
      for (uint32_t k = 0; k < K; ++k) {
        d0[k] = load2<ALIGNED>(col + STEP * k);
        if constexpr (YEXT) d1[k] = load2<ALIGNED>(col + count + STEP * k);
      }
    }
  };
  uint32_t pass = blockIdx.y;
  if constexpr (S::prefetch)
    if (pass * COLS < num_polys) load_pass(pass, y0, y1);
  for (; pass * COLS < num_polys; pass += gridDim.y) {
    const uint32_t c = pass * COLS + g;
    const bool valid = c < num_polys;
    const uint32_t next = pass + gridDim.y;
    U2 n0[K], n1[YEXT ? K : 1];
    if constexpr (S::prefetch) {
      if (next * COLS < num_polys) load_pass(next, n0, n1);
    } else {
      load_pass(pass, y0, y1);
    }
    // KIND 0: a0 = sum l0 y.  KIND 1: a0 = sum l0 y, a1 = sum l1 y.
    // KIND 2: a0 = sum l0 y0, a1 = sum (l0 y1 + l1 y0), a2 = sum l1 y1.
    gl::Acc a0, a1, a2;
    if (valid) {
#pragma unroll
      for (uint32_t k = 0; k < K; ++k) {
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
    for (int m = 1; m < (int)LPC; m <<= 1) {
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
    if constexpr (S::prefetch) {
#pragma unroll
      for (uint32_t k = 0; k < K; ++k) {
        y0[k] = n0[k];
        if constexpr (YEXT) y1[k] = n1[k];
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

// Tiles of a column (each gives one partial sum) for the shape that serves N
// rows; 0 selects the one-thread-per-column kernel.
template <uint32_t KIND> uint32_t tile_rows(uint64_t count) {
  using T = Shapes<KIND>;
  if (count >= BE_LARGE_COUNT && count >= T::large::rows) return T::large::rows;
  if (count >= T::mid::rows) return T::mid::rows;
  if (count >= T::small::rows) return T::small::rows;
  return 0;
}

uint64_t eval_tiles(uint32_t kind, uint64_t count) {
  const uint32_t rows = kind == BASE_AT_BASE  ? tile_rows<BASE_AT_BASE>(count)
                        : kind == BASE_AT_EXT ? tile_rows<BASE_AT_EXT>(count)
                                              : tile_rows<EXT_AT_EXT>(count);
  return rows ? count / rows : 0;
}

// Workgroups of one tile-kernel variant resident on the current device (CUs x
// workgroups per CU at its register use), queried once per device.
template <uint32_t KIND, class S> uint32_t resident_wgs() {
  constexpr int MAX_DEVICES = 64;
  static std::mutex mu;
  static uint32_t cache[MAX_DEVICES] = {};
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= MAX_DEVICES) return 768;
  std::lock_guard<std::mutex> lk(mu);
  if (!cache[dev]) {
    int cus = 0, per_cu = 0;
    const bool ok = hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) == hipSuccess &&
                    hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, bary_eval_tiles_kernel<KIND, true, S>,
                                                                 BE_NT, 0) == hipSuccess &&
                    cus > 0 && per_cu > 0;
    cache[dev] = ok ? (uint32_t)(cus * per_cu) : 768u;
  }
  return cache[dev];
}

// Launch grid: tiles x column-pass groups, about `target` workgroups (more
// for the short-column shapes, whose tiles are few).
template <uint32_t KIND, class S>
void launch_tiles(const uint64_t *ys, const uint64_t *coeffs, uint64_t count, uint32_t num_polys, uint64_t *partial,
                  uint64_t *evals, uint64_t target, hipStream_t s) {
  constexpr bool OUT_EXT = KIND != BASE_AT_BASE;
  const uint64_t tiles = count / S::rows;
  const uint64_t gx = (tiles + BE_WAVES - 1) / BE_WAVES;
  const uint32_t passes = (num_polys + S::cols - 1) / S::cols;
  uint64_t gy = (target + gx - 1) / gx;
  if (gy > passes) gy = passes;
  if (gy > 65535) gy = 65535;
  const dim3 grid((uint32_t)gx, (uint32_t)gy);
  const bool aligned = !(((uintptr_t)ys | (uintptr_t)coeffs) & 15);
  if (aligned)
    bary_eval_tiles_kernel<KIND, true, S><<<grid, BE_NT, 0, s>>>(ys, coeffs, count, num_polys, partial, tiles);
  else
    bary_eval_tiles_kernel<KIND, false, S><<<grid, BE_NT, 0, s>>>(ys, coeffs, count, num_polys, partial, tiles);
  bary_eval_final_kernel<OUT_EXT><<<dim3(num_polys), BE_NT, 0, s>>>(partial, tiles, evals);
}

template <uint32_t KIND>
void launch_eval(const uint64_t *ys, const uint64_t *coeffs, uint64_t count, uint32_t num_polys, uint64_t *partial,
                 uint64_t *evals, hipStream_t s) {
  using T = Shapes<KIND>;
  const uint32_t rows = tile_rows<KIND>(count);
  if (rows == 0)
    bary_eval_small_kernel<KIND><<<dim3((num_polys + 255) / 256), 256, 0, s>>>(ys, coeffs, (uint32_t)count,
                                                                               num_polys, evals);
  else if (rows == T::large::rows && count >= BE_LARGE_COUNT)  // about three whole rounds of resident workgroups
    launch_tiles<KIND, typename T::large>(ys, coeffs, count, num_polys, partial, evals,
                                          3ull * resident_wgs<KIND, typename T::large>(), s);
  else if (rows == T::mid::rows)
    launch_tiles<KIND, typename T::mid>(ys, coeffs, count, num_polys, partial, evals, 8192, s);
  else
    launch_tiles<KIND, typename T::small>(ys, coeffs, count, num_polys, partial, evals, 8192, s);
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
  // coset^N and f = coset / (N coset^N): host arithmetic on two scalars
  const uint64_t c = gl::canon(coset);
  const uint64_t cn = gl::sqn(c, log_count);
  const uint64_t f = gl::mul(c, gl::inv(gl::mul(count, cn)));
  const dim3 grid((uint32_t)((count + LC_NT * LC_B - 1) / (LC_NT * LC_B)));
  if (ext)
    bary_coeffs_kernel<true><<<grid, LC_NT, 0, s>>>(tab, x, common_factor, c, cn, f, coeffs, log_count, bit_reverse);
  else
    bary_coeffs_kernel<false><<<grid, LC_NT, 0, s>>>(tab, x, common_factor, c, cn, f, coeffs, log_count, bit_reverse);
  return finish();
}

extern "C" int bc_batch_eval_temp_sizes(uint32_t kind, uint64_t stride, uint32_t num_polys, size_t *partial_elems,
                                        size_t *final_bytes) {
  using namespace bc;
  if (kind > EXT_AT_EXT || !power_of_two(stride) || stride > (1ull << 31) || !partial_elems || !final_bytes)
    return BC_ERROR_INVALID_VALUE;
  const uint64_t tiles = eval_tiles(kind, stride);
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
  if (!ys || !coeffs || !evals || (eval_tiles(kind, stride) && (!temp_partial || temp_partial_elems < need)))
    return BC_ERROR_INVALID_VALUE;
  hipStream_t s = (hipStream_t)stream;
  switch (kind) {
  case BASE_AT_BASE: launch_eval<BASE_AT_BASE>(ys, coeffs, stride, num_polys, temp_partial, evals, s); break;
  case BASE_AT_EXT: launch_eval<BASE_AT_EXT>(ys, coeffs, stride, num_polys, temp_partial, evals, s); break;
  default: launch_eval<EXT_AT_EXT>(ys, coeffs, stride, num_polys, temp_partial, evals, s); break;
  }
  return finish();
}

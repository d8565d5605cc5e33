// Poseidon and Poseidon2 (Goldilocks, width 12 = rate 8 + capacity 4, 8 full
// and 22 partial rounds, S-box x^7) Merkle-tree hashing: leaf hashes of
// column-major matrices, node hashes of digest pairs, and whole trees.
//
// Reference: src/poseidon.rs:129-328 (argument checks, tree layout) and the
// single-thread kernels native/poseidon_single_thread.cu:14-48,
// native/poseidon_single_thread.cuh:18-145, native/poseidon2_single_thread.cu,
// native/poseidon2_single_thread.cuh, native/poseidon_utils.cuh:8-61.  Both
// hashes read the same 360 round constants.
//
// Round constants.  The reference compiles boojum's 360 textbook round
// constants into its kernels and runs the permutation in a fused form derived
// from them at build time (build/poseidon_constants.rs).  Here the caller
// supplies the 360 constants once (bc_poseidon_set_round_constants); the
// library derives the same fused form on the host and keeps it in a
// per-device table that every hashing kernel reads through the scalar cache:
//   - rounds 0-2 and 27-29: textbook (constants, S-box on all 12, MDS);
//   - round 3: its constants and S-box, then the fused constant layer and one
//     dense 12 x 12 matrix (round 3's MDS times the first partial round's
//     dense factor);
//   - 22 partial rounds: S-box on element 0 plus one fused constant, then a
//     sparse matrix (first row and first column plus the identity);
//   - round 26: S-box and MDS only (its constants were moved into the
//     partial rounds).
// Derivation: each partial round's MDS is N = B A with A = diag(1, N^)
// (commutes with the one-element S-box layer) and B = [[1, n^T N^-1],
// [w, I]] sparse; the round before then has N = A M.  Constants move back
// through M^-1: element 0 lands after the previous round's S-box, elements
// 1..11 pass that S-box untouched.
//
// One thread hashes one leaf or node: the permutation is a serial chain of
// field products with no data shared between hashes, and 12 state words fit
// in registers at full occupancy.  Bound: VALU issue (about 25k VALU per
// permutation against 64 bytes absorbed).
#include "../../include/boojum_ntt.h"
#include "gl_field.hpp"

#include <hip/hip_runtime.h>

#include <string.h>

#include <mutex>
#include <shared_mutex>
#include <vector>

namespace bc {
namespace {

constexpr int PT = 12, RATE = 8, CAP = 4, ROUNDS = 30, HALF_FULL = 4, PARTIAL = 22;
constexpr unsigned MDS_EXPS[PT] = {0, 0, 1, 0, 3, 5, 1, 8, 12, 3, 16, 10};

// Per-device table (canonical values).
struct PosTable {
  uint64_t rc_full[7][PT];  // textbook constants of rounds 0-3 and 27-29
  uint64_t rcf[PT];         // fused constant layer after round 3's S-box
  uint64_t dense[PT][PT];   // round 3's fused dense matrix
  uint64_t sbox_c[PARTIAL]; // constant after each partial round's S-box
  uint64_t vs[PARTIAL][PT - 1];     // first row (beyond the 1 at [0][0])
  uint64_t w_hats[PARTIAL][PT - 1]; // first column (below [0][0])
  // Poseidon2 (textbook form): constants of the full rounds 0-3 and 26-29,
  // and element 0's constant of each partial round
  uint64_t p2_full[8][PT];
  uint64_t p2_partial[PARTIAL];
};

__device__ PosTable pos_table;

// The table, through an address the optimizer cannot see through: a kernel
// that permutes in a loop then reloads the constants it needs in each pass
// (scalar-cache hits) instead of hoisting them all into more SGPRs than a
// wave has, which spilled 490 of them to VGPR lanes in the Poseidon leaf
// kernel (5% faster leaves without; Poseidon2, with fewer constants, keeps
// the direct table, 1% faster that way).
__device__ __forceinline__ const PosTable &table() {
  const PosTable *p = &pos_table;
  asm volatile("" : "+s"(p));
  return *p;
}

__device__ __forceinline__ uint64_t sbox(uint64_t x) {
  const uint64_t x2 = gl::mul(x, x);
  const uint64_t x3 = gl::mul(x2, x);
  const uint64_t x4 = gl::mul(x2, x2);
  return gl::mul(x3, x4);
}

// y_r = sum_i 2^e_i x_((r + i) mod 12): the shifted words summed into 96
// bits (12 terms below 2^80) and reduced once.
__device__ __forceinline__ void mds(uint64_t (&s)[PT]) {
  uint64_t y[PT];
#pragma unroll
  for (int r = 0; r < PT; ++r) {
    uint64_t lo = 0;
    uint32_t hi = 0;
#pragma unroll
    for (int i = 0; i < PT; ++i) {
      const uint64_t v = s[(r + i) % PT];
      const unsigned e = MDS_EXPS[i];
      const uint64_t t = e ? v << e : v;
      lo += t;
      hi += (e ? (uint32_t)(v >> (64 - e)) : 0u) + (lo < t ? 1u : 0u);
    }
    y[r] = gl::reduce_limbs((uint32_t)lo, (uint32_t)(lo >> 32), hi, 0u);
  }
#pragma unroll
  for (int r = 0; r < PT; ++r) s[r] = y[r];
}

__device__ __forceinline__ void full_round(uint64_t (&s)[PT], const uint64_t *rc) {
#pragma unroll
  for (int i = 0; i < PT; ++i) s[i] = sbox(rc ? gl::add(s[i], rc[i]) : s[i]);
}

__device__ __forceinline__ void permute(uint64_t (&s)[PT]) {
  const PosTable &t = table();
#pragma unroll 1
  for (int r = 0; r < HALF_FULL - 1; ++r) {
    full_round(s, t.rc_full[r]);
    mds(s);
  }
  full_round(s, t.rc_full[HALF_FULL - 1]);
  {
    uint64_t v[PT];
#pragma unroll
    for (int i = 0; i < PT; ++i) v[i] = gl::add(s[i], t.rcf[i]);
#pragma unroll
    for (int r = 0; r < PT; ++r) {
      gl::Acc acc;
#pragma unroll
      for (int c = 0; c < PT; ++c) gl::mac(acc, v[c], t.dense[r][c]);
      s[r] = gl::acc_reduce(acc);
    }
  }
#pragma unroll 1
  for (int k = 0; k < PARTIAL; ++k) {
    const uint64_t s0 = gl::add(sbox(s[0]), t.sbox_c[k]);
    gl::Acc acc;
    acc.l0 = (uint32_t)s0;
    acc.l1 = (uint32_t)(s0 >> 32);
#pragma unroll
    for (int i = 1; i < PT; ++i) gl::mac(acc, s[i], t.vs[k][i - 1]);
#pragma unroll
    for (int i = 1; i < PT; ++i) s[i] = gl::add(gl::mul(s0, t.w_hats[k][i - 1]), s[i]);
    s[0] = gl::acc_reduce(acc);
  }
  full_round(s, nullptr);
  mds(s);
#pragma unroll 1
  for (int r = 0; r < 3; ++r) {
    full_round(s, t.rc_full[HALF_FULL + r]);
    mds(s);
  }
}

// ---- Poseidon2 (native/poseidon2_single_thread.cu:7-24,
// native/poseidon2_single_thread.cuh:12-66): external matrix M_E = circ(2 M4,
// M4, M4) over the three 4-tiles, internal matrix M_I = J + diag(2^mu_i).
constexpr unsigned P2_MU[PT] = {4, 14, 11, 8, 0, 5, 2, 9, 13, 6, 3, 12};

// M4 = [[5,7,1,3],[4,6,1,1],[1,3,5,7],[1,1,4,6]] on one tile, by the
// addition chain of eprint 2023/323 Appendix B (8 additions, 4 doublings).
__device__ __forceinline__ void m4_tile(uint64_t *x) {
  const uint64_t t0 = gl::add(x[0], x[1]);
  const uint64_t t1 = gl::add(x[2], x[3]);
  const uint64_t t2 = gl::add(gl::add(x[1], x[1]), t1);
  const uint64_t t3 = gl::add(gl::add(x[3], x[3]), t0);
  const uint64_t t1x2 = gl::add(t1, t1), t0x2 = gl::add(t0, t0);
  const uint64_t t4 = gl::add(gl::add(t1x2, t1x2), t3);
  const uint64_t t5 = gl::add(gl::add(t0x2, t0x2), t2);
  x[0] = gl::add(t3, t5);
  x[1] = t5;
  x[2] = gl::add(t2, t4);
  x[3] = t4;
}

__device__ __forceinline__ void m_ext(uint64_t (&s)[PT]) {
  m4_tile(s);
  m4_tile(s + 4);
  m4_tile(s + 8);
#pragma unroll
  for (int j = 0; j < 4; ++j) {
    const uint64_t sum = gl::add(gl::add(s[j], s[4 + j]), s[8 + j]);
    s[j] = gl::add(s[j], sum);
    s[4 + j] = gl::add(s[4 + j], sum);
    s[8 + j] = gl::add(s[8 + j], sum);
  }
}

// y_i = sum + 2^mu_i x_i: the shifted word and the sum added in 96 bits,
// one reduction.
__device__ __forceinline__ void m_int(uint64_t (&s)[PT]) {
  uint64_t sum = s[0];
#pragma unroll
  for (int i = 1; i < PT; ++i) sum = gl::add(sum, s[i]);
#pragma unroll
  for (int i = 0; i < PT; ++i) {
    const unsigned e = P2_MU[i];
    uint64_t lo = e ? s[i] << e : s[i];
    uint32_t hi = e ? (uint32_t)(s[i] >> (64 - e)) : 0u;
    lo += sum;
    hi += lo < sum ? 1u : 0u;
    s[i] = gl::reduce_limbs((uint32_t)lo, (uint32_t)(lo >> 32), hi, 0u);
  }
}

__device__ __forceinline__ void permute2(uint64_t (&s)[PT]) {
  const PosTable &t = pos_table;  // its few constants hoist without spills
  m_ext(s);
#pragma unroll 1
  for (int r = 0; r < HALF_FULL; ++r) {
    full_round(s, t.p2_full[r]);
    m_ext(s);
  }
#pragma unroll 1
  for (int k = 0; k < PARTIAL; ++k) {
    s[0] = sbox(gl::add(s[0], t.p2_partial[k]));
    m_int(s);
  }
#pragma unroll 1
  for (int r = 0; r < HALF_FULL; ++r) {
    full_round(s, t.p2_full[HALF_FULL + r]);
    m_ext(s);
  }
}

// Hash selector of the kernels: 0 Poseidon, 1 Poseidon2.
template <int H> __device__ __forceinline__ void permute_h(uint64_t (&s)[PT]) {
  if constexpr (H == 0) permute(s);
  else permute2(s);
}

// native/poseidon_utils.cuh:8-42: leaf g absorbs the values (row r, column c)
// at values[g rows + r + c rows count] in the order c rows + r, 8 at a time
// into state[0..8] (the last chunk zero-padded), permuting after each chunk.
template <int H>
__global__ __launch_bounds__(256) void poseidon_leaves_kernel(const uint64_t *__restrict__ values,
                                                              uint64_t *__restrict__ results, uint32_t rows,
                                                              uint32_t cols, uint32_t count, bool load_intermediate,
                                                              bool store_intermediate) {
  const uint32_t g = blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= count) return;
  uint64_t s[PT];
#pragma unroll
  for (int i = 0; i < PT; ++i) s[i] = 0;
  if (load_intermediate) {
#pragma unroll
    for (int i = 0; i < CAP; ++i) s[RATE + i] = gl::canon(results[g + (uint64_t)i * count]);
  }
  const uint64_t col_stride = (uint64_t)rows * count;
  const uint64_t *base = values + (uint64_t)g * rows;
  const uint64_t total = (uint64_t)rows * cols;
  uint32_t r = 0, c = 0;  // position of `off` = c rows + r
  for (uint64_t off = 0; off < total; off += RATE) {
#pragma unroll
    for (int i = 0; i < RATE; ++i) {
      s[i] = c < cols ? gl::canon(base[r + c * col_stride]) : 0ull;
      if (++r == rows) {
        r = 0;
        ++c;
      }
    }
    permute_h<H>(s);
  }
#pragma unroll
  for (int i = 0; i < CAP; ++i) results[g + (uint64_t)i * count] = store_intermediate ? s[RATE + i] : s[i];
}

// native/poseidon_utils.cuh:44-61: node g hashes children 2g, 2g + 1 of the
// layer below (4 columns of 2 count digests): state[i] = left[i],
// state[4 + i] = right[i], capacity zero.
template <int H>
__global__ __launch_bounds__(256) void poseidon_nodes_kernel(const uint64_t *__restrict__ values,
                                                             uint64_t *__restrict__ results, uint32_t count) {
  const uint32_t g = blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= count) return;
  uint64_t s[PT];
#pragma unroll
  for (int i = 0; i < CAP; ++i) {
    const uint64_t *pair = values + 2 * ((uint64_t)g + (uint64_t)i * count);
    s[i] = gl::canon(pair[0]);
    s[CAP + i] = gl::canon(pair[1]);
    s[RATE + i] = 0;
  }
  permute_h<H>(s);
#pragma unroll
  for (int i = 0; i < CAP; ++i) results[g + (uint64_t)i * count] = s[i];
}

// ---- cooperative node hashing: one state over 16 lanes ------------------------
// A small layer is one dependent permutation chain per lane, about 28k
// instructions: ~91 us however few nodes the layer has.  Here lane j of a
// 16-lane group owns state element j (lanes 12-15 compute values nobody
// reads), so each lane runs one S-box and one matrix row per round; matrix
// rows gather the other elements with __shfl inside the group, and the
// constants are staged in LDS.  Used for layers of at most COOP_MAX_NODES
// nodes (the reference likewise switches to cooperative kernels for small
// layers at the same size, src/poseidon.rs:71,279-283).
constexpr uint32_t COOP_LANES = 16, COOP_MAX_NODES = 1u << 14;

__device__ __forceinline__ uint64_t grp(uint64_t v, uint32_t lane) { return __shfl(v, (int)lane, (int)COOP_LANES); }

// sum over the 16 lanes of the group (lanes 12-15 must contribute 0)
__device__ __forceinline__ uint64_t grp_sum(uint64_t v) {
#pragma unroll
  for (uint32_t m = 8; m >= 1; m >>= 1) v = gl::add(v, __shfl_xor(v, (int)m, (int)COOP_LANES));
  return v;
}

__device__ __forceinline__ uint64_t coop_mds_row(uint64_t x, uint32_t j) {
  uint64_t lo = 0;
  uint32_t hi = 0;
#pragma unroll
  for (int i = 0; i < PT; ++i) {
    const uint64_t v = grp(x, (j + (uint32_t)i) % PT);
    const unsigned e = MDS_EXPS[i];
    const uint64_t t = e ? v << e : v;
    lo += t;
    hi += (e ? (uint32_t)(v >> (64 - e)) : 0u) + (lo < t ? 1u : 0u);
  }
  return gl::reduce_limbs((uint32_t)lo, (uint32_t)(lo >> 32), hi, 0u);
}

__device__ __forceinline__ void coop_permute(uint64_t &x, uint32_t j, const PosTable &t) {
  const uint32_t jj = j < PT ? j : PT - 1;  // table row of this lane (lanes 12-15 reuse row 11)
  const uint32_t jm = jj ? jj - 1 : 0;      // index into the 11-entry sparse vectors
#pragma unroll 1
  for (int r = 0; r < HALF_FULL - 1; ++r) x = coop_mds_row(sbox(gl::add(x, t.rc_full[r][jj])), j);
  {
    const uint64_t v = gl::add(sbox(gl::add(x, t.rc_full[HALF_FULL - 1][jj])), t.rcf[jj]);
    gl::Acc acc;
#pragma unroll
    for (int c = 0; c < PT; ++c) gl::mac(acc, grp(v, (uint32_t)c), t.dense[jj][c]);
    x = gl::acc_reduce(acc);
  }
#pragma unroll 1
  for (int k = 0; k < PARTIAL; ++k) {
    const uint64_t s0 = grp(gl::add(sbox(x), t.sbox_c[k]), 0);
    const uint64_t term = j == 0 ? s0 : (j < PT ? gl::mul(x, t.vs[k][jm]) : 0ull);
    const uint64_t y0 = grp_sum(term);
    x = j == 0 ? y0 : gl::add(gl::mul(s0, t.w_hats[k][jm]), x);
  }
  x = coop_mds_row(sbox(x), j);
#pragma unroll 1
  for (int r = 0; r < 3; ++r) x = coop_mds_row(sbox(gl::add(x, t.rc_full[HALF_FULL + r][jj])), j);
}

// sum_c m_c v_c for small m_c (M4 entries, at most 7): 67-bit products
// summed in 96 bits, one reduction
__device__ __forceinline__ uint64_t small_dot4(const uint64_t *v, const uint32_t *m) {
  uint64_t lo = 0;
  uint32_t hi = 0;
#pragma unroll
  for (int c = 0; c < 4; ++c) {
    const uint64_t p = v[c] * m[c];
    const uint32_t ph = (uint32_t)__umul64hi(v[c], (uint64_t)m[c]);
    lo += p;
    hi += ph + (lo < p ? 1u : 0u);
  }
  return gl::reduce_limbs((uint32_t)lo, (uint32_t)(lo >> 32), hi, 0u);
}

__device__ __forceinline__ uint64_t coop_m_ext(uint64_t x, uint32_t j, const uint32_t *m4row) {
  const uint32_t tile = j & ~3u, q = j & 3u;
  uint64_t v[4];
#pragma unroll
  for (int c = 0; c < 4; ++c) v[c] = grp(x, tile + (uint32_t)c);
  const uint64_t y = small_dot4(v, m4row);
  const uint64_t tot = gl::add(gl::add(grp(y, q), grp(y, 4 + q)), grp(y, 8 + q));
  return gl::add(y, tot);
}

__device__ __forceinline__ uint64_t coop_m_int(uint64_t x, uint32_t j) {
  const uint64_t sum = grp_sum(j < PT ? x : 0ull);
  const unsigned e = P2_MU[j < PT ? j : 0];
  uint64_t lo = e ? x << e : x;
  uint32_t hi = e ? (uint32_t)(x >> (64 - e)) : 0u;
  lo += sum;
  hi += lo < sum ? 1u : 0u;
  return gl::reduce_limbs((uint32_t)lo, (uint32_t)(lo >> 32), hi, 0u);
}

__device__ __forceinline__ void coop_permute2(uint64_t &x, uint32_t j, const PosTable &t) {
  constexpr uint32_t M4[4][4] = {{5, 7, 1, 3}, {4, 6, 1, 1}, {1, 3, 5, 7}, {1, 1, 4, 6}};
  uint32_t m4row[4];
#pragma unroll
  for (int c = 0; c < 4; ++c) {
    const uint32_t q = j & 3u;
    m4row[c] = q == 0 ? M4[0][c] : (q == 1 ? M4[1][c] : (q == 2 ? M4[2][c] : M4[3][c]));
  }
  const uint32_t jj = j < PT ? j : PT - 1;
  x = coop_m_ext(x, j, m4row);
#pragma unroll 1
  for (int r = 0; r < HALF_FULL; ++r) x = coop_m_ext(sbox(gl::add(x, t.p2_full[r][jj])), j, m4row);
#pragma unroll 1
  for (int k = 0; k < PARTIAL; ++k) {
    const uint64_t s = sbox(gl::add(x, t.p2_partial[k]));
    x = coop_m_int(j == 0 ? s : x, j);
  }
#pragma unroll 1
  for (int r = 0; r < HALF_FULL; ++r) x = coop_m_ext(sbox(gl::add(x, t.p2_full[HALF_FULL + r][jj])), j, m4row);
}

template <int H>
__global__ __launch_bounds__(256) void poseidon_nodes_coop_kernel(const uint64_t *__restrict__ values,
                                                                  uint64_t *__restrict__ results, uint32_t count) {
  __shared__ PosTable lt;
  {
    const uint64_t *src = (const uint64_t *)&pos_table;
    uint64_t *dst = (uint64_t *)&lt;
    for (uint32_t i = threadIdx.x; i < sizeof(PosTable) / sizeof(uint64_t); i += blockDim.x) dst[i] = src[i];
  }
  __syncthreads();
  const uint32_t g = blockIdx.x * (blockDim.x / COOP_LANES) + threadIdx.x / COOP_LANES;
  const uint32_t j = threadIdx.x % COOP_LANES;
  if (g >= count) return;  // whole groups leave together
  // element j: left[j] (j < 4), right[j - 4] (4 <= j < 8), capacity 0
  uint64_t x = 0;
  if (j < 2 * CAP) x = gl::canon(values[2 * ((uint64_t)g + (uint64_t)(j & 3u) * count) + (j >> 2)]);
  if constexpr (H == 0) coop_permute(x, j, lt);
  else coop_permute2(x, j, lt);
  if (j < CAP) results[g + (uint64_t)j * count] = x;
}

// Cooperative leaves (src/poseidon.rs:194-216): the absorption of
// poseidon_leaves_kernel with one state over a 16-lane group; for each chunk,
// lane j < 8 loads position off + j (= c rows + r) into state element j.
template <int H>
__global__ __launch_bounds__(256) void poseidon_leaves_coop_kernel(const uint64_t *__restrict__ values,
                                                                   uint64_t *__restrict__ results,
                                                                   uint32_t log_rows, uint32_t cols, uint32_t count,
                                                                   bool load_intermediate, bool store_intermediate) {
  __shared__ PosTable lt;
  {
    const uint64_t *src = (const uint64_t *)&pos_table;
    uint64_t *dst = (uint64_t *)&lt;
    for (uint32_t i = threadIdx.x; i < sizeof(PosTable) / sizeof(uint64_t); i += blockDim.x) dst[i] = src[i];
  }
  __syncthreads();
  const uint32_t g = blockIdx.x * (blockDim.x / COOP_LANES) + threadIdx.x / COOP_LANES;
  const uint32_t j = threadIdx.x % COOP_LANES;
  if (g >= count) return;  // whole groups leave together
  const uint64_t rows = 1ull << log_rows;
  uint64_t x = 0;
  if (load_intermediate && j >= RATE && j < PT) x = gl::canon(results[g + (uint64_t)(j - RATE) * count]);
  const uint64_t col_stride = rows * count;
  const uint64_t *base = values + (uint64_t)g * rows;
  const uint64_t total = rows * cols;
  for (uint64_t off = 0; off < total; off += RATE) {
    if (j < RATE) {
      const uint64_t k = off + j;
      x = k < total ? gl::canon(base[(k & (rows - 1)) + (k >> log_rows) * col_stride]) : 0ull;
    }
    if constexpr (H == 0) coop_permute(x, j, lt);
    else coop_permute2(x, j, lt);
  }
  // the digest is state[0..4], the stored capacity state[8..12]
  if (store_intermediate) {
    if (j >= RATE && j < PT) results[g + (uint64_t)(j - RATE) * count] = x;
  } else if (j < CAP) {
    results[g + (uint64_t)j * count] = x;
  }
}

// ---- Merkle openings: row and path gathers (native/poseidon_common.cu:9-37) --

// result[col][i rpi + r] = values[col][indexes[i] rpi + r] for r < rpi =
// 2^log_rpi; rows past the source (an index out of range) are skipped, the
// destination left as it was.  x: destination rows, y: columns.
__global__ __launch_bounds__(256) void gather_rows_kernel(const uint32_t *__restrict__ indexes, uint64_t dst_rows,
                                                          uint32_t log_rpi, const uint64_t *__restrict__ values,
                                                          uint64_t values_stride, uint64_t values_rows,
                                                          uint64_t *__restrict__ results, uint64_t results_stride,
                                                          uint32_t cols) {
  const uint64_t dst = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (dst >= dst_rows) return;
  const uint64_t src = ((uint64_t)indexes[dst >> log_rpi] << log_rpi) + (dst & ((1ull << log_rpi) - 1));
  if (src >= values_rows) return;
  for (uint32_t c = blockIdx.y; c < cols; c += gridDim.y) results[c * results_stride + dst] = values[c * values_stride + src];
}

// Siblings along each index's path: for layer l < layers, column c (of 4) and
// index i, results[l count 4 + c count + i] = the layer-l digest element c of
// node (indexes[i] >> l) ^ 1.  Layer l of the tree buffer starts at
// 4 (2^(L+1) - 2^(L+1-l)) and holds 4 columns of 2^(L-l) (L = log2 leaves).
__global__ __launch_bounds__(256) void gather_merkle_paths_kernel(const uint32_t *__restrict__ indexes,
                                                                  uint32_t count, const uint64_t *__restrict__ values,
                                                                  uint32_t log_leaves, uint64_t *__restrict__ results) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= count) return;
  const uint32_t c = blockIdx.y, l = blockIdx.z;
  const uint64_t leaf = indexes[i];
  if (leaf >> log_leaves) return;  // index out of range: nothing to gather
  const uint64_t layer_off = ((uint64_t)CAP << (log_leaves + 1)) - ((uint64_t)CAP << (log_leaves + 1 - l));
  const uint64_t src = layer_off + ((uint64_t)c << (log_leaves - l)) + ((leaf >> l) ^ 1ull);
  results[(uint64_t)l * count * CAP + (uint64_t)count * c + i] = values[src];
}

// ---- host: the fused form ---------------------------------------------------

using Mat = std::vector<uint64_t>;  // row-major n x n

uint64_t mds_entry(int r, int c) { return 1ull << MDS_EXPS[((c - r) % PT + PT) % PT]; }

bool invert(Mat a, int n, Mat &out) {
  out.assign((size_t)n * n, 0);
  for (int i = 0; i < n; ++i) out[(size_t)i * n + i] = 1;
  for (int col = 0; col < n; ++col) {
    int piv = -1;
    for (int r = col; r < n && piv < 0; ++r)
      if (a[(size_t)r * n + col]) piv = r;
    if (piv < 0) return false;
    for (int j = 0; j < n; ++j) {
      std::swap(a[(size_t)col * n + j], a[(size_t)piv * n + j]);
      std::swap(out[(size_t)col * n + j], out[(size_t)piv * n + j]);
    }
    const uint64_t iv = gl::inv(a[(size_t)col * n + col]);
    for (int j = 0; j < n; ++j) {
      a[(size_t)col * n + j] = gl::mul(a[(size_t)col * n + j], iv);
      out[(size_t)col * n + j] = gl::mul(out[(size_t)col * n + j], iv);
    }
    for (int r = 0; r < n; ++r) {
      const uint64_t f = a[(size_t)r * n + col];
      if (r == col || !f) continue;
      for (int j = 0; j < n; ++j) {
        a[(size_t)r * n + j] = gl::sub(a[(size_t)r * n + j], gl::mul(f, a[(size_t)col * n + j]));
        out[(size_t)r * n + j] = gl::sub(out[(size_t)r * n + j], gl::mul(f, out[(size_t)col * n + j]));
      }
    }
  }
  return true;
}

void mat_vec(const Mat &m, const uint64_t *v, uint64_t *out) {
  for (int r = 0; r < PT; ++r) {
    uint64_t acc = 0;
    for (int c = 0; c < PT; ++c) acc = gl::add(acc, gl::mul(m[(size_t)r * PT + c], v[c]));
    out[r] = acc;
  }
}

bool derive(const uint64_t *rc_in, PosTable &t) {
  uint64_t rc[ROUNDS * PT];
  for (int i = 0; i < ROUNDS * PT; ++i) rc[i] = gl::canon(rc_in[i]);
  Mat m((size_t)PT * PT), minv;
  for (int r = 0; r < PT; ++r)
    for (int c = 0; c < PT; ++c) m[(size_t)r * PT + c] = mds_entry(r, c);
  if (!invert(m, PT, minv)) return false;
  for (int r = 0; r < HALF_FULL; ++r)
    for (int i = 0; i < PT; ++i) t.rc_full[r][i] = rc[r * PT + i];
  for (int r = 0; r < 3; ++r)
    for (int i = 0; i < PT; ++i) t.rc_full[HALF_FULL + r][i] = rc[(HALF_FULL + PARTIAL + 1 + r) * PT + i];
  for (int r = 0; r < HALF_FULL; ++r)
    for (int i = 0; i < PT; ++i) {
      t.p2_full[r][i] = rc[r * PT + i];
      t.p2_full[HALF_FULL + r][i] = rc[(HALF_FULL + PARTIAL + r) * PT + i];
    }
  for (int k = 0; k < PARTIAL; ++k) t.p2_partial[k] = rc[(HALF_FULL + k) * PT];
  // constants of round 26, then of each partial round, moved backwards
  uint64_t cur[PT], e[PT];
  for (int i = 0; i < PT; ++i) cur[i] = rc[(HALF_FULL + PARTIAL) * PT + i];
  for (int k = HALF_FULL + PARTIAL - 1; k >= HALF_FULL; --k) {
    mat_vec(minv, cur, e);
    t.sbox_c[k - HALF_FULL] = e[0];
    for (int i = 0; i < PT; ++i) cur[i] = i ? gl::add(rc[k * PT + i], e[i]) : rc[k * PT + i];
  }
  mat_vec(minv, cur, t.rcf);
  // matrices
  Mat n = m;
  for (int k = HALF_FULL + PARTIAL - 1; k >= HALF_FULL; --k) {
    const int pr = k - HALF_FULL;
    Mat nh((size_t)(PT - 1) * (PT - 1)), nhinv;
    for (int i = 0; i < PT - 1; ++i)
      for (int j = 0; j < PT - 1; ++j) nh[(size_t)i * (PT - 1) + j] = n[(size_t)(i + 1) * PT + j + 1];
    if (!invert(nh, PT - 1, nhinv)) return false;
    for (int j = 0; j < PT - 1; ++j) {
      uint64_t acc = 0;
      for (int i = 0; i < PT - 1; ++i) acc = gl::add(acc, gl::mul(n[i + 1], nhinv[(size_t)i * (PT - 1) + j]));
      t.vs[pr][j] = acc;
      t.w_hats[pr][j] = n[(size_t)(j + 1) * PT];
    }
    Mat next((size_t)PT * PT);
    for (int c = 0; c < PT; ++c) next[c] = m[c];
    for (int r = 1; r < PT; ++r)
      for (int c = 0; c < PT; ++c) {
        uint64_t acc = 0;
        for (int i = 1; i < PT; ++i)
          acc = gl::add(acc, gl::mul(nh[(size_t)(r - 1) * (PT - 1) + i - 1], m[(size_t)i * PT + c]));
        next[(size_t)r * PT + c] = acc;
      }
    n.swap(next);
  }
  for (int r = 0; r < PT; ++r)
    for (int c = 0; c < PT; ++c) t.dense[r][c] = n[(size_t)r * PT + c];
  return true;
}

constexpr int MAX_DEVICES = 64;
// Every hashing entry point holds this lock shared from its readiness check
// through its launches; replacing the table holds it exclusively.  So no
// hashing kernel is queued between the wait for queued work and the copy of
// the new table, on any stream or thread.
std::shared_mutex pos_mu;
bool pos_ready[MAX_DEVICES];

// Caller holds pos_mu (shared or exclusive).
int require_table() {
  int dev = 0;
  const hipError_t err = hipGetDevice(&dev);
  if (err != hipSuccess) return err;
  if (dev < 0 || dev >= MAX_DEVICES) return BC_ERROR_INVALID_VALUE;
  return pos_ready[dev] ? BC_SUCCESS : BC_ERROR_NOT_INITIALIZED;
}

using TableUse = std::shared_lock<std::shared_mutex>;

template <int H, bool COOP>
int launch_leaves(const uint64_t *values, uint64_t *results, uint32_t log_rows, uint32_t cols, uint32_t count,
                  bool load, bool store, hipStream_t s) {
  if (count == 0) return BC_SUCCESS;
  if constexpr (COOP) {
    constexpr uint32_t per_block = 256 / COOP_LANES;
    poseidon_leaves_coop_kernel<H><<<dim3((count + per_block - 1) / per_block), dim3(256), 0, s>>>(
        values, results, log_rows, cols, count, load, store);
  } else {
    poseidon_leaves_kernel<H><<<dim3((count + 255) / 256), dim3(256), 0, s>>>(values, results, 1u << log_rows, cols,
                                                                             count, load, store);
  }
  return hipGetLastError();
}

// COOP: always the cooperative kernel (the reference's
// launch_cooperative_nodes_kernel); otherwise the cooperative kernel for
// layers of at most COOP_MAX_NODES nodes and one thread per node above.
template <int H, bool COOP = false>
int launch_nodes(const uint64_t *values, uint64_t *results, uint32_t count, hipStream_t s) {
  if (count == 0) return BC_SUCCESS;
  if (COOP || count <= COOP_MAX_NODES) {
    constexpr uint32_t per_block = 256 / COOP_LANES;
    poseidon_nodes_coop_kernel<H><<<dim3((count + per_block - 1) / per_block), dim3(256), 0, s>>>(values, results,
                                                                                                 count);
  } else {
    poseidon_nodes_kernel<H><<<dim3((count + 255) / 256), dim3(256), 0, s>>>(values, results, count);
  }
  return hipGetLastError();
}

// The entry points of both hashes share these (src/poseidon.rs is generic
// over the hash in the same way).

template <int H, bool COOP = false>
int leaves_impl(const uint64_t *values, uint64_t values_len, uint64_t *results, uint64_t results_len,
                uint32_t log_rows_per_hash, bool load_intermediate, bool store_intermediate, void *stream) {
  // src/poseidon.rs:139-155
  if (log_rows_per_hash >= 32 || results_len % CAP != 0) return BC_ERROR_INVALID_VALUE;
  const uint64_t count = results_len / CAP;
  if (count == 0) return values_len == 0 ? BC_SUCCESS : BC_ERROR_INVALID_VALUE;
  if (!values || !results || count > 0xFFFFFFFFull) return BC_ERROR_INVALID_VALUE;
  const uint64_t rows = 1ull << log_rows_per_hash;
  if (values_len % (count * rows) != 0) return BC_ERROR_INVALID_VALUE;
  const uint64_t cols = values_len / (count * rows);
  if (cols > 0xFFFFFFFFull || rows * cols > 0xFFFFFFFFull) return BC_ERROR_INVALID_VALUE;
  if (store_intermediate && (rows * cols) % RATE != 0) return BC_ERROR_INVALID_VALUE;
  const int rc = require_table();
  if (rc != BC_SUCCESS) return rc;
  return launch_leaves<H, COOP>(values, results, log_rows_per_hash, (uint32_t)cols, (uint32_t)count,
                                load_intermediate, store_intermediate, (hipStream_t)stream);
}

template <int H, bool COOP = false>
int nodes_impl(const uint64_t *values, uint64_t values_len, uint64_t *results, uint64_t results_len, void *stream) {
  // src/poseidon.rs:224-232
  if (values_len % RATE != 0 || results_len % CAP != 0 || values_len != 2 * results_len)
    return BC_ERROR_INVALID_VALUE;
  const uint64_t count = results_len / CAP;
  if (count == 0) return BC_SUCCESS;
  if (!values || !results || count > 0xFFFFFFFFull) return BC_ERROR_INVALID_VALUE;
  const int rc = require_table();
  if (rc != BC_SUCCESS) return rc;
  return launch_nodes<H, COOP>(values, results, (uint32_t)count, (hipStream_t)stream);
}

template <int H>
int tree_nodes_impl(const uint64_t *values, uint64_t values_len, uint64_t *results, uint64_t results_len,
                    uint32_t layers_count, void *stream) {
  // src/poseidon.rs:262-286: each layer halves; its nodes fill the first half
  // of what remains of `results`
  if (layers_count == 0) return BC_SUCCESS;
  if (values_len % RATE != 0 || values_len != results_len) return BC_ERROR_INVALID_VALUE;
  const uint64_t leaves = values_len / RATE;
  if (leaves == 0 || (leaves & (leaves - 1)) != 0) return BC_ERROR_INVALID_VALUE;
  if (layers_count > 64 || (layers_count > 1 && (values_len >> (layers_count - 1)) < RATE))
    return BC_ERROR_INVALID_VALUE;
  const int rc = require_table();
  if (rc != BC_SUCCESS) return rc;
  const uint64_t *in = values;
  uint64_t len = values_len;
  uint64_t *out = results;
  for (uint32_t l = 0; l < layers_count; ++l) {
    const uint64_t half = len / 2;
    const int e = nodes_impl<H>(in, len, out, half, stream);
    if (e != BC_SUCCESS) return e;
    in = out;
    out += half;
    len = half;
  }
  return BC_SUCCESS;
}

template <int H>
int tree_impl(const uint64_t *values, uint64_t values_len, uint64_t *results, uint64_t results_len,
              uint32_t log_rows_per_hash, uint32_t layers_count, void *stream) {
  // src/poseidon.rs:311-328: the leaf digests fill the first half of
  // `results`, the node layers the rest
  if (layers_count == 0 || results_len % (2 * CAP) != 0) return BC_ERROR_INVALID_VALUE;
  const uint64_t leaves = results_len / (2 * CAP);
  if (leaves == 0 || layers_count - 1 >= 64 || (1ull << (layers_count - 1)) > leaves ||
      values_len % leaves != 0)
    return BC_ERROR_INVALID_VALUE;
  const uint64_t half = results_len / 2;
  const int e = leaves_impl<H>(values, values_len, results, half, log_rows_per_hash, false, false, stream);
  if (e != BC_SUCCESS) return e;
  return tree_nodes_impl<H>(results, half, results + half, half, layers_count - 1, stream);
}

}  // namespace

size_t poseidon_table_words() { return sizeof(PosTable) / sizeof(uint64_t); }

}  // namespace bc

extern "C" int bc_poseidon_set_round_constants(const uint64_t *round_constants) {
  using namespace bc;
  if (!round_constants) return BC_ERROR_INVALID_VALUE;
  int dev = 0;
  hipError_t err = hipGetDevice(&dev);
  if (err != hipSuccess) return err;
  if (dev < 0 || dev >= MAX_DEVICES) return BC_ERROR_INVALID_VALUE;
  PosTable t{};
  if (!derive(round_constants, t)) return BC_ERROR_INVALID_VALUE;
  std::unique_lock<std::shared_mutex> lk(pos_mu);
  // hashing kernels already queued on any stream of this device read the
  // table: let them finish before it changes (no new ones can be queued
  // while the lock is held)
  err = hipDeviceSynchronize();
  if (err != hipSuccess) return err;
  err = hipMemcpyToSymbol(HIP_SYMBOL(pos_table), &t, sizeof t, 0, hipMemcpyHostToDevice);
  if (err != hipSuccess) return err;
  err = hipDeviceSynchronize();
  if (err != hipSuccess) return err;
  pos_ready[dev] = true;
  return BC_SUCCESS;
}

extern "C" int bc_poseidon_fused_table(const uint64_t *round_constants, uint64_t *out, uint64_t out_words) {
  using namespace bc;
  if (!round_constants || !out || out_words < poseidon_table_words()) return BC_ERROR_INVALID_VALUE;
  PosTable t{};
  if (!derive(round_constants, t)) return BC_ERROR_INVALID_VALUE;
  memcpy(out, &t, sizeof t);
  return BC_SUCCESS;
}

extern "C" int bc_poseidon_leaves(const uint64_t *values, uint64_t values_len, uint64_t *results,
                                  uint64_t results_len, uint32_t log_rows_per_hash, bool load_intermediate,
                                  bool store_intermediate, void *stream) {
  const bc::TableUse use(bc::pos_mu);
  return bc::leaves_impl<0>(values, values_len, results, results_len, log_rows_per_hash, load_intermediate,
                            store_intermediate, stream);
}

extern "C" int bc_poseidon_nodes(const uint64_t *values, uint64_t values_len, uint64_t *results,
                                 uint64_t results_len, void *stream) {
  const bc::TableUse use(bc::pos_mu);
  return bc::nodes_impl<0>(values, values_len, results, results_len, stream);
}

extern "C" int bc_poseidon_build_merkle_tree_nodes(const uint64_t *values, uint64_t values_len, uint64_t *results,
                                                   uint64_t results_len, uint32_t layers_count, void *stream) {
  const bc::TableUse use(bc::pos_mu);
  return bc::tree_nodes_impl<0>(values, values_len, results, results_len, layers_count, stream);
}

extern "C" int bc_poseidon_build_merkle_tree(const uint64_t *values, uint64_t values_len, uint64_t *results,
                                             uint64_t results_len, uint32_t log_rows_per_hash,
                                             uint32_t layers_count, void *stream) {
  const bc::TableUse use(bc::pos_mu);
  return bc::tree_impl<0>(values, values_len, results, results_len, log_rows_per_hash, layers_count, stream);
}

extern "C" int bc_poseidon2_leaves(const uint64_t *values, uint64_t values_len, uint64_t *results,
                                   uint64_t results_len, uint32_t log_rows_per_hash, bool load_intermediate,
                                   bool store_intermediate, void *stream) {
  const bc::TableUse use(bc::pos_mu);
  return bc::leaves_impl<1>(values, values_len, results, results_len, log_rows_per_hash, load_intermediate,
                            store_intermediate, stream);
}

extern "C" int bc_poseidon2_nodes(const uint64_t *values, uint64_t values_len, uint64_t *results,
                                  uint64_t results_len, void *stream) {
  const bc::TableUse use(bc::pos_mu);
  return bc::nodes_impl<1>(values, values_len, results, results_len, stream);
}

extern "C" int bc_poseidon2_build_merkle_tree_nodes(const uint64_t *values, uint64_t values_len, uint64_t *results,
                                                    uint64_t results_len, uint32_t layers_count, void *stream) {
  const bc::TableUse use(bc::pos_mu);
  return bc::tree_nodes_impl<1>(values, values_len, results, results_len, layers_count, stream);
}

extern "C" int bc_poseidon2_build_merkle_tree(const uint64_t *values, uint64_t values_len, uint64_t *results,
                                              uint64_t results_len, uint32_t log_rows_per_hash,
                                              uint32_t layers_count, void *stream) {
  const bc::TableUse use(bc::pos_mu);
  return bc::tree_impl<1>(values, values_len, results, results_len, log_rows_per_hash, layers_count, stream);
}

extern "C" int bc_poseidon_leaves_cooperative(const uint64_t *values, uint64_t values_len, uint64_t *results,
                                              uint64_t results_len, uint32_t log_rows_per_hash,
                                              bool load_intermediate, bool store_intermediate, void *stream) {
  const bc::TableUse use(bc::pos_mu);
  return bc::leaves_impl<0, true>(values, values_len, results, results_len, log_rows_per_hash, load_intermediate,
                                  store_intermediate, stream);
}

extern "C" int bc_poseidon2_leaves_cooperative(const uint64_t *values, uint64_t values_len, uint64_t *results,
                                               uint64_t results_len, uint32_t log_rows_per_hash,
                                               bool load_intermediate, bool store_intermediate, void *stream) {
  const bc::TableUse use(bc::pos_mu);
  return bc::leaves_impl<1, true>(values, values_len, results, results_len, log_rows_per_hash, load_intermediate,
                                  store_intermediate, stream);
}

extern "C" int bc_poseidon_nodes_cooperative(const uint64_t *values, uint64_t values_len, uint64_t *results,
                                             uint64_t results_len, void *stream) {
  const bc::TableUse use(bc::pos_mu);
  return bc::nodes_impl<0, true>(values, values_len, results, results_len, stream);
}

extern "C" int bc_poseidon2_nodes_cooperative(const uint64_t *values, uint64_t values_len, uint64_t *results,
                                              uint64_t results_len, void *stream) {
  const bc::TableUse use(bc::pos_mu);
  return bc::nodes_impl<1, true>(values, values_len, results, results_len, stream);
}

extern "C" int bc_gather_rows(const uint32_t *indexes, uint32_t indexes_count, uint32_t log_rows_per_index,
                              const uint64_t *values, uint64_t values_stride, uint64_t values_rows, uint64_t *results,
                              uint64_t results_stride, uint64_t results_rows, uint32_t cols, void *stream) {
  using namespace bc;
  // src/poseidon.rs:337-346
  if (log_rows_per_index >= 32) return BC_ERROR_INVALID_VALUE;
  if (results_rows != ((uint64_t)indexes_count << log_rows_per_index)) return BC_ERROR_INVALID_VALUE;
  if (results_rows == 0 || cols == 0) return BC_SUCCESS;
  if (!indexes || !values || !results || values_stride < values_rows || results_stride < results_rows)
    return BC_ERROR_INVALID_VALUE;
  const uint64_t blocks = (results_rows + 255) / 256;
  if (blocks > 0x7FFFFFFFull) return BC_ERROR_INVALID_VALUE;
  gather_rows_kernel<<<dim3((uint32_t)blocks, cols < 65535 ? cols : 65535), dim3(256), 0, (hipStream_t)stream>>>(
      indexes, results_rows, log_rows_per_index, values, values_stride, values_rows, results, results_stride, cols);
  return hipGetLastError();
}

extern "C" int bc_gather_merkle_paths(const uint32_t *indexes, uint32_t indexes_count, const uint64_t *values,
                                      uint64_t values_len, uint64_t *results, uint64_t results_len,
                                      uint32_t layers_count, void *stream) {
  using namespace bc;
  // src/poseidon.rs:367-382
  if (values_len % CAP != 0) return BC_ERROR_INVALID_VALUE;
  const uint64_t values_count = values_len / CAP;
  if (values_count == 0 || (values_count & (values_count - 1)) != 0) return BC_ERROR_INVALID_VALUE;
  const uint32_t log_values = (uint32_t)__builtin_ctzll(values_count);
  if (log_values == 0) return BC_ERROR_INVALID_VALUE;
  const uint32_t log_leaves = log_values - 1;
  if (layers_count >= log_leaves) return BC_ERROR_INVALID_VALUE;
  if ((uint64_t)indexes_count * layers_count * CAP != results_len) return BC_ERROR_INVALID_VALUE;
  if (indexes_count == 0 || layers_count == 0) return BC_SUCCESS;
  if (!indexes || !values || !results) return BC_ERROR_INVALID_VALUE;
  gather_merkle_paths_kernel<<<dim3((indexes_count + 255) / 256, CAP, layers_count), dim3(256), 0,
                               (hipStream_t)stream>>>(indexes, indexes_count, values, log_leaves, results);
  return hipGetLastError();
}

// Standalone launches of the NTT pass kernels (one kernel per pass); the
// pass bodies and their math are in ntt_passes.hpp.
#include "ntt_passes.hpp"

#include <mutex>

namespace bc {
namespace {
using namespace pass;

// Register bounds of three workgroups per CU for the DIF (n2b) variants and
// two for the DIT (b2n) ones, except four for the plain and n^-1 12-bit b2n
// kernels.  With their LDS slots addressed as one base plus immediate
// offsets, the 12-bit n2b kernels need 124 VGPRs without spills and run 4
// workgroups per CU (LDS allows 4).  The 12-bit b2n kernels need 142 (3 per
// CU); held to 128 they spill 14 dwords and still ran 1.5-2% faster at 4 per
// CU on 2^16-2^21 columns.  Their coset variants spill 20-26 dwords at 128
// and were not faster, so they keep the bound of two.
template <int LH, bool INV, bool DIT, int MODE, bool MC>
__global__ __launch_bounds__(NT, DIT && LH == 12 && !MC && (MODE == MODE_NONE || MODE == MODE_NINV) ? 4
                                 : DIT                                                           ? 2
                                                                                                 : 3) void
    ntt_contig_kernel(FastArgs a) {
  __shared__ PassLds sh;
  contig_run<LH, INV, DIT, MODE, MC>(a, sh, blockIdx.x, a.tiles, gridDim.x);
}

// Three workgroups per CU as well (the coset-scaling variants included: their
// factors live in the pass and level twists and a 16-entry LDS table, not in
// per-thread registers), except where 168 VGPRs spill heavily: h = 6 (four
// radix-4 groups per thread in the second level; 70-76 spilled dwords at
// three per CU, 28-38% slower on 512 x 2^18 and 8 x 2^24 n2b calls) and b2n
// h = 5 (16-22 dwords, 3-4% slower).  The other b2n variants spill 4-8
// dwords at three per CU and still ran faster there than at two.
template <int LH, bool INV, bool DIT, int MODE>
__global__ __launch_bounds__(NT, LH == 6 || (LH == 5 && DIT) ? 2 : 3) void ntt_strided_kernel(FastArgs a) {
  __shared__ PassLds sh;
  const uint64_t total = (uint64_t)a.num_cols << (a.log_n - a.log_N);
  const uint64_t p_begin = (uint64_t)blockIdx.y * a.blocks_per_wg;
  const uint64_t p_end = p_begin + a.blocks_per_wg < total ? p_begin + a.blocks_per_wg : total;
  strided_run<LH, INV, DIT, MODE>(a, sh, blockIdx.x, p_begin, p_end);
}

// -------------------------------------------------------------------------
// launch dispatch (only the (direction, mode) pairs the planner can ask for)

template <int LH, bool INV, bool DIT, bool MC>
hipError_t contig_mode(const FastArgs &a, int mode, dim3 grid, hipStream_t s) {
  switch (mode) {
  case MODE_NONE: ntt_contig_kernel<LH, INV, DIT, MODE_NONE, MC><<<grid, NT, 0, s>>>(a); break;
  case MODE_NINV: ntt_contig_kernel<LH, INV, DIT, MODE_NINV, MC><<<grid, NT, 0, s>>>(a); break;
  case MODE_LDE_IN: ntt_contig_kernel<LH, INV, DIT, MODE_LDE_IN, MC><<<grid, NT, 0, s>>>(a); break;
  case MODE_LDE_OUT: ntt_contig_kernel<LH, INV, DIT, MODE_LDE_OUT, MC><<<grid, NT, 0, s>>>(a); break;
  default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

template <int LH, bool MC>
hipError_t contig_dir(const FastArgs &a, bool inv, bool dit, int mode, dim3 grid, hipStream_t s) {
  // forward transforms never scale outputs; inverse ones never scale inputs
  if (inv ? mode == MODE_LDE_IN : (mode == MODE_NINV || mode == MODE_LDE_OUT)) return hipErrorInvalidValue;
  if (inv) return dit ? contig_mode<LH, true, true, MC>(a, mode, grid, s) : contig_mode<LH, true, false, MC>(a, mode, grid, s);
  return dit ? contig_mode<LH, false, true, MC>(a, mode, grid, s) : contig_mode<LH, false, false, MC>(a, mode, grid, s);
}

template <int LH, bool INV, bool DIT>
hipError_t strided_mode(const FastArgs &a, int mode, dim3 grid, hipStream_t s) {
  switch (mode) {
  case MODE_NONE: ntt_strided_kernel<LH, INV, DIT, MODE_NONE><<<grid, NT, 0, s>>>(a); break;
  case MODE_NINV:
    if constexpr (INV) { ntt_strided_kernel<LH, INV, DIT, MODE_NINV><<<grid, NT, 0, s>>>(a); break; }
    return hipErrorInvalidValue;
  case MODE_LDE_IN:
    if constexpr (!INV && !DIT) { ntt_strided_kernel<LH, INV, DIT, MODE_LDE_IN><<<grid, NT, 0, s>>>(a); break; }
    return hipErrorInvalidValue;
  case MODE_LDE_OUT:
    if constexpr (INV && DIT) { ntt_strided_kernel<LH, INV, DIT, MODE_LDE_OUT><<<grid, NT, 0, s>>>(a); break; }
    return hipErrorInvalidValue;
  default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

template <int LH>
hipError_t strided_dir(const FastArgs &a, bool inv, bool dit, int mode, dim3 grid, hipStream_t s) {
  if (inv) return dit ? strided_mode<LH, true, true>(a, mode, grid, s) : strided_mode<LH, true, false>(a, mode, grid, s);
  return dit ? strided_mode<LH, false, true>(a, mode, grid, s) : strided_mode<LH, false, false>(a, mode, grid, s);
}

// Default grid of the contiguous pass: two whole rounds of resident
// workgroups.  The tile loop is grid-strided, so a grid that is a whole
// number of rounds leaves no partly filled last round (2048 workgroups on the
// 768 slots of 256 CUs x 3 ran 2.67 rounds and were 4% slower on 256 x 2^20).
// Queried once per device and direction: the n2b and b2n 12-bit kernels
// differ in register use and so in workgroups per CU.
constexpr int MAX_DEVICES = 64;
std::mutex g_grid_mu;
uint32_t g_grid[MAX_DEVICES][2] = {};

uint32_t default_contig_grid(bool dit) {
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= MAX_DEVICES) return 2048;
  std::lock_guard<std::mutex> lk(g_grid_mu);
  uint32_t &g = g_grid[dev][dit ? 1 : 0];
  if (!g) {
    int cus = 0, per_cu = 0;
    const bool ok =
        hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) == hipSuccess &&
        (dit ? hipOccupancyMaxActiveBlocksPerMultiprocessor(
                   &per_cu, ntt_contig_kernel<12, false, true, MODE_NONE, false>, NT, 0)
             : hipOccupancyMaxActiveBlocksPerMultiprocessor(
                   &per_cu, ntt_contig_kernel<12, false, false, MODE_NONE, false>, NT, 0)) == hipSuccess &&
        cus > 0 && per_cu > 0;
    g = ok ? (uint32_t)(2 * cus * per_cu) : 2048u;
  }
  return g;
}

}  // namespace

hipError_t launch_contig(const FastArgs &a, int log_h, bool inverse, bool dit, int mode, uint32_t max_wgs,
                         hipStream_t stream) {
  if (a.tiles == 0) return hipSuccess;
  if (max_wgs == 0) max_wgs = default_contig_grid(dit);
  const dim3 grid(a.tiles < max_wgs ? a.tiles : max_wgs);
  if (a.log_n < TILE_LOG) {  // single pass over whole columns, several per tile
    switch (log_h) {
    case 8: return contig_dir<8, true>(a, inverse, dit, mode, grid, stream);
    case 9: return contig_dir<9, true>(a, inverse, dit, mode, grid, stream);
    case 10: return contig_dir<10, true>(a, inverse, dit, mode, grid, stream);
    case 11: return contig_dir<11, true>(a, inverse, dit, mode, grid, stream);
    default: return hipErrorInvalidValue;
    }
  }
  switch (log_h) {
  case 9: return contig_dir<9, false>(a, inverse, dit, mode, grid, stream);
  case 10: return contig_dir<10, false>(a, inverse, dit, mode, grid, stream);
  case 11: return contig_dir<11, false>(a, inverse, dit, mode, grid, stream);
  case 12: return contig_dir<12, false>(a, inverse, dit, mode, grid, stream);
  default: return hipErrorInvalidValue;
  }
}

hipError_t launch_strided(const FastArgs &a, int log_h, bool inverse, bool dit, int mode, uint32_t positions,
                          uint32_t groups, hipStream_t stream) {
  if (positions == 0 || groups == 0) return hipSuccess;
  if (groups > 65535) return hipErrorInvalidConfiguration;
  const dim3 grid(positions, groups);
  switch (log_h) {
  case 4: return strided_dir<4>(a, inverse, dit, mode, grid, stream);
  case 5: return strided_dir<5>(a, inverse, dit, mode, grid, stream);
  case 6: return strided_dir<6>(a, inverse, dit, mode, grid, stream);
  case 7: return strided_dir<7>(a, inverse, dit, mode, grid, stream);
  case 8: return strided_dir<8>(a, inverse, dit, mode, grid, stream);
  default: return hipErrorInvalidValue;
  }
}

}  // namespace bc

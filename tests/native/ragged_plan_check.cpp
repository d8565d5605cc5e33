// Host-side invariants of the ragged-batch planner (ragged_build in
// era-boojum-cuda_amd/csrc/ntt_ragged.hip) on a 1000-job plan drawn like the
// ragged benchmark (log_n uniform in [10, 23], op uniform over NTT, iNTT,
// LDE x2, LDE x4), plus small edge plans.  Built host-only (no device code)
// under AddressSanitizer and UndefinedBehaviorSanitizer; no HIP call is made.
// Prints "OK <segments> <levels> <items>" on success.
#include "../../era-boojum-cuda_amd/csrc/ntt_ragged.hpp"

#include <stdio.h>
#include <stdlib.h>

#include <algorithm>
#include <random>
#include <vector>

using namespace bc;

#define REQUIRE(c)                                                        \
  do {                                                                    \
    if (!(c)) {                                                           \
      printf("FAIL %s:%d: %s\n", __FILE__, __LINE__, #c);                 \
      exit(1);                                                            \
    }                                                                     \
  } while (0)

struct Buf {
  uint64_t lo, hi;  // byte address range of one job buffer
};

// Every plan invariant the persistent kernel relies on.
static void check_plan(const std::vector<bc_ragged_job> &jobs, const RaggedHostPlan &hp) {
  REQUIRE(hp.levels >= 1 && hp.levels <= RAGGED_MAX_LEVELS);
  REQUIRE(hp.level_seg.size() == hp.levels + 1 && hp.level_item.size() == hp.levels + 1);
  REQUIRE(hp.level_seg[0] == 0 && hp.level_seg[hp.levels] == hp.segs.size());
  REQUIRE(hp.level_item[0] == 0 && hp.level_item[hp.levels] == hp.total_items);
  for (uint32_t l = 0; l < hp.levels; ++l) {
    REQUIRE(hp.level_seg[l] <= hp.level_seg[l + 1]);
    REQUIRE(hp.level_item[l] <= hp.level_item[l + 1]);
    // level l's items are exactly its segments' units
    if (hp.level_seg[l] < hp.segs.size()) REQUIRE(hp.segs[hp.level_seg[l]].first_item == hp.level_item[l]);
  }
  // queue positions: consecutive, every unit covers real work
  uint64_t item = 0;
  size_t cols_used = 0;
  for (size_t i = 0; i < hp.segs.size(); ++i) {
    const RaggedSeg &s = hp.segs[i];
    REQUIRE(s.first_item == item);
    REQUIRE(s.units >= 1 && s.per_unit >= 1 && s.total >= 1);
    if (s.positions == 0) {  // contiguous: tiles / tiles per unit
      REQUIRE(s.units == (s.total + s.per_unit - 1) / s.per_unit);
      REQUIRE(s.a.tiles == s.total);
    } else {  // strided: pair groups x r-ranges
      REQUIRE(s.per_unit <= s.total);
      REQUIRE((uint64_t)s.units == ((s.total + s.per_unit - 1) / s.per_unit) * s.positions);
    }
    REQUIRE(s.a.in == nullptr && s.a.out == nullptr);  // columns come from the table
    REQUIRE(s.col_off + s.a.num_cols <= hp.cols.size());
    cols_used = std::max<size_t>(cols_used, s.col_off + s.a.num_cols);
    item += s.units;
  }
  REQUIRE(item == hp.total_items);
  REQUIRE(cols_used == hp.cols.size());
  // chunk index: entry c names the segment holding item c << RAGGED_CHUNK_LOG
  const uint64_t chunks = (hp.total_items >> RAGGED_CHUNK_LOG) + 1;
  REQUIRE(hp.chunk_seg.size() == chunks + 1);
  for (uint64_t c = 0; c < chunks; ++c) {
    const uint64_t it = c << RAGGED_CHUNK_LOG;
    const uint32_t sg = hp.chunk_seg[c];
    REQUIRE(sg < hp.segs.size());
    if (it < hp.total_items) {
      REQUIRE(hp.segs[sg].first_item <= it && it < hp.segs[sg].first_item + hp.segs[sg].units);
    }
  }
  REQUIRE(hp.chunk_seg[chunks] == hp.segs.size() - 1);
  // every column pointer lies inside a buffer of the job that owns it
  std::vector<Buf> bufs;
  for (const bc_ragged_job &j : jobs) {
    const uint64_t n = 1ull << j.log_n;
    bufs.push_back({(uint64_t)j.inputs, (uint64_t)(j.inputs + j.stride_in * (j.num_cols - 1) + n)});
    const uint64_t out_len = j.op == BC_OP_LDE ? (n << j.log_blowup) : n;
    bufs.push_back({(uint64_t)j.outputs, (uint64_t)(j.outputs + j.stride_out * (j.num_cols - 1) + out_len)});
    if (j.op == BC_OP_LDE)
      bufs.push_back({(uint64_t)j.coeffs, (uint64_t)(j.coeffs + j.stride_coeffs * (j.num_cols - 1) + n)});
  }
  std::sort(bufs.begin(), bufs.end(), [](const Buf &a, const Buf &b) { return a.lo < b.lo; });
  auto inside = [&](uint64_t p) {
    auto it = std::upper_bound(bufs.begin(), bufs.end(), p, [](uint64_t v, const Buf &b) { return v < b.lo; });
    if (it == bufs.begin()) return false;
    --it;
    return p >= it->lo && p < it->hi;
  };
  for (const ColPtrs &c : hp.cols) REQUIRE(inside((uint64_t)c.in) && inside((uint64_t)c.out));
  // queue order: by level; inside a level by kernel variant
  for (uint32_t l = 0; l < hp.levels; ++l)
    for (uint32_t i = hp.level_seg[l] + 1; i < hp.level_seg[l + 1]; ++i)
      REQUIRE(hp.segs[i - 1].variant <= hp.segs[i].variant);
}

int main() {
  const Tables tab{};  // planning stores the table pointers; it never reads them
  // fake, non-overlapping device addresses: job k's buffers start at (k + 1) << 40
  std::mt19937_64 rng(4);
  std::vector<bc_ragged_job> jobs;
  uint32_t n_lde = 0;
  for (uint32_t k = 0; k < 1000; ++k) {
    bc_ragged_job j{};
    j.log_n = 10 + (uint32_t)(rng() % 14);
    const uint32_t op = (uint32_t)(rng() % 4);
    const uint64_t n = 1ull << j.log_n;
    j.num_cols = 1;
    j.stride_in = n;
    j.inputs = (const uint64_t *)((uint64_t)(k + 1) << 40);
    j.outputs = (uint64_t *)(((uint64_t)(k + 1) << 40) + (1ull << 38));
    if (op < 2) {
      j.op = op == 0 ? BC_OP_NTT : BC_OP_INTT;
      j.stride_out = n;
    } else {
      j.op = BC_OP_LDE;
      j.log_blowup = op == 2 ? 1 : 2;
      j.coeffs = (uint64_t *)(((uint64_t)(k + 1) << 40) + (1ull << 39));
      j.stride_coeffs = n;
      j.stride_out = n << j.log_blowup;
      ++n_lde;
    }
    jobs.push_back(j);
  }
  RaggedHostPlan hp;
  REQUIRE(ragged_build(jobs.data(), (uint32_t)jobs.size(), tab, hp) == BC_SUCCESS);
  check_plan(jobs, hp);
  REQUIRE(n_lde > 0);

  // multi-column jobs with padded strides
  std::vector<bc_ragged_job> wide;
  for (uint32_t k = 0; k < 40; ++k) {
    bc_ragged_job j = jobs[k];
    j.num_cols = 1 + k % 7;
    const uint64_t n = 1ull << j.log_n;
    j.stride_in = n + 8;
    j.stride_out = (j.op == BC_OP_LDE ? (n << j.log_blowup) : n) + 16;
    if (j.op == BC_OP_LDE) j.stride_coeffs = n + 24;
    wide.push_back(j);
  }
  RaggedHostPlan hw;
  REQUIRE(ragged_build(wide.data(), (uint32_t)wide.size(), tab, hw) == BC_SUCCESS);
  check_plan(wide, hw);

  // one job of each kind alone, at the smallest and largest log_n
  for (uint32_t ln : {8u, 10u, 12u, 23u, 24u})
    for (uint32_t op = 0; op < 3; ++op) {
      bc_ragged_job j = jobs[0];
      j.log_n = ln;
      const uint64_t n = 1ull << ln;
      j.stride_in = n;
      j.op = op;
      j.log_blowup = op == BC_OP_LDE ? 2 : 0;
      j.coeffs = (uint64_t *)(3ull << 40);
      j.stride_coeffs = n;
      j.stride_out = n << j.log_blowup;
      RaggedHostPlan h1;
      REQUIRE(ragged_build(&j, 1, tab, h1) == BC_SUCCESS);
      check_plan({j}, h1);
    }

  // invalid jobs are refused before anything is planned
  bc_ragged_job bad = jobs[0];
  bad.log_n = 7;
  REQUIRE(ragged_build(&bad, 1, tab, hw) == BC_ERROR_INVALID_VALUE);
  bad = jobs[0];
  bad.stride_in = (1ull << bad.log_n) - 1;
  REQUIRE(ragged_build(&bad, 1, tab, hw) == BC_ERROR_INVALID_VALUE);

  // packed device image: pack into a host buffer (under the sanitizers) and
  // check every segment's column table lands inside it
  const size_t table_bytes = ragged_table_bytes(hp);
  std::vector<uint64_t> packed((table_bytes + 7) / 8 + 1, 0);
  const uint64_t fake_base = 0x100000000ull;
  ragged_pack_table(hp, packed.data(), (const void *)(uintptr_t)fake_base);
  const RaggedSeg *ps = (const RaggedSeg *)packed.data();
  for (size_t i = 0; i < hp.segs.size(); ++i) {
    const uint64_t col_addr = (uint64_t)(uintptr_t)ps[i].a.cols;
    REQUIRE(col_addr >= fake_base + ragged_chunk_offset(hp));
    REQUIRE(col_addr + (uint64_t)ps[i].a.num_cols * sizeof(ColPtrs) <= fake_base + table_bytes);
  }
  printf("OK %zu %u %llu %zu\n", hp.segs.size(), hp.levels, (unsigned long long)hp.total_items, table_bytes);
  return 0;
}

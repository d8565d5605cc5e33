// Ragged batches: one persistent kernel runs many jobs of different sizes and
// operations (NTT, iNTT, coset LDE) from a device work queue.
//
// Planning (host).  Every job becomes a chain of passes (describe_fast_passes):
// an LDE is an inverse transform followed by 2^k coset transforms that all
// depend on it.  Each (job, stage, pass) is a segment of independent work
// units (groups of contiguous tiles, or (r-range, pair group) units of a
// strided pass).  A segment's dependency level is its pass index along its
// job's chain, and every segment depends only on segments of lower levels.
// Segments are ordered by level, then by kernel variant, then by job, so the
// units of one level are one contiguous range of the queue and concurrently
// running workgroups mostly share code.
//
// Execution (device).  The persistent kernel is launched once per level (at
// most 16 launches for any batch).  Its workgroups pop unit indices from the
// level's atomic counter, find the unit's segment by binary search and run it
// with the same pass bodies as the standalone kernels.  Units of one level
// never wait for each other, so there are no spins and no device-scope
// fences inside a launch: on gfx950 an agent-scope fence writes back and
// invalidates the L2 of the XCD, which per unit costs more than the unit.
#include "ntt_ragged.hpp"
#include "ntt_passes.hpp"

#include <stddef.h>
#include <string.h>

#include <algorithm>
#include <atomic>
#include <map>
#include <mutex>
#include <vector>

namespace bc {

namespace {
using namespace pass;

// LDS of the persistent kernel (file scope so the out-of-line unit functions
// below address it directly).
__shared__ PassLds ragged_lds;

// One out-of-line function per variant: each gets its own register
// allocation instead of the maximum over every variant inlined into the
// persistent loop.
template <int LH, bool MC, uint32_t DM>
__device__ __noinline__ void contig_unit(const RaggedSeg *sg, uint32_t b, uint32_t e) {
  constexpr bool INV = DM >= 2;
  constexpr int MODE = DM == 1 ? MODE_LDE_IN : (DM == 3 ? MODE_NINV : MODE_NONE);
  contig_run<LH, INV, INV, MODE, MC, true>(sg->a, ragged_lds, b, e, 1);
}

template <int LH, uint32_t DM>
__device__ __noinline__ void strided_unit(const RaggedSeg *sg, uint32_t rho, uint64_t b, uint64_t e) {
  constexpr bool INV = DM >= 2;
  constexpr int MODE = DM == 1 ? MODE_LDE_IN : (DM == 3 ? MODE_NINV : MODE_NONE);
  strided_run<LH, INV, INV, MODE, true>(sg->a, ragged_lds, rho, b, e);
}

template <int LH, bool MC>
__device__ __forceinline__ void run_contig_dm(const RaggedSeg *sg, uint32_t dm, uint32_t b, uint32_t e) {
  switch (dm) {
  case 0: contig_unit<LH, MC, 0>(sg, b, e); break;
  case 1: contig_unit<LH, MC, 1>(sg, b, e); break;
  case 2: contig_unit<LH, MC, 2>(sg, b, e); break;
  default: contig_unit<LH, MC, 3>(sg, b, e); break;
  }
}

template <int LH>
__device__ __forceinline__ void run_strided_dm(const RaggedSeg *sg, uint32_t dm, uint32_t rho, uint64_t b,
                                               uint64_t e) {
  switch (dm) {
  case 0: strided_unit<LH, 0>(sg, rho, b, e); break;
  case 1: strided_unit<LH, 1>(sg, rho, b, e); break;
  case 2: strided_unit<LH, 2>(sg, rho, b, e); break;
  default: strided_unit<LH, 3>(sg, rho, b, e); break;
  }
}

__device__ __forceinline__ void run_unit(const RaggedSeg *sg, uint32_t variant, uint64_t total, uint32_t per_unit,
                                         uint32_t positions, uint64_t unit) {
  const uint32_t v = variant;
  const uint32_t dm = v & 3u, h = (v >> 4) & 15u;
  const bool mc = (v >> 2) & 1u;
  if ((v >> 8) == 0) {
    const uint64_t b64 = unit * per_unit;
    const uint32_t b = (uint32_t)b64;
    const uint32_t e = (uint32_t)(b64 + per_unit < total ? b64 + per_unit : total);
    if (mc) {
      switch (h) {
      case 8: run_contig_dm<8, true>(sg, dm, b, e); break;
      case 9: run_contig_dm<9, true>(sg, dm, b, e); break;
      case 10: run_contig_dm<10, true>(sg, dm, b, e); break;
      default: run_contig_dm<11, true>(sg, dm, b, e); break;
      }
    } else {
      switch (h) {
      case 9: run_contig_dm<9, false>(sg, dm, b, e); break;
      case 10: run_contig_dm<10, false>(sg, dm, b, e); break;
      case 11: run_contig_dm<11, false>(sg, dm, b, e); break;
      default: run_contig_dm<12, false>(sg, dm, b, e); break;
      }
    }
  } else {
    const uint32_t rho = (uint32_t)(unit % positions);
    const uint64_t b = (unit / positions) * per_unit;
    const uint64_t e = b + per_unit < total ? b + per_unit : total;
    switch (h) {
    case 4: run_strided_dm<4>(sg, dm, rho, b, e); break;
    case 5: run_strided_dm<5>(sg, dm, rho, b, e); break;
    case 6: run_strided_dm<6>(sg, dm, rho, b, e); break;
    case 7: run_strided_dm<7>(sg, dm, rho, b, e); break;
    default: run_strided_dm<8>(sg, dm, rho, b, e); break;
    }
  }
}

struct QueueState {
  unsigned long long head[RAGGED_MAX_LEVELS];  // units popped so far, per level
  unsigned arrive[RAGGED_MAX_LEVELS];          // single launch: workgroups done with each level
  unsigned timed_out;                          // single launch: a level wait hit its spin limit
};

// Clears the queue state with ordinary stores (a fill kernel, so a captured
// launch sequence holds no memset node).
__global__ void ragged_reset_kernel(QueueState *st) {
  const uint32_t i = threadIdx.x;
  if (i < RAGGED_MAX_LEVELS) {
    st->head[i] = 0;
    st->arrive[i] = 0;
  }
  if (i == 0) st->timed_out = 0;
}

// Level ranges of a plan, passed by value to the single-launch kernel.
struct RaggedLevels {
  uint64_t item[RAGGED_MAX_LEVELS + 1];
  uint32_t seg[RAGGED_MAX_LEVELS + 1];
  uint32_t count;
};

// Polls of one level wait before it gives up: each is a 512-clock sleep and
// an uncached load, about 1-2 us, so 2^22 polls are seconds, against at most
// tens of milliseconds for a level.  A host call can lower it to force the
// timeout path.
constexpr uint32_t RAGGED_DEFAULT_SPIN_LIMIT = 1u << 22;

// Drains the units of one dependency level: segments [seg_begin, seg_end)
// hold queue items [item_begin, item_end), and no unit of a level depends on
// another unit of the same level, so units need no waits and no device-scope
// fences.  Between levels, the launch boundary (one launch per level) or
// grid_wait (the single launch) orders them.
//
// Loop shape: the only single-thread block that touches the queue sits at the
// end of the body and is followed by a barrier in uniform control flow.  With
// a pop at the top of the body as well, the compiler merges two
// `threadIdx.x == 0` blocks across the back edge, which leaves lane 0 alone on
// the path to the pop and the rest of the workgroup in an inner loop that
// re-reads the old unit index forever.
__device__ __forceinline__ void drain_level(const RaggedSeg *__restrict__ segs,
                                            const uint32_t *__restrict__ chunk_seg, uint32_t seg_begin,
                                            uint32_t seg_end, uint64_t item_begin, uint64_t item_end,
                                            unsigned long long *head, unsigned long long &s_item) {
  if (threadIdx.x == 0) s_item = item_begin + atomicAdd(head, 1ull);
  __syncthreads();
  for (;;) {
    const uint64_t item = s_item;
    __syncthreads();  // every thread has read s_item before it is overwritten
    if (item >= item_end) break;
    // segment: the last one whose first_item <= item, between the segments
    // holding the first items of this chunk and of the next one
    const uint64_t c = item >> RAGGED_CHUNK_LOG;
    uint32_t lo = max(seg_begin, chunk_seg[c]);
    uint32_t hi = min(seg_end, chunk_seg[c + 1] + 1);
    while (hi - lo > 1) {
      const uint32_t mid = (lo + hi) >> 1;
      if (segs[mid].first_item <= item) lo = mid;
      else hi = mid;
    }
    lo = __builtin_amdgcn_readfirstlane(lo);
    const RaggedSeg &sg = segs[lo];
    run_unit(&sg, sg.variant, sg.total, sg.per_unit, sg.positions, item - sg.first_item);
    __syncthreads();  // the unit's LDS use is over before the next unit starts
    if (threadIdx.x == 0) s_item = item_begin + atomicAdd(head, 1ull);
    __syncthreads();
  }
}

__global__ __launch_bounds__(NT, 2) void ntt_ragged_kernel(const RaggedSeg *__restrict__ segs,
                                                        const uint32_t *__restrict__ chunk_seg, uint32_t seg_begin,
                                                        uint32_t seg_end, uint64_t item_begin, uint64_t item_end,
                                                        unsigned long long *head) {
  __shared__ unsigned long long s_item;
  drain_level(segs, chunk_seg, seg_begin, seg_end, item_begin, item_end, head, s_item);
}

// Waits until every workgroup of the grid has arrived at this level's
// counter.  Every wave first waits for its own memory operations, so the
// workgroup's stores are in its XCD's L2.  Then one wave releases them
// device-wide (an agent-scope release fence writes that L2 back), arrives,
// polls the counter with relaxed agent-scope loads (coherent, no cache
// maintenance), and after the last arrival takes one acquire fence, which
// invalidates the CU's L1 and the XCD's L2 for the whole workgroup before any
// wave reads what other workgroups wrote.  Measured on 512 workgroups: fences
// in every wave cost 0.4 ms per level, and acquire polls (an L2 invalidation
// each, under the workgroups still working) 0.24 ms.  The spin is bounded:
// past spin_limit polls the wait records a timeout and returns, so every wave
// reaches the end of the kernel.
__device__ __forceinline__ void grid_wait(unsigned *arrive, unsigned *timed_out, uint32_t spin_limit) {
  __builtin_amdgcn_s_waitcnt(0);
  __syncthreads();
  if (threadIdx.x < 64) {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
    if (threadIdx.x == 0) {
      __hip_atomic_fetch_add(arrive, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      uint32_t spins = 0;
      while (__hip_atomic_load(arrive, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < gridDim.x) {
        if (++spins > spin_limit) {
          __hip_atomic_store(timed_out, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
          break;
        }
        __builtin_amdgcn_s_sleep(8);
      }
    }
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
  }
  __syncthreads();
}

// The whole plan in one launch (cooperative, so that every workgroup is
// resident): the workgroups drain level 0's units, wait for each other, drain
// level 1's, and so on.  A level's units depend only on earlier levels, so
// the wait between levels is the only synchronization.
__global__ __launch_bounds__(NT, 2) void ntt_ragged_all_kernel(const RaggedSeg *__restrict__ segs,
                                                            const uint32_t *__restrict__ chunk_seg, RaggedLevels lv,
                                                            QueueState *st, uint32_t spin_limit) {
  __shared__ unsigned long long s_item;
  for (uint32_t l = 0; l < lv.count; ++l) {
    drain_level(segs, chunk_seg, lv.seg[l], lv.seg[l + 1], lv.item[l], lv.item[l + 1], &st->head[l], s_item);
    if (l + 1 < lv.count) grid_wait(&st->arrive[l], &st->timed_out, spin_limit);
  }
}

}  // namespace

Tables current_tables(bool &ok);  // ntt_api.hip

size_t ragged_state_bytes() { return sizeof(QueueState); }

uint32_t ragged_workgroups(int dev) {
  int cus = 0, per_cu = 0, per_cu_all = 0;
  if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || cus <= 0) cus = 256;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, ntt_ragged_kernel, NT, 0) != hipSuccess || per_cu <= 0)
    per_cu = 1;
  // the single-launch kernel needs the whole grid resident at once
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu_all, ntt_ragged_all_kernel, NT, 0) == hipSuccess &&
      per_cu_all > 0)
    per_cu = std::min(per_cu, per_cu_all);
  return (uint32_t)(cus * per_cu);
}

static std::atomic<uint32_t> g_spin_limit{RAGGED_DEFAULT_SPIN_LIMIT};

size_t ragged_timeout_offset() { return offsetof(QueueState, timed_out); }

int ragged_read_wait_status(const void *d_state) {
  unsigned flag = 0;
  const hipError_t err =
      hipMemcpy(&flag, (const uint8_t *)d_state + ragged_timeout_offset(), sizeof(flag), hipMemcpyDeviceToHost);
  if (err != hipSuccess) return err;
  return flag ? BC_ERROR_WAIT_TIMEOUT : BC_SUCCESS;
}

hipError_t ragged_launch(const void *d_table, const RaggedHostPlan &hp, void *d_state, uint32_t workgroups,
                         hipStream_t s, bool single_launch) {
  QueueState *st = (QueueState *)d_state;
  const RaggedSeg *d_segs = (const RaggedSeg *)d_table;
  const uint32_t *d_chunks = (const uint32_t *)((const uint8_t *)d_table + ragged_chunk_offset(hp));
  ragged_reset_kernel<<<dim3(1), dim3(64), 0, s>>>(st);
  hipError_t err = hipGetLastError();
  if (err != hipSuccess || hp.levels == 0) return err;
  // a stream under graph capture gets plain launches (one per level)
  if (single_launch) {
    hipStreamCaptureStatus cap = hipStreamCaptureStatusNone;
    if (hipStreamIsCapturing(s, &cap) != hipSuccess || cap != hipStreamCaptureStatusNone) {
      (void)hipGetLastError();
      single_launch = false;
    }
  }
  if (single_launch && hp.levels <= RAGGED_MAX_LEVELS) {
    RaggedLevels lv{};
    for (uint32_t l = 0; l <= hp.levels; ++l) {
      lv.item[l] = hp.level_item[l];
      lv.seg[l] = hp.level_seg[l];
    }
    lv.count = hp.levels;
    uint32_t spin_limit = g_spin_limit.load();
    uint32_t wgs = std::max(1u, workgroups);
    void *args[] = {(void *)&d_segs, (void *)&d_chunks, (void *)&lv, (void *)&st, (void *)&spin_limit};
    err = hipLaunchCooperativeKernel((const void *)ntt_ragged_all_kernel, dim3(wgs), dim3(NT), args, 0, s);
    if (err == hipSuccess) return hipSuccess;
    // the runtime cannot make the grid resident at once: one launch per level
    (void)hipGetLastError();
    if (err != hipErrorCooperativeLaunchTooLarge && err != hipErrorNotSupported) return err;
    err = hipSuccess;
  }
  for (uint32_t l = 0; l < hp.levels && err == hipSuccess; ++l) {
    const uint64_t units = hp.level_item[l + 1] - hp.level_item[l];
    if (units == 0) continue;
    const uint32_t wgs = (uint32_t)std::min<uint64_t>(std::max(1u, workgroups), units);
    ntt_ragged_kernel<<<dim3(wgs), dim3(NT), 0, s>>>(d_segs, d_chunks, hp.level_seg[l], hp.level_seg[l + 1],
                                                     hp.level_item[l], hp.level_item[l + 1], &st->head[l]);
    err = hipGetLastError();
  }
  return err;
}

}  // namespace bc

struct bc_ragged_plan {
  int device;
  uint32_t workgroups;
  bc::RaggedHostPlan hp;  // level ranges (the packed table itself is on the device)
  void *d_table;
  void *d_state;
};

extern "C" int bc_ragged_plan_create(const bc_ragged_job *jobs, uint32_t num_jobs, bc_ragged_plan **out) {
  using namespace bc;
  if (!out || (!jobs && num_jobs)) return BC_ERROR_INVALID_VALUE;
  *out = nullptr;
  for (uint32_t j = 0; j < num_jobs; ++j) {
    const int rc = ragged_validate_job(jobs[j]);
    if (rc != BC_SUCCESS) return rc;
  }
  bool ok = false;
  const Tables tab = current_tables(ok);
  if (!ok) return BC_ERROR_NOT_INITIALIZED;
  RaggedHostPlan hp;
  int rc = ragged_build(jobs, num_jobs, tab, hp);
  if (rc != BC_SUCCESS) return rc;

  int dev = 0;
  hipError_t err = hipGetDevice(&dev);
  if (err != hipSuccess) return err;
  bc_ragged_plan *p = new bc_ragged_plan{};
  p->device = dev;
  p->workgroups = ragged_workgroups(dev);
  const size_t table_bytes = ragged_table_bytes(hp);
  std::vector<uint64_t> packed((table_bytes + 7) / 8);
  err = hipMalloc(&p->d_table, table_bytes);
  if (err == hipSuccess) err = hipMalloc(&p->d_state, ragged_state_bytes());
  if (err == hipSuccess) {
    ragged_pack_table(hp, packed.data(), p->d_table);
    err = hipMemcpy(p->d_table, packed.data(), table_bytes, hipMemcpyHostToDevice);
  }
  p->hp = std::move(hp);
  if (err != hipSuccess) {
    (void)hipFree(p->d_table);
    (void)hipFree(p->d_state);
    delete p;
    return err;
  }
  *out = p;
  return BC_SUCCESS;
}

extern "C" int bc_ragged_run(bc_ragged_plan *p, void *stream) {
  if (!p) return BC_ERROR_INVALID_VALUE;
  return bc::ragged_launch(p->d_table, p->hp, p->d_state, p->workgroups, (hipStream_t)stream, true);
}

extern "C" int bc_ragged_wait_status(const bc_ragged_plan *p) {
  if (!p) return BC_ERROR_INVALID_VALUE;
  return bc::ragged_read_wait_status(p->d_state);
}

extern "C" int bc_ragged_set_wait_limit(uint32_t spins) {
  bc::g_spin_limit.store(spins ? spins : bc::RAGGED_DEFAULT_SPIN_LIMIT);
  return BC_SUCCESS;
}

extern "C" int bc_ragged_plan_info(const bc_ragged_plan *p, uint64_t *units, uint32_t *levels, uint32_t *workgroups) {
  if (!p) return BC_ERROR_INVALID_VALUE;
  if (units) *units = p->hp.total_items;
  if (levels) *levels = p->hp.levels;
  if (workgroups) *workgroups = p->workgroups;
  return BC_SUCCESS;
}

extern "C" int bc_ragged_plan_destroy(bc_ragged_plan *p) {
  if (!p) return BC_ERROR_INVALID_VALUE;
  hipError_t e1 = hipFree(p->d_table);
  hipError_t e2 = hipFree(p->d_state);
  delete p;
  return e1 != hipSuccess ? e1 : e2;
}

extern "C" int bc_ragged_run_plain(const bc_ragged_job *jobs, uint32_t num_jobs, void *stream) {
  if (!jobs && num_jobs) return BC_ERROR_INVALID_VALUE;
  for (uint32_t j = 0; j < num_jobs; ++j) {
    const bc_ragged_job &jb = jobs[j];
    int rc = bc::ragged_validate_job(jb);
    if (rc != BC_SUCCESS) return rc;
    if (jb.op == BC_OP_NTT)
      rc = bc_batch_ntt_internal(jb.inputs, jb.outputs, jb.log_n, jb.num_cols, jb.stride_in, jb.stride_out, false,
                                 false, 0, 0, stream);
    else if (jb.op == BC_OP_INTT)
      rc = bc_batch_ntt_internal(jb.inputs, jb.outputs, jb.log_n, jb.num_cols, jb.stride_in, jb.stride_out, true,
                                 true, 0, 0, stream);
    else
      rc = bc_batch_lde(jb.inputs, jb.coeffs, jb.outputs, jb.log_n, jb.num_cols, jb.stride_in, jb.stride_coeffs,
                        jb.stride_out, jb.log_blowup, true, stream);
    if (rc != BC_SUCCESS) return rc;
  }
  return BC_SUCCESS;
}

// ---- Allocation-free variant (bc_ragged_run_no_alloc) ----
//
// The packed plan table and the queue state live in one file-scope device
// array instead of hipMalloc'd buffers.  Its host-side name only stands in for
// the symbol: kernels and copies use the per-device address that
// hipGetSymbolAddress returns (each GPU holds its own copy).
namespace bc {
namespace {
constexpr size_t NOALLOC_SCRATCH_BYTES = (size_t)8 << 20;
constexpr int NOALLOC_MAX_DEVICES = 64;
}  // namespace
}  // namespace bc

static __device__ __align__(256) uint8_t scratch[bc::NOALLOC_SCRATCH_BYTES];

namespace {
std::mutex g_noalloc_mu;
uint8_t *g_noalloc_ptr[bc::NOALLOC_MAX_DEVICES] = {};
uint32_t g_noalloc_wgs[bc::NOALLOC_MAX_DEVICES] = {};

// Device address of `scratch` on the current device `dev` (looked up once per
// device, with that device current) and the persistent grid size there.
int noalloc_scratch(int dev, uint8_t **ptr, uint32_t *wgs) {
  if (dev < 0 || dev >= bc::NOALLOC_MAX_DEVICES) return BC_ERROR_INVALID_VALUE;
  std::lock_guard<std::mutex> lk(g_noalloc_mu);
  if (!g_noalloc_ptr[dev]) {
    void *p = nullptr;
    const hipError_t err = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (err != hipSuccess) return err;
    if (!p) return BC_ERROR_NOT_INITIALIZED;
    g_noalloc_wgs[dev] = bc::ragged_workgroups(dev);
    g_noalloc_ptr[dev] = (uint8_t *)p;
  }
  *ptr = g_noalloc_ptr[dev];
  *wgs = g_noalloc_wgs[dev];
  return BC_SUCCESS;
}
}  // namespace

extern "C" int bc_ragged_run_no_alloc(const bc_ragged_job *jobs, uint32_t num_jobs, void *stream) {
  using namespace bc;
  if (!jobs && num_jobs) return BC_ERROR_INVALID_VALUE;
  for (uint32_t j = 0; j < num_jobs; ++j) {
    const int rc = ragged_validate_job(jobs[j]);
    if (rc != BC_SUCCESS) return rc;
  }
  bool ok = false;
  const Tables tab = current_tables(ok);
  if (!ok) return BC_ERROR_NOT_INITIALIZED;
  RaggedHostPlan hp;
  int rc = ragged_build(jobs, num_jobs, tab, hp);
  if (rc != BC_SUCCESS) return rc;
  // the queue state at the start, then the table on the next 256-byte boundary
  const size_t table_bytes = ragged_table_bytes(hp);
  const size_t table_off = (ragged_state_bytes() + 255) & ~(size_t)255;
  if (table_off + table_bytes > NOALLOC_SCRATCH_BYTES) return BC_ERROR_OUT_OF_MEMORY;
  int dev = 0;
  hipError_t err = hipGetDevice(&dev);
  if (err != hipSuccess) return err;
  uint8_t *base = nullptr;
  uint32_t wgs = 0;
  rc = noalloc_scratch(dev, &base, &wgs);
  if (rc != BC_SUCCESS) return rc;
  // The staging copy is pageable host memory, so HIP performs the upload
  // synchronously: the vector may be released when the call returns.
  std::vector<uint64_t> packed((table_bytes + 7) / 8);
  uint8_t *table = base + table_off;
  ragged_pack_table(hp, packed.data(), table);
  const hipStream_t s = (hipStream_t)stream;
  if (table_bytes) err = hipMemcpyAsync(table, packed.data(), table_bytes, hipMemcpyHostToDevice, s);
  if (err != hipSuccess) return err;
  // plain kernel launches, one per level (no cooperative launch)
  return ragged_launch(table, hp, base, wgs, s, false);
}

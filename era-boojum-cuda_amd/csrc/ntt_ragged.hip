// Ragged batches: one persistent kernel runs many jobs of different sizes and
// operations (NTT, iNTT, coset LDE) from a device work queue.
//
// Planning (host).  Every job becomes a chain of passes (describe_fast_passes):
// an LDE is an inverse transform followed by 2^k coset transforms that all
// depend on it.  Each (job, stage, pass) is a segment of independent work
// units (groups of contiguous tiles, or (r-range, pair group) units of a
// strided pass).  Segments are ordered by dependency level, then by kernel
// variant, then by job, so a unit's predecessors always have lower queue
// positions and concurrently running workgroups mostly share code.
//
// Execution (device).  Workgroups pop unit indices from an atomic counter,
// find their segment by binary search, wait (bounded) until the segment it
// depends on has completed all its units, run the unit with the same pass
// bodies as the standalone kernels, then publish their stores with a
// device-scope release and bump the segment's completion counter.  Waiting
// only ever targets units already taken by running workgroups, so the queue
// always drains; the bound turns a planning error into a flag, not a hang.
#include "ntt_ragged.hpp"
#include "ntt_passes.hpp"

#include <algorithm>
#include <vector>

namespace bc {

namespace {
using namespace pass;

// variant = kind << 8 | h << 4 | mc << 2 | dm, dm: 0 fwd DIF, 1 fwd DIF + coset
// scaling of the inputs, 2 inv DIT, 3 inv DIT + n^-1
constexpr uint32_t variant_id(bool strided, uint32_t h, bool mc, uint32_t dm) {
  return ((strided ? 1u : 0u) << 8) | (h << 4) | ((mc ? 1u : 0u) << 2) | dm;
}

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
  contig_run<LH, INV, INV, MODE, MC>(sg->a, ragged_lds, b, e, 1);
}

template <int LH, uint32_t DM>
__device__ __noinline__ void strided_unit(const RaggedSeg *sg, uint32_t rho, uint64_t b, uint64_t e) {
  constexpr bool INV = DM >= 2;
  constexpr int MODE = DM == 1 ? MODE_LDE_IN : (DM == 3 ? MODE_NINV : MODE_NONE);
  strided_run<LH, INV, INV, MODE>(sg->a, ragged_lds, rho, b, e);
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
  unsigned long long head;
  uint32_t error;
  uint32_t pad;
  uint32_t done[1];  // one completion counter per segment (allocated to size)
};

__global__ void ragged_reset_kernel(QueueState *st, uint32_t nsegs) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i == 0) {
    st->head = 0;
    st->error = 0;
  }
  for (uint32_t k = i; k < nsegs; k += gridDim.x * blockDim.x) st->done[k] = 0;
}

// Loop shape: the only single-thread block that touches the queue sits at the
// end of the body (publish completion + pop the next unit) and is followed by
// a barrier in uniform control flow.  With a pop at the top of the body
// instead, the compiler merges the two `threadIdx.x == 0` blocks across the
// back edge, which leaves lane 0 alone on the path to the pop and the rest of
// the workgroup in an inner loop that re-reads the old unit index forever.
__global__ __launch_bounds__(NT, 2) void ntt_ragged_kernel(const RaggedSeg *__restrict__ segs, uint32_t nsegs,
                                                        uint64_t total_items, QueueState *st, uint32_t spin_limit) {
  __shared__ unsigned long long s_item;
  if (threadIdx.x == 0) s_item = atomicAdd(&st->head, 1ull);
  __syncthreads();
  for (;;) {
    const uint64_t item = s_item;
    __syncthreads();  // every thread has read s_item before it is overwritten
    if (item >= total_items) break;
    // segment: the last one whose first_item <= item
    uint32_t lo = 0, hi = nsegs;
    while (hi - lo > 1) {
      const uint32_t mid = (lo + hi) >> 1;
      if (segs[mid].first_item <= item) lo = mid;
      else hi = mid;
    }
    lo = __builtin_amdgcn_readfirstlane(lo);
    const RaggedSeg &sg = segs[lo];
    if (sg.dep >= 0) {
      if (threadIdx.x == 0) {
        uint32_t spins = 0;
        while (__hip_atomic_load(&st->done[sg.dep], __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT) < sg.dep_units) {
          __builtin_amdgcn_s_sleep(8);
          // give up at the bound, or at once when another unit already has
          if (++spins > spin_limit ||
              __hip_atomic_load(&st->error, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0) {
            __hip_atomic_store(&st->error, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            break;
          }
        }
      }
      __syncthreads();
      __threadfence();  // acquire: the predecessor's stores are visible to every wave
    }
    run_unit(&sg, sg.variant, sg.total, sg.per_unit, sg.positions, item - sg.first_item);
    __threadfence();  // release this thread's stores device-wide
    __syncthreads();
    if (threadIdx.x == 0) {
      __threadfence();
      __hip_atomic_fetch_add(&st->done[lo], 1u, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
      s_item = atomicAdd(&st->head, 1ull);
    }
    __syncthreads();
  }
}

uint32_t brev_bits_host(uint32_t x, uint32_t bits) {
  uint32_t r = 0;
  for (uint32_t i = 0; i < bits; ++i) r |= ((x >> i) & 1u) << (bits - 1 - i);
  return r;
}

struct PlanSeg {
  RaggedSeg s;
  uint32_t level, job, stage, pass;
  int32_t dep_key;  // index into the unsorted list
};

}  // namespace

Tables current_tables(bool &ok);  // ntt_api.hip

int ragged_validate_job(const bc_ragged_job &j) {
  if (j.log_n < 8 || j.log_n >= 32 || !j.inputs || !j.outputs) return BC_ERROR_INVALID_VALUE;
  const uint64_t n = 1ull << j.log_n;
  if (j.stride_in < n) return BC_ERROR_INVALID_VALUE;
  if (j.op == BC_OP_NTT || j.op == BC_OP_INTT) {
    if (j.stride_out < n) return BC_ERROR_INVALID_VALUE;
  } else if (j.op == BC_OP_LDE) {
    if (!j.coeffs || j.stride_coeffs < n || j.log_blowup < 1 || j.log_blowup > 8) return BC_ERROR_INVALID_VALUE;
    if ((uint64_t)j.log_n + j.log_blowup > OMEGA_LOG_ORDER) return BC_ERROR_INVALID_VALUE;
    if ((uint64_t)j.stride_out < (n << j.log_blowup)) return BC_ERROR_INVALID_VALUE;
  } else {
    return BC_ERROR_INVALID_VALUE;
  }
  return BC_SUCCESS;
}

int ragged_build(const bc_ragged_job *jobs, uint32_t num_jobs, const Tables &tab, RaggedHostPlan &out) {
  out.segs.clear();
  out.total_items = 0;
  out.levels = 0;
  for (uint32_t j = 0; j < num_jobs; ++j) {
    const int rc = ragged_validate_job(jobs[j]);
    if (rc != BC_SUCCESS) return rc;
  }
  std::vector<PlanSeg> segs;
  for (uint32_t j = 0; j < num_jobs; ++j) {
    const bc_ragged_job &jb = jobs[j];
    if (jb.num_cols == 0) continue;
    struct Stage {
      const uint64_t *in;
      uint64_t *out;
      uint64_t si, so;
      bool bitrev, inv;
      uint32_t ext, coset;
    };
    std::vector<Stage> stages;
    const uint64_t n = 1ull << jb.log_n;
    if (jb.op == BC_OP_NTT) stages.push_back({jb.inputs, jb.outputs, jb.stride_in, jb.stride_out, false, false, 0, 0});
    else if (jb.op == BC_OP_INTT) stages.push_back({jb.inputs, jb.outputs, jb.stride_in, jb.stride_out, true, true, 0, 0});
    else {
      stages.push_back({jb.inputs, jb.coeffs, jb.stride_in, jb.stride_coeffs, true, true, 0, 0});
      for (uint32_t c = 0; c < (1u << jb.log_blowup); ++c)
        stages.push_back({jb.coeffs, jb.outputs + (uint64_t)brev_bits_host(c, jb.log_blowup) * n, jb.stride_coeffs,
                          jb.stride_out, false, false, jb.log_blowup, c});
    }
    const TuneParams tp = tuned(jb.log_n, jb.num_cols);
    int32_t stage0_last = -1;
    uint32_t stage0_passes = 0;
    for (uint32_t st = 0; st < stages.size(); ++st) {
      const Stage &sgd = stages[st];
      PassDesc pd[8];
      const int np = describe_fast_passes(tab, sgd.in, sgd.out, jb.log_n, jb.num_cols, sgd.si, sgd.so, sgd.bitrev,
                                          sgd.inv, sgd.ext, sgd.coset, pd);
      if (np <= 0) return BC_ERROR_INVALID_VALUE;
      const uint32_t base_level = st == 0 ? 0 : stage0_passes;
      int32_t prev = st == 0 ? -1 : stage0_last;
      for (int k = 0; k < np; ++k) {
        PlanSeg ps{};
        ps.level = base_level + k;
        ps.job = j;
        ps.stage = st;
        ps.pass = k;
        ps.dep_key = prev;
        ps.s.a = pd[k].a;
        uint32_t dm;
        if (!sgd.inv && !sgd.bitrev) dm = pd[k].mode == MODE_LDE_IN ? 1u : 0u;
        else if (sgd.inv && sgd.bitrev) dm = pd[k].mode == MODE_NINV ? 3u : 2u;
        else return BC_ERROR_INVALID_VALUE;
        if ((dm == 0 || dm == 2) && pd[k].mode != MODE_NONE) return BC_ERROR_INVALID_VALUE;
        if (pd[k].contig) {
          const bool mc = jb.log_n < TILE_LOG;
          ps.s.variant = variant_id(false, pd[k].log_h, mc, dm);
          ps.s.total = pd[k].a.tiles;
          ps.s.per_unit = tp.unit_contig ? tp.unit_contig : 2;
          ps.s.units = (uint32_t)((ps.s.total + ps.s.per_unit - 1) / ps.s.per_unit);
          ps.s.positions = 0;
        } else {
          ps.s.variant = variant_id(true, pd[k].log_h, false, dm);
          ps.s.total = pd[k].pairs;
          ps.s.per_unit = (uint32_t)std::min<uint64_t>(pd[k].pairs, tp.unit_strided ? tp.unit_strided : 4);
          ps.s.positions = (uint32_t)pd[k].positions;
          const uint64_t groups = (ps.s.total + ps.s.per_unit - 1) / ps.s.per_unit;
          ps.s.units = (uint32_t)(groups * ps.s.positions);
        }
        prev = (int32_t)segs.size();
        segs.push_back(ps);
      }
      if (st == 0) {
        stage0_last = prev;
        stage0_passes = (uint32_t)np;
      }
    }
  }

  // order: level, then variant, then job/stage; remap dependencies
  std::vector<uint32_t> order(segs.size());
  for (uint32_t i = 0; i < order.size(); ++i) order[i] = i;
  std::stable_sort(order.begin(), order.end(), [&](uint32_t x, uint32_t y) {
    const PlanSeg &a = segs[x], &b = segs[y];
    if (a.level != b.level) return a.level < b.level;
    if (a.s.variant != b.s.variant) return a.s.variant < b.s.variant;
    if (a.job != b.job) return a.job < b.job;
    return a.stage < b.stage;
  });
  std::vector<uint32_t> pos_of(segs.size());
  for (uint32_t i = 0; i < order.size(); ++i) pos_of[order[i]] = i;
  out.segs.resize(segs.size());
  uint64_t item = 0;
  for (uint32_t i = 0; i < order.size(); ++i) {
    const PlanSeg &ps = segs[order[i]];
    RaggedSeg r = ps.s;
    r.first_item = item;
    item += r.units;
    if (ps.dep_key >= 0) {
      r.dep = (int32_t)pos_of[ps.dep_key];
      r.dep_units = segs[ps.dep_key].s.units;
      if ((uint32_t)r.dep >= i) return BC_ERROR_INVALID_VALUE;  // ordering invariant
    } else {
      r.dep = -1;
      r.dep_units = 0;
    }
    out.levels = std::max(out.levels, ps.level + 1);
    out.segs[i] = r;
  }
  out.total_items = item;
  return BC_SUCCESS;
}

size_t ragged_state_bytes(uint32_t nsegs) {
  return sizeof(QueueState) + std::max<size_t>(1, nsegs) * sizeof(uint32_t);
}

uint32_t ragged_workgroups(int dev) {
  int cus = 0, per_cu = 0;
  if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || cus <= 0) cus = 256;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, ntt_ragged_kernel, NT, 0) != hipSuccess || per_cu <= 0)
    per_cu = 1;
  return (uint32_t)(cus * per_cu);
}

hipError_t ragged_launch(const RaggedSeg *d_segs, uint32_t nsegs, uint64_t total_items, void *d_state,
                         uint32_t workgroups, hipStream_t s) {
  QueueState *st = (QueueState *)d_state;
  ragged_reset_kernel<<<dim3((nsegs + 255) / 256 + 1), dim3(256), 0, s>>>(st, nsegs);
  hipError_t err = hipGetLastError();
  if (err != hipSuccess) return err;
  if (total_items == 0) return hipSuccess;
  const uint32_t wgs = (uint32_t)std::min<uint64_t>(std::max(1u, workgroups), total_items);
  // ~0.2 us per sleep: the bound is seconds, far beyond any real wait
  ntt_ragged_kernel<<<dim3(wgs), dim3(NT), 0, s>>>(d_segs, nsegs, total_items, st, 1u << 24);
  return hipGetLastError();
}

uint32_t *ragged_error_ptr(void *d_state) { return &((QueueState *)d_state)->error; }

hipError_t ragged_read_error(const void *d_state, uint32_t *error) {
  return hipMemcpy(error, &((const QueueState *)d_state)->error, sizeof(uint32_t), hipMemcpyDeviceToHost);
}

}  // namespace bc

struct bc_ragged_plan {
  int device;
  uint32_t nsegs;
  uint64_t total_items;
  uint32_t levels;
  uint32_t workgroups;
  bc::RaggedSeg *d_segs;
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
  p->nsegs = (uint32_t)hp.segs.size();
  p->total_items = hp.total_items;
  p->levels = hp.levels;
  p->workgroups = ragged_workgroups(dev);
  const size_t seg_bytes = std::max<size_t>(1, hp.segs.size()) * sizeof(RaggedSeg);
  err = hipMalloc((void **)&p->d_segs, seg_bytes);
  if (err == hipSuccess) err = hipMalloc(&p->d_state, ragged_state_bytes(p->nsegs));
  if (err == hipSuccess && !hp.segs.empty())
    err = hipMemcpy(p->d_segs, hp.segs.data(), hp.segs.size() * sizeof(RaggedSeg), hipMemcpyHostToDevice);
  if (err != hipSuccess) {
    (void)hipFree(p->d_segs);
    (void)hipFree(p->d_state);
    delete p;
    return err;
  }
  *out = p;
  return BC_SUCCESS;
}

extern "C" int bc_ragged_run(bc_ragged_plan *p, void *stream) {
  if (!p) return BC_ERROR_INVALID_VALUE;
  return bc::ragged_launch(p->d_segs, p->nsegs, p->total_items, p->d_state, p->workgroups, (hipStream_t)stream);
}

extern "C" int bc_ragged_status(bc_ragged_plan *p, int *failed) {
  if (!p || !failed) return BC_ERROR_INVALID_VALUE;
  uint32_t e = 0;
  hipError_t err = bc::ragged_read_error(p->d_state, &e);
  if (err != hipSuccess) return err;
  *failed = e ? 1 : 0;
  return BC_SUCCESS;
}

extern "C" int bc_ragged_plan_info(const bc_ragged_plan *p, uint64_t *units, uint32_t *levels, uint32_t *workgroups) {
  if (!p) return BC_ERROR_INVALID_VALUE;
  if (units) *units = p->total_items;
  if (levels) *levels = p->levels;
  if (workgroups) *workgroups = p->workgroups;
  return BC_SUCCESS;
}

extern "C" int bc_ragged_plan_destroy(bc_ragged_plan *p) {
  if (!p) return BC_ERROR_INVALID_VALUE;
  hipError_t e1 = hipFree(p->d_segs);
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

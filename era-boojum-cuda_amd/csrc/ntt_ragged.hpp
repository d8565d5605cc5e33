// Ragged-batch planning and launch, shared by the bc_ragged_* entry points
// and the serving loop (ntt_serve.hip).
#pragma once
#include "../../include/boojum_ntt.h"
#include "ntt_internal.hpp"

#include <vector>

namespace bc {

// One (job, stage, pass) of a ragged batch: `units` independent work units.
struct RaggedSeg {
  FastArgs a;
  uint64_t first_item;  // queue position of the first unit
  uint64_t total;       // contiguous: tiles; strided: (column, block) pairs
  uint32_t units;
  uint32_t variant;
  int32_t dep;          // segment that must be complete first, -1 if none
  uint32_t dep_units;
  uint32_t per_unit;    // contiguous: tiles per unit; strided: pairs per unit
  uint32_t positions;   // strided: r-ranges
};

struct RaggedHostPlan {
  std::vector<RaggedSeg> segs;  // in queue order
  uint64_t total_items = 0;
  uint32_t levels = 0;
};

int ragged_validate_job(const bc_ragged_job &j);
// Host planning only (no HIP calls).
int ragged_build(const bc_ragged_job *jobs, uint32_t num_jobs, const Tables &tab, RaggedHostPlan &out);
// Device bytes of the queue state (head, error flag, one counter per segment).
size_t ragged_state_bytes(uint32_t nsegs);
// Persistent workgroups that fit on device `dev` at once.
uint32_t ragged_workgroups(int dev);
// Resets the queue state and launches the persistent kernel on `s`; the
// segments must already be in device memory.  Allocates nothing.
hipError_t ragged_launch(const RaggedSeg *d_segs, uint32_t nsegs, uint64_t total_items, void *d_state,
                         uint32_t workgroups, hipStream_t s);
// Device address of the error flag inside the queue state.
uint32_t *ragged_error_ptr(void *d_state);
// Synchronous read of the error flag of a completed run.
hipError_t ragged_read_error(const void *d_state, uint32_t *error);

}  // namespace bc

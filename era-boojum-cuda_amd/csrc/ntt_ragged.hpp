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
  uint32_t per_unit;    // contiguous: tiles per unit; strided: pairs per unit
  uint32_t positions;   // strided: r-ranges
  uint64_t col_off;     // first entry of this segment's column table (a.cols is set when packed)
};

// variant = kind << 8 | h << 4 | mc << 2 | dm, dm: 0 fwd DIF, 1 fwd DIF + coset
// scaling of the inputs, 2 inv DIT, 3 inv DIT + n^-1
__host__ __device__ constexpr uint32_t variant_id(bool strided, uint32_t h, bool mc, uint32_t dm) {
  return ((strided ? 1u : 0u) << 8) | (h << 4) | ((mc ? 1u : 0u) << 2) | dm;
}

// Dependency levels of a plan: pass k of a job's first stage is level k; the
// coset stages of an LDE follow the inverse stage's last level.
constexpr uint32_t RAGGED_MAX_LEVELS = 16;
// Queue items per entry of the chunk index (segment lookup without a full
// binary search).
constexpr uint32_t RAGGED_CHUNK_LOG = 6;

struct RaggedHostPlan {
  std::vector<RaggedSeg> segs;        // in queue order: by level, then kernel variant, then job
  std::vector<uint32_t> level_seg;    // levels + 1 entries: first segment of each level, then segs.size()
  std::vector<uint64_t> level_item;   // levels + 1 entries: first queue item of each level, then total_items
  std::vector<uint32_t> chunk_seg;    // entry c: segment holding item c << RAGGED_CHUNK_LOG; last entry: segs.size() - 1
  std::vector<ColPtrs> cols;          // column tables of all segments
  uint64_t total_items = 0;
  uint32_t levels = 0;
};

int ragged_validate_job(const bc_ragged_job &j);
// Host planning only (no HIP calls).
int ragged_build(const bc_ragged_job *jobs, uint32_t num_jobs, const Tables &tab, RaggedHostPlan &out);
// Device bytes of the queue state (one head counter per level).
size_t ragged_state_bytes();
// Persistent workgroups that fit on device `dev` at once.
uint32_t ragged_workgroups(int dev);
// Device image of a plan: the segment table, the chunk index and the column
// tables, in one buffer of ragged_table_bytes(hp) bytes.  Packing writes the
// device addresses of the column tables, so it needs the buffer's device
// address (device_buf) as well as the host staging copy.
size_t ragged_table_bytes(const RaggedHostPlan &hp);
void ragged_pack_table(const RaggedHostPlan &hp, void *host_buf, const void *device_buf);
// Byte offset of the chunk index inside the packed table.
size_t ragged_chunk_offset(const RaggedHostPlan &hp);
// Resets the queue state and launches the persistent kernel on `s`: with
// single_launch, once for the whole plan (a cooperative launch whose
// workgroups wait for each other between levels; if the runtime cannot make
// the grid resident at once, it falls back to) once per dependency level.
// The packed table must already be in device memory at d_table.  Allocates
// nothing.
hipError_t ragged_launch(const void *d_table, const RaggedHostPlan &hp, void *d_state, uint32_t workgroups,
                         hipStream_t s, bool single_launch);
// BC_ERROR_WAIT_TIMEOUT if a single launch's wait between levels gave up
// (results invalid), else BC_SUCCESS; reads the flag with a blocking copy.
int ragged_read_wait_status(const void *d_state);

}  // namespace bc

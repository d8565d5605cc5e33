// Serving loop: NTT / iNTT / coset-LDE requests arrive at any time, are split
// by columns into bounded parts, admitted first-come first-served against an
// HBM arena of fixed size, gathered into ragged batches and launched on one of
// several streams, so requests that arrive while others run are batched and
// overlap with them on the device (continuous batching).  Also the seeded
// input generator the serving tests and benchmarks use.
//
// Threads: callers submit / wait / release under one mutex; a scheduler thread
// owns every launch.  Per iteration it retires finished batches (event
// query), admits queued parts in arrival order while their buffers fit, and
// launches the admitted parts as one persistent ragged launch on a free
// stream.  Each stream slot owns preallocated segment and queue-state memory,
// so nothing is allocated or freed on the device after bc_server_create.
#include "admission.hpp"
#include "arena.hpp"
#include "ntt_ragged.hpp"

#include <string.h>

#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <deque>
#include <memory>
#include <mutex>
#include <thread>
#include <unordered_map>
#include <vector>

namespace bc {

Tables current_tables(bool &ok);  // ntt_api.hip

namespace {

constexpr uint64_t GAMMA = 0x9E3779B97F4A7C15ull;  // splitmix64 increment

__device__ __forceinline__ uint64_t splitmix_mix(uint64_t z) {
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

// Element i of column c is draw i + 1 of splitmix64 started at
// seed * GAMMA + first_col + c: the state after i + 1 steps is start + (i + 1) GAMMA.
__global__ void fill_seeded_kernel(uint64_t *out, uint32_t log_n, uint32_t cols, uint64_t stride, uint64_t seed,
                                   uint32_t first_col) {
  const uint64_t n = 1ull << log_n;
  const uint64_t total = n * cols;
  const uint64_t step = (uint64_t)gridDim.x * blockDim.x;
  for (uint64_t k = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; k < total; k += step) {
    const uint64_t c = k >> log_n, i = k & (n - 1);
    out[c * stride + i] = splitmix_mix(seed * GAMMA + first_col + c + (i + 1) * GAMMA);
  }
}

}  // namespace

hipError_t fill_seeded(uint64_t *out, uint32_t log_n, uint32_t cols, uint64_t stride, uint64_t seed,
                       uint32_t first_col, hipStream_t s) {
  const uint64_t total = (uint64_t)cols << log_n;
  if (total == 0) return hipSuccess;
  const uint64_t blocks = std::min<uint64_t>((total + 255) / 256, 16384);
  fill_seeded_kernel<<<dim3((uint32_t)blocks), dim3(256), 0, s>>>(out, log_n, cols, stride, seed, first_col);
  return hipGetLastError();
}

namespace {

using Clock = std::chrono::steady_clock;
int64_t now_ns() {
  return std::chrono::duration_cast<std::chrono::nanoseconds>(Clock::now().time_since_epoch()).count();
}

constexpr uint64_t ARENA_ALIGN = 64 * 1024;
constexpr uint32_t SLOT_MAX_SEGS = 8192;
constexpr uint64_t SLOT_COL_CAP = 1u << 19;  // column-table entries per batch
// segment table, a chunk index of up to 2^18 entries (2^24 queue items) and
// the column tables
constexpr size_t SLOT_TABLE_BYTES = SLOT_MAX_SEGS * sizeof(RaggedSeg) + (1u << 18) * sizeof(uint32_t) + 16 +
                                    SLOT_COL_CAP * sizeof(ColPtrs);

struct Request;

struct Part {
  Request *req = nullptr;
  uint32_t index = 0, first_col = 0, cols = 0;
  PartMem mem;             // arena placement; tmp: seeded inputs and LDE coefficients
  uint64_t work = 0;       // transformed elements
  uint32_t seg_bound = 0;  // upper bound on its ragged segments
  uint64_t col_bound = 0;  // upper bound on its column-table entries
};

struct Request {
  uint64_t id = 0;
  bc_request r{};
  uint64_t out_col = 0, tmp_col = 0, work_col = 0, stages = 0;  // bytes / elements / stages per column
  std::deque<Part> parts;  // stable addresses: parts split at admission are appended
  uint32_t parts_done = 0;
  int64_t t_submit = 0, t_first_launch = 0, t_done = 0;
  int32_t status = 0;
  uint64_t kept_reserved = 0;
};

struct Slot {
  hipStream_t stream = nullptr;
  hipEvent_t done = nullptr;
  void *d_table = nullptr;      // packed segment table and chunk index
  void *d_state = nullptr;
  void *h_table = nullptr;      // pinned staging of the table
  bool busy = false;
  std::vector<Part *> parts;
};

uint64_t lde_blowup(const bc_request &r) { return r.op == BC_OP_LDE ? (1ull << r.log_blowup) : 1ull; }

}  // namespace
}  // namespace bc

struct bc_server {
  int device = 0;
  bc_server_config cfg{};
  uint8_t *arena_base = nullptr;
  std::unique_ptr<bc::Arena> arena;
  std::vector<bc::Slot> slots;
  uint32_t workgroups = 0;
  std::mutex mu;
  std::condition_variable cv_work, cv_done;
  std::deque<bc::Part *> pending;   // queued, buffers not yet allocated
  std::deque<bc::Part *> admitted;  // buffers allocated, not yet launched
  std::unordered_map<uint64_t, std::unique_ptr<bc::Request>> reqs;
  uint64_t next_id = 1;
  uint64_t open_requests = 0;
  uint64_t kept_reserved = 0;  // arena bytes promised to kept outputs
  bool stop = false;
  bc_server_stats stats{};
  std::thread worker;

  void finish_part(bc::Part *p, int64_t t, int32_t err);
  void finish_parts(std::vector<bc::Part *> &parts, int32_t err);
  bool launch_batch(bc::Slot &sl);
  bc::Part *split(bc::Part *p);
  void loop();
};

void bc_server::finish_part(bc::Part *p, int64_t t, int32_t err) {
  bc::Request *r = p->req;
  bc::free_tmp(*arena, p->mem);
  if (!r->r.keep_outputs || err) bc::free_out(*arena, p->mem);
  ++r->parts_done;
  ++stats.parts_done;
  if (err && r->status == 0) r->status = -err;
  if (r->parts_done == r->parts.size()) {
    r->t_done = t;
    if (r->status == 0) r->status = 1;
    ++stats.requests_done;
    --open_requests;
    cv_done.notify_all();
  }
}

void bc_server::finish_parts(std::vector<bc::Part *> &parts, int32_t err) {
  const int64_t t = bc::now_ns();
  for (bc::Part *p : parts) finish_part(p, t, err);
  parts.clear();
}

// Halves a queued part that cannot be placed while nothing runs (see
// admit_queued); the second half is appended to the request's parts.
bc::Part *bc_server::split(bc::Part *p) {
  using namespace bc;
  Request *r = p->req;
  uint64_t max_queued_need = 0;
  for (const Part *q : pending) max_queued_need = std::max(max_queued_need, part_need(q->mem));
  PartMem upper;
  if (!split_part(*arena, r->out_col, r->tmp_col, p->cols, max_queued_need, p->mem, upper, kept_reserved,
                  r->kept_reserved))
    return nullptr;
  const uint32_t c1 = p->cols / 2;
  r->parts.emplace_back();
  Part &q = r->parts.back();
  q.req = r;
  q.index = (uint32_t)(r->parts.size() - 1);
  q.first_col = p->first_col + c1;
  q.cols = p->cols - c1;
  q.mem = upper;
  q.work = r->work_col * q.cols;
  q.seg_bound = p->seg_bound;
  q.col_bound = 8 * r->stages * q.cols;
  p->cols = c1;
  p->work = r->work_col * c1;
  p->col_bound = 8 * r->stages * c1;
  return &q;
}

// Takes admitted parts in order (bounded by parts, work and segment table
// size) and launches them as one ragged batch on the slot's stream.
bool bc_server::launch_batch(bc::Slot &sl) {
  using namespace bc;
  std::vector<Part *> parts;
  uint64_t work = 0, segs_bound = 0, cols_bound = 0;
  const uint64_t work_cap = 2 * cfg.max_part_elems;
  while (!admitted.empty() && parts.size() < cfg.max_parts_per_batch) {
    Part *p = admitted.front();
    if (!parts.empty() && (work + p->work > work_cap || segs_bound + p->seg_bound > SLOT_MAX_SEGS ||
                           cols_bound + p->col_bound > SLOT_COL_CAP))
      break;
    parts.push_back(p);
    admitted.pop_front();
    work += p->work;
    segs_bound += p->seg_bound;
    cols_bound += p->col_bound;
  }
  if (parts.empty()) return false;

  std::vector<bc_ragged_job> jobs(parts.size());
  for (size_t k = 0; k < parts.size(); ++k) {
    const Part &p = *parts[k];
    const bc_request &r = p.req->r;
    const uint64_t n = 1ull << r.log_n;
    uint64_t *out = (uint64_t *)(arena_base + p.mem.out_off);
    uint64_t *tmp = (uint64_t *)(arena_base + p.mem.tmp_off);
    bc_ragged_job &j = jobs[k];
    j.op = r.op;
    j.log_n = r.log_n;
    j.num_cols = p.cols;
    j.log_blowup = r.op == BC_OP_LDE ? r.log_blowup : 0;
    j.outputs = out;
    j.stride_out = (uint32_t)(n * lde_blowup(r));
    if (r.inputs) {
      j.inputs = r.inputs + (uint64_t)p.first_col * r.stride_in;
      j.stride_in = (uint32_t)r.stride_in;
    } else {
      j.inputs = tmp;
      j.stride_in = (uint32_t)n;
    }
    if (r.op == BC_OP_LDE) {
      j.coeffs = tmp + (r.inputs ? 0 : n * p.cols);
      j.stride_coeffs = (uint32_t)n;
    }
  }
  bool ok = false;
  const Tables tab = current_tables(ok);
  if (!ok) {
    finish_parts(parts, BC_ERROR_NOT_INITIALIZED);
    return true;
  }
  RaggedHostPlan hp;
  int rc = ragged_build(jobs.data(), (uint32_t)jobs.size(), tab, hp);
  if (rc == BC_SUCCESS && ragged_table_bytes(hp) > SLOT_TABLE_BYTES) rc = BC_ERROR_INVALID_VALUE;
  if (rc != BC_SUCCESS) {
    finish_parts(parts, rc);
    return true;
  }
  const size_t table_bytes = ragged_table_bytes(hp);
  ragged_pack_table(hp, sl.h_table, sl.d_table);
  hipError_t err = hipMemcpyAsync(sl.d_table, sl.h_table, table_bytes, hipMemcpyHostToDevice, sl.stream);
  for (size_t k = 0; k < parts.size() && err == hipSuccess; ++k) {
    const Part &p = *parts[k];
    if (!p.req->r.inputs)
      err = fill_seeded((uint64_t *)(arena_base + p.mem.tmp_off), p.req->r.log_n, p.cols, 1ull << p.req->r.log_n,
                        p.req->r.seed, p.first_col, sl.stream);
  }
  if (err == hipSuccess)
    // one launch per level: batches overlap on the server's streams, which a
    // cooperative launch (the whole grid resident at once) would serialize
    err = ragged_launch(sl.d_table, hp, sl.d_state, workgroups, sl.stream, false);
  if (err == hipSuccess) err = hipEventRecord(sl.done, sl.stream);
  if (err != hipSuccess) {
    (void)hipStreamSynchronize(sl.stream);
    finish_parts(parts, (int32_t)err);
    return true;
  }
  const int64_t t = now_ns();
  for (Part *p : parts)
    if (!p->req->t_first_launch) p->req->t_first_launch = t;
  sl.parts = std::move(parts);
  sl.busy = true;
  ++stats.batches_launched;
  return true;
}

void bc_server::loop() {
  using namespace bc;
  (void)hipSetDevice(device);
  std::unique_lock<std::mutex> lk(mu);
  for (;;) {
    bool progress = false;
    // retire finished batches
    for (Slot &sl : slots) {
      if (!sl.busy) continue;
      const hipError_t q = hipEventQuery(sl.done);
      if (q == hipErrorNotReady) continue;
      const int32_t e = q != hipSuccess ? (int32_t)q : 0;
      finish_parts(sl.parts, e);  // e == 0: success
      sl.busy = false;
      progress = true;
    }
    // admit in arrival order while the next part's buffers fit
    bool running = !admitted.empty();
    for (const Slot &sl : slots) running = running || sl.busy;
    const int64_t t_adm = now_ns();
    progress |= admit_queued(
        *arena, pending, running, [&](Part *p) { admitted.push_back(p); }, [&](Part *p) { return split(p); },
        [&](Part *p) { finish_part(p, t_adm, hipErrorOutOfMemory); });
    // launch on a free stream
    uint32_t busy = 0;
    for (Slot &sl : slots) {
      if (!sl.busy && !admitted.empty() && launch_batch(sl)) progress = true;
      busy += sl.busy ? 1u : 0u;
    }
    stats.max_batches_in_flight_seen = std::max(stats.max_batches_in_flight_seen, busy);
    if (stop && pending.empty() && admitted.empty() && busy == 0) break;
    if (!progress) cv_work.wait_for(lk, std::chrono::microseconds(busy ? 20 : 500));
  }
}

namespace {

void destroy_server(bc_server *s) {
  for (bc::Slot &sl : s->slots) {
    if (sl.stream) (void)hipStreamSynchronize(sl.stream);
    if (sl.done) (void)hipEventDestroy(sl.done);
    if (sl.stream) (void)hipStreamDestroy(sl.stream);
    if (sl.d_table) (void)hipFree(sl.d_table);
    if (sl.d_state) (void)hipFree(sl.d_state);
    if (sl.h_table) (void)hipHostFree(sl.h_table);
  }
  if (s->arena_base) (void)hipFree(s->arena_base);
  delete s;
}

}  // namespace

extern "C" int bc_fill_seeded(uint64_t *out, uint32_t log_n, uint32_t num_cols, uint64_t stride, uint64_t seed,
                              uint32_t first_col, void *stream) {
  if ((!out && num_cols) || log_n >= 32 || (num_cols && stride < (1ull << log_n))) return BC_ERROR_INVALID_VALUE;
  return bc::fill_seeded(out, log_n, num_cols, stride, seed, first_col, (hipStream_t)stream);
}

extern "C" int bc_server_create(const bc_server_config *config, bc_server **out) {
  using namespace bc;
  if (!config || !out) return BC_ERROR_INVALID_VALUE;
  *out = nullptr;
  if (config->hbm_budget_bytes < ARENA_ALIGN) return BC_ERROR_INVALID_VALUE;
  bool ok = false;
  (void)current_tables(ok);
  if (!ok) return BC_ERROR_NOT_INITIALIZED;
  bc_server *s = new bc_server();
  s->cfg = *config;
  if (!s->cfg.max_part_elems) s->cfg.max_part_elems = 1ull << 28;
  if (!s->cfg.max_batches_in_flight) s->cfg.max_batches_in_flight = 3;
  if (!s->cfg.max_parts_per_batch) s->cfg.max_parts_per_batch = 64;
  s->cfg.max_batches_in_flight = std::min(s->cfg.max_batches_in_flight, 8u);
  hipError_t err = hipGetDevice(&s->device);
  const uint64_t cap = s->cfg.hbm_budget_bytes / ARENA_ALIGN * ARENA_ALIGN;
  if (err == hipSuccess) err = hipMalloc((void **)&s->arena_base, cap);
  if (err == hipSuccess) s->arena.reset(new Arena(cap, ARENA_ALIGN));
  s->slots.resize(s->cfg.max_batches_in_flight);
  for (Slot &sl : s->slots) {
    if (err == hipSuccess) err = hipStreamCreateWithFlags(&sl.stream, hipStreamNonBlocking);
    if (err == hipSuccess) err = hipEventCreateWithFlags(&sl.done, hipEventDisableTiming);
    if (err == hipSuccess) err = hipMalloc(&sl.d_table, SLOT_TABLE_BYTES);
    if (err == hipSuccess) err = hipMalloc(&sl.d_state, ragged_state_bytes());
    if (err == hipSuccess) err = hipHostMalloc(&sl.h_table, SLOT_TABLE_BYTES, 0);
  }
  if (err != hipSuccess) {
    destroy_server(s);
    return err;
  }
  s->workgroups = ragged_workgroups(s->device);
  s->worker = std::thread([s] { s->loop(); });
  *out = s;
  return BC_SUCCESS;
}

extern "C" int bc_server_submit(bc_server *s, const bc_request *rq, uint64_t *id) {
  using namespace bc;
  if (!s || !rq || !id) return BC_ERROR_INVALID_VALUE;
  const bc_request &r = *rq;
  if (r.op != BC_OP_NTT && r.op != BC_OP_INTT && r.op != BC_OP_LDE) return BC_ERROR_INVALID_VALUE;
  if (r.log_n < 8 || r.num_cols == 0) return BC_ERROR_INVALID_VALUE;
  const uint32_t lb = r.op == BC_OP_LDE ? r.log_blowup : 0;
  if (r.op == BC_OP_LDE && (lb < 1 || lb > 8)) return BC_ERROR_INVALID_VALUE;
  if (r.log_n > 31 - lb) return BC_ERROR_INVALID_VALUE;  // column strides are 32-bit (lb <= 8 here)
  const uint64_t n = 1ull << r.log_n;
  if (r.inputs && (r.stride_in < n || r.stride_in > 0xFFFFFFFFull)) return BC_ERROR_INVALID_VALUE;
  const uint64_t B = lde_blowup(r);
  const uint64_t out_col = 8 * n * B;
  const uint64_t tmp_col = 8 * n * ((r.inputs ? 0 : 1) + (r.op == BC_OP_LDE ? 1 : 0));
  const uint64_t work_col = n * (r.op == BC_OP_LDE ? 1 + B : 1);
  const uint64_t stages = r.op == BC_OP_LDE ? 1 + B : 1;

  std::unique_lock<std::mutex> lk(s->mu);
  Arena &ar = *s->arena;
  uint64_t cols_pp = std::max<uint64_t>(1, std::min<uint64_t>(r.num_cols, s->cfg.max_part_elems / work_col));
  // column tables: at most 8 passes per stage, and room for two parts per batch
  cols_pp = std::max<uint64_t>(1, std::min<uint64_t>(cols_pp, SLOT_COL_CAP / (2 * 8 * stages)));
  uint64_t max_queued_need = 0;
  for (const Part *q : s->pending) max_queued_need = std::max(max_queued_need, part_need(q->mem));
  RequestSizing sz;
  if (!size_request(ar, s->kept_reserved, max_queued_need, out_col, tmp_col, r.keep_outputs != 0, r.num_cols,
                    cols_pp, &sz))
    return hipErrorOutOfMemory;
  cols_pp = sz.cols_per_part;
  const uint32_t nparts = (uint32_t)((r.num_cols + cols_pp - 1) / cols_pp);

  std::unique_ptr<Request> req(new Request());
  req->r = r;
  req->out_col = out_col;
  req->tmp_col = tmp_col;
  req->work_col = work_col;
  req->stages = stages;
  req->parts.resize(nparts);
  for (uint32_t k = 0; k < nparts; ++k) {
    Part &p = req->parts[k];
    p.req = req.get();
    p.index = k;
    p.first_col = (uint32_t)(k * cols_pp);
    p.cols = (uint32_t)std::min<uint64_t>(cols_pp, r.num_cols - p.first_col);
    p.mem.keep = r.keep_outputs != 0;
    p.mem.out_bytes = ar.round(out_col * p.cols);
    p.mem.tmp_bytes = tmp_col ? ar.round(tmp_col * p.cols) : 0;
    p.work = work_col * p.cols;
    p.seg_bound = (uint32_t)(8 * stages);
    p.col_bound = 8 * stages * p.cols;
  }
  s->kept_reserved += sz.kept_bytes;
  req->kept_reserved = sz.kept_bytes;
  req->id = s->next_id++;
  req->t_submit = now_ns();
  for (Part &p : req->parts) s->pending.push_back(&p);
  *id = req->id;
  ++s->open_requests;
  s->reqs.emplace(req->id, std::move(req));
  lk.unlock();
  s->cv_work.notify_one();
  return BC_SUCCESS;
}

extern "C" int bc_server_wait(bc_server *s, uint64_t id) {
  if (!s) return BC_ERROR_INVALID_VALUE;
  std::unique_lock<std::mutex> lk(s->mu);
  auto it = s->reqs.find(id);
  if (it == s->reqs.end()) return BC_ERROR_INVALID_VALUE;
  bc::Request *r = it->second.get();
  s->cv_done.wait(lk, [&] { return r->status != 0; });
  return r->status == 1 ? BC_SUCCESS : -r->status;
}

extern "C" int bc_server_request_info(bc_server *s, uint64_t id, bc_request_info *info) {
  if (!s || !info) return BC_ERROR_INVALID_VALUE;
  std::lock_guard<std::mutex> lk(s->mu);
  auto it = s->reqs.find(id);
  if (it == s->reqs.end()) return BC_ERROR_INVALID_VALUE;
  const bc::Request &r = *it->second;
  info->t_submit_ns = r.t_submit;
  info->t_first_launch_ns = r.t_first_launch;
  info->t_done_ns = r.t_done;
  info->parts = (uint32_t)r.parts.size();
  info->parts_done = r.parts_done;
  info->status = r.status;
  info->reserved = 0;
  return BC_SUCCESS;
}

extern "C" int bc_server_part(bc_server *s, uint64_t id, uint32_t part, bc_part_info *info) {
  if (!s || !info) return BC_ERROR_INVALID_VALUE;
  std::lock_guard<std::mutex> lk(s->mu);
  auto it = s->reqs.find(id);
  if (it == s->reqs.end()) return BC_ERROR_INVALID_VALUE;
  const bc::Request &r = *it->second;
  if (r.status != 1 || part >= r.parts.size() || !r.parts[part].mem.out_held) return BC_ERROR_INVALID_VALUE;
  const bc::Part &p = r.parts[part];
  info->outputs = (uint64_t *)(s->arena_base + p.mem.out_off);
  info->stride_out = (1ull << r.r.log_n) * bc::lde_blowup(r.r);
  info->first_col = p.first_col;
  info->num_cols = p.cols;
  return BC_SUCCESS;
}

extern "C" int bc_server_release(bc_server *s, uint64_t id) {
  if (!s) return BC_ERROR_INVALID_VALUE;
  {
    std::lock_guard<std::mutex> lk(s->mu);
    auto it = s->reqs.find(id);
    if (it == s->reqs.end() || it->second->status == 0) return BC_ERROR_INVALID_VALUE;
    bc::Request &r = *it->second;
    for (bc::Part &p : r.parts) bc::free_out(*s->arena, p.mem);
    s->kept_reserved -= r.kept_reserved;
    s->reqs.erase(it);
  }
  s->cv_work.notify_one();
  return BC_SUCCESS;
}

extern "C" int bc_server_drain(bc_server *s) {
  if (!s) return BC_ERROR_INVALID_VALUE;
  std::unique_lock<std::mutex> lk(s->mu);
  s->cv_done.wait(lk, [&] { return s->open_requests == 0; });
  return BC_SUCCESS;
}

extern "C" int bc_server_stats_get(bc_server *s, bc_server_stats *out) {
  if (!s || !out) return BC_ERROR_INVALID_VALUE;
  std::lock_guard<std::mutex> lk(s->mu);
  *out = s->stats;
  out->pending_parts = (uint32_t)(s->pending.size() + s->admitted.size());
  out->arena_high_water_bytes = s->arena->high_water();
  out->arena_bytes_in_use = s->arena->in_use();
  return BC_SUCCESS;
}

extern "C" int bc_server_destroy(bc_server *s) {
  if (!s) return BC_ERROR_INVALID_VALUE;
  {
    std::lock_guard<std::mutex> lk(s->mu);
    s->stop = true;
  }
  s->cv_work.notify_one();
  if (s->worker.joinable()) s->worker.join();
  destroy_server(s);
  return BC_SUCCESS;
}

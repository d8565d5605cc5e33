// CPU check of the serving loop's memory policy (era-boojum-cuda_amd/csrc/
// admission.hpp) on a simulated server: requests of random sizes, kept or
// streamed outputs, parts completing in random order and kept outputs
// released late or never.  Checks after every step:
//   - the arena's own invariants;
//   - kept_reserved + (largest need of a queued part) <= capacity;
//   - after an admission round, parts are queued only while parts run, so a
//     caller that waits for every request before releasing anything never
//     waits forever (the reported hang: a streamed part sized against the
//     whole arena behind kept outputs).
// Prints "OK ..." with the request, failure and split counts on success.
#include "../../era-boojum-cuda_amd/csrc/admission.hpp"

#include <stdio.h>

#include <algorithm>
#include <deque>
#include <memory>
#include <random>
#include <vector>

struct Req;
struct Part {
  Req *req = nullptr;
  bc::PartMem mem;
  uint64_t cols = 0;
  bool done = false;
};
struct Req {
  bool keep = false, failed = false, released = false;
  uint64_t kept = 0, out_col = 0, tmp_col = 0;
  std::deque<Part> parts;  // stable addresses: parts split later are appended
  size_t done = 0;
};

struct Sim {
  bc::Arena ar;
  uint64_t kept_reserved = 0;
  std::deque<Part *> pending;
  std::vector<Part *> running;
  std::vector<std::unique_ptr<Req>> reqs;
  uint64_t frag_failures = 0, splits = 0;
  explicit Sim(uint64_t cap) : ar(cap, 64 * 1024) {}

  // returns false when the request is refused
  bool submit(uint64_t out_col, uint64_t tmp_col, bool keep, uint64_t cols, uint64_t max_cpp) {
    uint64_t maxq = 0;
    for (const Part *q : pending) maxq = std::max(maxq, bc::part_need(q->mem));
    bc::RequestSizing sz;
    if (!bc::size_request(ar, kept_reserved, maxq, out_col, tmp_col, keep, cols, max_cpp, &sz)) return false;
    std::unique_ptr<Req> r(new Req());
    r->keep = keep;
    r->kept = sz.kept_bytes;
    r->out_col = out_col;
    r->tmp_col = tmp_col;
    const uint64_t n = (cols + sz.cols_per_part - 1) / sz.cols_per_part;
    r->parts.resize(n);
    uint64_t sum_out = 0;
    for (uint64_t k = 0; k < n; ++k) {
      const uint64_t c = std::min(sz.cols_per_part, cols - k * sz.cols_per_part);
      Part &p = r->parts[k];
      p.req = r.get();
      p.cols = c;
      p.mem.keep = keep;
      p.mem.out_bytes = ar.round(out_col * c);
      p.mem.tmp_bytes = tmp_col ? ar.round(tmp_col * c) : 0;
      sum_out += p.mem.out_bytes;
    }
    if (keep && sum_out != sz.kept_bytes) return printf("FAIL kept bytes %llu != %llu\n",
                                                        (unsigned long long)sum_out,
                                                        (unsigned long long)sz.kept_bytes), false;
    kept_reserved += sz.kept_bytes;
    for (Part &p : r->parts) pending.push_back(&p);
    reqs.push_back(std::move(r));
    return true;
  }

  void finish(Part *p, bool err) {
    bc::free_tmp(ar, p->mem);
    if (!p->req->keep || err) bc::free_out(ar, p->mem);
    p->done = true;
    if (err) p->req->failed = true;
    ++p->req->done;
  }

  Part *split(Part *p) {
    uint64_t maxq = 0;
    for (const Part *q : pending) maxq = std::max(maxq, bc::part_need(q->mem));
    bc::PartMem upper;
    Req *r = p->req;
    if (!bc::split_part(ar, r->out_col, r->tmp_col, p->cols, maxq, p->mem, upper, kept_reserved, r->kept))
      return nullptr;
    r->parts.emplace_back();
    Part &q = r->parts.back();
    q.req = r;
    q.mem = upper;
    q.cols = p->cols - p->cols / 2;
    p->cols /= 2;
    ++splits;
    return &q;
  }

  void admit() {
    bc::admit_queued(
        ar, pending, !running.empty(), [&](Part *p) { running.push_back(p); },
        [&](Part *p) { return split(p); },
        [&](Part *p) {
          ++frag_failures;
          finish(p, true);
        });
  }

  void complete_one(std::mt19937_64 &rng) {
    if (running.empty()) return;
    const size_t k = rng() % running.size();
    Part *p = running[k];
    running.erase(running.begin() + (long)k);
    finish(p, false);
  }

  void release_one(std::mt19937_64 &rng) {
    std::vector<Req *> done;
    for (auto &r : reqs)
      if (!r->released && r->done == r->parts.size()) done.push_back(r.get());
    if (done.empty()) return;
    Req *r = done[rng() % done.size()];
    for (Part &p : r->parts) bc::free_out(ar, p.mem);
    kept_reserved -= r->kept;
    r->released = true;
  }

  bool invariants(const char **what) {
    if (!ar.check()) return *what = "arena", false;
    uint64_t maxq = 0;
    for (const Part *q : pending) maxq = std::max(maxq, bc::part_need(q->mem));
    if (kept_reserved + maxq > ar.capacity()) return *what = "reserve", false;
    return true;
  }
};

static int fail(const char *what, long step) {
  printf("FAIL %s at step %ld\n", what, step);
  return 1;
}

int main() {
  const uint64_t U = 64 * 1024;
  // the reported case: kept outputs fill most of the arena, then a streamed
  // request whose single-column part fits the whole arena but not beside
  // them; the caller waits for it before releasing anything
  {
    Sim s(100 * U);
    if (!s.submit(60 * U, 10 * U, true, 1, 1)) return fail("keep submit", -1);
    std::mt19937_64 rng1(1);
    s.admit();
    s.complete_one(rng1);
    if (!s.running.empty() || !s.pending.empty()) return fail("keep run", -1);
    // 8 columns of 8 U out + 4 U tmp: one part of all 8 (96 U) fits the arena
    // but not the 40 U beside the kept 60 U; parts of 2 columns (24 U) do
    if (!s.submit(8 * U, 4 * U, false, 8, 8)) return fail("stream submit", -1);
    std::mt19937_64 rng(2);
    for (int it = 0; it < 1000 && (!s.pending.empty() || !s.running.empty()); ++it) {
      s.admit();
      if (!s.pending.empty() && s.running.empty()) return fail("stalled behind kept outputs", it);
      s.complete_one(rng);
    }
    if (!s.pending.empty() || !s.running.empty() || s.reqs[1]->failed) return fail("stream did not finish", -1);
    if (s.reqs[1]->parts.size() < 4) return fail("stream was not split", -1);
    // a kept request that cannot fit beside the kept 60 U is refused up front
    if (s.submit(50 * U, 0, true, 1, 1)) return fail("over-reserve accepted", -1);
  }
  // random workload
  std::mt19937_64 rng(11);
  uint64_t frag = 0, requests = 0, splits = 0, failed_requests = 0;
  for (int round = 0; round < 16; ++round) {
    Sim s((200 + rng() % 800) * U);
    long step = 0;
    for (; step < 20000; ++step) {
      const int op = (int)(rng() % 10);
      if (op < 3) {
        const uint64_t out_col = (1 + rng() % 32) * U / (1 + rng() % 4);
        const uint64_t tmp_col = (rng() % 3) ? out_col / (1 + rng() % 4) : 0;
        const bool keep = rng() % 2;
        requests += s.submit(out_col, tmp_col, keep, 1 + rng() % 64, 1 + rng() % 16);
      } else if (op < 8) {
        s.complete_one(rng);
      } else if (op == 8 && (rng() % 4 == 0)) {
        s.release_one(rng);
      }
      s.admit();
      const char *what = nullptr;
      if (!s.invariants(&what)) return fail(what, step);
      if (!s.pending.empty() && s.running.empty()) return fail("queued with nothing running", step);
    }
    // drain without releasing anything: every request must finish or fail
    for (long it = 0; !s.pending.empty() || !s.running.empty(); ++it) {
      if (it > 10000000) return fail("drain did not end", it);
      s.admit();
      if (!s.pending.empty() && s.running.empty()) return fail("drain stalled", it);
      s.complete_one(rng);
    }
    uint64_t kept_sum = 0;
    for (auto &r : s.reqs) {
      if (r->done != r->parts.size()) return fail("request left unfinished", -1);
      failed_requests += r->failed;
      if (!r->released) kept_sum += r->kept;
    }
    if (kept_sum != s.kept_reserved) return fail("kept accounting", -1);
    frag += s.frag_failures;
    splits += s.splits;
  }
  printf("OK %llu requests, %llu failed (%llu parts), %llu splits\n", (unsigned long long)requests,
         (unsigned long long)failed_requests, (unsigned long long)frag, (unsigned long long)splits);
  return 0;
}

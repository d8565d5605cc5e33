// Memory policy of the serving loop (host bookkeeping only; no device calls):
// how a request is split into parts, and how a part's buffers are placed in
// the arena.
//
// Every queued part P needs need(P) more bytes before it can run: its
// temporaries, plus its outputs unless they are kept (kept outputs are
// reserved for the whole request when it is submitted).  The server keeps
//
//     kept_reserved + max over queued parts of need(P) <= capacity
//
// at every submit, so once nothing is running the head part always fits by
// size.  What could still stop it is fragmentation: kept outputs are placed
// from the top of the arena and everything transient from the bottom, which
// keeps the transient space in one piece; a head part that still cannot be
// placed while nothing runs is split into smaller parts, and one that cannot
// be split further is failed instead of waited for (nothing would ever free
// memory for it).
#pragma once
#include <stdint.h>

#include <algorithm>
#include <deque>

#include "arena.hpp"

namespace bc {

struct PartMem {
  uint64_t out_bytes = 0, tmp_bytes = 0;  // rounded to the arena's alignment
  uint64_t out_off = 0, tmp_off = 0;
  bool keep = false;                      // outputs stay until the request is released
  bool out_held = false, tmp_held = false;
};

inline uint64_t part_need(const PartMem &m) { return m.tmp_bytes + (m.keep ? 0 : m.out_bytes); }

struct RequestSizing {
  uint64_t cols_per_part = 0;
  uint64_t kept_bytes = 0;  // outputs reserved for the whole request (keep only)
};

// Columns per part for a request of num_cols columns whose columns take
// out_col output and tmp_col temporary bytes: the largest value, halving from
// max_cols_per_part, that keeps the invariant above with kept_reserved
// already promised and queued parts needing up to max_queued_need.  False:
// the request cannot run even one column at a time.
inline bool size_request(const Arena &ar, uint64_t kept_reserved, uint64_t max_queued_need, uint64_t out_col,
                         uint64_t tmp_col, bool keep, uint64_t num_cols, uint64_t max_cols_per_part,
                         RequestSizing *out) {
  typedef unsigned __int128 u128;
  const u128 cap = ar.capacity();
  const u128 align = ar.round(1);
  auto round = [&](u128 b) { return (b + align - 1) / align * align; };
  uint64_t cpp = std::max<uint64_t>(1, std::min(num_cols, max_cols_per_part));
  for (;;) {
    const uint64_t full = num_cols / cpp, rest = num_cols % cpp;
    const u128 out_p = round((u128)out_col * cpp);
    const u128 tmp_p = tmp_col ? round((u128)tmp_col * cpp) : 0;
    const u128 kept = keep ? (u128)full * out_p + (rest ? round((u128)out_col * rest) : 0) : 0;
    const u128 need = tmp_p + (keep ? 0 : out_p);  // a full part's need, the largest of the request
    const u128 others = keep ? (u128)max_queued_need : 0;
    if ((u128)kept_reserved + kept + std::max(need, others) <= cap && out_p + tmp_p <= cap) {
      out->cols_per_part = cpp;
      out->kept_bytes = (uint64_t)kept;
      return true;
    }
    if (cpp == 1) return false;
    cpp /= 2;
  }
}

// Places a part's buffers: kept outputs from the top of the arena, transient
// outputs and temporaries from the bottom.  All or nothing.
inline bool place_part(Arena &ar, PartMem &m) {
  uint64_t off = 0;
  if (m.out_bytes) {
    if (!(m.keep ? ar.alloc_high(m.out_bytes, &off) : ar.alloc(m.out_bytes, &off))) return false;
    m.out_off = off;
    m.out_held = true;
  }
  if (m.tmp_bytes) {
    if (!ar.alloc(m.tmp_bytes, &off)) {
      if (m.out_held) {
        ar.release(m.out_off, m.out_bytes);
        m.out_held = false;
      }
      return false;
    }
    m.tmp_off = off;
    m.tmp_held = true;
  }
  return true;
}

// Splits a queued part of `cols` >= 2 columns (out_col / tmp_col bytes per
// column) into its first cols / 2 columns, left in m, and the rest, written
// to upper.  Rounding each piece to the arena's alignment can add up to one
// unit per piece to a kept part's outputs; that is added to kept_reserved and
// to the request's kept_bytes, and the split is refused (false, nothing
// changed) if the total would break the reservation invariant with queued
// parts needing up to max_queued_need.
inline bool split_part(const Arena &ar, uint64_t out_col, uint64_t tmp_col, uint64_t cols, uint64_t max_queued_need,
                       PartMem &m, PartMem &upper, uint64_t &kept_reserved, uint64_t &req_kept) {
  if (cols < 2) return false;
  const uint64_t c1 = cols / 2, c2 = cols - c1;
  PartMem lo = m, hi = m;
  lo.out_bytes = ar.round(out_col * c1);
  hi.out_bytes = ar.round(out_col * c2);
  lo.tmp_bytes = tmp_col ? ar.round(tmp_col * c1) : 0;
  hi.tmp_bytes = tmp_col ? ar.round(tmp_col * c2) : 0;
  const uint64_t grow = m.keep ? lo.out_bytes + hi.out_bytes - m.out_bytes : 0;
  if (grow && kept_reserved + grow + std::max(max_queued_need, std::max(part_need(lo), part_need(hi))) >
                  ar.capacity())
    return false;
  kept_reserved += grow;
  req_kept += grow;
  m = lo;
  upper = hi;
  return true;
}

// One admission round over the queue (parts with a PartMem `mem` and a
// request pointer `req`; a request's parts are queued together): places
// parts in arrival order, handing each to admit(), until one does not fit.
// While anything runs, that part waits for memory to be freed.  If nothing
// runs (`running` false and nothing admitted this round), nothing would ever
// free memory for it: split(p) may then halve it (returning the new second
// half, queued right behind it, or nullptr), and a part that can be split no
// further has its request's queued parts go to fail(), and the round goes on
// with the next request.  Returns whether anything changed.
template <class Part, class Admit, class Split, class Fail>
bool admit_queued(Arena &ar, std::deque<Part *> &pending, bool running, Admit admit, Split split, Fail fail) {
  bool progress = false;
  while (!pending.empty()) {
    Part *p = pending.front();
    if (place_part(ar, p->mem)) {
      pending.pop_front();
      admit(p);
      running = true;
      progress = true;
      continue;
    }
    if (running) break;
    if (Part *q = split(p)) {
      pending.insert(pending.begin() + 1, q);
      progress = true;
      continue;
    }
    const auto *r = p->req;
    while (!pending.empty() && pending.front()->req == r) {
      Part *q = pending.front();
      pending.pop_front();
      fail(q);
    }
    progress = true;
  }
  return progress;
}

inline void free_tmp(Arena &ar, PartMem &m) {
  if (m.tmp_held) {
    ar.release(m.tmp_off, m.tmp_bytes);
    m.tmp_held = false;
  }
}

inline void free_out(Arena &ar, PartMem &m) {
  if (m.out_held) {
    ar.release(m.out_off, m.out_bytes);
    m.out_held = false;
  }
}

}  // namespace bc

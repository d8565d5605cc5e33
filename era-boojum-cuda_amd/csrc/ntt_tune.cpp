// Tuning table: per (log_n, log2 of the column count) launch parameters that
// replace the built-in pass split and launch heuristics.  Entries come from
// the autotuner (boojum_cuda_amd/autotune.py) through bc_tune_set or a text
// database (bc_tune_load / bc_tune_save); bc_context_create loads the
// database named by $BC_TUNE_DB, else tuning/mi355x.txt next to this library,
// when it exists.  Host code only.
#include "../../include/boojum_ntt.h"
#include "ntt_internal.hpp"

#include <dlfcn.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <map>
#include <mutex>
#include <string>

namespace bc {
namespace {

std::mutex g_tune_mu;
std::map<uint32_t, TuneParams> g_tune;  // key: log_n << 8 | log_cols

uint32_t tune_key(uint32_t log_n, uint32_t log_cols) { return log_n << 8 | log_cols; }

uint32_t log_cols_of(uint32_t num_cols) {
  uint32_t l = 0;
  while (l < 31 && (2u << l) <= num_cols) ++l;
  return l;
}

}  // namespace

bool tune_params_valid(uint32_t log_n, const TuneParams &p) {
  if (p.contig_log) {
    if (log_n <= TILE_LOG) {
      if (p.contig_log != log_n || p.num_strided) return false;
    } else {
      if (p.contig_log < 9 || p.contig_log > TILE_LOG || p.num_strided < 1 || p.num_strided > 7) return false;
      const uint32_t s = log_n - p.contig_log;
      if (s < 4 * p.num_strided || s > 8 * p.num_strided) return false;
    }
  } else if (p.num_strided) {
    return false;
  }
  if (p.tile_log != 0 && p.tile_log != 8 && p.tile_log != TILE_LOG) return false;
  if (p.tile_log == 8 && log_n != SMALL_LOG_N) return false;
  return p.strided_bpw <= 64 && p.contig_grid <= 65535 && p.unit_contig <= 64 && p.unit_strided <= 64;
}

TuneParams tuned(uint32_t log_n, uint32_t num_cols) {
  const uint32_t lc = log_cols_of(num_cols ? num_cols : 1);
  std::lock_guard<std::mutex> lk(g_tune_mu);
  auto it = g_tune.find(tune_key(log_n, lc));
  if (it != g_tune.end()) return it->second;
  // nearest column bucket tuned for this log_n (ties: the smaller bucket)
  const TuneParams *best = nullptr;
  uint32_t best_d = ~0u;
  for (auto i = g_tune.lower_bound(tune_key(log_n, 0)); i != g_tune.end() && (i->first >> 8) == log_n; ++i) {
    const uint32_t b = i->first & 0xFF;
    const uint32_t d = b > lc ? b - lc : lc - b;
    if (d < best_d) {
      best_d = d;
      best = &i->second;
    }
  }
  return best ? *best : TuneParams{};
}

int tune_load_file(const char *path) {
  FILE *f = fopen(path, "r");
  if (!f) return BC_ERROR_INVALID_VALUE;
  char line[512];
  std::map<uint32_t, TuneParams> add;
  int rc = BC_SUCCESS;
  while (fgets(line, sizeof line, f)) {
    const char *p = line;
    while (*p == ' ' || *p == '\t') ++p;
    if (*p == '#' || *p == '\n' || *p == '\0') continue;
    unsigned v[9] = {};
    const int got =
        sscanf(p, "%u %u %u %u %u %u %u %u %u", &v[0], &v[1], &v[2], &v[3], &v[4], &v[5], &v[6], &v[7], &v[8]);
    if (got != 8 && got != 9) {
      rc = BC_ERROR_INVALID_VALUE;
      break;
    }
    const TuneParams t{v[2], v[3], v[4], v[5], v[6], v[7], v[8]};
    if (v[0] < 1 || v[0] > OMEGA_LOG_ORDER || v[1] > 31 || !tune_params_valid(v[0], t)) {
      rc = BC_ERROR_INVALID_VALUE;
      break;
    }
    add[tune_key(v[0], v[1])] = t;
  }
  fclose(f);
  if (rc != BC_SUCCESS) return rc;
  std::lock_guard<std::mutex> lk(g_tune_mu);
  for (const auto &kv : add) g_tune[kv.first] = kv.second;
  return BC_SUCCESS;
}

// $BC_TUNE_DB, else tuning/mi355x.txt beside this shared library.  A missing
// default database is not an error; a malformed one is ignored.
void tune_autoload() {
  static std::once_flag once;
  std::call_once(once, [] {
    const char *env = getenv("BC_TUNE_DB");
    if (env && *env) {
      (void)tune_load_file(env);
      return;
    }
    Dl_info info;
    if (!dladdr((void *)&tune_autoload, &info) || !info.dli_fname) return;
    std::string dir(info.dli_fname);
    const size_t slash = dir.rfind('/');
    dir = slash == std::string::npos ? std::string(".") : dir.substr(0, slash);
    (void)tune_load_file((dir + "/tuning/mi355x.txt").c_str());
  });
}

}  // namespace bc

extern "C" int bc_tune_set(uint32_t log_n, uint32_t log_cols, const bc_tune_params *p) {
  using namespace bc;
  if (log_n < 1 || log_n > OMEGA_LOG_ORDER || log_cols > 31) return BC_ERROR_INVALID_VALUE;
  std::lock_guard<std::mutex> lk(g_tune_mu);
  if (!p) {
    g_tune.erase(tune_key(log_n, log_cols));
    return BC_SUCCESS;
  }
  const TuneParams t{p->contig_log,         p->num_strided,         p->strided_bpw, p->contig_grid,
                     p->ragged_unit_contig, p->ragged_unit_strided, p->tile_log};
  if (!tune_params_valid(log_n, t)) return BC_ERROR_INVALID_VALUE;
  g_tune[tune_key(log_n, log_cols)] = t;
  return BC_SUCCESS;
}

extern "C" int bc_tune_get(uint32_t log_n, uint32_t num_cols, bc_tune_params *p) {
  if (!p) return BC_ERROR_INVALID_VALUE;
  const bc::TuneParams t = bc::tuned(log_n, num_cols);
  p->contig_log = t.contig_log;
  p->num_strided = t.num_strided;
  p->strided_bpw = t.strided_bpw;
  p->contig_grid = t.contig_grid;
  p->ragged_unit_contig = t.unit_contig;
  p->ragged_unit_strided = t.unit_strided;
  p->tile_log = t.tile_log;
  return BC_SUCCESS;
}

extern "C" int bc_tune_clear(void) {
  std::lock_guard<std::mutex> lk(bc::g_tune_mu);
  bc::g_tune.clear();
  return BC_SUCCESS;
}

extern "C" int bc_tune_load(const char *path) {
  if (!path) return BC_ERROR_INVALID_VALUE;
  return bc::tune_load_file(path);
}

extern "C" int bc_tune_save(const char *path) {
  using namespace bc;
  if (!path) return BC_ERROR_INVALID_VALUE;
  FILE *f = fopen(path, "w");
  if (!f) return BC_ERROR_INVALID_VALUE;
  fprintf(f, "# log_n log_cols contig_log num_strided strided_bpw contig_grid ragged_unit_contig "
             "ragged_unit_strided tile_log\n");
  {
    std::lock_guard<std::mutex> lk(g_tune_mu);
    for (const auto &kv : g_tune) {
      const TuneParams &t = kv.second;
      fprintf(f, "%u %u %u %u %u %u %u %u %u\n", kv.first >> 8, kv.first & 0xFF, t.contig_log, t.num_strided,
              t.strided_bpw, t.contig_grid, t.unit_contig, t.unit_strided, t.tile_log);
    }
  }
  return fclose(f) == 0 ? BC_SUCCESS : BC_ERROR_INVALID_VALUE;
}

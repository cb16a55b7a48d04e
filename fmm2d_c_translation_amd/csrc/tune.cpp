// tune.cpp -- the library's autotuner for the translation kernels' tile
// shapes, and the table it fills.
//
// Choices: the row tiles per wave of the all-levels M2L kernels for each
// order's row-tile count, Laplace and Helmholtz, FP64 and FP32 apart (more
// row tiles per wave reuse each loaded operand more; fewer keep more waves
// resident), and the level size below which M2M/L2L launches split their
// slots over wave groups.  fmm2dhip_autotune times the FMM steps these
// kernels run (the solve's own step events) on the caller's problem under
// every candidate and keeps the fastest; fmm2dhip_autotune_batch does the
// same on each fused group of a batch, solved as one forest.  The table is
// process-wide, can be saved and loaded, and is loaded at start-up from the
// file FMM2DHIP_TUNE_FILE names.
#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <mutex>

#include "engine.h"

namespace fmm {

TuneTable::TuneTable() {
  for (auto& fam : m2l_mw)
    for (auto& row : fam)
      for (auto& v : row) v.store(0, std::memory_order_relaxed);
  xlate_waves.store(0, std::memory_order_relaxed);
}

namespace {
std::mutex g_tune_mu;

const char* kFamily[2] = {"laplace", "helmholtz"};

int load_table(TuneTable& t, const char* path) {
  FILE* f = std::fopen(path, "r");
  if (!f) return -1;
  char kind[32], fam[16], prec[16];
  int a = 0, b = 0, n = 0;
  char line[256];
  while (std::fgets(line, sizeof line, f)) {
    if (std::sscanf(line, "%31s %15s %15s %d %d", kind, fam, prec, &a, &b) == 5 &&
        !std::strcmp(kind, "m2l_mw") && a >= 1 && a <= 8 && b >= 0 && b <= 4 &&
        (!std::strcmp(fam, kFamily[0]) || !std::strcmp(fam, kFamily[1])) &&
        (!std::strcmp(prec, "fp64") || !std::strcmp(prec, "fp32"))) {
      t.m2l_mw[std::strcmp(fam, kFamily[0]) ? 1 : 0][std::strcmp(prec, "fp32") ? 0 : 1][a].store(
          b, std::memory_order_relaxed);
      ++n;
    } else if (std::sscanf(line, "%31s %d", kind, &a) == 2 && !std::strcmp(kind, "xlate_waves") &&
               a >= 0 && a <= (1 << 20)) {
      t.xlate_waves.store(a, std::memory_order_relaxed);
      ++n;
    }
  }
  std::fclose(f);
  return n;
}
}  // namespace

TuneTable& tune_table() {
  static TuneTable* t = [] {
    auto* p = new TuneTable();
    if (const char* path = std::getenv("FMM2DHIP_TUNE_FILE")) (void)load_table(*p, path);
    return p;
  }();
  return *t;
}

int m2l_tiles(int MT, bool fp32, int family) {
  const int v = MT >= 1 && MT <= 8 && (family == 0 || family == 1)
                    ? tune_table().m2l_mw[family][fp32 ? 1 : 0][MT].load(std::memory_order_relaxed)
                    : 0;
  if (v >= 1 && v <= 4 && v <= MT) return v;
  return MT <= 4 ? MT : (MT + 1) / 2;
}

int xlate_waves() {
  const int v = tune_table().xlate_waves.load(std::memory_order_relaxed);
  return v > 0 ? v : 4096;
}

namespace {
// best device time of steps (bit k: timeinfo step k) over nrep solves (of
// a problem, or of a fused group as one forest)
template <class Solve>
double time_steps(Solve solve, unsigned steps, int nrep, int* err) {
  double best = 1e30;
  for (int r = 0; r < nrep; ++r) {
    if (int e = solve()) {
      *err = e;
      return 1e30;
    }
    const SolveStats& st = last_stats();
    double t = 0;
    for (int k = 0; k < 8; ++k)
      if (steps >> k & 1) t += st.step_s[k];
    best = std::min(best, t);
  }
  return best;
}
}  // namespace

}  // namespace fmm

extern "C" {

}  // extern "C"

namespace fmm {
namespace {
// Tunes the entries the solve uses; returns the number changed or -ier.
template <class Solve>
int autotune(Solve solve, int nrep) {
  nrep = std::max(1, std::min(nrep, 20));
  int err = 0, changed = 0;
  TuneTable& t = tune_table();
  (void)time_steps(solve, 0, 1, &err);  // warm: workspace, tables, caches
  if (err) return -err;
  const uint64_t mask = last_stats().m2l_mt_mask;
  // a candidate replaces the current choice only when at least 2% faster
  for (int fam = 0; fam < 2; ++fam)
    for (int f32 = 0; f32 < 2; ++f32)
      for (int MT = 1; MT <= 8; ++MT) {
        if (!(mask >> ((fam * 2 + f32) * 9 + MT) & 1)) continue;
        std::atomic<int>& slot = t.m2l_mw[fam][f32][MT];
        const int cur = m2l_tiles(MT, f32 != 0, fam);
        int best = cur;
        double tbest = time_steps(solve, 1u << 3, nrep, &err);
        for (int c = 1; c <= std::min(4, MT) && !err; ++c) {
          if (c == cur) continue;
          slot.store(c, std::memory_order_relaxed);
          const double tc = time_steps(solve, 1u << 3, nrep, &err);
          if (tc < 0.98 * tbest) {
            tbest = tc;
            best = c;
          }
        }
        slot.store(best, std::memory_order_relaxed);
        changed += best != cur;
        if (err) return -err;
      }
  {
    const int cur = xlate_waves();
    int best = cur;
    const unsigned steps = (1u << 2) | (1u << 4);  // M2M and L2L
    double tbest = time_steps(solve, steps, nrep, &err);
    for (int c : {1024, 2048, 4096, 8192}) {
      if (c == cur || err) continue;
      t.xlate_waves.store(c, std::memory_order_relaxed);
      const double tc = time_steps(solve, steps, nrep, &err);
      if (tc < 0.98 * tbest) {
        tbest = tc;
        best = c;
      }
    }
    t.xlate_waves.store(best, std::memory_order_relaxed);
    changed += best != cur;
    if (err) return -err;
  }
  return changed;
}
}  // namespace
}  // namespace fmm

extern "C" {

int fmm2dhip_autotune(fmm2dhip_problem* p, int nrep) {
  if (!p) return -fmm::IER_ARGS;
  std::lock_guard<std::mutex> g(fmm::g_tune_mu);
  return fmm::autotune([&] { return fmm::solve_problem(*p); }, nrep);
}

int fmm2dhip_autotune_batch(int n, fmm2dhip_problem* probs, int nrep) {
  if (n <= 0 || !probs) return -fmm::IER_ARGS;
  std::lock_guard<std::mutex> g(fmm::g_tune_mu);
  int changed = 0;
  for (const std::vector<int>& grp : fmm::fused_groups(n, probs)) {
    const int c = fmm::autotune([&] { return fmm::solve_forest(probs, grp) ? fmm::IER_HIP : 0; },
                                nrep);
    if (c < 0) return c;
    changed += c;
  }
  return changed;
}

int fmm2dhip_tune_get(int* out, int n) {
  int v[37];
  for (int fam = 0; fam < 2; ++fam)
    for (int f = 0; f < 2; ++f)
      for (int MT = 0; MT <= 8; ++MT)
        v[9 * (2 * fam + f) + MT] = MT ? fmm::m2l_tiles(MT, f != 0, fam) : 0;
  v[36] = fmm::xlate_waves();
  for (int k = 0; k < n && k < 37; ++k) out[k] = v[k];
  return 37;
}

int fmm2dhip_tune_save(const char* path) {
  if (!path) return -1;
  FILE* f = std::fopen(path, "w");
  if (!f) return -1;
  std::fprintf(f, "# fmm2dhip tile choices (fmm2dhip_autotune)\n");
  for (int fam = 0; fam < 2; ++fam)
    for (int f32 = 0; f32 < 2; ++f32)
      for (int MT = 1; MT <= 8; ++MT)
        std::fprintf(f, "m2l_mw %s %s %d %d\n", fmm::kFamily[fam], f32 ? "fp32" : "fp64", MT,
                     fmm::m2l_tiles(MT, f32 != 0, fam));
  std::fprintf(f, "xlate_waves %d\n", fmm::xlate_waves());
  return std::fclose(f) == 0 ? 0 : -1;
}

int fmm2dhip_tune_load(const char* path) {
  if (!path) return -1;
  std::lock_guard<std::mutex> g(fmm::g_tune_mu);
  return fmm::load_table(fmm::tune_table(), path);
}

}  // extern "C"

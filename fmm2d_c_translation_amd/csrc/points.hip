// points.hip -- device front end: bounding box, Morton keys, radix sort,
// in-leaf ordering and gather/scatter of point data; the tree itself is
// built by dtree.hip.
#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <chrono>
#include <cstdlib>
#include <cstring>
#include <atomic>
#include <thread>
#include <vector>

#include "engine.h"
#include "rsort.h"

namespace fmm {

static thread_local StepTimer g_timer;
double wall() {
  return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

namespace {
thread_local HostKeeper* g_keeper = nullptr;
thread_local bool g_capture = false;
}  // namespace

bool capturing() { return g_capture; }
CaptureScope::CaptureScope(HostKeeper* keeper) {
  g_keeper = keeper;
  g_capture = true;
}
CaptureScope::~CaptureScope() {
  g_keeper = nullptr;
  g_capture = false;
}
int HostKeeper::reserve(size_t bytes) {
  if (pool) return IER_OK;
  if (hipHostMalloc(reinterpret_cast<void**>(&pool), bytes, hipHostMallocDefault) != hipSuccess) {
    pool = nullptr;
    return IER_ALLOC;
  }
  cap = bytes;
  used = 0;
  return IER_OK;
}
HostKeeper::~HostKeeper() {
  if (pool) (void)hipHostFree(pool);
}

void StepTimer::mark(hipStream_t st) {
  if (ok && n < 9 && !g_capture) (void)hipEventRecord(ev[n++], st);
}

StepTimer* step_timer_peek() { return &g_timer; }
StepTimer* step_timer() {
  StepTimer& t = g_timer;
  if (!t.ok) {
    bool good = true;
    for (auto& e : t.ev) good = good && hipEventCreate(&e) == hipSuccess;
    t.ok = good;
  }
  t.n = 0;
  return &t;
}

void finish_timer(double* timeinfo) {
  StepTimer& t = *step_timer_peek();
  SolveStats& st = last_stats();
  for (int k = 0; k < 8; ++k) {
    float ms = 0.f;
    if (t.ok && t.n >= k + 2 && hipEventElapsedTime(&ms, t.ev[k], t.ev[k + 1]) == hipSuccess) {
      st.step_s[t.order[k]] = ms * 1e-3;
      if (timeinfo) timeinfo[t.order[k]] = ms * 1e-3;
    }
  }
}

SolveStats& last_stats() {
  static thread_local SolveStats s;
  return s;
}

int& thread_precision() {
  static const int initial = [] {
    const char* v = std::getenv("FMM2DHIP_PRECISION");
    return v && std::strcmp(v, "mixed") == 0 ? (int)PREC_MIXED : (int)PREC_FP64;
  }();
  static thread_local int mode = initial;
  return mode;
}

namespace {
struct ThreadStream {
  hipStream_t st = nullptr;
  int dev = -1;
};
thread_local ThreadStream g_stream;
struct Pinned {
  void* buf[8] = {};
  size_t cap[8] = {};
};
thread_local Pinned g_pinned;
// the thread's staged DMAs in flight (the staging, pinned slot 0, is reused)
thread_local hipEvent_t g_stage_done = nullptr;
thread_local bool g_stage_pending = false;
}  // namespace

namespace {
struct ArenaOverride {
  char* base = nullptr;
  size_t cap = 0, used = 0, high = 0;
  hipStream_t st = nullptr;
  bool on = false, overflow = false;
};
thread_local ArenaOverride g_fixed;
}  // namespace

FixedArena::FixedArena(void* base, size_t cap, hipStream_t stream) {
  g_fixed = ArenaOverride{static_cast<char*>(base), cap, 0, 0, stream, true, false};
}
FixedArena::~FixedArena() { g_fixed = ArenaOverride{}; }
size_t FixedArena::used() const { return g_fixed.high; }
bool FixedArena::overflowed() const { return g_fixed.overflow; }

int get_ctx(Ctx& c) {
  ThreadStream& T = g_stream;
  int d = 0;
  if (hipGetDevice(&d) != hipSuccess) return IER_NOGPU;
  if (g_fixed.on) {
    c.st = g_fixed.st;
    c.device = d;
    return IER_OK;
  }
  if (!T.st || T.dev != d) {
    if (hipStreamCreateWithFlags(&T.st, hipStreamNonBlocking) != hipSuccess) return IER_NOGPU;
    T.dev = d;
  }
  c.st = T.st;
  c.device = d;
  return IER_OK;
}

namespace {
constexpr int kMaxDevices = 16;
struct Arena {
  std::vector<std::pair<char*, size_t>> blocks;  // the last one is being filled
  size_t used = 0;   // bytes taken from the last block
  size_t total = 0;  // bytes taken by the current solve
  int depth = 0;     // live Scratch objects
};
thread_local Arena g_arena[kMaxDevices];
constexpr size_t kBlockMin = size_t(64) << 20;
}  // namespace

Scratch::Scratch(hipStream_t s) : st_(s) {
  if (hipGetDevice(&dev_) != hipSuccess || dev_ < 0 || dev_ >= kMaxDevices) dev_ = 0;
  g_arena[dev_].depth++;
}

void* Scratch::take(size_t bytes) {
  bytes = (bytes + 255) & ~size_t(255);
  if (g_fixed.on) {  // a FixedArena: bump only, never allocate
    ArenaOverride& F = g_fixed;
    if (bytes > F.cap - F.used) {
      F.overflow = true;
      return nullptr;
    }
    void* p = F.base + F.used;
    F.used += bytes;
    F.high = std::max(F.high, F.used);
    return p;
  }
  Arena& A = g_arena[dev_];
  A.total += bytes;
  if (!A.blocks.empty() && A.used + bytes <= A.blocks.back().second) {
    void* p = A.blocks.back().first + A.used;
    A.used += bytes;
    return p;
  }
  size_t want = std::max(bytes, kBlockMin);
  char* p = nullptr;
  if (hipMalloc(&p, want) != hipSuccess) return nullptr;
  A.blocks.push_back({p, want});
  A.used = bytes;
  return p;
}

Scratch::~Scratch() {
  Arena& A = g_arena[dev_];
  if (--A.depth > 0) return;
  (void)hipStreamSynchronize(st_);
  if (g_fixed.on) {
    last_stats().workspace_bytes = (double)g_fixed.high;
    g_fixed.used = 0;
    return;
  }
  if (A.blocks.size() > 1) {
    size_t want = std::max(A.total + A.total / 4, kBlockMin);
    for (auto& b : A.blocks) (void)hipFree(b.first);
    A.blocks.clear();
    char* p = nullptr;
    if (hipMalloc(&p, want) == hipSuccess) A.blocks.push_back({p, want});
  }
  last_stats().workspace_bytes = (double)A.total;
  A.used = 0;
  A.total = 0;
}

size_t workspace_capacity() {
  int d = 0;
  if (hipGetDevice(&d) != hipSuccess || d < 0 || d >= kMaxDevices) return 0;
  size_t n = 0;
  for (auto& b : g_arena[d].blocks) n += b.second;
  return n;
}

void release_workspace() {
  int d = 0;
  if (hipGetDevice(&d) != hipSuccess || d < 0 || d >= kMaxDevices) return;
  Arena& A = g_arena[d];
  if (A.depth > 0) return;
  (void)hipDeviceSynchronize();
  for (auto& b : A.blocks) (void)hipFree(b.first);
  A.blocks.clear();
  A.used = A.total = 0;
}

void release_thread() {
  release_workspace();
  if (g_timer.ok) {
    for (auto& e : g_timer.ev) (void)hipEventDestroy(e);
    g_timer = StepTimer{};
  }
  if (g_stage_pending) (void)hipEventSynchronize(g_stage_done);
  g_stage_pending = false;
  if (g_stage_done) (void)hipEventDestroy(g_stage_done);
  g_stage_done = nullptr;
  for (int k = 0; k < 8; ++k) {
    if (g_pinned.buf[k]) (void)hipHostFree(g_pinned.buf[k]);
    g_pinned.buf[k] = nullptr;
    g_pinned.cap[k] = 0;
  }
  if (g_stream.st) {
    (void)hipStreamSynchronize(g_stream.st);
    (void)hipStreamDestroy(g_stream.st);
  }
  g_stream = ThreadStream{};
}

namespace {
__global__ void k_fill32(uint32_t* p, uint32_t v, size_t n) {
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (size_t)gridDim.x * blockDim.x)
    p[i] = v;
}
}  // namespace

void fill_async(void* p, uint32_t v, size_t n, hipStream_t st) {
  if (!n) return;
  const size_t blocks = std::min<size_t>((n + 255) / 256, 8192);
  k_fill32<<<(unsigned)blocks, 256, 0, st>>>(static_cast<uint32_t*>(p), v, n);
}

int upload_small(void* d, const void* h, size_t bytes, hipStream_t st) {
  if (!bytes) return IER_OK;
  if (g_keeper) {
    HostKeeper& K = *g_keeper;
    const size_t at = (K.used + 255) & ~size_t(255);
    if (!K.pool || at + bytes > K.cap) return IER_ALLOC;
    std::memcpy(K.pool + at, h, bytes);
    K.used = at + bytes;
    h = K.pool + at;
  }
  HIP_TRY(hipMemcpyAsync(d, h, bytes, hipMemcpyHostToDevice, st));
  return IER_OK;
}

void detach_workspace(std::vector<void*>& out) {
  int d = 0;
  if (hipGetDevice(&d) != hipSuccess || d < 0 || d >= kMaxDevices) return;
  Arena& A = g_arena[d];
  for (auto& b : A.blocks) out.push_back(b.first);
  A.blocks.clear();
  A.used = 0;
  A.total = 0;
}

namespace {

// memcpy of a list of (src, dst, bytes) between the callers' arrays and the
// staging (array i at stage + at[i], at[n] the staging's extent) on up to
// eight threads, each taking a contiguous share of the staging's extent
// (pieces of 4 MiB at least); small totals stay on the calling thread
void par_copy(const HostXfer* x, int n, bool from_staging, const char* stage,
              const size_t* at) {
  const size_t total = at[n];
  const size_t kPiece = size_t(4) << 20;
  const unsigned hw = std::max(1u, std::thread::hardware_concurrency());
  // helper threads across all callers (the batch workers copy at once) stay
  // within the host's cores
  static std::atomic<int> busy{0};
  int want = (int)std::min<size_t>({8, hw, std::max<size_t>(1, total / kPiece)}) - 1;
  const int cap = std::max(0, (int)std::min(16u, hw) - busy.load());
  want = std::min(want, cap);
  busy += want;
  const int nth = 1 + want;
  // bytes [b0, b1) of the concatenation
  auto run = [&](size_t b0, size_t b1) {
    for (int i = 0; i < n && b0 < b1; ++i) {
      const size_t lo = at[i], hi = at[i] + x[i].bytes;
      if (hi <= b0 || lo >= b1) continue;
      const size_t a = std::max(lo, b0), b = std::min(hi, b1);
      if (from_staging)
        std::memcpy(static_cast<char*>(x[i].dst) + (a - lo), stage + a, b - a);
      else
        std::memcpy(const_cast<char*>(stage) + a, static_cast<const char*>(x[i].src) + (a - lo),
                    b - a);
    }
  };
  if (nth <= 1) {
    run(0, total);
    return;
  }
  std::vector<std::thread> th;
  const size_t share = (total + nth - 1) / nth;
  for (int k = 1; k < nth; ++k)
    th.emplace_back(run, std::min(total, k * share), std::min(total, (k + 1) * share));
  run(0, std::min(total, share));
  for (auto& t : th) t.join();
  busy -= want;
}

int stage_wait() {
  if (g_stage_pending) {
    HIP_TRY(hipEventSynchronize(g_stage_done));
    g_stage_pending = false;
  }
  return IER_OK;
}

}  // namespace

// Transfers whose staging would exceed this (or that a stream capture
// records) go as plain copies from and to the callers' memory.
constexpr size_t kStageMax = size_t(1) << 30;

int h2d_staged(hipStream_t st, const HostXfer* x, int n) {
  std::vector<size_t> at(n + 1, 0);
  for (int i = 0; i < n; ++i) at[i + 1] = at[i] + ((x[i].bytes + 255) & ~size_t(255));
  if (!at[n]) return IER_OK;
  char* stage = nullptr;
  if (!capturing() && at[n] <= kStageMax) {
    int e = stage_wait();
    if (e) return e;
    stage = static_cast<char*>(pinned(at[n], 0));
  }
  if (!stage) {  // plain copies (no staging, or none to be had)
    for (int i = 0; i < n; ++i)
      if (x[i].bytes) HIP_TRY(hipMemcpyAsync(x[i].dst, x[i].src, x[i].bytes, hipMemcpyHostToDevice, st));
    return IER_OK;
  }
  par_copy(x, n, false, stage, at.data());
  for (int i = 0; i < n; ++i)
    if (x[i].bytes) HIP_TRY(hipMemcpyAsync(x[i].dst, stage + at[i], x[i].bytes, hipMemcpyHostToDevice, st));
  if (!g_stage_done) HIP_TRY(hipEventCreateWithFlags(&g_stage_done, hipEventDisableTiming));
  HIP_TRY(hipEventRecord(g_stage_done, st));
  g_stage_pending = true;
  return IER_OK;
}

int d2h_staged(hipStream_t st, const HostXfer* x, int n) {
  std::vector<size_t> at(n + 1, 0);
  for (int i = 0; i < n; ++i) at[i + 1] = at[i] + ((x[i].bytes + 255) & ~size_t(255));
  if (!at[n]) return IER_OK;
  char* stage = nullptr;
  if (!capturing() && at[n] <= kStageMax) {
    int e = stage_wait();
    if (e) return e;
    stage = static_cast<char*>(pinned(at[n], 0));
  }
  if (!stage) {  // plain copies (no staging, or none to be had)
    for (int i = 0; i < n; ++i)
      if (x[i].bytes) HIP_TRY(hipMemcpyAsync(x[i].dst, x[i].src, x[i].bytes, hipMemcpyDeviceToHost, st));
    if (!capturing()) HIP_TRY(hipStreamSynchronize(st));
    return IER_OK;
  }
  for (int i = 0; i < n; ++i)
    if (x[i].bytes) HIP_TRY(hipMemcpyAsync(stage + at[i], x[i].src, x[i].bytes, hipMemcpyDeviceToHost, st));
  HIP_TRY(hipStreamSynchronize(st));
  par_copy(x, n, true, stage, at.data());
  return IER_OK;
}

void* pinned(size_t bytes, int slot) {
  void** buf = g_pinned.buf;
  size_t* cap = g_pinned.cap;
  if (slot < 0 || slot >= 8) return nullptr;
  if (bytes > cap[slot]) {
    if (buf[slot]) (void)hipHostFree(buf[slot]);
    size_t want = std::max(bytes, cap[slot] * 3 / 2);
    if (hipHostMalloc(&buf[slot], want, hipHostMallocDefault) != hipSuccess) {
      buf[slot] = nullptr;
      cap[slot] = 0;
      return nullptr;
    }
    cap[slot] = want;
  }
  return buf[slot];
}

namespace {

constexpr int kBB = 256;

__global__ void k_bbox_partial(const double2* a, int na, const double2* b, int nb,
                               double4* part) {
  __shared__ double4 sh[kBB];
  double4 m = make_double4(INFINITY, -INFINITY, INFINITY, -INFINITY);
  int n = na + nb;
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    double2 p = i < na ? a[i] : b[i - na];
    m.x = fmin(m.x, p.x); m.y = fmax(m.y, p.x);
    m.z = fmin(m.z, p.y); m.w = fmax(m.w, p.y);
  }
  sh[threadIdx.x] = m;
  __syncthreads();
  for (int s = kBB / 2; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) {
      double4 o = sh[threadIdx.x + s];
      m.x = fmin(m.x, o.x); m.y = fmax(m.y, o.y);
      m.z = fmin(m.z, o.z); m.w = fmax(m.w, o.w);
      sh[threadIdx.x] = m;
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) part[blockIdx.x] = sh[0];
}

// the partial boxes reduced (min and max are exact in any order) to the
// root box: geo = {centre x, centre y, size}, as pts_tree2d.c:493-516
__global__ void __launch_bounds__(1024) k_bbox_final(const double4* part, int nbk, double* geo) {
  __shared__ double4 sh[1024];
  double4 m = make_double4(INFINITY, -INFINITY, INFINITY, -INFINITY);
  for (int i = threadIdx.x; i < nbk; i += blockDim.x) {
    const double4 o = part[i];
    m.x = fmin(m.x, o.x); m.y = fmax(m.y, o.y);
    m.z = fmin(m.z, o.z); m.w = fmax(m.w, o.w);
  }
  sh[threadIdx.x] = m;
  __syncthreads();
  for (int s = blockDim.x / 2; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) {
      const double4 o = sh[threadIdx.x + s];
      m.x = fmin(m.x, o.x); m.y = fmax(m.y, o.y);
      m.z = fmin(m.z, o.z); m.w = fmax(m.w, o.w);
      sh[threadIdx.x] = m;
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    double bs0 = m.y - m.x;
    if (m.w - m.z > bs0) bs0 = m.w - m.z;
    geo[0] = (m.x + m.y) / 2;
    geo[1] = (m.z + m.w) / 2;
    geo[2] = bs0;
  }
}

// Keys of both point sets in one array, targets flagged just above the
// 2 kMaxDepth key bits, so that one sort leaves the sources and then the
// targets, each in key order; idx is the index within the point's own set.
// geo = {centre x, centre y, size} of the root box, on the device.
constexpr uint64_t kTargetFlag = uint64_t(1) << (2 * kMaxDepth);
__global__ void k_keys(const double2* s, int ns, const double2* t, int nt, const double* geo,
                       uint64_t* key, int* idx) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= ns + nt) return;
  const bool tg = i >= ns;
  const double2 q = tg ? t[i - ns] : s[i];
  key[i] = point_key(q.x, q.y, geo[0], geo[1], geo[2]) | (tg ? kTargetFlag : 0);
  idx[i] = tg ? i - ns : i;
}
// Keys of a forest's points: a point of problem p takes the path key of its
// own root box (cx, cy, R) -- the key it has alone -- below the Morton key
// of p's root cell (rx, ry) at level m; with a != 0 the point's normalised
// coordinates (x - cx) a go to norm (indexed as the input).
__device__ inline int forest_prob_of(const int* off, int np, int i) {
  int lo = 0, hi = np;  // largest p with off[p] <= i
  while (hi - lo > 1) {
    const int mid = (lo + hi) >> 1;
    if (off[mid] <= i) lo = mid;
    else hi = mid;
  }
  return lo;
}
__global__ void k_keys_forest(const double2* s, int ns, const double2* t, int nt,
                              const ForestProb* prob, const int* soff, const int* toff, int np,
                              uint64_t* key, int* idx, double2* s_norm, double2* t_norm) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= ns + nt) return;
  const bool tg = i >= ns;
  const int j = tg ? i - ns : i;
  const ForestProb P = prob[forest_prob_of(tg ? toff : soff, np, j)];
  const double2 q = tg ? t[j] : s[j];
  const uint64_t local = point_key(q.x, q.y, P.cx, P.cy, P.R);
  uint64_t root = 0;
  for (int b = 0; b < P.m; ++b)
    root |= ((uint64_t)((P.rx >> b) & 1) << (2 * b)) | ((uint64_t)((P.ry >> b) & 1) << (2 * b + 1));
  key[i] = (root << (2 * (kMaxDepth - P.m))) | (local >> (2 * P.m)) | (tg ? kTargetFlag : 0);
  idx[i] = j;
  double2* norm = tg ? t_norm : s_norm;
  if (norm && P.a != 0.0) norm[j] = make_double2((q.x - P.cx) * P.a, (q.y - P.cy) * P.a);
}

// Box centres and sizes of a forest in each problem's own kernel
// coordinates: for a box at or below its problem's root, the reference's
// centre recurrence from the problem's root box there (fcx, fcy, fR: the
// centre and size the box has when the problem is solved alone, or its
// normalised image); other boxes keep the unit square's.
__global__ void k_forest_frames(int nb, const uint64_t* key, const int* level, ForestDev F,
                                const ForestProb* prob, double2* ctr, double* bsb) {
  const int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= nb) return;
  const int l = level[b];
  const uint64_t k = key[b];
  ForestSlot sl;
  if (!forest_slot(F, k, l, sl) || l < sl.m) return;
  const ForestProb P = prob[sl.prob];
  double cx = P.fcx, cy = P.fcy, bs = P.fR;
  for (int j = sl.m + 1; j <= l; ++j) {
    const int q = (int)(k >> (2 * (l - j))) & 3;
    bs = bs / 2;
    cx = child_coord(cx, q & 1, bs);
    cy = child_coord(cy, q & 2, bs);
  }
  ctr[b] = make_double2(cx, cy);
  bsb[b] = bs;
}

__global__ void k_unflag(uint64_t* key, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) key[i] &= kTargetFlag - 1;
}

// Inside every leaf, restore original-index order (the reference's
// partitions are stable).  One thread per leaf, insertion sort.
// original indices of each leaf's points in increasing order: a wave per
// leaf; a leaf of at most 64 points ranks its (distinct) indices through
// shuffles, a larger one is sorted in place by one lane
__global__ void k_leaf_sort(int* idx_s, const int2* rng_s, const int* leaves_s, int nls,
                            int* idx_t, const int2* rng_t, const int* leaves_t, int nlt) {
  int w = blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);  // one wave per leaf
  if (w >= nls + nlt) return;
  const bool tg = w >= nls;  // source leaves first, then target leaves
  if (tg) w -= nls;
  int* idx = tg ? idx_t : idx_s;
  const int2* rng = tg ? rng_t : rng_s;
  const int* leaves = tg ? leaves_t : leaves_s;
  const int lane = threadIdx.x & 63;
  const int2 r = rng[leaves[w]];
  const int n = r.y - r.x;
  if (n <= 1) return;
  if (n <= 64) {
    const int v = lane < n ? idx[r.x + lane] : 0x7fffffff;
    int rank = 0;
    for (int j = 0; j < n; ++j) rank += __shfl(v, j) < v ? 1 : 0;
    if (lane < n) idx[r.x + rank] = v;
    return;
  }
  if (lane != 0) return;
  for (int i = r.x + 1; i < r.y; ++i) {
    int v = idx[i], j = i - 1;
    while (j >= r.x && idx[j] > v) {
      idx[j + 1] = idx[j];
      --j;
    }
    idx[j + 1] = v;
  }
}

__global__ void k_gather2(const double2* in, double2* out, const int* perm, int n) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) out[i] = in[perm[i]];
}

__global__ void k_gather_ch(const double2* in, double2* out, const int* perm, int n, int nd,
                            int ncomp) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  int src = perm[i];
  for (int c = 0; c < ncomp; ++c)
    for (int d = 0; d < nd; ++d)
      out[(size_t)(c * nd + d) * n + i] = in[((size_t)src * ncomp + c) * nd + d];
}

// Helmholtz source strengths in one pass over the permutation: charges and
// dipole strengths (complex, nd per point) and dipole vectors (2 nd reals),
// each to channel-major sorted order (as k_gather_ch / k_gather_re)
__global__ void k_gather_src(const double2* chg, double2* chg1, const double2* dip, double2* dip1,
                             const double* dv, double* dv1, const int* perm, int n, int nd) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const size_t s = (size_t)perm[i];
  for (int d = 0; d < nd; ++d) {
    if (chg1) chg1[(size_t)d * n + i] = chg[s * nd + d];
    if (dip1) dip1[(size_t)d * n + i] = dip[s * nd + d];
  }
  if (dv1) {
    if (nd == 1 && (reinterpret_cast<uintptr_t>(dv) & 15) == 0) {  // one 16-byte load
      const double2 v = reinterpret_cast<const double2*>(dv)[s];
      dv1[i] = v.x;
      dv1[(size_t)n + i] = v.y;
    } else {
      for (int c = 0; c < 2 * nd; ++c) dv1[(size_t)c * n + i] = dv[s * 2 * nd + c];
    }
  }
}

__global__ void k_gather_re(const double* in, double* out, const int* perm, int n, int nd,
                            int ncomp) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  int src = perm[i];
  for (int c = 0; c < ncomp; ++c)
    for (int d = 0; d < nd; ++d)
      out[(size_t)(c * nd + d) * n + i] = in[((size_t)src * ncomp + c) * nd + d];
}

__global__ void k_scatter_ch(const double2* in, double2* out, const int* perm, int n, int nd,
                             int ncomp) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  int dst = perm[i];
  for (int c = 0; c < ncomp; ++c)
    for (int d = 0; d < nd; ++d)
      out[((size_t)dst * ncomp + c) * nd + d] = in[(size_t)(c * nd + d) * n + i];
}

inline int nblk(int n, int b) { return (n + b - 1) / b; }

// Keys and key-sorted permutations of both point sets (one stable LSD radix
// sort on key bits [begin_bit, 2 kMaxDepth) and the target flag):
// perm[0 .. ns) for the sources, perm[ns .. ns + nt) for the targets; *ks and
// *kt are the sorted keys of each set (on the device).
int sort_points(Ctx& c, Scratch& S, const double2* src, int ns, const double2* trg, int nt,
                const double* geo, int begin_bit, uint32_t* err, int* perm, uint64_t** ks,
                uint64_t** kt, const ForestIn* fin, double2* s_norm, double2* t_norm) {
  const int n = ns + nt;
  uint64_t* k0 = S.get<uint64_t>(n);
  uint64_t* k1 = S.get<uint64_t>(n);
  int* i0 = S.get<int>(n);
  if (S.err) return S.err;
  *ks = k1;
  *kt = k1 + ns;
  if (n == 0) return IER_OK;
  if (fin)
    k_keys_forest<<<nblk(n, 256), 256, 0, c.st>>>(src, ns, trg, nt, fin->prob, fin->soff,
                                                  fin->toff, fin->np, k0, i0, s_norm, t_norm);
  else
    k_keys<<<nblk(n, 256), 256, 0, c.st>>>(src, ns, trg, nt, geo, k0, i0);
  int e = rsort::radix_sort(c, S, k0, k1, i0, perm, n, begin_bit, 2 * kMaxDepth + 1, err);
  if (e) return e;
  if (nt) k_unflag<<<nblk(nt, 256), 256, 0, c.st>>>(k1 + ns, nt);
  return IER_OK;
}

}  // namespace

void gather_channels(hipStream_t st, const double2* in, double2* out, const int* perm, int n,
                     int nd, int ncomp) {
  if (n > 0) k_gather_ch<<<nblk(n, 256), 256, 0, st>>>(in, out, perm, n, nd, ncomp);
}
void gather_sources(hipStream_t st, const double2* chg, double2* chg1, const double2* dip,
                    double2* dip1, const double* dv, double* dv1, const int* perm, int n,
                    int nd) {
  if (n > 0) k_gather_src<<<nblk(n, 256), 256, 0, st>>>(chg, chg1, dip, dip1, dv, dv1, perm, n, nd);
}
void gather_real(hipStream_t st, const double* in, double* out, const int* perm, int n, int nd,
                 int ncomp) {
  if (n > 0) k_gather_re<<<nblk(n, 256), 256, 0, st>>>(in, out, perm, n, nd, ncomp);
}
void scatter_channels(hipStream_t st, const double2* in, double2* out, const int* perm, int n,
                      int nd, int ncomp) {
  if (n > 0) k_scatter_ch<<<nblk(n, 256), 256, 0, st>>>(in, out, perm, n, nd, ncomp);
}

int build_plan(Ctx& c, Scratch& S, const double* d_src, int ns, const double* d_trg, int nt,
               int idivflag, int ndiv, SortedPoints& P, Plan& plan, bool want_lap_work,
               int ifpgh, int ifpghtarg, const ForestIn* fin) {
  (void)want_lap_work;
  const double2* src = reinterpret_cast<const double2*>(d_src);
  const double2* trg = reinterpret_cast<const double2*>(d_trg);
  P.ns = ns;
  P.nt = nt;
  last_stats() = SolveStats();
  double th[7];
  th[0] = wall();
  // The root box stays on the device for the keys; its host copy is valid
  // after the tree's first read-back synchronises the stream.
  double* d_geo = S.get<double>(3);
  double* h_geo = static_cast<double*>(pinned(sizeof(double) * 3, 2));
  if (S.err) return S.err;
  if (!h_geo) return IER_ALLOC;
  // a forest's normalised coordinates (input order), when its kernels use them
  double2* s_norm = nullptr;
  double2* t_norm = nullptr;
  if (fin) {
    // ---- a forest: the unit square holds every problem's root cell
    h_geo[0] = 0.5;
    h_geo[1] = 0.5;
    h_geo[2] = 1.0;
    if (fin->normalise) {
      s_norm = S.get<double2>(ns);
      t_norm = S.get<double2>(nt);
      if (S.err) return S.err;
    }
  } else {
    // ---- bounding box of sources and targets (pts_tree2d.c:493-516)
    int nbk = std::min(1024, std::max(1, nblk(ns + nt, kBB)));
    double4* part = S.get<double4>(nbk);
    if (S.err) return S.err;
    k_bbox_partial<<<nbk, kBB, 0, c.st>>>(src, ns, trg, nt, part);
    k_bbox_final<<<1, 1024, 0, c.st>>>(part, nbk, d_geo);
    HIP_TRY(hipMemcpyAsync(h_geo, d_geo, sizeof(double) * 3, hipMemcpyDeviceToHost, c.st));
  }
  th[1] = wall();

  // ---- keys + sort.  Boxes down to level D only need the keys ordered by
  // their top 2D bits (box ranges and the depth search compare prefixes;
  // order inside a leaf is restored below), so the sort first covers
  // kSortLevels levels and is redone on every bit when the tree turns out
  // deeper.
  plan.sort_err = static_cast<uint32_t*>(pinned(sizeof(uint32_t), 6));
  if (!plan.sort_err) return IER_ALLOC;
  *plan.sort_err = 0;  // the previous solve on this thread has synchronised
  P.perm_s = S.get<int>((size_t)ns + nt);
  if (S.err) return S.err;
  P.perm_t = P.perm_s + ns;
  double tt[3] = {0, 0, 0};
  int e = IER_OK;
  // (a forest's trees sit below its root cells: the forest's estimate)
  for (int levels : {fin ? std::min(fin->sort_levels, kMaxDepth) : kSortLevels, kMaxDepth}) {
    const int begin_bit = 2 * (kMaxDepth - levels);
    uint64_t *ks = nullptr, *kt = nullptr;
    e = sort_points(c, S, src, ns, trg, nt, d_geo, begin_bit, plan.sort_err, P.perm_s, &ks,
                    &kt, fin, s_norm, t_norm);
    if (e) return e;
    th[2] = wall();
    // ---- tree, lists and work lists on the device
    e = build_dtree(c, S, ks, ns, kt, nt, h_geo, idivflag, ndiv, ifpgh, ifpghtarg,
                    levels, plan, tt, fin ? &fin->dev : nullptr, fin ? fin->lfirst - 2 : 0);
    if (e != IER_RESORT) break;
  }
  if (e) return e;
  P.c0x = h_geo[0];  // the tree's read-backs have synchronised the stream
  P.c0y = h_geo[1];
  P.bs0 = h_geo[2];
  if (fin) {
    plan.lfirst = fin->lfirst;
    if (fin->frames) {
      P.bs0 = fin->min_bs0;
      k_forest_frames<<<nblk(plan.nb, 256), 256, 0, c.st>>>(
          plan.nb, plan.D.key, plan.D.level, fin->dev, fin->prob,
          const_cast<double2*>(plan.D.ctr), const_cast<double*>(plan.D.bsb));
    }
    if (s_norm) src = s_norm;
    if (t_norm) trg = t_norm;
  }
  th[3] = th[2] + tt[0];
  th[4] = th[3] + tt[1];
  const DTree& D = plan.D;
  th[5] = wall();
  const Span lss = plan.leaves_s.all(), lst = plan.leaves_t.all();
  // ---- original-index order inside leaves, then sorted coordinates
  if (lss.n + lst.n)
    k_leaf_sort<<<nblk(lss.n + lst.n, 4), 256, 0, c.st>>>(
        P.perm_s, D.srng, plan.dwork + lss.off, lss.n, P.perm_t, D.trng, plan.dwork + lst.off,
        lst.n);
  P.src = S.get<double2>(ns);
  P.trg = S.get<double2>(nt);
  if (S.err) return S.err;
  if (ns) k_gather2<<<nblk(ns, 256), 256, 0, c.st>>>(src, P.src, P.perm_s, ns);
  if (nt) k_gather2<<<nblk(nt, 256), 256, 0, c.st>>>(trg, P.trg, P.perm_t, nt);
  HIP_TRY(hipGetLastError());
  th[6] = wall();
  SolveStats& st = last_stats();
  for (int k = 0; k < 6; ++k) st.host_s[k] = th[k + 1] - th[k];
  return IER_OK;
}

}  // namespace fmm

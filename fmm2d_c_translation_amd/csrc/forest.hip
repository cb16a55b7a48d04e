// forest.hip -- fused batches: many independent problems of one kind solved
// as one tree, so that a whole batch runs as one level-synchronous launch
// sequence (one sort, one tree, one launch per FMM step and level) instead
// of one solve, with its host round trips and ~100 launches, per problem.
//
// Layout.  The problems of a group get slots, the cells of level M of the
// unit square (4^M >= group size).  Problem i's root box becomes the cell
// (rx, ry) = (slot corner) + (1, 1) of level m >= M + 2, so the eight cells
// around it at level m lie in its slot and hold no points.  A point's key
// is its problem's root-cell key followed by the key the point has when the
// problem is solved alone (ForestProb, points.hip), so below its root every
// problem has exactly its own tree: refinement counts its own points, and
// level restriction runs inside its root (dtree.hip).  Lists of a box below
// a root stay in its problem or reach only empty cells; the cells at or
// above the roots do no FMM work (their own level in the problem, below 2).
//
// Kernels.  Each problem keeps coordinates of its own (box centres from
// the centre recurrence of its root box, k_forest_frames), so differences of
// points and centres keep full relative precision however deep its root
// cell lies.  Laplace (Cauchy, log and real kernels) is solved in the
// problem's input coordinates: every root sits at one level m, centres and
// box sizes are those of the solve alone, and only M2L's log(box size) term
// is added per target box.  So a problem's result equals its solve alone up
// to rounding.  Helmholtz needs one wavenumber per level: each problem is
// scaled by a = zk / K about its root centre, K common, its root box
// enlarged from the points' box b0 to R in [b0, 2 b0) with K R a = zk R =
// K 2^-m.  Box sizes at a level are then the same number of wavelengths in
// every problem, and the operators of a level are shared.  Dipole strengths are multiplied by a
// before the solve, gradients by a and Hessians by a^2 after it; the
// potential is unchanged.  The enlarged root gives a different tree from
// the solve alone, so the two agree within the requested accuracy.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <map>
#include <memory>
#include <tuple>

#include "engine.h"

namespace fmm {
namespace {

inline int nblk(long n, int b) { return (int)((n + b - 1) / b); }

// Root box of every problem (pts_tree2d.c:493-516): {centre x, centre y,
// size, unused}.  The points of problem p (its sources, then its targets)
// are cut into chunks of kBBChunk, one block each (blocks pfirst[p] to
// pfirst[p+1] - 1), so a large problem does not hold the launch up; a
// second kernel takes the extremes over each problem's chunks.
constexpr int kBBChunk = 16384;

__device__ inline void bb_add(double4& m, double2 q) {
  m.x = fmin(m.x, q.x); m.y = fmax(m.y, q.x); m.z = fmin(m.z, q.y); m.w = fmax(m.w, q.y);
}
__device__ inline void bb_merge(double4& m, const double4& o) {
  m.x = fmin(m.x, o.x); m.y = fmax(m.y, o.y); m.z = fmin(m.z, o.z); m.w = fmax(m.w, o.w);
}

__global__ void __launch_bounds__(256) k_bbox_parts(const double2* src, const double2* trg,
                                                    const int* soff, const int* toff,
                                                    const int* pfirst, int np, double4* part) {
  const int b = blockIdx.x;
  int lo = 0, hi = np;  // the problem whose chunks hold block b
  while (hi - lo > 1) {
    const int mid = (lo + hi) >> 1;
    if (pfirst[mid] <= b) lo = mid;
    else hi = mid;
  }
  const int p = lo;
  const long s0 = soff[p], ns = soff[p + 1] - s0, t0 = toff[p];
  const long n = ns + (toff[p + 1] - t0);
  const long i0 = (long)(b - pfirst[p]) * kBBChunk, i1 = min(n, i0 + kBBChunk);
  double4 m = make_double4(INFINITY, -INFINITY, INFINITY, -INFINITY);
  for (long i = i0 + threadIdx.x; i < i1; i += blockDim.x)
    bb_add(m, i < ns ? src[s0 + i] : trg[t0 + (i - ns)]);
  __shared__ double4 sh[256];
  sh[threadIdx.x] = m;
  __syncthreads();
  for (int st = 128; st > 0; st >>= 1) {
    if ((int)threadIdx.x < st) {
      bb_merge(m, sh[threadIdx.x + st]);
      sh[threadIdx.x] = m;
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) part[b] = m;
}

__global__ void __launch_bounds__(64) k_bbox_probs(const double4* part, const int* pfirst,
                                                   double4* out) {
  const int p = blockIdx.x;
  double4 m = make_double4(INFINITY, -INFINITY, INFINITY, -INFINITY);
  for (int j = pfirst[p] + (int)threadIdx.x; j < pfirst[p + 1]; j += 64) bb_merge(m, part[j]);
  for (int o = 32; o > 0; o >>= 1) {
    double4 y;
    y.x = __shfl_xor(m.x, o); y.y = __shfl_xor(m.y, o);
    y.z = __shfl_xor(m.z, o); y.w = __shfl_xor(m.w, o);
    bb_merge(m, y);
  }
  if (threadIdx.x == 0) {
    double bs0 = m.y - m.x;
    if (m.w - m.z > bs0) bs0 = m.w - m.z;
    out[p] = make_double4((m.x + m.y) / 2, (m.z + m.w) / 2, bs0, 0.0);
  }
}

// Copies of many arrays in one launch: segment blockIdx.y, 8-byte words.
struct Seg {
  const double* src;
  double* dst;
  long long n;  // words
};
__global__ void k_copy_segs(const Seg* seg) {
  const Seg s = seg[blockIdx.y];
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < s.n;
       i += (long long)gridDim.x * blockDim.x)
    s.dst[i] = s.src[i];
}

__device__ inline int prob_of(const int* off, int np, int i) {
  int lo = 0, hi = np;  // largest p with off[p] <= i
  while (hi - lo > 1) {
    const int mid = (lo + hi) >> 1;
    if (off[mid] <= i) lo = mid;
    else hi = mid;
  }
  return lo;
}

// v[i * w + k] *= scale(problem of i)^pw for every point i (w complex values
// per point), scale from the problems' normalisation factors.
__global__ void k_scale_points(double2* v, int n, int w, const int* off, int np,
                               const ForestProb* prob, int pw) {
  const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (long long)n * w) return;
  const int i = (int)(t / w);
  const double a = prob[prob_of(off, np, i)].a;
  const double f = pw == 2 ? a * a : a;
  double2 x = v[t];
  x.x *= f;
  x.y *= f;
  v[t] = x;
}

struct GroupKey {
  int kernel, nd, ifc, ifd, pg, pgt, hastarg, ndiv, prec, dev;
  double eps;
  bool operator<(const GroupKey& o) const {
    return std::tie(kernel, nd, ifc, ifd, pg, pgt, hastarg, ndiv, prec, dev, eps) <
           std::tie(o.kernel, o.nd, o.ifc, o.ifd, o.pg, o.pgt, o.hastarg, o.ndiv, o.prec, o.dev,
                    o.eps);
  }
};

int leaf_size(const fmm2dhip_problem& p) {
  return p.ndiv > 0 ? p.ndiv : ndiv_rule(p.eps, p.ns, p.nt);
}

// a forest's problems have their leaf size's worth of sources at least once
// over (their root is split, so no point of theirs sits in a shared cell)
bool eligible(const fmm2dhip_problem& p) {
  if (p.kernel < FMM2DHIP_CAUCHY || p.kernel > FMM2DHIP_REAL) return false;
  if (p.nd < 1 || p.ns < 1 || p.nt < 0 || p.ifpgh < 0 || p.ifpgh > 3 || p.ifpghtarg < 0 ||
      p.ifpghtarg > 3)
    return false;
  if (!p.ifcharge && !p.ifdipole) return false;
  if (p.device_ptrs != 0 && p.device_ptrs != 1) return false;
  if (p.kernel == FMM2DHIP_HELMHOLTZ && !(p.zk_im == 0.0 && p.zk_re > 0.0)) return false;
  if (!(p.eps > 0.0)) return false;
  return p.ns > leaf_size(p);
}

// Bytes per point of each array in the problem layouts (include/fmm2d_hip.h).
struct Widths {
  size_t chg, dip, dv, pot, grad, hess;
};
Widths widths(const fmm2dhip_problem& p) {
  const size_t nd = p.nd;
  const size_t e = p.kernel == FMM2DHIP_REAL ? sizeof(double) : sizeof(double2);
  const bool split = p.kernel != FMM2DHIP_CAUCHY;  // gradient (nd,2,n), Hessian (nd,3,n)
  Widths w;
  w.chg = nd * e;
  w.dip = nd * e;
  w.dv = p.kernel == FMM2DHIP_CAUCHY ? 0 : 2 * nd * sizeof(double);
  w.pot = nd * e;
  w.grad = (split ? 2 : 1) * nd * e;
  w.hess = (split ? 3 : 1) * nd * e;
  return w;
}

// The largest Helmholtz zk * (root box) a forest takes: its boxes two levels
// below a root (R <= 2 b0, so at most b0 / 2) stay within 16 wavelengths,
// where the low-frequency translations apply.
constexpr double kMaxZkB0 = 150.0;
// Deepest root level: the problem keeps 29 - m levels of its own.
constexpr int kMaxRootLevel = 16;

class Copier {
 public:
  Copier(hipStream_t st, bool device) : st_(st), device_(device) {}
  // dst <- src, one of them the caller's array (in device memory, or in
  // host memory: then staged through pinned memory, h2d_staged and
  // d2h_staged); to_host: dst is the caller's
  int add(const void* src, void* dst, size_t bytes, bool to_host) {
    if (!bytes) return IER_OK;
    if (device_) {
      segs_.push_back({static_cast<const double*>(src), static_cast<double*>(dst),
                       (long long)(bytes / sizeof(double))});
      return IER_OK;
    }
    (to_host ? down_ : up_).push_back(HostXfer{src, dst, bytes});
    return IER_OK;
  }
  int flush(Scratch& S) {
    int e = IER_OK;
    if (!up_.empty() && (e = h2d_staged(st_, up_.data(), (int)up_.size()))) return e;
    if (!down_.empty() && (e = d2h_staged(st_, down_.data(), (int)down_.size()))) return e;
    up_.clear();
    down_.clear();
    if (segs_.empty()) return IER_OK;
    for (size_t i0 = 0; i0 < segs_.size(); i0 += 65535) {
      const size_t n = std::min<size_t>(65535, segs_.size() - i0);
      // the pinned slot is reused: its previous copy has to have landed
      HIP_TRY(hipStreamSynchronize(st_));
      Seg* h = static_cast<Seg*>(pinned(sizeof(Seg) * n, 7));
      Seg* d = S.get<Seg>(n);
      if (!h) return IER_ALLOC;
      if (S.err) return S.err;
      std::memcpy(h, segs_.data() + i0, sizeof(Seg) * n);
      HIP_TRY(hipMemcpyAsync(d, h, sizeof(Seg) * n, hipMemcpyHostToDevice, st_));
      k_copy_segs<<<dim3(64, (unsigned)n), 256, 0, st_>>>(d);
    }
    segs_.clear();
    HIP_TRY(hipGetLastError());
    return IER_OK;
  }

 private:
  hipStream_t st_;
  bool device_;
  std::vector<Seg> segs_;
  std::vector<HostXfer> up_, down_;
};

// k_copy_segs over a device table of n segments, in launches of at most
// 65535 (stream work only: graph captures record it).
void copy_segs(hipStream_t st, const Seg* d, size_t n) {
  for (size_t i0 = 0; i0 < n; i0 += 65535)
    k_copy_segs<<<dim3(64, (unsigned)std::min<size_t>(65535, n - i0)), 256, 0, st>>>(d + i0);
}

}  // namespace

// One unit of a plan: a forest, or a problem alone, whose front end ran once
// and whose evaluation -- strengths copied in, the FMM, outputs copied out --
// is a captured graph, replayed by every execute.  It owns the device memory
// the graph reads and writes (detached from the thread's workspace) and the
// pinned pool holding the host data its uploads read (4 MiB).
struct PlanUnit {
  std::vector<int> members;
  bool helm = false;
  LapReq lr;
  HelmReq hr;
  ForestIn fin;
  Prepared q;
  HostKeeper keeper;
  std::vector<void*> blocks;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  ~PlanUnit() {
    if (exec) (void)hipGraphExecDestroy(exec);
    if (graph) (void)hipGraphDestroy(graph);
    for (void* b : blocks) (void)hipFree(b);
  }
};

namespace {

// Captures pre(); the unit's evaluation; post() into unit->exec.  The
// capture is always ended, also after an error, so the stream stays usable.
template <class Pre, class Post>
int capture_unit(PlanUnit* u, Scratch& S, Pre pre, Post post) {
  const hipStream_t st = u->q.c.st;
  // the host data of the captured uploads (level orders, operator jobs)
  if (int e = u->keeper.reserve(size_t(4) << 20)) return e;
  HIP_TRY(hipStreamSynchronize(st));
  HIP_TRY(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
  int e = IER_OK;
  {
    CaptureScope cs(&u->keeper);
    pre();
    e = u->helm ? helm_eval(u->hr, S, u->q) : lap_eval(u->lr, S, u->q);
    if (!e) post();
  }
  hipGraph_t g = nullptr;
  const hipError_t ce = hipStreamEndCapture(st, &g);
  u->graph = g;
  if (e) return e;
  if (ce != hipSuccess || !g) return IER_HIP;
  HIP_TRY(hipGraphInstantiate(&u->exec, g, nullptr, nullptr, 0));
  return IER_OK;
}

int run_forest(fmm2dhip_problem* probs, const std::vector<int>& g, std::vector<int>& alone,
               PlanUnit* unit) {
  Ctx c;
  int e = get_ctx(c);
  if (e) return e;
  Scratch S(c.st);
  const fmm2dhip_problem& p0 = probs[g[0]];
  const int np = (int)g.size();
  const bool helm = p0.kernel == FMM2DHIP_HELMHOLTZ;
  const bool dev = p0.device_ptrs == 1;
  const Widths W = widths(p0);
  const int ndiv = leaf_size(p0);
  std::vector<int> soff(np + 1, 0), toff(np + 1, 0);
  for (int i = 0; i < np; ++i) {
    soff[i + 1] = soff[i] + probs[g[i]].ns;
    toff[i + 1] = toff[i] + probs[g[i]].nt;
  }
  const long long NSl = soff[np], NTl = toff[np];
  if (NSl + NTl >= (1LL << 30)) return IER_ARGS;
  const int NS = (int)NSl, NT = (int)NTl;
  // ---- concatenated inputs
  auto get = [&](size_t bytes) { return S.get<char>(std::max<size_t>(bytes, 16)); };
  char* src = get(16 * (size_t)NS);
  char* trg = get(16 * (size_t)NT);
  char* chg = p0.ifcharge ? get(W.chg * NS) : nullptr;
  char* dip = p0.ifdipole ? get(W.dip * NS) : nullptr;
  char* dv = p0.ifdipole && W.dv ? get(W.dv * NS) : nullptr;
  // chunks of the root-box search: blocks pfirst[p] .. pfirst[p + 1] - 1
  std::vector<int> pfirst(np + 1, 0);
  for (int i = 0; i < np; ++i) {
    const long long n = (long long)probs[g[i]].ns + probs[g[i]].nt;
    pfirst[i + 1] = pfirst[i] + (int)std::max<long long>(1, (n + kBBChunk - 1) / kBBChunk);
  }
  int* d_off = S.get<int>(3 * (size_t)(np + 1));
  double4* d_bb = S.get<double4>(np);
  double4* d_bbp = S.get<double4>(pfirst[np]);
  if (S.err) return S.err;
  Copier in(c.st, dev);
  for (int i = 0; i < np; ++i) {
    const fmm2dhip_problem& p = probs[g[i]];
    const size_t s = soff[i], t = toff[i];
    if ((e = in.add(p.sources, src + 16 * s, 16 * (size_t)p.ns, false))) return e;
    if (p.nt && (e = in.add(p.targ, trg + 16 * t, 16 * (size_t)p.nt, false))) return e;
    if (chg && (e = in.add(p.charge, chg + W.chg * s, W.chg * p.ns, false))) return e;
    if (dip && (e = in.add(p.dipstr, dip + W.dip * s, W.dip * p.ns, false))) return e;
    if (dv && (e = in.add(p.dipvec, dv + W.dv * s, W.dv * p.ns, false))) return e;
  }
  if ((e = in.flush(S))) return e;
  {
    int* h = static_cast<int*>(pinned(sizeof(int) * 3 * (np + 1), 4));
    if (!h) return IER_ALLOC;
    std::memcpy(h, soff.data(), sizeof(int) * (np + 1));
    std::memcpy(h + np + 1, toff.data(), sizeof(int) * (np + 1));
    std::memcpy(h + 2 * (np + 1), pfirst.data(), sizeof(int) * (np + 1));
    HIP_TRY(hipMemcpyAsync(d_off, h, sizeof(int) * 3 * (np + 1), hipMemcpyHostToDevice, c.st));
  }
  const int* d_soff = d_off;
  const int* d_toff = d_off + np + 1;
  const int* d_pfirst = d_off + 2 * (np + 1);
  k_bbox_parts<<<pfirst[np], 256, 0, c.st>>>(reinterpret_cast<const double2*>(src),
                                             reinterpret_cast<const double2*>(trg), d_soff, d_toff,
                                             d_pfirst, np, d_bbp);
  k_bbox_probs<<<np, 64, 0, c.st>>>(d_bbp, d_pfirst, d_bb);
  std::vector<double4> bb(np);
  {
    double4* h = static_cast<double4*>(pinned(sizeof(double4) * np, 3));
    if (!h) return IER_ALLOC;
    HIP_TRY(hipMemcpyAsync(h, d_bb, sizeof(double4) * np, hipMemcpyDeviceToHost, c.st));
    HIP_TRY(hipStreamSynchronize(c.st));
    std::memcpy(bb.data(), h, sizeof(double4) * np);
  }
  // ---- layout: slots, root levels and cells, scales
  std::vector<char> ok(np, 1);
  double zkb_max = 0.0;
  for (int i = 0; i < np; ++i) {
    const double b0 = bb[i].z;
    if (!(b0 > 0.0) || !std::isfinite(b0)) ok[i] = 0;
    if (helm && ok[i]) {
      const double zb = probs[g[i]].zk_re * b0;
      if (zb > kMaxZkB0) ok[i] = 0;
      else zkb_max = std::max(zkb_max, zb);
    }
  }
  int nok = 0;
  for (int i = 0; i < np; ++i) nok += ok[i];
  if (nok < np) {
    // members unfit for the forest are solved alone; the rest again as one
    std::vector<int> rest;
    for (int i = 0; i < np; ++i) (ok[i] ? rest : alone).push_back(g[i]);
    if (rest.size() >= 2) return run_forest(probs, rest, alone, unit);
    for (int i : rest) alone.push_back(i);
    return IER_OK;
  }
  int M = 0;
  while ((1LL << (2 * M)) < np) ++M;
  const double K = helm ? std::ldexp(zkb_max, M + 2) * (1.0 + 0x1p-20) : 0.0;
  std::vector<ForestProb> fp(np);
  std::vector<ForestSlot> slots((size_t)1 << (2 * M), ForestSlot{-1, 0, 0, -1});
  int mmin = 64;
  double min_b0 = INFINITY;
  for (int i = 0; i < np; ++i) {
    const fmm2dhip_problem& p = probs[g[i]];
    ForestProb& f = fp[i];
    const double b0 = bb[i].z;
    f.cx = bb[i].x;
    f.cy = bb[i].y;
    int m = M + 2;
    double R = b0;
    if (helm) {
      // the largest m with R = K 2^-m / zk >= b0 (root at least the points' box)
      const double zk = p.zk_re;
      m = std::ilogb(K / (zk * b0));
      R = std::ldexp(K / zk, -m);
      while (R < b0 * (1.0 + 0x1p-30)) {
        --m;
        R *= 2;
      }
      while (R / 2 >= b0 * (1.0 + 0x1p-30)) {
        ++m;
        R /= 2;
      }
      if (m < M + 2) return IER_HIP;  // K was chosen so that this cannot happen
      if (m > kMaxRootLevel) {
        alone.push_back(g[i]);  // the deepest roots leave too few levels
        std::vector<int> rest;
        for (int j = 0; j < np; ++j)
          if (j != i) rest.push_back(g[j]);
        if (rest.size() >= 2) return run_forest(probs, rest, alone, unit);
        for (int j : rest) alone.push_back(j);
        return IER_OK;
      }
    }
    f.R = R;
    f.m = m;
    auto compact = [](uint32_t v) {
      v &= 0x55555555u;
      v = (v | (v >> 1)) & 0x33333333u;
      v = (v | (v >> 2)) & 0x0f0f0f0fu;
      v = (v | (v >> 4)) & 0x00ff00ffu;
      v = (v | (v >> 8)) & 0x0000ffffu;
      return (int)v;
    };
    const int cx = compact((uint32_t)i), cy = compact((uint32_t)i >> 1);
    f.rx = (cx << (m - M)) + 1;
    f.ry = (cy << (m - M)) + 1;
    // kernel coordinates: Laplace the problem's own; Helmholtz scaled by
    // a about the root centre, so the root is the cell's size at the origin
    const double cell = std::ldexp(1.0, -m);
    f.a = helm ? cell / R : 0.0;
    f.fcx = helm ? 0.0 : f.cx;
    f.fcy = helm ? 0.0 : f.cy;
    f.fR = helm ? cell : R;
    f.soff = soff[i];
    f.toff = toff[i];
    slots[i] = ForestSlot{m, f.rx, f.ry, i};
    mmin = std::min(mmin, m);
    min_b0 = std::min(min_b0, helm ? cell : b0);
  }
  ForestProb* d_fp = S.get<ForestProb>(np);
  ForestSlot* d_slots = S.get<ForestSlot>(slots.size());
  if (S.err) return S.err;
  {
    const size_t b1 = sizeof(ForestProb) * np, b2 = sizeof(ForestSlot) * slots.size();
    char* h = static_cast<char*>(pinned(b1 + b2, 3));
    if (!h) return IER_ALLOC;
    std::memcpy(h, fp.data(), b1);
    std::memcpy(h + b1, slots.data(), b2);
    HIP_TRY(hipMemcpyAsync(d_fp, h, b1, hipMemcpyHostToDevice, c.st));
    HIP_TRY(hipMemcpyAsync(d_slots, h + b1, b2, hipMemcpyHostToDevice, c.st));
  }
  ForestIn fin;
  fin.np = np;
  fin.prob = d_fp;
  fin.soff = d_soff;
  fin.toff = d_toff;
  fin.normalise = helm;
  fin.frames = true;
  fin.min_bs0 = min_b0;
  fin.lfirst = mmin + 2;
  // depth estimate for the first point sort: each problem's root level,
  // plus the levels a uniform tree of its points would have, plus five for
  // clustering; rounded up to what the same number of 8-bit passes covers
  // (2 levels + the target flag: 15, 19, 23, 27 levels in 4..7 passes).  A
  // deeper tree sorts again on every bit.
  {
    int est = 0;
    for (int i = 0; i < np; ++i) {
      const double n = (double)probs[g[i]].ns + (double)probs[g[i]].nt;
      const int lu = (int)std::ceil(std::log(std::max(n / std::max(ndiv, 1), 1.0)) / std::log(4.0));
      est = std::max(est, fp[i].m + lu + 1 + 5);
    }
    fin.sort_levels = est <= 15 ? 15 : est <= 19 ? 19 : est <= 23 ? 23 : est <= 27 ? 27 : kMaxDepth;
  }
  fin.dev.M = M;
  fin.dev.slot = d_slots;
  // dipole strengths of normalised problems times a
  if (helm && dip)
    k_scale_points<<<nblk((long)NS * p0.nd, 256), 256, 0, c.st>>>(
        reinterpret_cast<double2*>(dip), NS, p0.nd, d_soff, np, d_fp, 1);
  // ---- outputs of the concatenated problem
  fmm2dhip_problem q = p0;
  q.ns = NS;
  q.nt = NT;
  q.sources = reinterpret_cast<const double*>(src);
  q.targ = reinterpret_cast<const double*>(trg);
  q.charge = chg;
  q.dipstr = dip;
  q.dipvec = reinterpret_cast<const double*>(dv);
  q.device_ptrs = 1;
  q.ndiv = ndiv;
  if (helm) {
    q.zk_re = K;
    q.zk_im = 0.0;
  }
  auto out = [&](int flag, int need, size_t w, int n) -> void* {
    return flag >= need ? static_cast<void*>(get(w * (size_t)n)) : nullptr;
  };
  q.pot = out(p0.ifpgh, 1, W.pot, NS);
  q.grad = out(p0.ifpgh, 2, W.grad, NS);
  q.hess = out(p0.ifpgh, 3, W.hess, NS);
  q.pottarg = out(p0.ifpghtarg, 1, W.pot, NT);
  q.gradtarg = out(p0.ifpghtarg, 2, W.grad, NT);
  q.hesstarg = out(p0.ifpghtarg, 3, W.hess, NT);
  if (S.err) return S.err;
  if (unit) {
    // ---- a plan: front end now, the rest captured (strengths copied in and
    // outputs copied out by the graph from device tables made here)
    unit->members = g;
    unit->helm = helm;
    unit->fin = fin;
    unit->q.c = c;
    std::vector<Seg> sin, sout;
    for (int i = 0; i < np; ++i) {
      const fmm2dhip_problem& p = probs[g[i]];
      const size_t s0 = soff[i], t0 = toff[i];
      auto seg = [](std::vector<Seg>& v, const void* from, void* to, size_t bytes) {
        if (bytes) v.push_back({static_cast<const double*>(from), static_cast<double*>(to),
                                (long long)(bytes / sizeof(double))});
      };
      if (chg) seg(sin, p.charge, chg + W.chg * s0, W.chg * p.ns);
      if (dip) seg(sin, p.dipstr, dip + W.dip * s0, W.dip * p.ns);
      if (dv) seg(sin, p.dipvec, dv + W.dv * s0, W.dv * p.ns);
      auto oseg = [&](const void* from, void* to, size_t w, size_t at, int n) {
        if (from && n) seg(sout, static_cast<const char*>(from) + w * at, to, w * n);
      };
      oseg(q.pot, p.pot, W.pot, s0, p.ns);
      oseg(q.grad, p.grad, W.grad, s0, p.ns);
      oseg(q.hess, p.hess, W.hess, s0, p.ns);
      oseg(q.pottarg, p.pottarg, W.pot, t0, p.nt);
      oseg(q.gradtarg, p.gradtarg, W.grad, t0, p.nt);
      oseg(q.hesstarg, p.hesstarg, W.hess, t0, p.nt);
    }
    Seg* d_in = S.get<Seg>(sin.size());
    Seg* d_out = S.get<Seg>(sout.size());
    if (S.err) return S.err;
    if (!sin.empty())
      HIP_TRY(hipMemcpyAsync(d_in, sin.data(), sizeof(Seg) * sin.size(), hipMemcpyHostToDevice, c.st));
    if (!sout.empty())
      HIP_TRY(hipMemcpyAsync(d_out, sout.data(), sizeof(Seg) * sout.size(), hipMemcpyHostToDevice, c.st));
    HIP_TRY(hipStreamSynchronize(c.st));  // the host tables die here
    if (helm) helm_req_of(q, &unit->fin, unit->hr);
    else lap_req_of(q, &unit->fin, unit->lr);
    e = helm ? helm_setup(unit->hr, S, unit->q) : lap_setup(unit->lr, S, unit->q);
    if (e) return e;
    const int nd = p0.nd;
    e = capture_unit(
        unit, S,
        [&] {
          copy_segs(c.st, d_in, sin.size());
          if (helm && dip)
            k_scale_points<<<nblk((long)NS * nd, 256), 256, 0, c.st>>>(
                reinterpret_cast<double2*>(dip), NS, nd, d_soff, np, d_fp, 1);
        },
        [&] {
          if (helm) {
            if (q.grad) k_scale_points<<<nblk((long)NS * 2 * nd, 256), 256, 0, c.st>>>(static_cast<double2*>(q.grad), NS, 2 * nd, d_soff, np, d_fp, 1);
            if (q.hess) k_scale_points<<<nblk((long)NS * 3 * nd, 256), 256, 0, c.st>>>(static_cast<double2*>(q.hess), NS, 3 * nd, d_soff, np, d_fp, 2);
            if (NT && q.gradtarg) k_scale_points<<<nblk((long)NT * 2 * nd, 256), 256, 0, c.st>>>(static_cast<double2*>(q.gradtarg), NT, 2 * nd, d_toff, np, d_fp, 1);
            if (NT && q.hesstarg) k_scale_points<<<nblk((long)NT * 3 * nd, 256), 256, 0, c.st>>>(static_cast<double2*>(q.hesstarg), NT, 3 * nd, d_toff, np, d_fp, 2);
          }
          copy_segs(c.st, d_out, sout.size());
        });
    if (!e) detach_workspace(unit->blocks);  // before S rewinds the workspace
    return e;
  }
  e = solve_problem(q, &fin);
  if (e) return e;
  if (helm) {  // gradients times a, Hessians times a^2 (per point, nd (2|3) values)
    const int nd = p0.nd;
    if (q.grad) k_scale_points<<<nblk((long)NS * 2 * nd, 256), 256, 0, c.st>>>(static_cast<double2*>(q.grad), NS, 2 * nd, d_soff, np, d_fp, 1);
    if (q.hess) k_scale_points<<<nblk((long)NS * 3 * nd, 256), 256, 0, c.st>>>(static_cast<double2*>(q.hess), NS, 3 * nd, d_soff, np, d_fp, 2);
    if (NT && q.gradtarg) k_scale_points<<<nblk((long)NT * 2 * nd, 256), 256, 0, c.st>>>(static_cast<double2*>(q.gradtarg), NT, 2 * nd, d_toff, np, d_fp, 1);
    if (NT && q.hesstarg) k_scale_points<<<nblk((long)NT * 3 * nd, 256), 256, 0, c.st>>>(static_cast<double2*>(q.hesstarg), NT, 3 * nd, d_toff, np, d_fp, 2);
  }
  // ---- each problem's slice of the outputs to its own arrays
  Copier outc(c.st, dev);
  for (int i = 0; i < np; ++i) {
    const fmm2dhip_problem& p = probs[g[i]];
    const size_t s = soff[i], t = toff[i];
    auto put = [&](const void* from, void* to, size_t w, size_t at, int n) {
      return from && n ? outc.add(static_cast<const char*>(from) + w * at, to, w * n, true) : 0;
    };
    if ((e = put(q.pot, p.pot, W.pot, s, p.ns))) return e;
    if ((e = put(q.grad, p.grad, W.grad, s, p.ns))) return e;
    if ((e = put(q.hess, p.hess, W.hess, s, p.ns))) return e;
    if ((e = put(q.pottarg, p.pottarg, W.pot, t, p.nt))) return e;
    if ((e = put(q.gradtarg, p.gradtarg, W.grad, t, p.nt))) return e;
    if ((e = put(q.hesstarg, p.hesstarg, W.hess, t, p.nt))) return e;
  }
  if ((e = outc.flush(S))) return e;
  HIP_TRY(hipGetLastError());
  HIP_TRY(hipStreamSynchronize(c.st));
  return IER_OK;
}

}  // namespace

std::vector<std::vector<int>> fused_groups(int n, const fmm2dhip_problem* probs) {
  std::map<GroupKey, std::vector<int>> by;
  for (int i = 0; i < n; ++i) {
    const fmm2dhip_problem& p = probs[i];
    if (!eligible(p)) continue;
    GroupKey k{p.kernel, p.nd, p.ifcharge ? 1 : 0, p.ifdipole ? 1 : 0, p.ifpgh, p.ifpghtarg,
               p.nt > 0 ? 1 : 0, leaf_size(p), p.precision == PREC_MIXED ? 1 : 0,
               p.device_ptrs, p.eps};
    by[k].push_back(i);
  }
  std::vector<std::vector<int>> out;
  for (auto& kv : by) {
    // at most 2^28 points and 4^8 problems per forest
    std::vector<int> cur;
    long long pts = 0;
    for (int i : kv.second) {
      const long long q = (long long)probs[i].ns + probs[i].nt;
      if (!cur.empty() && (pts + q >= (1LL << 28) || cur.size() >= 65536)) {
        if (cur.size() >= 2) out.push_back(cur);
        cur.clear();
        pts = 0;
      }
      cur.push_back(i);
      pts += q;
    }
    if (cur.size() >= 2) out.push_back(cur);
  }
  return out;
}

int solve_forest(fmm2dhip_problem* probs, const std::vector<int>& group) {
  if (group.empty()) return 0;
  const double t0 = wall();
  int& mode = thread_precision();
  const int saved = mode;
  mode = probs[group[0]].precision == PREC_MIXED ? PREC_MIXED : PREC_FP64;
  std::vector<int> alone;
  int e;
  try {
    e = run_forest(probs, group, alone, nullptr);
  } catch (const std::bad_alloc&) {
    e = IER_ALLOC;
  } catch (...) {
    e = IER_HIP;
  }
  mode = saved;
  const double t1 = wall();
  int failed = 0;
  for (int i : group) {
    if (std::find(alone.begin(), alone.end(), i) != alone.end()) continue;
    probs[i].ier = e;
    probs[i].seconds = t1 - t0;
    if (e) ++failed;
  }
  for (int i : alone)
    if (solve_problem(probs[i]) != 0) ++failed;
  return failed;
}

namespace {

// A unit for one problem solved alone (not fused).
int alone_unit(fmm2dhip_problem& p, int index, PlanUnit* u) {
  Ctx c;
  int e = get_ctx(c);
  if (e) return e;
  Scratch S(c.st);
  u->members = {index};
  u->helm = p.kernel == FMM2DHIP_HELMHOLTZ;
  u->q.c = c;
  if (u->helm) helm_req_of(p, nullptr, u->hr);
  else lap_req_of(p, nullptr, u->lr);
  e = u->helm ? helm_setup(u->hr, S, u->q) : lap_setup(u->lr, S, u->q);
  if (e) return e;
  e = capture_unit(u, S, [] {}, [] {});
  if (!e) detach_workspace(u->blocks);
  return e;
}

// Runs f with the thread's precision mode set to p's.
template <class F>
int with_precision(const fmm2dhip_problem& p, F f) {
  int& mode = thread_precision();
  const int saved = mode;
  mode = p.precision == PREC_MIXED ? PREC_MIXED : PREC_FP64;
  int e;
  try {
    e = f();
  } catch (const std::bad_alloc&) {
    e = IER_ALLOC;
  } catch (...) {
    e = IER_HIP;
  }
  mode = saved;
  return e;
}

}  // namespace
}  // namespace fmm

struct fmm2dhip_plan {
  std::vector<std::unique_ptr<fmm::PlanUnit>> units;
  fmm2dhip_problem* probs = nullptr;
  int n = 0;
};

extern "C" {

fmm2dhip_plan* fmm2dhip_plan_create(int n, fmm2dhip_problem* probs, int* ier) {
  int e = fmm::IER_OK;
  fmm2dhip_plan* pl = nullptr;
  try {
    if (n <= 0 || !probs) e = fmm::IER_ARGS;
    for (int i = 0; !e && i < n; ++i)
      if (probs[i].device_ptrs != 1) e = fmm::IER_ARGS;  // plans take device arrays only
    if (!e) {
      pl = new fmm2dhip_plan;
      pl->probs = probs;
      pl->n = n;
      const std::vector<std::vector<int>> groups = fmm::fused_groups(n, probs);
      std::vector<char> grouped(n, 0);
      std::vector<int> alone;
      for (const auto& g : groups) {
        for (int i : g) grouped[i] = 1;
        // a first solve sets up what the capture must find ready (cached
        // operator tables, kernel attributes) and checks the group
        if (fmm::solve_forest(probs, g) != 0) {
          e = fmm::IER_HIP;
          for (int i : g)
            if (probs[i].ier) e = probs[i].ier;
          break;
        }
        auto u = std::make_unique<fmm::PlanUnit>();
        std::vector<int> out;
        e = fmm::with_precision(probs[g[0]], [&] {
          return fmm::run_forest(probs, g, out, u.get());
        });
        if (e) break;
        alone.insert(alone.end(), out.begin(), out.end());
        if (!u->members.empty()) pl->units.push_back(std::move(u));
      }
      for (int i = 0; !e && i < n; ++i)
        if (!grouped[i]) alone.push_back(i);
      for (size_t k = 0; !e && k < alone.size(); ++k) {
        const int i = alone[k];
        if (fmm::solve_problem(probs[i]) != 0) {
          e = probs[i].ier;
          break;
        }
        auto u = std::make_unique<fmm::PlanUnit>();
        e = fmm::with_precision(probs[i], [&] { return fmm::alone_unit(probs[i], i, u.get()); });
        if (!e) pl->units.push_back(std::move(u));
      }
    }
  } catch (const std::bad_alloc&) {
    e = fmm::IER_ALLOC;
  } catch (...) {
    e = fmm::IER_HIP;
  }
  if (e) {
    if (pl) {
      (void)hipDeviceSynchronize();
      delete pl;
    }
    pl = nullptr;
  }
  if (ier) *ier = e;
  return pl;
}

int fmm2dhip_plan_execute(fmm2dhip_plan* pl) {
  if (!pl) return fmm::IER_ARGS;
  fmm::Ctx c;
  int e = fmm::get_ctx(c);
  if (e) return e;
  const double t0 = fmm::wall();
  for (auto& u : pl->units)
    if (hipGraphLaunch(u->exec, c.st) != hipSuccess) e = fmm::IER_HIP;
  if (hipStreamSynchronize(c.st) != hipSuccess) e = fmm::IER_HIP;
  const double t1 = fmm::wall();
  for (int i = 0; i < pl->n; ++i) {
    pl->probs[i].ier = e;
    pl->probs[i].seconds = t1 - t0;
  }
  return e;
}

void fmm2dhip_plan_destroy(fmm2dhip_plan* pl) {
  if (!pl) return;
  (void)hipDeviceSynchronize();
  delete pl;
}

}  // extern "C"

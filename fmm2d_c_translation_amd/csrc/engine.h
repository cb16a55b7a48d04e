// engine.h -- internal interface between the C-ABI layer and the device engine.
#pragma once

#include <hip/hip_runtime.h>

#include <atomic>
#include <complex>
#include <cstdint>
#include <vector>

#include "../../include/fmm2d_hip.h"
#include "common.h"
#include "tree.h"

namespace fmm {

// Device memory of one solve, carved from a per-thread, per-device
// workspace that persists across solves: allocation is a pointer bump, and
// the last Scratch of a solve to die synchronises its stream and rewinds the
// workspace.  A solve that outgrows the workspace takes extra blocks; at its
// end they are merged into one block sized for that solve.
class Scratch {
 public:
  explicit Scratch(hipStream_t s);
  ~Scratch();
  Scratch(const Scratch&) = delete;
  Scratch& operator=(const Scratch&) = delete;
  template <class T>
  T* get(size_t n) {
    void* p = take((n ? n : 1) * sizeof(T));
    if (!p) {
      err = IER_ALLOC;
      return nullptr;
    }
    return static_cast<T*>(p);
  }
  int err = IER_OK;

 private:
  void* take(size_t bytes);
  hipStream_t st_;
  int dev_ = 0;
};
// Frees the calling thread's workspace on the current device.
void release_workspace();
// Device bytes held by the calling thread's workspace on the current device.
size_t workspace_capacity();
// Frees everything the calling thread holds: its workspaces, pinned staging
// and stream (for worker threads that are about to exit).
void release_thread();

struct Ctx {
  hipStream_t st = nullptr;
  int device = 0;
};
int get_ctx(Ctx& c);

// While a FixedArena is alive on a thread, that thread's Scratch takes its
// device memory from [base, base + cap) by pointer bump and never calls
// hipMalloc (a request beyond cap fails as IER_ALLOC), and get_ctx hands
// out `stream` instead of creating the thread's own.  For callers that
// cannot allocate; the arena's previous contents are garbage.
class FixedArena {
 public:
  FixedArena(void* base, size_t cap, hipStream_t stream);
  ~FixedArena();
  FixedArena(const FixedArena&) = delete;
  FixedArena& operator=(const FixedArena&) = delete;
  size_t used() const;      // high-water mark of the bump pointer, bytes
  bool overflowed() const;  // a request did not fit
};

// h2d_directcdg_ (helmkernels2d.c) with its return code (direct.hip)
int h2d_direct_cdg(int nd, const fmm2d_cplx* zk, const double* s, int ns, const fmm2d_cplx* c,
                   const fmm2d_cplx* q, const double* v, const double* t, int nt,
                   fmm2d_cplx* pot, fmm2d_cplx* grad, double th);

// A forest: independent problems solved as one tree.  Problem i's root box
// is the cell (rx, ry) of level m of the unit square, inside a slot (a cell
// of level M) that holds nothing else; the root sits one cell in from the
// slot's corner, so the eight cells around it at level m hold no points.
// Below its root a problem's tree is the one it has alone (its keys are
// its own, prefixed by the root cell's), and the cells around the roots
// keep every list of a box below a root inside that problem; the boxes at
// or above the roots do no FMM work.
struct ForestSlot {
  int m, rx, ry, prob;  // m < 0: an empty slot
};
struct ForestDev {
  int M = -1;                        // slot level; -1: not a forest
  const ForestSlot* slot = nullptr;  // 4^M entries, by the Morton index of the level-M cell
};
// The slot of box (key, level l), false when the box lies above the slots
// or in an empty one.
FMM_HD bool forest_slot(const ForestDev& F, uint64_t key, int l, ForestSlot& s) {
  if (l < F.M) return false;
  s = F.slot[key >> (2 * (l - F.M))];
  return s.m >= 0;
}
// Level of box (key, l) in its own problem's tree: l outside a forest, -1
// for a box above every root or in an empty slot.
FMM_HD int forest_plev(const ForestDev& F, uint64_t key, int l) {
  if (F.M < 0) return l;
  ForestSlot s;
  if (!forest_slot(F, key, l, s) || l < s.m) return -1;
  return l - s.m;
}
// One problem of a forest (device copy read by the key and frame kernels).
struct ForestProb {
  double cx, cy, R;     // root box in the problem's own coordinates (keys)
  double a;             // normalised coordinates (x - cx) a; a = 0: none
  double fcx, fcy, fR;  // root box in the coordinates the kernels see
  int m, rx, ry;        // root cell (rx, ry) of level m
  int soff, toff;       // first source / target in the concatenated arrays
};
// Forest input of the point front end (build_plan).
struct ForestIn {
  int np = 0;
  const ForestProb* prob = nullptr;  // device, np entries
  const int* soff = nullptr;         // device, np + 1 prefix offsets of the sources
  const int* toff = nullptr;         // and of the targets
  bool normalise = false;  // the kernels see normalised coordinates (Helmholtz)
  // box centres and sizes in each problem's own kernel coordinates, so
  // that point-to-centre and point-to-point differences keep their relative
  // precision whatever the problem's place in the unit square
  bool frames = false;
  double min_bs0 = 0;      // smallest root box, kernel coordinates (the near field's skip threshold)
  int lfirst = 2;          // 2 below the highest root (the smallest m)
  // levels the first point sort orders (then every level if the tree turns
  // out deeper): an estimate of the forest's depth, kMaxDepth for all bits
  int sort_levels = 29;
  ForestDev dev;
};

// Tree arrays on the device (0-based box ids, lists in CSR form).
struct DTree {
  int nb = 0;
  const uint64_t* key = nullptr;  // Morton prefix of each box at its level
  const double2* ctr = nullptr;
  const int* level = nullptr;
  const double* bs = nullptr;    // box size per level
  const double* bsb = nullptr;   // box size per box (in a forest, in its problem's coordinates)
  const int* nchild = nullptr;
  const int* parent = nullptr;
  const int4* child = nullptr;
  const int2* srng = nullptr;  // sorted source range [x, y)
  const int2* trng = nullptr;  // sorted target range
  const int* nbr = nullptr;    // colleagues nbr[9*b + (dy+1)*3 + (dx+1)], -1 none
  const int *o1 = nullptr, *l1 = nullptr, *o2 = nullptr, *l2 = nullptr;
  const int *o3 = nullptr, *l3 = nullptr, *o4 = nullptr, *l4 = nullptr;
};

// Output of the point-sorting front end shared by every kernel family.
struct SortedPoints {
  int ns = 0, nt = 0;
  double bs0 = 0, c0x = 0, c0y = 0;
  int* perm_s = nullptr;  // sorted position -> original index
  int* perm_t = nullptr;
  double2* src = nullptr;  // sorted coordinates
  double2* trg = nullptr;
};

// Device tree + work lists.  Box ids are level-major (Morton order inside a
// level); every work list holds box ids in increasing order, hence sorted by
// level, with lev[l] the first entry at level >= l (lev[L+1] = n).
struct Span {
  int off = 0, n = 0;
};
struct WorkList {
  int off = 0, n = 0;    // entries dwork[off .. off+n)
  std::vector<int> lev;  // size L+2
  Span all() const { return {off, n}; }
  Span at(int l) const { return {off + lev[l], lev[l + 1] - lev[l]}; }
  Span from(int l) const {
    if (l < 0) l = 0;
    if (l >= (int)lev.size()) return {off + n, 0};
    return {off + lev[l], n - lev[l]};
  }
};
struct Plan {
  int nb = 0, L = 0;            // boxes, deepest level
  std::vector<double> bs;       // box size per level
  std::vector<int> loff;        // first box of each level, loff[L+1] = nb
  DTree D;
  int* dwork = nullptr;         // device: all work lists
  WorkList formmp;              // leaves at levels >= 2 with sources
  WorkList m2m;                 // boxes with children and sources
  WorkList m2l;                 // levels >= 2: boxes with output points and list 2
  WorkList l2l;                 // boxes with output points and children
  WorkList l2l_kids;            // the children of those (lev[] by parent level)
  WorkList list4, list3_s, list3_t, leaves_s, leaves_t, p2p_s, p2p_t;
  std::vector<double> m2l_pairs;  // list-2 pairs of the m2l boxes, per level
  // largest distance between a point of a leaf with points and a point of
  // a box with sources in its list 1 (from box centres and sizes): bounds
  // r over every near-field pair
  double p2p_rmax = 0;
  // the same bound per level of the target leaves (0: no list-1 pair there)
  std::vector<double> p2p_rmax_lev;
  // first level whose boxes can take part in M2L, list 3/4 or carry
  // expansions: 2, or in a forest 2 below the highest problem root
  int lfirst = 2;
  ForestDev forest;
  // pinned host word, cleared before the front end's sorts and set by a sort
  // whose look-back wait gave up; read after the stream synchronises
  uint32_t* sort_err = nullptr;
};
// IER_HIP when a front-end sort of this plan reported a failed wait (call
// after the stream has synchronised).
inline int sort_status(const Plan& p) { return p.sort_err && *p.sort_err ? IER_HIP : IER_OK; }

// Device tree and lists from keys sorted on their top 2 sorted_levels bits
// (dtree.hip).  Returns IER_RESORT, before building anything, when the tree
// needs more levels than the keys are sorted on.  geo: pinned host copy of
// the root box {centre x, centre y, size}, written by a copy already queued
// on the stream (read after the first read-back).  host_s receives the wall
// time of refinement, level restriction and boxes+lists.
constexpr int kSortLevels = 15;  // with the target flag, 31 bits: four 8-bit passes
constexpr int IER_RESORT = -1;  // internal: sort on more key bits and retry
int build_dtree(Ctx& c, Scratch& S, const uint64_t* ks, int ns, const uint64_t* kt, int nt,
                const double* geo, int idivflag, int ndiv, int ifpgh, int ifpghtarg,
                int sorted_levels, Plan& plan, double* host_s,
                const ForestDev* forest = nullptr, int forest_mlo = 0);

// Pinned host staging reused across the calls of one thread (slot < 8).
void* pinned(size_t bytes, int slot);

// Caller arrays in pageable host memory to and from the device through the
// thread's pinned staging (slot 0): the host side copied by up to eight
// threads, then one DMA per array from or to pinned memory (about 57 GB/s
// on MI355X against 29 GB/s in and 18 GB/s out from pageable memory).
// h2d_staged returns once the host data is in the staging buffer (the
// caller's arrays may change after it); the DMAs stay on the stream, and
// the next staged transfer on the thread waits for them.  d2h_staged
// synchronises the stream and returns with the data in the caller's arrays.
// Under a stream capture, above 1 GiB of staging, or without pinned memory
// to be had, both fall back on plain copies.
struct HostXfer {
  const void* src;
  void* dst;
  size_t bytes;
};
int h2d_staged(hipStream_t st, const HostXfer* x, int n);
int d2h_staged(hipStream_t st, const HostXfer* x, int n);

// Stream work of the FMM steps that a graph capture can record and replay:
// fill_async sets n 32-bit words to v with a kernel (memset nodes in a
// replayed graph are avoided); upload_small copies host data to the device,
// through a pinned copy in the thread's HostKeeper while one is installed (a
// captured copy reads its source again at every replay, and no pinned
// allocation may happen while a stream captures: the keeper's pool is
// reserved before).
void fill_async(void* p, uint32_t v, size_t n, hipStream_t st);
int upload_small(void* d, const void* h, size_t bytes, hipStream_t st);
struct HostKeeper {
  char* pool = nullptr;  // pinned, freed with the keeper
  size_t cap = 0, used = 0;
  int reserve(size_t bytes);
  ~HostKeeper();
};
// While a CaptureScope is alive on a thread, its uploads go through keeper
// and the step timer records nothing (events stay out of the graph).
class CaptureScope {
 public:
  explicit CaptureScope(HostKeeper* keeper);
  ~CaptureScope();
  CaptureScope(const CaptureScope&) = delete;
  CaptureScope& operator=(const CaptureScope&) = delete;
};
bool capturing();
// Moves the calling thread's workspace blocks on the current device to out
// (their memory then belongs to the caller, to free with hipFree); the
// workspace starts empty.  Call while the allocations of interest are live.
void detach_workspace(std::vector<void*>& out);

// Boundaries of the eight FMM steps, recorded on the solve's stream:
// ev[k] .. ev[k+1] bracket step order[k]; indices follow the reference's
// timeinfo (0 form mp, 1 form lo, 2 M2M, 3 M2L, 4 L2L, 5 list 3, 6 local
// evaluation, 7 near field).
struct StepTimer {
  hipEvent_t ev[9] = {};
  int order[8] = {0, 2, 1, 3, 4, 5, 6, 7};
  int n = 0;
  bool ok = false;
  void mark(hipStream_t st);
};
StepTimer* step_timer();       // per thread, reset
StepTimer* step_timer_peek();  // per thread, as left by the last solve
void finish_timer(double* timeinfo);  // after the stream is synchronised

// Arithmetic of the list-1 and M2L steps.  PREC_FP64 keeps every step in
// FP64.  PREC_MIXED runs those two steps in FP32 whenever eps >= 1e-6
// (operands rounded from FP64 after the point differences are taken in FP64,
// sums carried into FP64 accumulators); tighter eps stays FP64.  The mode is
// per thread; its initial value is PREC_MIXED when the environment variable
// FMM2DHIP_PRECISION is "mixed", else PREC_FP64.
enum Precision { PREC_FP64 = 0, PREC_MIXED = 1 };
int& thread_precision();
inline bool mixed_for(double eps) { return thread_precision() == PREC_MIXED && eps >= 1e-6; }

// Work counts of the last solve on this thread (for roofline reporting).
struct SolveStats {
  double p2p_pairs_src = 0, p2p_pairs_trg = 0;  // sum over targets of list-1 sources
  // sum over the leaves holding sources (targets) of their list-1 source
  // counts: the source reads of the near field's algorithmic traffic
  double p2p_reads_src = 0, p2p_reads_trg = 0;
  double m2l_pairs = 0, m2l_flops = 0;          // list-2 pairs, 8 * ncoef^2 per pair
  double nboxes = 0, nlevels = 0, pmax = 0;
  double step_s[8] = {0, 0, 0, 0, 0, 0, 0, 0};
  // host wall time of the front end: 0 bbox, 1 keys+sort, 2 adaptive
  // refinement, 3 level restriction, 4 boxes+lists+work lists, 5 leaf
  // order+gather, 6 operator setup, 7 whole call
  double host_s[8] = {0, 0, 0, 0, 0, 0, 0, 0};
  double workspace_bytes = 0;  // device workspace the solve took
  // M2L row-tile counts MT used: bit (family * 2 + fp32) * 9 + MT
  uint64_t m2l_mt_mask = 0;
};

// Tile choices of the translation kernels that the autotuner sets
// (tune.cpp); 0 leaves the built-in rule.  Process-wide, read at launch.
struct TuneTable {
  // [family: 0 Laplace (and biharmonic), 1 Helmholtz][0 FP64 / 1 FP32]
  // [row tiles MT of the order]: row tiles per wave of the all-levels M2L
  // kernels (1..min(4, MT))
  std::atomic<int> m2l_mw[2][2][9];
  // M2M/L2L: a level with fewer 16-row wave tiles than this takes one row
  // tile per wave and splits the slots over wave groups
  std::atomic<int> xlate_waves;
  TuneTable();
};
TuneTable& tune_table();
int m2l_tiles(int MT, bool fp32, int family);  // row tiles per wave for MT row tiles
int xlate_waves();
double wall();  // seconds, monotonic
SolveStats& last_stats();


// Point front end: bounding box, keys, sort, tree, lists, sorted coordinates.
// With a forest (fin), the points are the concatenated points of its
// problems, keyed below their root cells; the root box is the unit square.
int build_plan(Ctx& c, Scratch& S, const double* d_src, int ns, const double* d_trg, int nt,
               int idivflag, int ndiv, SortedPoints& P, Plan& plan, bool want_lap_work,
               int ifpgh, int ifpghtarg, const ForestIn* fin = nullptr);

// Gathers (nd-fastest, original order) strengths into channel-major sorted
// order: out[(c*nd + d)*n + i] = in[(perm[i]*ncomp + c)*nd + d].
void gather_channels(hipStream_t st, const double2* in, double2* out, const int* perm, int n,
                     int nd, int ncomp);
void gather_real(hipStream_t st, const double* in, double* out, const int* perm, int n, int nd,
                 int ncomp);
// Charges, dipole strengths and dipole vectors (each may be null) in one
// pass: chg1/dip1 as gather_channels(ncomp 1), dv1 as gather_real(ncomp 2).
void gather_sources(hipStream_t st, const double2* chg, double2* chg1, const double2* dip,
                    double2* dip1, const double* dv, double* dv1, const int* perm, int n, int nd);
// Inverse: out[(perm[i]*ncomp + c)*nd + d] = in[(c*nd + d)*n + i].
void scatter_channels(hipStream_t st, const double2* in, double2* out, const int* perm, int n,
                      int nd, int ncomp);

// Final outputs of one point set in input order, for the near-field kernel
// to write (pot null: not used).  kind 0: Cauchy, complex, one component
// each (scatter_channels layout); 1: log kernel with complex densities,
// sorted channel 2d + h giving the real (h = 0) or imaginary (h = 1) part of
// the caller's channel d (lfmm2d unpack); 2: log kernel with real outputs
// (rfmm2d unpack).  nd: the caller's channels.
struct LapOut {
  int kind = 0, nd = 1;
  const int* perm = nullptr;
  void *pot = nullptr, *grad = nullptr, *hess = nullptr;
};

// Cauchy-kernel FMM on sorted, device-resident data (cfmm2dmain semantics).
struct LapArgs {
  int nd = 1;
  double eps = 1e-6;
  int ifcharge = 0, ifdipole = 0, ifpgh = 0, ifpghtarg = 0, ifnear = 1;
  bool mixed = false;                                 // FP32 list 1 and M2L
  const double2 *chg = nullptr, *dip = nullptr;      // sorted, channel-major
  double2 *pot = nullptr, *grad = nullptr, *hess = nullptr;   // sorted sources
  double2 *pott = nullptr, *gradt = nullptr, *hesst = nullptr;  // sorted targets
  // The near field is the last step to touch a point and covers every
  // point once (each leaf's list 1 holds the leaf): where it runs it writes
  // sorted value + near field straight to out_s / out_t, and fused_s /
  // fused_t say so; the caller unpacks the other sets.
  LapOut out_s, out_t;
  mutable bool fused_s = false, fused_t = false;
};
int lap_run(Ctx& c, Scratch& S, const SortedPoints& P, const Plan& plan, const LapArgs& a);

// Helmholtz FMM on sorted, device-resident data (hfmm2dmain semantics, low
// frequency).  Sorted arrays are [component][channel][point].
struct HelmArgs {
  int nd = 1;
  double eps = 1e-6;
  std::complex<double> zk = 1.0;
  int ifcharge = 0, ifdipole = 0, ifpgh = 0, ifpghtarg = 0, ifnear = 1;
  bool mixed = false;                                 // FP32 M2L
  const double2 *chg = nullptr, *dip = nullptr;
  const double* dipvec = nullptr;                            // [2][nd][ns]
  double2 *pot = nullptr, *grad = nullptr, *hess = nullptr;   // sorted sources
  double2 *pott = nullptr, *gradt = nullptr, *hesst = nullptr;  // sorted targets
  // Outputs in input order (scatter_channels layout, ncomp 1/2/3), or null.
  // The near field is the last step to touch a point and covers every
  // point once (each leaf's list 1 holds the leaf), so when it runs it
  // writes sorted value + near field there directly; fused_s / fused_t
  // report for which set it did, and the caller scatters the others.
  double2 *upot = nullptr, *ugrad = nullptr, *uhess = nullptr;
  double2 *upott = nullptr, *ugradt = nullptr, *uhesst = nullptr;
  mutable bool fused_s = false, fused_t = false;
};
int helm_run(Ctx& c, Scratch& S, const SortedPoints& P, const Plan& plan, const HelmArgs& a);
int h2dterms_host(double bs, std::complex<double> zk, double eps, int* ier);

// Diagonal-form (high-frequency) translations of one level (wideband.hip):
// Y[work[w]] += sum over slots s with map[w][s] >= 0 of the translation of
// X[map[w][s]], through signatures of length nsig.  kind 0: M2M from the
// children (slot = quadrant) into their parents at box size bs; 1: M2L at
// box size bs (49 list-2 offsets); 2: L2L from parents of box size bs into
// the children (slot = the child's quadrant).  Input boxes are the id range
// in_first .. in_first + in_count - 1 (one level).  Expansions as
// everywhere: [box][channel][order + p], scale 1.
struct HfJob {
  int kind = 0;
  int pin = 0, pout = 0, nsig = 0;
  const int* work = nullptr;
  int nwork = 0;
  const int* map = nullptr;
  int nslot = 0;
  int in_first = 0, in_count = 0;
  double bs = 0;
};
int hf_translate(hipStream_t st, Scratch& S, std::complex<double> zk, const HfJob& j,
                 const double2* X, size_t xs, double2* Y, size_t ys, int nd);
// the reference's FFT length rule (next235.c:10-32)
int next235(double base);
// unnormalised DFTs of nrows rows of n (2^a 3^b 5^c) values in place; sign
// -1 forward (FFTPACK zfftf), +1 backward (zfftb)
int fft_rows_device(hipStream_t st, Scratch& S, double2* rows, int nrows, int n, int sign);

int l2dterms(double eps);
int ndiv_rule(double eps, int ns, int nt);

// A drop-in solve's arguments: Laplace (cfmm2d, lfmm2d, rfmm2d) and
// Helmholtz (hfmm2d), as the C-ABI entry points hand them over (abi.hip).
struct LapReq {
  bool logk = false;  // lfmm (log kernel, complex densities) vs cfmm (Cauchy)
  bool realk = false;  // rfmm (log kernel, real densities and outputs)
  int nd = 1;
  double eps = 1e-6;
  int ns = 0, nt = 0;
  const double *src = nullptr, *trg = nullptr, *dipvec = nullptr;
  int ifc = 0, ifd = 0;
  const void *chg = nullptr, *dip = nullptr;
  int pg = 0, pgt = 0;
  void *pot = nullptr, *grad = nullptr, *hess = nullptr;
  void *pott = nullptr, *gradt = nullptr, *hesst = nullptr;
  int ndiv = 0, idivflag = 0, ifnear = 1;
  bool device = false;
  hipStream_t stream = nullptr;
  double* timeinfo = nullptr;
  const ForestIn* forest = nullptr;  // the points are a forest's (forest.hip)
};

struct HelmReq {
  int nd = 1;
  double eps = 1e-6;
  std::complex<double> zk = 1.0;
  int ns = 0, nt = 0;
  const double *src = nullptr, *trg = nullptr, *dipvec = nullptr;
  int ifc = 0, ifd = 0;
  const void *chg = nullptr, *dip = nullptr;
  int pg = 0, pgt = 0;
  void *pot = nullptr, *grad = nullptr, *hess = nullptr;
  void *pott = nullptr, *gradt = nullptr, *hesst = nullptr;
  int ndiv = 0, idivflag = 0, ifnear = 1;
  bool device = false;
  hipStream_t stream = nullptr;
  double* timeinfo = nullptr;
  const ForestIn* forest = nullptr;  // the points are a forest's (forest.hip)
};

// A solve between its front end and its evaluation: the device inputs, the
// point sort and the tree.  lap_setup / helm_setup run the front end (it
// reads sizes back from the device); lap_eval / helm_eval queue the rest --
// strengths in sorted order, the FMM steps, the outputs in the caller's
// layout -- as stream work alone, which a graph capture can record.
struct Prepared {
  Ctx c;
  SortedPoints P;
  Plan plan;
  const double *src = nullptr, *trg = nullptr, *dv = nullptr;
  const void *chg = nullptr, *dip = nullptr;
};
int lap_setup(const LapReq& r, Scratch& S, Prepared& q);
int lap_eval(const LapReq& r, Scratch& S, Prepared& q);
int helm_setup(const HelmReq& r, Scratch& S, Prepared& q);
int helm_eval(const HelmReq& r, Scratch& S, Prepared& q);
// The drop-in request of a batched problem (leaf size by the eps rule
// unless the problem sets one).
void lap_req_of(const fmm2dhip_problem& p, const ForestIn* forest, LapReq& r);
void helm_req_of(const fmm2dhip_problem& p, const ForestIn* forest, HelmReq& r);

// One problem of the batched/served interface, solved on the calling
// thread (abi.hip); writes p.ier and p.seconds, returns p.ier.  With a
// forest, p holds the concatenated device arrays of the forest's problems.
int solve_problem(fmm2dhip_problem& p, const ForestIn* forest = nullptr);

// Fused batches (forest.hip).  fused_groups: the groups of at least two
// problems of probs[0..n) that can be solved together as one forest (same
// kernel, flags, channels, eps, leaf size, precision and memory kind, more
// sources than the leaf size, a real positive Helmholtz wavenumber), as
// indices; every other problem is in no group.  solve_forest: solves one
// group on the calling thread (a member whose root box turns out unfit for
// the forest is solved alone); writes each member's ier and seconds and
// returns the number of members whose ier is not 0.
std::vector<std::vector<int>> fused_groups(int n, const fmm2dhip_problem* probs);
int solve_forest(fmm2dhip_problem* probs, const std::vector<int>& group);

}  // namespace fmm

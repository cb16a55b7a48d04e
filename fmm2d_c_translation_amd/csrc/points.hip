// points.hip -- device front end: bounding box, Morton keys, radix sort,
// in-leaf ordering, gather/scatter of point data, and the upload of the
// box-level tree built by tree.cpp.
#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <chrono>
#include <cstring>

#include "engine.h"

namespace fmm {

static thread_local StepTimer g_timer;
double wall() {
  return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
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

int get_ctx(Ctx& c) {
  static thread_local hipStream_t st = nullptr;
  static thread_local int dev = -1;
  int d = 0;
  if (hipGetDevice(&d) != hipSuccess) return IER_NOGPU;
  if (!st || dev != d) {
    if (hipStreamCreateWithFlags(&st, hipStreamNonBlocking) != hipSuccess) return IER_NOGPU;
    dev = d;
    hipMemPool_t pool;
    if (hipDeviceGetDefaultMemPool(&pool, d) == hipSuccess) {
      uint64_t keep = UINT64_MAX;
      (void)hipMemPoolSetAttribute(pool, hipMemPoolAttrReleaseThreshold, &keep);
    }
  }
  c.st = st;
  c.device = d;
  return IER_OK;
}

namespace {

// pinned host staging reused across calls of one thread
void* pinned(size_t bytes, int slot) {
  static thread_local void* buf[4] = {nullptr, nullptr, nullptr, nullptr};
  static thread_local size_t cap[4] = {0, 0, 0, 0};
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

__global__ void k_keys(const double2* p, int n, double cx, double cy, double bs0,
                       uint64_t* key, int* idx) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  double2 q = p[i];
  key[i] = point_key(q.x, q.y, cx, cy, bs0);
  idx[i] = i;
}

// Inside every leaf, restore original-index order (the reference's
// partitions are stable).  One thread per leaf, insertion sort.
__global__ void k_leaf_sort(int* idx, const int2* rng, const int* leaves, int nleaves) {
  int w = blockIdx.x * blockDim.x + threadIdx.x;
  if (w >= nleaves) return;
  int2 r = rng[leaves[w]];
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

// sort keys/indices of one point set; returns sorted keys on the host
int sort_set(Ctx& c, Scratch& S, const double2* pts, int n, double cx, double cy, double bs0,
             int* perm, std::vector<uint64_t>& hkeys, int slot) {
  hkeys.resize(n);
  if (n == 0) return IER_OK;
  uint64_t* k0 = S.get<uint64_t>(n);
  uint64_t* k1 = S.get<uint64_t>(n);
  int* i0 = S.get<int>(n);
  if (S.err) return S.err;
  k_keys<<<nblk(n, 256), 256, 0, c.st>>>(pts, n, cx, cy, bs0, k0, i0);
  size_t tb = 0;
  HIP_TRY(hipcub::DeviceRadixSort::SortPairs(nullptr, tb, k0, k1, i0, perm, n, 0, 2 * kMaxDepth,
                                             c.st));
  void* tmp = S.get<char>(tb);
  if (S.err) return S.err;
  HIP_TRY(hipcub::DeviceRadixSort::SortPairs(tmp, tb, k0, k1, i0, perm, n, 0, 2 * kMaxDepth,
                                             c.st));
  void* h = pinned(sizeof(uint64_t) * n, slot);
  if (!h) return IER_ALLOC;
  HIP_TRY(hipMemcpyAsync(h, k1, sizeof(uint64_t) * n, hipMemcpyDeviceToHost, c.st));
  HIP_TRY(hipStreamSynchronize(c.st));
  std::memcpy(hkeys.data(), h, sizeof(uint64_t) * n);
  return IER_OK;
}

void csr(const std::vector<int>& cnt, const std::vector<int>& fixed, int w,
         std::vector<int>& off, std::vector<int>& val) {
  int n = (int)cnt.size();
  off.assign(n + 1, 0);
  for (int b = 0; b < n; ++b) off[b + 1] = off[b] + std::min(cnt[b], w);
  val.resize(off[n]);
  for (int b = 0; b < n; ++b)
    for (int i = 0; i < std::min(cnt[b], w); ++i) val[off[b] + i] = fixed[(size_t)w * b + i];
}

}  // namespace

void gather_channels(hipStream_t st, const double2* in, double2* out, const int* perm, int n,
                     int nd, int ncomp) {
  if (n > 0) k_gather_ch<<<nblk(n, 256), 256, 0, st>>>(in, out, perm, n, nd, ncomp);
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
               int ifpgh, int ifpghtarg) {
  (void)want_lap_work;
  const double2* src = reinterpret_cast<const double2*>(d_src);
  const double2* trg = reinterpret_cast<const double2*>(d_trg);
  P.ns = ns;
  P.nt = nt;
  double th[7];
  th[0] = wall();
  // ---- bounding box of sources and targets (pts_tree2d.c:493-516)
  int nbk = std::min(1024, std::max(1, nblk(ns + nt, kBB)));
  double4* part = S.get<double4>(nbk);
  if (S.err) return S.err;
  k_bbox_partial<<<nbk, kBB, 0, c.st>>>(src, ns, trg, nt, part);
  double4* hp = static_cast<double4*>(pinned(sizeof(double4) * nbk, 2));
  if (!hp) return IER_ALLOC;
  HIP_TRY(hipMemcpyAsync(hp, part, sizeof(double4) * nbk, hipMemcpyDeviceToHost, c.st));
  HIP_TRY(hipStreamSynchronize(c.st));
  double xmin = hp[0].x, xmax = hp[0].y, ymin = hp[0].z, ymax = hp[0].w;
  for (int i = 1; i < nbk; ++i) {
    xmin = std::min(xmin, hp[i].x); xmax = std::max(xmax, hp[i].y);
    ymin = std::min(ymin, hp[i].z); ymax = std::max(ymax, hp[i].w);
  }
  P.bs0 = xmax - xmin;
  if (ymax - ymin > P.bs0) P.bs0 = ymax - ymin;
  P.c0x = (xmin + xmax) / 2;
  P.c0y = (ymin + ymax) / 2;
  th[1] = wall();

  // ---- keys + sort
  P.perm_s = S.get<int>(ns);
  P.perm_t = S.get<int>(nt);
  if (S.err) return S.err;
  std::vector<uint64_t> ks, kt;
  int e = sort_set(c, S, src, ns, P.c0x, P.c0y, P.bs0, P.perm_s, ks, 0);
  if (e) return e;
  e = sort_set(c, S, trg, nt, P.c0x, P.c0y, P.bs0, P.perm_t, kt, 1);
  if (e) return e;

  // ---- box-level tree and lists on the host
  th[2] = wall();
  Tree& T = plan.T;
  e = build_tree(ks.data(), ns, kt.data(), nt, P.c0x, P.c0y, P.bs0, idivflag, ndiv, T);
  if (e) return e;
  th[3] = wall();
  build_lists(T);
  th[4] = wall();
  const int nb = T.nboxes, L = T.nlevels;
  for (int b = 0; b < nb; ++b)
    if (T.n1[b] > 13 || T.n2[b] > 27 || T.n3[b] > 20 || T.n4[b] > 5) return IER_ARGS;

  // ---- work lists
  std::vector<int>& W = plan.work;
  W.clear();
  auto span = [&](auto&& pred, int lo, int hi) {
    Plan::Span s;
    s.off = (int)W.size();
    for (int b = lo; b < hi; ++b)
      if (pred(b)) W.push_back(b);
    s.n = (int)W.size() - s.off;
    return s;
  };
  auto ns_ = [&](int b) { return T.s1[b] - T.s0[b]; };
  auto nt_ = [&](int b) { return T.t1[b] - T.t0[b]; };
  auto npts = [&](int b) { return (ifpghtarg > 0 ? nt_(b) : 0) + (ifpgh > 0 ? ns_(b) : 0); };
  plan.formmp = span([&](int b) { return T.level[b] >= 2 && T.nchild[b] == 0 && ns_(b) > 0; }, 0, nb);
  plan.m2m.assign(L + 1, {});
  plan.m2l.assign(L + 1, {});
  plan.l2l.assign(L + 1, {});
  plan.l2l_kids.assign(L + 1, {});
  for (int l = 0; l <= L; ++l) {
    int f = T.lfirst[l], n = T.lcount[l];
    plan.m2m[l] = span([&](int b) { return T.nchild[b] > 0; }, f, f + n);
    plan.m2l[l] = span([&](int b) { return l >= 2 && npts(b) > 0 && T.n2[b] > 0; }, f, f + n);
    plan.l2l[l] = span([&](int b) { return npts(b) > 0 && T.nchild[b] > 0; }, f, f + n);
    Plan::Span k;
    k.off = (int)W.size();
    for (int i = plan.l2l[l].off; i < plan.l2l[l].off + plan.l2l[l].n; ++i)
      for (int q = 0; q < 4; ++q) W.push_back(T.child[4 * W[i] + q]);
    k.n = (int)W.size() - k.off;
    plan.l2l_kids[l] = k;
  }
  plan.list4 = span([&](int b) { return T.level[b] >= 2 && npts(b) > 0 && T.n4[b] > 0; }, 0, nb);
  plan.list3_s = span([&](int b) { return T.n3[b] > 0 && ns_(b) > 0; }, 0, nb);
  plan.list3_t = span([&](int b) { return T.n3[b] > 0 && nt_(b) > 0; }, 0, nb);
  plan.leaves_s = span([&](int b) { return T.nchild[b] == 0 && ns_(b) > 0; }, 0, nb);
  plan.leaves_t = span([&](int b) { return T.nchild[b] == 0 && nt_(b) > 0; }, 0, nb);
  plan.p2p_s = span([&](int b) { return T.n1[b] > 0 && ns_(b) > 0; }, 0, nb);
  plan.p2p_t = span([&](int b) { return T.n1[b] > 0 && nt_(b) > 0; }, 0, nb);
  plan.leafsort_s = plan.leaves_s;
  plan.leafsort_t = plan.leaves_t;

  // ---- work counts of this solve
  {
    SolveStats& st = last_stats();
    st = SolveStats();
    for (int b = 0; b < nb; ++b) {
      double srcs = 0;
      for (int k = 0; k < std::min(T.n1[b], 13); ++k) srcs += ns_(T.l1[13 * b + k]);
      st.p2p_pairs_src += srcs * ns_(b);
      st.p2p_pairs_trg += srcs * nt_(b);
    }
    st.nboxes = nb;
    st.nlevels = L;
  }

  // ---- upload: box arrays, CSR lists, work lists
  std::vector<int> o1, v1, o2, v2, o3, v3, o4, v4;
  csr(T.n1, T.l1, 13, o1, v1);
  csr(T.n2, T.l2, 27, o2, v2);
  csr(T.n3, T.l3, 20, o3, v3);
  csr(T.n4, T.l4, 5, o4, v4);
  size_t ints = (size_t)nb * (1 + 4 + 2 + 2) + o1.size() + v1.size() + o2.size() + v2.size() +
                o3.size() + v3.size() + o4.size() + v4.size() + W.size();
  size_t bytes = sizeof(double2) * nb + sizeof(double) * (L + 1) + sizeof(int) * (ints + 2 * nb) + 64 * 16;
  char* hbuf = static_cast<char*>(pinned(bytes, 3));
  char* dbuf = S.get<char>(bytes);
  if (!hbuf || S.err) return IER_ALLOC;
  size_t off = 0;
  auto put = [&](const void* p, size_t n, size_t align) -> size_t {
    off = (off + align - 1) / align * align;
    size_t at = off;
    if (n) std::memcpy(hbuf + at, p, n);
    off += n;
    return at;
  };
  std::vector<double2> ctr(nb);
  std::vector<int> child(4 * (size_t)nb);
  std::vector<int2> sr(nb), tr(nb);
  for (int b = 0; b < nb; ++b) {
    ctr[b] = make_double2(T.cx[b], T.cy[b]);
    for (int q = 0; q < 4; ++q) child[4 * b + q] = T.child[4 * b + q];
    sr[b] = make_int2(T.s0[b], T.s1[b]);
    tr[b] = make_int2(T.t0[b], T.t1[b]);
  }
  size_t a_ctr = put(ctr.data(), sizeof(double2) * nb, 16);
  size_t a_bs = put(T.bs.data(), sizeof(double) * (L + 1), 16);
  size_t a_lev = put(T.level.data(), sizeof(int) * nb, 16);
  size_t a_nch = put(T.nchild.data(), sizeof(int) * nb, 16);
  size_t a_par = put(T.parent.data(), sizeof(int) * nb, 16);
  size_t a_chi = put(child.data(), sizeof(int) * 4 * nb, 16);
  size_t a_sr = put(sr.data(), sizeof(int2) * nb, 16);
  size_t a_tr = put(tr.data(), sizeof(int2) * nb, 16);
  size_t a_o1 = put(o1.data(), sizeof(int) * o1.size(), 16), a_v1 = put(v1.data(), sizeof(int) * v1.size(), 16);
  size_t a_o2 = put(o2.data(), sizeof(int) * o2.size(), 16), a_v2 = put(v2.data(), sizeof(int) * v2.size(), 16);
  size_t a_o3 = put(o3.data(), sizeof(int) * o3.size(), 16), a_v3 = put(v3.data(), sizeof(int) * v3.size(), 16);
  size_t a_o4 = put(o4.data(), sizeof(int) * o4.size(), 16), a_v4 = put(v4.data(), sizeof(int) * v4.size(), 16);
  size_t a_w = put(W.data(), sizeof(int) * W.size(), 16);
  if (off > bytes) return IER_ARGS;
  HIP_TRY(hipMemcpyAsync(dbuf, hbuf, off, hipMemcpyHostToDevice, c.st));
  DTree& D = plan.D;
  D.nb = nb;
  D.ctr = reinterpret_cast<const double2*>(dbuf + a_ctr);
  D.bs = reinterpret_cast<const double*>(dbuf + a_bs);
  D.level = reinterpret_cast<const int*>(dbuf + a_lev);
  D.nchild = reinterpret_cast<const int*>(dbuf + a_nch);
  D.parent = reinterpret_cast<const int*>(dbuf + a_par);
  D.child = reinterpret_cast<const int4*>(dbuf + a_chi);
  D.srng = reinterpret_cast<const int2*>(dbuf + a_sr);
  D.trng = reinterpret_cast<const int2*>(dbuf + a_tr);
  D.o1 = reinterpret_cast<const int*>(dbuf + a_o1); D.l1 = reinterpret_cast<const int*>(dbuf + a_v1);
  D.o2 = reinterpret_cast<const int*>(dbuf + a_o2); D.l2 = reinterpret_cast<const int*>(dbuf + a_v2);
  D.o3 = reinterpret_cast<const int*>(dbuf + a_o3); D.l3 = reinterpret_cast<const int*>(dbuf + a_v3);
  D.o4 = reinterpret_cast<const int*>(dbuf + a_o4); D.l4 = reinterpret_cast<const int*>(dbuf + a_v4);
  plan.dwork = reinterpret_cast<int*>(dbuf + a_w);

  th[5] = wall();
  // ---- original-index order inside leaves, then sorted coordinates
  if (plan.leafsort_s.n)
    k_leaf_sort<<<nblk(plan.leafsort_s.n, 64), 64, 0, c.st>>>(P.perm_s, D.srng,
                                                              plan.dwork + plan.leafsort_s.off,
                                                              plan.leafsort_s.n);
  if (plan.leafsort_t.n)
    k_leaf_sort<<<nblk(plan.leafsort_t.n, 64), 64, 0, c.st>>>(P.perm_t, D.trng,
                                                              plan.dwork + plan.leafsort_t.off,
                                                              plan.leafsort_t.n);
  P.src = S.get<double2>(ns);
  P.trg = S.get<double2>(nt);
  if (S.err) return S.err;
  if (ns) k_gather2<<<nblk(ns, 256), 256, 0, c.st>>>(src, P.src, P.perm_s, ns);
  if (nt) k_gather2<<<nblk(nt, 256), 256, 0, c.st>>>(trg, P.trg, P.perm_t, nt);
  HIP_TRY(hipGetLastError());
  // the pinned upload buffer is reused by the next call: finish the copy now
  HIP_TRY(hipStreamSynchronize(c.st));
  th[6] = wall();
  SolveStats& st = last_stats();
  for (int k = 0; k < 6; ++k) st.host_s[k] = th[k + 1] - th[k];
  return IER_OK;
}

}  // namespace fmm

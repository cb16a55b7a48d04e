// abi.hip -- the C-ABI layer: gfortran-ABI drop-in entry points and the
// device-resident variants.  Host pointers are staged through HBM; all
// compute runs in the device engine (points.hip, dtree.hip, lap.hip,
// helm.hip); tree.cpp serves only the host tree test hook.
#include <algorithm>
#include <cstring>
#include <new>
#include <numeric>

#include "../../include/fmm2d_hip.h"
#include "engine.h"
#include "rsort.h"

namespace fmm {
namespace {

inline int nblk(int n, int b) { return (n + b - 1) / b; }

// Complex-density log kernel as two Cauchy channels per density
// (lfmm2d_ndiv.c:139-161): charge1 = (Re c, Im c), dipstr1 = -(v1 + i v2) (Re d, Im d).
__global__ void k_lfmm_prep(const double2* chg, const double2* dip, const double* dipvec,
                            const int* perm, int n, int nd, double2* chg1, double2* dip1) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  int s = perm[i];
  for (int d = 0; d < nd; ++d) {
    if (chg1) {
      double2 c = chg[(size_t)s * nd + d];
      chg1[(size_t)(2 * d) * n + i] = make_double2(c.x, 0.0);
      chg1[(size_t)(2 * d + 1) * n + i] = make_double2(c.y, 0.0);
    }
    if (dip1) {
      double2 q = dip[(size_t)s * nd + d];
      double v1 = dipvec[((size_t)s * 2 + 0) * nd + d], v2 = dipvec[((size_t)s * 2 + 1) * nd + d];
      dip1[(size_t)(2 * d) * n + i] = make_double2(-q.x * v1, -q.x * v2);
      dip1[(size_t)(2 * d + 1) * n + i] = make_double2(-q.y * v1, -q.y * v2);
    }
  }
}

// unpack (lfmm2d_ndiv.c:171-243): pot = Re ch1 + i Re ch2; grad = (Re, -Im) of
// d/dz; hess = (Re, -Im, -Re) of d2/dz2; scattered to original order
__global__ void k_lfmm_unpack(const double2* P, const double2* G, const double2* H,
                              const int* perm, int n, int nd, int pg, double2* pot,
                              double2* grad, double2* hess) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  int t = perm[i];
  for (int d = 0; d < nd; ++d) {
    size_t a = (size_t)(2 * d) * n + i, b = (size_t)(2 * d + 1) * n + i;
    pot[(size_t)t * nd + d] = make_double2(P[a].x, P[b].x);
    if (pg >= 2) {
      grad[((size_t)t * 2 + 0) * nd + d] = make_double2(G[a].x, G[b].x);
      grad[((size_t)t * 2 + 1) * nd + d] = make_double2(-G[a].y, -G[b].y);
    }
    if (pg >= 3) {
      hess[((size_t)t * 3 + 0) * nd + d] = make_double2(H[a].x, H[b].x);
      hess[((size_t)t * 3 + 1) * nd + d] = make_double2(-H[a].y, -H[b].y);
      hess[((size_t)t * 3 + 2) * nd + d] = make_double2(-H[a].x, -H[b].x);
    }
  }
}

// rfmm inputs (rfmm2d_ndiv.c:130-150): charge1 = charge, dipstr1 =
// dipstr * (-(v1 + i v2)), gathered channel-major into sorted order
__global__ void k_rfmm_prep(const double* chg, const double* dip, const double* dipvec,
                            const int* perm, int n, int nd, double2* chg1, double2* dip1) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  int s = perm[i];
  for (int d = 0; d < nd; ++d) {
    if (chg1) chg1[(size_t)d * n + i] = make_double2(chg[(size_t)s * nd + d], 0.0);
    if (dip1) {
      double q = dip[(size_t)s * nd + d];
      double v1 = dipvec[((size_t)s * 2 + 0) * nd + d], v2 = dipvec[((size_t)s * 2 + 1) * nd + d];
      dip1[(size_t)d * n + i] = make_double2(-(q * v1), -(q * v2));
    }
  }
}

// rfmm unpack (rfmm2d_ndiv.c:156-214): pot = Re; grad = (Re, -Im) of d/dz;
// hess = (Re, -Im, -Re) of d2/dz2; real outputs scattered to original order
__global__ void k_rfmm_unpack(const double2* P, const double2* G, const double2* H,
                              const int* perm, int n, int nd, int pg, double* pot, double* grad,
                              double* hess) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  int t = perm[i];
  for (int d = 0; d < nd; ++d) {
    size_t a = (size_t)d * n + i;
    pot[(size_t)t * nd + d] = P[a].x;
    if (pg >= 2) {
      grad[((size_t)t * 2 + 0) * nd + d] = G[a].x;
      grad[((size_t)t * 2 + 1) * nd + d] = -G[a].y;
    }
    if (pg >= 3) {
      hess[((size_t)t * 3 + 0) * nd + d] = H[a].x;
      hess[((size_t)t * 3 + 1) * nd + d] = -H[a].y;
      hess[((size_t)t * 3 + 2) * nd + d] = -H[a].x;
    }
  }
}


template <class T>
int upload(Scratch& S, hipStream_t st, const T* h, size_t n, const T** d) {
  T* p = S.get<T>(n);
  if (S.err) return S.err;
  if (n) HIP_TRY(hipMemcpyAsync(p, h, sizeof(T) * n, hipMemcpyHostToDevice, st));
  *d = p;
  return IER_OK;
}

// A solve's input arrays in host memory, moved together through the
// thread's pinned staging (h2d_staged); add points *d at the device copy.
struct Uploads {
  std::vector<HostXfer> x;
  template <class T>
  int add(Scratch& S, const T* h, size_t n, const T** d) {
    T* p = S.get<T>(n);
    if (S.err) return S.err;
    if (n) x.push_back(HostXfer{h, p, sizeof(T) * n});
    *d = p;
    return IER_OK;
  }
  int flush(hipStream_t st) { return h2d_staged(st, x.data(), (int)x.size()); }
};
// Output arrays to the callers' host memory, moved together (d2h_staged).
struct Downloads {
  std::vector<HostXfer> x;
  void add(const void* d, void* h, size_t bytes) {
    if (bytes) x.push_back(HostXfer{d, h, bytes});
  }
  int flush(hipStream_t st) { return d2h_staged(st, x.data(), (int)x.size()); }
};

}  // namespace

// Laplace front end: arguments, host inputs to the device, point sort, tree
// and lists (the part of a solve that reads sizes back).
int lap_setup(const LapReq& r, Scratch& S, Prepared& q) {
  Ctx& c = q.c;
  if (r.nd < 1 || r.ns < 1 || r.nt < 0 || r.pg < 0 || r.pg > 3 || r.pgt < 0 || r.pgt > 3)
    return IER_ARGS;
  int e = IER_OK;
  const int nd = r.nd, ns = r.ns, nt = r.nt;
  const bool split = r.logk || r.realk;
  q.src = r.src;
  q.trg = r.trg;
  q.dv = r.dipvec;
  q.chg = r.chg;
  q.dip = r.dip;
  if (!r.device) {
    const size_t esz = r.realk ? sizeof(double) : sizeof(double2);
    Uploads up;
    if ((e = up.add(S, r.src, 2 * (size_t)ns, &q.src))) return e;
    if ((e = up.add(S, r.trg, 2 * (size_t)nt, &q.trg))) return e;
    const char *cc = static_cast<const char*>(r.chg), *qq = static_cast<const char*>(r.dip);
    if (r.ifc && (e = up.add(S, cc, esz * nd * ns, &cc))) return e;
    if (r.ifd && (e = up.add(S, qq, esz * nd * ns, &qq))) return e;
    q.chg = cc;
    q.dip = qq;
    if (split && r.ifd && (e = up.add(S, r.dipvec, 2 * (size_t)nd * ns, &q.dv))) return e;
    if ((e = up.flush(c.st))) return e;
  }
  return build_plan(c, S, q.src, ns, q.trg, nt, r.idivflag, r.ndiv, q.P, q.plan, true, r.pg,
                    r.pgt, r.forest);
}

// Laplace evaluation after lap_setup: strengths in sorted order, the FMM
// steps, outputs in the caller's layout (and to the host for host
// outputs); stream work only, no synchronisation.
int lap_eval(const LapReq& r, Scratch& S, Prepared& q) {
  Ctx& c = q.c;
  SortedPoints& P = q.P;
  Plan& plan = q.plan;
  int e = IER_OK;
  const int nd = r.nd, ns = r.ns, nt = r.nt;
  const bool split = r.logk || r.realk;  // gradient (nd,2,n), Hessian (nd,3,n)
  const int ncomp_g = split ? 2 : 1, ncomp_h = split ? 3 : 1;
  const size_t esz = r.realk ? sizeof(double) : sizeof(double2);  // user element
  const double* dv = q.dv;
  const double2 *chg = static_cast<const double2*>(q.chg), *dip = static_cast<const double2*>(q.dip);
  const double *rchg = static_cast<const double*>(q.chg), *rdip = static_cast<const double*>(q.dip);
  const int ndc = r.logk ? 2 * nd : nd;
  LapArgs a;
  a.nd = ndc;
  a.eps = r.eps;
  a.ifcharge = r.ifc;
  a.ifdipole = r.ifd;
  a.ifpgh = r.pg;
  a.ifpghtarg = r.pgt;
  a.ifnear = r.ifnear;
  a.mixed = mixed_for(r.eps);
  double2* cs = r.ifc ? S.get<double2>((size_t)ndc * ns) : nullptr;
  double2* qs = r.ifd ? S.get<double2>((size_t)ndc * ns) : nullptr;
  if (S.err) return S.err;
  if (r.realk) {
    k_rfmm_prep<<<nblk(ns, 256), 256, 0, c.st>>>(rchg, rdip, dv, P.perm_s, ns, nd, cs, qs);
  } else if (r.logk) {
    k_lfmm_prep<<<nblk(ns, 256), 256, 0, c.st>>>(chg, dip, dv, P.perm_s, ns, nd, cs, qs);
  } else {
    if (cs) gather_channels(c.st, chg, cs, P.perm_s, ns, nd, 1);
    if (qs) gather_channels(c.st, dip, qs, P.perm_s, ns, nd, 1);
  }
  a.chg = cs;
  a.dip = qs;
  auto zeros = [&](int flag, int need, int n) -> double2* {
    if (flag < need) return nullptr;
    size_t cnt = (size_t)ndc * std::max(n, 1);
    double2* p = S.get<double2>(cnt);
    if (p) fill_async(p, 0u, 4 * cnt, c.st);
    return p;
  };
  a.pot = zeros(r.pg, 1, ns); a.grad = zeros(r.pg, 2, ns); a.hess = zeros(r.pg, 3, ns);
  a.pott = zeros(r.pgt, 1, nt); a.gradt = zeros(r.pgt, 2, nt); a.hesst = zeros(r.pgt, 3, nt);
  if (S.err) return S.err;
  // results in original order, in device memory (caller's, or staging),
  // written by the near field where it runs (LapArgs::out_s / out_t)
  struct Out { int flag, n; const int* perm; double2 *sp, *sg, *sh; void *up, *ug, *uh;
               void *dp, *dg, *dh; };
  Out outs[2] = {{r.pg, ns, P.perm_s, a.pot, a.grad, a.hess, r.pot, r.grad, r.hess},
                 {r.pgt, nt, P.perm_t, a.pott, a.gradt, a.hesst, r.pott, r.gradt, r.hesst}};
  LapOut* lo[2] = {&a.out_s, &a.out_t};
  for (int k = 0; k < 2; ++k) {
    Out& o = outs[k];
    o.dp = o.dg = o.dh = nullptr;
    if (o.flag <= 0 || o.n == 0) continue;
    const size_t np = (size_t)nd * o.n, ng = np * ncomp_g, nh = np * ncomp_h;
    o.dp = o.up;
    o.dg = o.flag >= 2 ? o.ug : nullptr;
    o.dh = o.flag >= 3 ? o.uh : nullptr;
    if (!r.device) {
      o.dp = S.get<char>(esz * np);
      o.dg = o.flag >= 2 ? S.get<char>(esz * ng) : nullptr;
      o.dh = o.flag >= 3 ? S.get<char>(esz * nh) : nullptr;
      if (S.err) return S.err;
    }
    lo[k]->kind = r.realk ? 2 : r.logk ? 1 : 0;
    lo[k]->nd = nd;
    lo[k]->perm = o.perm;
    lo[k]->pot = o.dp;
    lo[k]->grad = o.dg;
    lo[k]->hess = o.dh;
  }
  e = lap_run(c, S, P, plan, a);
  if (e) return e;
  const bool fused[2] = {a.fused_s, a.fused_t};
  Downloads down;
  for (int k = 0; k < 2; ++k) {
    const Out& o = outs[k];
    if (o.flag <= 0 || o.n == 0) continue;
    const size_t np = (size_t)nd * o.n, ng = np * ncomp_g, nh = np * ncomp_h;
    if (fused[k]) {
      // written by the near field
    } else if (r.realk) {
      k_rfmm_unpack<<<nblk(o.n, 256), 256, 0, c.st>>>(
          o.sp, o.sg, o.sh, o.perm, o.n, nd, o.flag, static_cast<double*>(o.dp),
          static_cast<double*>(o.dg), static_cast<double*>(o.dh));
    } else if (r.logk) {
      k_lfmm_unpack<<<nblk(o.n, 256), 256, 0, c.st>>>(
          o.sp, o.sg, o.sh, o.perm, o.n, nd, o.flag, static_cast<double2*>(o.dp),
          static_cast<double2*>(o.dg), static_cast<double2*>(o.dh));
    } else {
      scatter_channels(c.st, o.sp, static_cast<double2*>(o.dp), o.perm, o.n, nd, 1);
      if (o.flag >= 2) scatter_channels(c.st, o.sg, static_cast<double2*>(o.dg), o.perm, o.n, nd, 1);
      if (o.flag >= 3) scatter_channels(c.st, o.sh, static_cast<double2*>(o.dh), o.perm, o.n, nd, 1);
    }
    if (!r.device) {
      down.add(o.dp, o.up, esz * np);
      if (o.flag >= 2) down.add(o.dg, o.ug, esz * ng);
      if (o.flag >= 3) down.add(o.dh, o.uh, esz * nh);
    }
  }
  if (!r.device && (e = down.flush(c.st))) return e;
  HIP_TRY(hipGetLastError());
  return IER_OK;
}

namespace {

int run_lap(const LapReq& r) {
  Prepared q;
  int e = get_ctx(q.c);
  if (e) return e;
  if (r.stream) q.c.st = r.stream;
  Scratch S(q.c.st);
  if ((e = lap_setup(r, S, q))) return e;
  if ((e = lap_eval(r, S, q))) return e;
  HIP_TRY(hipStreamSynchronize(q.c.st));
  finish_timer(r.timeinfo);
  return sort_status(q.plan);
}

int guarded(const LapReq& r) {
  try {
    const double t0 = wall();
    int e = run_lap(r);
    last_stats().host_s[7] = wall() - t0;
    return e;
  } catch (const std::bad_alloc&) {
    return IER_ALLOC;
  } catch (...) {
    return IER_HIP;
  }
}



}  // namespace

// Helmholtz front end (as lap_setup).
int helm_setup(const HelmReq& r, Scratch& S, Prepared& q) {
  Ctx& c = q.c;
  if (r.nd < 1 || r.ns < 1 || r.nt < 0 || r.pg < 0 || r.pg > 3 || r.pgt < 0 || r.pgt > 3)
    return IER_ARGS;
  int e = IER_OK;
  const int nd = r.nd, ns = r.ns, nt = r.nt;
  q.src = r.src;
  q.trg = r.trg;
  q.dv = r.dipvec;
  q.chg = r.chg;
  q.dip = r.dip;
  if (!r.device) {
    const double2 *cc = static_cast<const double2*>(r.chg), *qq = static_cast<const double2*>(r.dip);
    Uploads up;
    if ((e = up.add(S, r.src, 2 * (size_t)ns, &q.src))) return e;
    if ((e = up.add(S, r.trg, 2 * (size_t)nt, &q.trg))) return e;
    if (r.ifc && (e = up.add(S, cc, (size_t)nd * ns, &cc))) return e;
    if (r.ifd && (e = up.add(S, qq, (size_t)nd * ns, &qq))) return e;
    if (r.ifd && (e = up.add(S, r.dipvec, 2 * (size_t)nd * ns, &q.dv))) return e;
    if ((e = up.flush(c.st))) return e;
    q.chg = cc;
    q.dip = qq;
  }
  return build_plan(c, S, q.src, ns, q.trg, nt, r.idivflag, r.ndiv, q.P, q.plan, true, r.pg,
                    r.pgt, r.forest);
}

// Helmholtz evaluation after helm_setup (as lap_eval).
int helm_eval(const HelmReq& r, Scratch& S, Prepared& q) {
  Ctx& c = q.c;
  SortedPoints& P = q.P;
  Plan& plan = q.plan;
  int e = IER_OK;
  const int nd = r.nd, ns = r.ns, nt = r.nt;
  const double* dv = q.dv;
  const double2 *chg = static_cast<const double2*>(q.chg), *dip = static_cast<const double2*>(q.dip);
  HelmArgs a;
  a.nd = nd; a.eps = r.eps; a.zk = r.zk;
  a.ifcharge = r.ifc; a.ifdipole = r.ifd; a.ifpgh = r.pg; a.ifpghtarg = r.pgt; a.ifnear = r.ifnear;
  a.mixed = mixed_for(r.eps);
  double2* cs = r.ifc ? S.get<double2>((size_t)nd * ns) : nullptr;
  double2* qs = r.ifd ? S.get<double2>((size_t)nd * ns) : nullptr;
  double* vs = r.ifd ? S.get<double>(2 * (size_t)nd * ns) : nullptr;
  if (S.err) return S.err;
  gather_sources(c.st, chg, cs, dip, qs, dv, vs, P.perm_s, ns, nd);
  a.chg = cs; a.dip = qs; a.dipvec = vs;
  auto zeros = [&](int flag, int need, int n, int ncomp) -> double2* {
    if (flag < need) return nullptr;
    size_t cnt = (size_t)nd * ncomp * std::max(n, 1);
    double2* p = S.get<double2>(cnt);
    if (p) fill_async(p, 0u, 4 * cnt, c.st);
    return p;
  };
  a.pot = zeros(r.pg, 1, ns, 1); a.grad = zeros(r.pg, 2, ns, 2); a.hess = zeros(r.pg, 3, ns, 3);
  a.pott = zeros(r.pgt, 1, nt, 1); a.gradt = zeros(r.pgt, 2, nt, 2); a.hesst = zeros(r.pgt, 3, nt, 3);
  if (S.err) return S.err;
  // results in original order, in device memory (caller's, or staging),
  // written by the near field where it runs (HelmArgs::upot)
  struct Out { int flag, n; const int* perm; double2 *sp, *sg, *sh; void *up, *ug, *uh;
               double2 *dp, *dg, *dh; };
  Out outs[2] = {{r.pg, ns, P.perm_s, a.pot, a.grad, a.hess, r.pot, r.grad, r.hess},
                 {r.pgt, nt, P.perm_t, a.pott, a.gradt, a.hesst, r.pott, r.gradt, r.hesst}};
  for (auto& o : outs) {
    o.dp = o.dg = o.dh = nullptr;
    if (o.flag <= 0 || o.n == 0) continue;
    const size_t np = (size_t)nd * o.n;
    o.dp = static_cast<double2*>(o.up);
    o.dg = o.flag >= 2 ? static_cast<double2*>(o.ug) : nullptr;
    o.dh = o.flag >= 3 ? static_cast<double2*>(o.uh) : nullptr;
    if (!r.device) {
      o.dp = S.get<double2>(np);
      o.dg = o.flag >= 2 ? S.get<double2>(2 * np) : nullptr;
      o.dh = o.flag >= 3 ? S.get<double2>(3 * np) : nullptr;
      if (S.err) return S.err;
    }
  }
  a.upot = outs[0].dp; a.ugrad = outs[0].dg; a.uhess = outs[0].dh;
  a.upott = outs[1].dp; a.ugradt = outs[1].dg; a.uhesst = outs[1].dh;
  e = helm_run(c, S, P, plan, a);
  if (e) return e;
  const bool fused[2] = {a.fused_s, a.fused_t};
  Downloads down;
  for (int k = 0; k < 2; ++k) {
    const Out& o = outs[k];
    if (o.flag <= 0 || o.n == 0) continue;
    const size_t np = (size_t)nd * o.n;
    if (!fused[k]) {
      scatter_channels(c.st, o.sp, o.dp, o.perm, o.n, nd, 1);
      if (o.flag >= 2) scatter_channels(c.st, o.sg, o.dg, o.perm, o.n, nd, 2);
      if (o.flag >= 3) scatter_channels(c.st, o.sh, o.dh, o.perm, o.n, nd, 3);
    }
    if (!r.device) {
      down.add(o.dp, o.up, sizeof(double2) * np);
      if (o.flag >= 2) down.add(o.dg, o.ug, sizeof(double2) * 2 * np);
      if (o.flag >= 3) down.add(o.dh, o.uh, sizeof(double2) * 3 * np);
    }
  }
  if (!r.device && (e = down.flush(c.st))) return e;
  HIP_TRY(hipGetLastError());
  return IER_OK;
}

namespace {

int run_helm(const HelmReq& r) {
  Prepared q;
  int e = get_ctx(q.c);
  if (e) return e;
  if (r.stream) q.c.st = r.stream;
  Scratch S(q.c.st);
  if ((e = helm_setup(r, S, q))) return e;
  if ((e = helm_eval(r, S, q))) return e;
  HIP_TRY(hipStreamSynchronize(q.c.st));
  finish_timer(r.timeinfo);
  return sort_status(q.plan);
}

int guarded_h(const HelmReq& r) {
  try {
    const double t0 = wall();
    int e = run_helm(r);
    last_stats().host_s[7] = wall() - t0;
    return e;
  } catch (const std::bad_alloc&) {
    return IER_ALLOC;
  } catch (...) {
    return IER_HIP;
  }
}

void zero8(double* t) {
  if (t)
    for (int i = 0; i < 8; ++i) t[i] = 0.0;
}

}  // namespace

// One problem of the batched/served interface on the calling thread, with
// the problem's precision mode for the duration of the solve; with a
// forest, the problem is the concatenation of the forest's problems.
void helm_req_of(const fmm2dhip_problem& p, const ForestIn* forest, HelmReq& r) {
  r.nd = p.nd; r.eps = p.eps; r.zk = std::complex<double>(p.zk_re, p.zk_im);
  r.ns = p.ns; r.nt = p.nt; r.src = p.sources; r.trg = p.targ;
  r.ifc = p.ifcharge; r.ifd = p.ifdipole; r.chg = p.charge; r.dip = p.dipstr; r.dipvec = p.dipvec;
  r.pg = p.ifpgh; r.pgt = p.ifpghtarg;
  r.pot = p.pot; r.grad = p.grad; r.hess = p.hess;
  r.pott = p.pottarg; r.gradt = p.gradtarg; r.hesst = p.hesstarg;
  r.ndiv = p.ndiv > 0 ? p.ndiv : ndiv_rule(p.eps, p.ns, p.nt);
  r.device = p.device_ptrs == 1;
  r.forest = forest;
}

void lap_req_of(const fmm2dhip_problem& p, const ForestIn* forest, LapReq& r) {
  r.logk = p.kernel == FMM2DHIP_LOG;
  r.realk = p.kernel == FMM2DHIP_REAL;
  r.nd = p.nd; r.eps = p.eps; r.ns = p.ns; r.nt = p.nt; r.src = p.sources; r.trg = p.targ;
  r.ifc = p.ifcharge; r.ifd = p.ifdipole; r.chg = p.charge; r.dip = p.dipstr; r.dipvec = p.dipvec;
  r.pg = p.ifpgh; r.pgt = p.ifpghtarg;
  r.pot = p.pot; r.grad = p.grad; r.hess = p.hess;
  r.pott = p.pottarg; r.gradt = p.gradtarg; r.hesst = p.hesstarg;
  r.ndiv = p.ndiv > 0 ? p.ndiv : ndiv_rule(p.eps, p.ns, p.nt);
  r.device = p.device_ptrs == 1;
  r.forest = forest;
}

int solve_problem(fmm2dhip_problem& p, const ForestIn* forest) {
  const double t0 = wall();
  int& mode = thread_precision();
  const int saved = mode;
  mode = p.precision == PREC_MIXED ? PREC_MIXED : PREC_FP64;
  int e;
  if (p.kernel == FMM2DHIP_HELMHOLTZ) {
    HelmReq r;
    helm_req_of(p, forest, r);
    e = guarded_h(r);
  } else if (p.kernel == FMM2DHIP_CAUCHY || p.kernel == FMM2DHIP_LOG ||
             p.kernel == FMM2DHIP_REAL) {
    LapReq r;
    lap_req_of(p, forest, r);
    e = guarded(r);
  } else {
    e = IER_ARGS;
  }
  mode = saved;
  p.ier = e;
  p.seconds = wall() - t0;
  return e;
}

}  // namespace fmm

using fmm::LapReq;

extern "C" {

void cfmm2d_ndiv_(const fmm2d_int* nd, const double* eps, const fmm2d_int* ns,
                  const double* sources, const fmm2d_int* ifcharge, const fmm2d_cplx* charge,
                  const fmm2d_int* ifdipole, const fmm2d_cplx* dipstr, fmm2d_int* iper,
                  const fmm2d_int* ifpgh, fmm2d_cplx* pot, fmm2d_cplx* grad, fmm2d_cplx* hess,
                  const fmm2d_int* nt, const double* targ, const fmm2d_int* ifpghtarg,
                  fmm2d_cplx* pottarg, fmm2d_cplx* gradtarg, fmm2d_cplx* hesstarg,
                  const fmm2d_int* ndiv, const fmm2d_int* idivflag, const fmm2d_int* ifnear,
                  double* timeinfo, fmm2d_int* ier) {
  *iper = 0;
  fmm::zero8(timeinfo);
  LapReq r;
  r.nd = *nd; r.eps = *eps; r.ns = *ns; r.nt = *nt; r.src = sources; r.trg = targ;
  r.ifc = *ifcharge; r.ifd = *ifdipole; r.chg = charge; r.dip = dipstr;
  r.pg = *ifpgh; r.pgt = *ifpghtarg;
  r.pot = pot; r.grad = grad; r.hess = hess; r.pott = pottarg; r.gradt = gradtarg; r.hesst = hesstarg;
  r.ndiv = *ndiv; r.idivflag = *idivflag; r.ifnear = *ifnear; r.timeinfo = timeinfo;
  *ier = fmm::guarded(r);
}

void cfmm2d_(const fmm2d_int* nd, const double* eps, const fmm2d_int* ns, const double* sources,
             const fmm2d_int* ifcharge, const fmm2d_cplx* charge, const fmm2d_int* ifdipole,
             const fmm2d_cplx* dipstr, fmm2d_int* iper, const fmm2d_int* ifpgh, fmm2d_cplx* pot,
             fmm2d_cplx* grad, fmm2d_cplx* hess, const fmm2d_int* nt, const double* targ,
             const fmm2d_int* ifpghtarg, fmm2d_cplx* pottarg, fmm2d_cplx* gradtarg,
             fmm2d_cplx* hesstarg, fmm2d_int* ier) {
  fmm2d_int ndiv = fmm::ndiv_rule(*eps, *ns, *nt), idiv = 0, ifnear = 1;
  double ti[8];
  cfmm2d_ndiv_(nd, eps, ns, sources, ifcharge, charge, ifdipole, dipstr, iper, ifpgh, pot, grad,
               hess, nt, targ, ifpghtarg, pottarg, gradtarg, hesstarg, &ndiv, &idiv, &ifnear, ti,
               ier);
}

void lfmm2d_ndiv_(const fmm2d_int* nd, const double* eps, const fmm2d_int* ns,
                  const double* sources, const fmm2d_int* ifcharge, const fmm2d_cplx* charge,
                  const fmm2d_int* ifdipole, const fmm2d_cplx* dipstr, const double* dipvec,
                  fmm2d_int* iper, const fmm2d_int* ifpgh, fmm2d_cplx* pot, fmm2d_cplx* grad,
                  fmm2d_cplx* hess, const fmm2d_int* nt, const double* targ,
                  const fmm2d_int* ifpghtarg, fmm2d_cplx* pottarg, fmm2d_cplx* gradtarg,
                  fmm2d_cplx* hesstarg, const fmm2d_int* ndiv, const fmm2d_int* idivflag,
                  const fmm2d_int* ifnear, double* timeinfo, fmm2d_int* ier) {
  *iper = 0;
  fmm::zero8(timeinfo);
  LapReq r;
  r.logk = true;
  r.nd = *nd; r.eps = *eps; r.ns = *ns; r.nt = *nt; r.src = sources; r.trg = targ;
  r.ifc = *ifcharge; r.ifd = *ifdipole; r.chg = charge; r.dip = dipstr; r.dipvec = dipvec;
  r.pg = *ifpgh; r.pgt = *ifpghtarg;
  r.pot = pot; r.grad = grad; r.hess = hess; r.pott = pottarg; r.gradt = gradtarg; r.hesst = hesstarg;
  r.ndiv = *ndiv; r.idivflag = *idivflag; r.ifnear = *ifnear; r.timeinfo = timeinfo;
  *ier = fmm::guarded(r);
}

void lfmm2d_(const fmm2d_int* nd, const double* eps, const fmm2d_int* ns, const double* sources,
             const fmm2d_int* ifcharge, const fmm2d_cplx* charge, const fmm2d_int* ifdipole,
             const fmm2d_cplx* dipstr, const double* dipvec, fmm2d_int* iper,
             const fmm2d_int* ifpgh, fmm2d_cplx* pot, fmm2d_cplx* grad, fmm2d_cplx* hess,
             const fmm2d_int* nt, const double* targ, const fmm2d_int* ifpghtarg,
             fmm2d_cplx* pottarg, fmm2d_cplx* gradtarg, fmm2d_cplx* hesstarg, fmm2d_int* ier) {
  fmm2d_int ndiv = fmm::ndiv_rule(*eps, *ns, *nt), idiv = 0, ifnear = 1;
  double ti[8];
  lfmm2d_ndiv_(nd, eps, ns, sources, ifcharge, charge, ifdipole, dipstr, dipvec, iper, ifpgh, pot,
               grad, hess, nt, targ, ifpghtarg, pottarg, gradtarg, hesstarg, &ndiv, &idiv,
               &ifnear, ti, ier);
}


void rfmm2d_ndiv_(const fmm2d_int* nd, const double* eps, const fmm2d_int* ns,
                  const double* sources, const fmm2d_int* ifcharge, const double* charge,
                  const fmm2d_int* ifdipole, const double* dipstr, const double* dipvec,
                  fmm2d_int* iper, const fmm2d_int* ifpgh, double* pot, double* grad,
                  double* hess, const fmm2d_int* nt, const double* targ,
                  const fmm2d_int* ifpghtarg, double* pottarg, double* gradtarg,
                  double* hesstarg, const fmm2d_int* ndiv, const fmm2d_int* idivflag,
                  const fmm2d_int* ifnear, double* timeinfo, fmm2d_int* ier) {
  *iper = 0;
  fmm::zero8(timeinfo);
  LapReq r;
  r.realk = true;
  r.nd = *nd; r.eps = *eps; r.ns = *ns; r.nt = *nt; r.src = sources; r.trg = targ;
  r.ifc = *ifcharge; r.ifd = *ifdipole; r.chg = charge; r.dip = dipstr; r.dipvec = dipvec;
  r.pg = *ifpgh; r.pgt = *ifpghtarg;
  r.pot = pot; r.grad = grad; r.hess = hess; r.pott = pottarg; r.gradt = gradtarg; r.hesst = hesstarg;
  r.ndiv = *ndiv; r.idivflag = *idivflag; r.ifnear = *ifnear; r.timeinfo = timeinfo;
  *ier = fmm::guarded(r);
}

void rfmm2d_(const fmm2d_int* nd, const double* eps, const fmm2d_int* ns, const double* sources,
             const fmm2d_int* ifcharge, const double* charge, const fmm2d_int* ifdipole,
             const double* dipstr, const double* dipvec, fmm2d_int* iper, const fmm2d_int* ifpgh,
             double* pot, double* grad, double* hess, const fmm2d_int* nt, const double* targ,
             const fmm2d_int* ifpghtarg, double* pottarg, double* gradtarg, double* hesstarg,
             fmm2d_int* ier) {
  fmm2d_int ndiv = fmm::ndiv_rule(*eps, *ns, *nt), idiv = 0, ifnear = 1;
  double ti[8];
  rfmm2d_ndiv_(nd, eps, ns, sources, ifcharge, charge, ifdipole, dipstr, dipvec, iper, ifpgh, pot,
               grad, hess, nt, targ, ifpghtarg, pottarg, gradtarg, hesstarg, &ndiv, &idiv,
               &ifnear, ti, ier);
}

void hfmm2d_ndiv_(const fmm2d_int* nd, const double* eps, const fmm2d_cplx* zk, const fmm2d_int* ns,
                  const double* sources, const fmm2d_int* ifcharge, const fmm2d_cplx* charge,
                  const fmm2d_int* ifdipole, const fmm2d_cplx* dipstr, const double* dipvec,
                  fmm2d_int* iper, const fmm2d_int* ifpgh, fmm2d_cplx* pot, fmm2d_cplx* grad,
                  fmm2d_cplx* hess, const fmm2d_int* nt, const double* targ,
                  const fmm2d_int* ifpghtarg, fmm2d_cplx* pottarg, fmm2d_cplx* gradtarg,
                  fmm2d_cplx* hesstarg, const fmm2d_int* ndiv, const fmm2d_int* idivflag,
                  const fmm2d_int* ifnear, double* timeinfo, fmm2d_int* ier) {
  *iper = 0;
  fmm::zero8(timeinfo);
  fmm::HelmReq r;
  r.nd = *nd; r.eps = *eps; r.zk = *zk; r.ns = *ns; r.nt = *nt; r.src = sources; r.trg = targ;
  r.ifc = *ifcharge; r.ifd = *ifdipole; r.chg = charge; r.dip = dipstr; r.dipvec = dipvec;
  r.pg = *ifpgh; r.pgt = *ifpghtarg;
  r.pot = pot; r.grad = grad; r.hess = hess; r.pott = pottarg; r.gradt = gradtarg; r.hesst = hesstarg;
  r.ndiv = *ndiv; r.idivflag = *idivflag; r.ifnear = *ifnear; r.timeinfo = timeinfo;
  *ier = fmm::guarded_h(r);
}

void hfmm2d_(const fmm2d_int* nd, const double* eps, const fmm2d_cplx* zk, const fmm2d_int* ns,
             const double* sources, const fmm2d_int* ifcharge, const fmm2d_cplx* charge,
             const fmm2d_int* ifdipole, const fmm2d_cplx* dipstr, const double* dipvec,
             fmm2d_int* iper, const fmm2d_int* ifpgh, fmm2d_cplx* pot, fmm2d_cplx* grad,
             fmm2d_cplx* hess, const fmm2d_int* nt, const double* targ, const fmm2d_int* ifpghtarg,
             fmm2d_cplx* pottarg, fmm2d_cplx* gradtarg, fmm2d_cplx* hesstarg, fmm2d_int* ier) {
  fmm2d_int ndiv = fmm::ndiv_rule(*eps, *ns, *nt), idiv = 0, ifnear = 1;
  double ti[8];
  hfmm2d_ndiv_(nd, eps, zk, ns, sources, ifcharge, charge, ifdipole, dipstr, dipvec, iper, ifpgh,
               pot, grad, hess, nt, targ, ifpghtarg, pottarg, gradtarg, hesstarg, &ndiv, &idiv,
               &ifnear, ti, ier);
}

void h2dterms_(const double* bsize, const fmm2d_cplx* zk, const double* eps, fmm2d_int* nterms,
               fmm2d_int* ier) {
  int e = 0;
  *nterms = fmm::h2dterms_host(*bsize, *zk, *eps, &e);
  *ier = e;
}

void fmm2dhip_hfmm2d_dev(int nd, double eps, double zk_re, double zk_im, int ns,
                         const double* sources, int ifcharge, const void* charge, int ifdipole,
                         const void* dipstr, const double* dipvec, int ifpgh, void* pot,
                         void* grad, void* hess, int nt, const double* targ, int ifpghtarg,
                         void* pottarg, void* gradtarg, void* hesstarg, int ndiv, int idivflag,
                         int ifnear, double* timeinfo, void* stream, int* ier) {
  fmm::zero8(timeinfo);
  fmm::HelmReq r;
  r.device = true;
  r.stream = static_cast<hipStream_t>(stream);
  r.nd = nd; r.eps = eps; r.zk = std::complex<double>(zk_re, zk_im); r.ns = ns; r.nt = nt;
  r.src = sources; r.trg = targ;
  r.ifc = ifcharge; r.ifd = ifdipole; r.chg = charge; r.dip = dipstr; r.dipvec = dipvec;
  r.pg = ifpgh; r.pgt = ifpghtarg;
  r.pot = pot; r.grad = grad; r.hess = hess; r.pott = pottarg; r.gradt = gradtarg; r.hesst = hesstarg;
  r.ndiv = ndiv; r.idivflag = idivflag; r.ifnear = ifnear; r.timeinfo = timeinfo;
  *ier = fmm::guarded_h(r);
}

void hndiv2d_(const double* eps, const fmm2d_int* ns, const fmm2d_int* nt, const fmm2d_int*,
              const fmm2d_int*, const fmm2d_int*, const fmm2d_int*, fmm2d_int* ndiv,
              fmm2d_int* idivflag) {
  *idivflag = 0;
  *ndiv = fmm::ndiv_rule(*eps, *ns, *nt);
}

void lndiv2d_(const double* eps, const fmm2d_int* ns, const fmm2d_int* nt, const fmm2d_int*,
              const fmm2d_int*, const fmm2d_int*, const fmm2d_int*, fmm2d_int* ndiv,
              fmm2d_int* idivflag) {
  *idivflag = 0;
  *ndiv = fmm::ndiv_rule(*eps, *ns, *nt);
}

void l2dterms_(const double* eps, fmm2d_int* nterms, fmm2d_int* ier) {
  *ier = 0;
  *nterms = fmm::l2dterms(*eps);
}

void fmm2dhip_cfmm2d_dev(int nd, double eps, int ns, const double* sources, int ifcharge,
                         const void* charge, int ifdipole, const void* dipstr, int ifpgh,
                         void* pot, void* grad, void* hess, int nt, const double* targ,
                         int ifpghtarg, void* pottarg, void* gradtarg, void* hesstarg, int ndiv,
                         int idivflag, int ifnear, double* timeinfo, void* stream, int* ier) {
  fmm::zero8(timeinfo);
  LapReq r;
  r.device = true;
  r.stream = static_cast<hipStream_t>(stream);
  r.nd = nd; r.eps = eps; r.ns = ns; r.nt = nt; r.src = sources; r.trg = targ;
  r.ifc = ifcharge; r.ifd = ifdipole; r.chg = charge; r.dip = dipstr;
  r.pg = ifpgh; r.pgt = ifpghtarg;
  r.pot = pot; r.grad = grad; r.hess = hess; r.pott = pottarg; r.gradt = gradtarg; r.hesst = hesstarg;
  r.ndiv = ndiv; r.idivflag = idivflag; r.ifnear = ifnear; r.timeinfo = timeinfo;
  *ier = fmm::guarded(r);
}

void fmm2dhip_lfmm2d_dev(int nd, double eps, int ns, const double* sources, int ifcharge,
                         const void* charge, int ifdipole, const void* dipstr,
                         const double* dipvec, int ifpgh, void* pot, void* grad, void* hess,
                         int nt, const double* targ, int ifpghtarg, void* pottarg,
                         void* gradtarg, void* hesstarg, int ndiv, int idivflag, int ifnear,
                         double* timeinfo, void* stream, int* ier) {
  fmm::zero8(timeinfo);
  LapReq r;
  r.logk = true;
  r.device = true;
  r.stream = static_cast<hipStream_t>(stream);
  r.nd = nd; r.eps = eps; r.ns = ns; r.nt = nt; r.src = sources; r.trg = targ;
  r.ifc = ifcharge; r.ifd = ifdipole; r.chg = charge; r.dip = dipstr; r.dipvec = dipvec;
  r.pg = ifpgh; r.pgt = ifpghtarg;
  r.pot = pot; r.grad = grad; r.hess = hess; r.pott = pottarg; r.gradt = gradtarg; r.hesst = hesstarg;
  r.ndiv = ndiv; r.idivflag = idivflag; r.ifnear = ifnear; r.timeinfo = timeinfo;
  *ier = fmm::guarded(r);
}

int fmm2dhip_test_tree(const double* src, int ns, const double* targ, int nt, int idivflag,
                       int ndiv, int* nlevels, int* nboxes, int* itree, double* centers,
                       double* boxsize, int* isrc, int* isrcse, int* itarg, int* itargse,
                       int* nlist, int* lists) {
  try {
    double xmin = src[0], xmax = src[0], ymin = src[1], ymax = src[1];
    for (int pass = 0; pass < 2; ++pass) {
      const double* a = pass ? targ : src;
      int n = pass ? nt : ns;
      for (int i = 0; i < n; ++i) {
        xmin = std::min(xmin, a[2 * i]); xmax = std::max(xmax, a[2 * i]);
        ymin = std::min(ymin, a[2 * i + 1]); ymax = std::max(ymax, a[2 * i + 1]);
      }
    }
    double bs0 = std::max(xmax - xmin, ymax - ymin), cx = (xmin + xmax) / 2, cy = (ymin + ymax) / 2;
    auto sorted = [&](const double* a, int n, std::vector<uint64_t>& k, std::vector<int>& idx) {
      std::vector<uint64_t> raw(n);
      for (int i = 0; i < n; ++i) raw[i] = fmm::point_key(a[2 * i], a[2 * i + 1], cx, cy, bs0);
      idx.resize(n);
      std::iota(idx.begin(), idx.end(), 0);
      std::stable_sort(idx.begin(), idx.end(), [&](int x, int y) { return raw[x] < raw[y]; });
      k.resize(n);
      for (int i = 0; i < n; ++i) k[i] = raw[idx[i]];
    };
    std::vector<uint64_t> ks, kt;
    std::vector<int> is, it;
    sorted(src, ns, ks, is);
    sorted(targ, nt, kt, it);
    fmm::Tree T;
    int e = fmm::build_tree(ks.data(), ns, kt.data(), nt, cx, cy, bs0, idivflag, ndiv, T);
    if (e) return e;
    *nlevels = T.nlevels;
    *nboxes = T.nboxes;
    if (!itree) return 0;
    fmm::build_lists(T);
    const int n = T.nboxes, L = T.nlevels;
    int k = 0;
    auto id1 = [](int b) { return b >= 0 ? b + 1 : -1; };
    for (int l = 0; l <= L; ++l) {
      itree[k++] = T.lfirst[l] + 1;
      itree[k++] = T.lfirst[l] + T.lcount[l];
    }
    for (int b = 0; b < n; ++b) itree[k++] = T.level[b];
    for (int b = 0; b < n; ++b) itree[k++] = id1(T.parent[b]);
    for (int b = 0; b < n; ++b) itree[k++] = T.nchild[b];
    for (int b = 0; b < n; ++b)
      for (int q = 0; q < 4; ++q) itree[k++] = id1(T.child[4 * b + q]);
    for (int b = 0; b < n; ++b) itree[k++] = T.nnb[b];
    for (int b = 0; b < n; ++b)
      for (int i = 0; i < 9; ++i) itree[k++] = id1(T.nbr[9 * b + i]);
    for (int b = 0; b < n; ++b) {
      centers[2 * b] = T.cx[b];
      centers[2 * b + 1] = T.cy[b];
      isrcse[2 * b] = T.s0[b] + 1;
      isrcse[2 * b + 1] = T.s1[b];
      itargse[2 * b] = T.t0[b] + 1;
      itargse[2 * b + 1] = T.t1[b];
      nlist[4 * b + 0] = T.n1[b]; nlist[4 * b + 1] = T.n2[b];
      nlist[4 * b + 2] = T.n3[b]; nlist[4 * b + 3] = T.n4[b];
      int* o = lists + 65 * b;
      for (int i = 0; i < 13; ++i) o[i] = i < T.n1[b] ? id1(T.l1[13 * b + i]) : 0;
      for (int i = 0; i < 27; ++i) o[13 + i] = i < T.n2[b] ? id1(T.l2[27 * b + i]) : 0;
      for (int i = 0; i < 20; ++i) o[40 + i] = i < T.n3[b] ? id1(T.l3[20 * b + i]) : 0;
      for (int i = 0; i < 5; ++i) o[60 + i] = i < T.n4[b] ? id1(T.l4[5 * b + i]) : 0;
    }
    for (int l = 0; l <= L; ++l) boxsize[l] = T.bs[l];
    // original-index order inside each leaf
    for (int b = 0; b < n; ++b) {
      if (T.nchild[b]) continue;
      std::sort(is.begin() + T.s0[b], is.begin() + T.s1[b]);
      std::sort(it.begin() + T.t0[b], it.begin() + T.t1[b]);
    }
    for (int i = 0; i < ns; ++i) isrc[i] = is[i] + 1;
    for (int i = 0; i < nt; ++i) itarg[i] = it[i] + 1;
    return 0;
  } catch (...) {
    return fmm::IER_ALLOC;
  }
}

int fmm2dhip_test_devtree(const double* src, int ns, const double* targ, int nt, int idivflag,
                          int ndiv, int cap_boxes, int cap_lists, int* nlevels, int* nboxes,
                          int* nlists, int* boxinfo, double* centers, int* lists, int* isrc,
                          int* itarg) {
  try {
    fmm::Ctx c;
    int e = fmm::get_ctx(c);
    if (e) return e;
    if (ns < 1 || nt < 0) return fmm::IER_ARGS;
    int out = 0;
    {
      fmm::Scratch S(c.st);
      const double *ds = nullptr, *dt = nullptr;
      if ((e = fmm::upload(S, c.st, src, 2 * (size_t)ns, &ds))) return e;
      if ((e = fmm::upload(S, c.st, targ, 2 * (size_t)nt, &dt))) return e;
      fmm::SortedPoints P;
      fmm::Plan plan;
      e = fmm::build_plan(c, S, ds, ns, dt, nt, idivflag, ndiv, P, plan, true, 1, 1);
      if (e) return e;
      const fmm::DTree& D = plan.D;
      const int nb = plan.nb;
      std::vector<int> o(4 * ((size_t)nb + 1));
      const int* od[4] = {D.o1, D.o2, D.o3, D.o4};
      for (int k = 0; k < 4; ++k)
        HIP_TRY(hipMemcpyAsync(o.data() + (size_t)k * (nb + 1), od[k], sizeof(int) * (nb + 1),
                               hipMemcpyDeviceToHost, c.st));
      HIP_TRY(hipStreamSynchronize(c.st));
      if ((e = fmm::sort_status(plan))) return e;
      int tot = 0;
      for (int k = 0; k < 4; ++k) tot += o[(size_t)k * (nb + 1) + nb];
      *nlevels = plan.L;
      *nboxes = nb;
      *nlists = tot;
      if (nb > cap_boxes || tot > cap_lists) {
        out = fmm::IER_ARGS;
      } else {
        std::vector<uint64_t> key(nb);
        std::vector<int> level(nb), nch(nb);
        std::vector<int2> sr(nb), tr(nb);
        std::vector<double2> ctr(nb);
        std::vector<std::vector<int>> lv(4);
        const int* ld[4] = {D.l1, D.l2, D.l3, D.l4};
        for (int k = 0; k < 4; ++k) {
          int n = o[(size_t)k * (nb + 1) + nb];
          lv[k].resize(std::max(n, 1));
          HIP_TRY(hipMemcpyAsync(lv[k].data(), ld[k], sizeof(int) * n, hipMemcpyDeviceToHost, c.st));
        }
        HIP_TRY(hipMemcpyAsync(key.data(), D.key, sizeof(uint64_t) * nb, hipMemcpyDeviceToHost, c.st));
        HIP_TRY(hipMemcpyAsync(level.data(), D.level, sizeof(int) * nb, hipMemcpyDeviceToHost, c.st));
        HIP_TRY(hipMemcpyAsync(nch.data(), D.nchild, sizeof(int) * nb, hipMemcpyDeviceToHost, c.st));
        HIP_TRY(hipMemcpyAsync(sr.data(), D.srng, sizeof(int2) * nb, hipMemcpyDeviceToHost, c.st));
        HIP_TRY(hipMemcpyAsync(tr.data(), D.trng, sizeof(int2) * nb, hipMemcpyDeviceToHost, c.st));
        HIP_TRY(hipMemcpyAsync(ctr.data(), D.ctr, sizeof(double2) * nb, hipMemcpyDeviceToHost, c.st));
        HIP_TRY(hipMemcpyAsync(isrc, P.perm_s, sizeof(int) * ns, hipMemcpyDeviceToHost, c.st));
        if (nt) HIP_TRY(hipMemcpyAsync(itarg, P.perm_t, sizeof(int) * nt, hipMemcpyDeviceToHost, c.st));
        HIP_TRY(hipStreamSynchronize(c.st));
        int* lp = lists;
        for (int b = 0; b < nb; ++b) {
          int ix = 0, iy = 0;
          for (int j = 0; j < level[b]; ++j) {
            ix |= (int)((key[b] >> (2 * j)) & 1) << j;
            iy |= (int)((key[b] >> (2 * j + 1)) & 1) << j;
          }
          int* bi = boxinfo + 12 * (size_t)b;
          bi[0] = level[b]; bi[1] = ix; bi[2] = iy; bi[3] = nch[b];
          bi[4] = sr[b].x; bi[5] = sr[b].y; bi[6] = tr[b].x; bi[7] = tr[b].y;
          for (int k = 0; k < 4; ++k) {
            int a = o[(size_t)k * (nb + 1) + b], z = o[(size_t)k * (nb + 1) + b + 1];
            bi[8 + k] = z - a;
            for (int i = a; i < z; ++i) *lp++ = lv[k][i];
          }
          centers[2 * b] = ctr[b].x;
          centers[2 * b + 1] = ctr[b].y;
        }
      }
    }
    HIP_TRY(hipStreamSynchronize(c.st));
    return out;
  } catch (...) {
    return fmm::IER_ALLOC;
  }
}

int fmm2dhip_test_fft(int n, int nrows, int sign, const double* in, double* out) {
  try {
    if (n < 1 || nrows < 1 || (sign != 1 && sign != -1) || fmm::next235(n) != n)
      return fmm::IER_ARGS;
    fmm::Ctx c;
    int e = fmm::get_ctx(c);
    if (e) return e;
    fmm::Scratch S(c.st);
    const size_t m = (size_t)n * nrows;
    double2* d = S.get<double2>(m);
    if (S.err) return S.err;
    HIP_TRY(hipMemcpyAsync(d, in, sizeof(double2) * m, hipMemcpyHostToDevice, c.st));
    if ((e = fmm::fft_rows_device(c.st, S, d, nrows, n, sign))) return e;
    HIP_TRY(hipMemcpyAsync(out, d, sizeof(double2) * m, hipMemcpyDeviceToHost, c.st));
    HIP_TRY(hipStreamSynchronize(c.st));
    return fmm::IER_OK;
  } catch (...) {
    return fmm::IER_ALLOC;
  }
}

int fmm2dhip_test_sort(const void* keys, int n, int key_bytes, int begin, int end,
                       void* keys_out, int* vals_out) {
  try {
    if (n < 0 || (key_bytes != 4 && key_bytes != 8)) return fmm::IER_ARGS;
    if (n == 0) return fmm::IER_OK;
    fmm::Ctx c;
    int e = fmm::get_ctx(c);
    if (e) return e;
    uint32_t* flag = static_cast<uint32_t*>(fmm::pinned(sizeof(uint32_t), 6));
    if (!flag) return fmm::IER_ALLOC;
    *flag = 0;
    {
      fmm::Scratch S(c.st);
      int* vi = vals_out ? S.get<int>(n) : nullptr;
      int* vo = vals_out ? S.get<int>(n) : nullptr;
      if (S.err) return S.err;
      if (vi) {
        std::vector<int> iota(n);
        std::iota(iota.begin(), iota.end(), 0);
        HIP_TRY(hipMemcpyAsync(vi, iota.data(), sizeof(int) * n, hipMemcpyHostToDevice, c.st));
        HIP_TRY(hipStreamSynchronize(c.st));
      }
      if (key_bytes == 8) {
        const uint64_t* kin = nullptr;
        if ((e = fmm::upload(S, c.st, static_cast<const uint64_t*>(keys), n, &kin))) return e;
        uint64_t* kout = S.get<uint64_t>(n);
        if (S.err) return S.err;
        if ((e = fmm::rsort::radix_sort(c, S, kin, kout, vi, vo, n, begin, end, flag))) return e;
        HIP_TRY(hipMemcpyAsync(keys_out, kout, 8 * (size_t)n, hipMemcpyDeviceToHost, c.st));
      } else {
        const uint32_t* kin = nullptr;
        if ((e = fmm::upload(S, c.st, static_cast<const uint32_t*>(keys), n, &kin))) return e;
        uint32_t* kout = S.get<uint32_t>(n);
        if (S.err) return S.err;
        if ((e = fmm::rsort::radix_sort(c, S, kin, kout, vi, vo, n, begin, end, flag))) return e;
        HIP_TRY(hipMemcpyAsync(keys_out, kout, 4 * (size_t)n, hipMemcpyDeviceToHost, c.st));
      }
      if (vo) HIP_TRY(hipMemcpyAsync(vals_out, vo, sizeof(int) * n, hipMemcpyDeviceToHost, c.st));
      HIP_TRY(hipStreamSynchronize(c.st));
    }
    return *flag ? fmm::IER_HIP : fmm::IER_OK;
  } catch (...) {
    return fmm::IER_ALLOC;
  }
}

void fmm2dhip_release_workspace(void) { fmm::release_workspace(); }

int fmm2dhip_set_precision(int mode) {
  int& m = fmm::thread_precision();
  const int old = m;
  if (mode == fmm::PREC_FP64 || mode == fmm::PREC_MIXED) m = mode;
  return old;
}

int fmm2dhip_get_precision(void) { return fmm::thread_precision(); }

int fmm2dhip_last_stats(double* out, int n) {
  const fmm::SolveStats& s = fmm::last_stats();
  double v[26] = {s.p2p_pairs_src, s.p2p_pairs_trg, s.m2l_pairs, s.m2l_flops, s.nboxes,
                  s.nlevels, s.pmax};
  for (int k = 0; k < 8; ++k) v[7 + k] = s.step_s[k];
  for (int k = 0; k < 8; ++k) v[15 + k] = s.host_s[k];
  v[23] = s.workspace_bytes;
  v[24] = s.p2p_reads_src;
  v[25] = s.p2p_reads_trg;
  for (int k = 0; k < n && k < 26; ++k) out[k] = v[k];
  return 26;
}

int fmm2dhip_device_info(char* name, int len) {
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess || n < 1) return fmm::IER_NOGPU;
  hipDeviceProp_t p;
  int d = 0;
  if (hipGetDevice(&d) != hipSuccess || hipGetDeviceProperties(&p, d) != hipSuccess)
    return fmm::IER_NOGPU;
  if (name && len > 0) {
    std::strncpy(name, p.gcnArchName, (size_t)len - 1);
    name[len - 1] = 0;
  }
  return std::strncmp(p.gcnArchName, "gfx950", 6) == 0 ? 0 : fmm::IER_NOGPU;
}

}  // extern "C"

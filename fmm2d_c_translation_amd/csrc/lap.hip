// lap.hip -- Cauchy/Laplace FMM kernels (FP64) and their launch sequence.
//
// Expansion conventions are those of c_translation/src/laprouts2d.c:
//   multipole about c, scale r:  phi(z) = mp_0 log|z-c| + sum_n mp_n (r/(z-c))^n
//   local about c, scale r:      phi(z) = sum_n lo_n ((z-c)/r)^n
// with r the box size of the level (cfmm2d.c:1206-1211).  Expansions live in
// HBM box-major: [box][channel][0..p] complex.  Sorted point data is
// channel-major ([channel][point]).
#include <algorithm>
#include <cmath>
#include <complex>
#include <vector>
#include <cstring>
#include <map>
#include <mutex>

#include "engine.h"
#include "translate.h"

namespace fmm {

int l2dterms(double eps) {
  // smallest j >= 2 with (sqrt(2)/2)^j / 1.5^(j+1) < eps (l2dterms.c:41-66)
  for (int j = 2; j <= 100; ++j) {
    double v = std::pow(std::sqrt(2.0) / 2.0, (double)j) * (1.0 / std::pow(1.5, (double)(j + 1)));
    if (v < eps) return j;
  }
  return 1;
}

int ndiv_rule(double eps, int ns, int nt) {
  // leaf size table (src/laplace/lndiv2d.f:17-35, hndiv2d.c:31-47)
  static const double lim[] = {0.5, 0.5e-1, 0.5e-2, 0.5e-3, 0.5e-6, 0.5e-9, 0.5e-12, 0.5e-15};
  static const int val[] = {3, 5, 8, 10, 15, 20, 25, 45};
  for (int i = 0; i < 8; ++i)
    if (eps >= lim[i]) return val[i];
  return ns + nt;
}

namespace {

constexpr int WAVE = 64;
inline int nblk(int n, int b) { return (n + b - 1) / b; }

// ---------------------------------------------------------------------
// translation operators: scale-free matrices T[slot][j][i] (column j =
// input coefficient, i = output coefficient), so lane i of a wave reads a
// contiguous row segment for each j.
// ---------------------------------------------------------------------
struct LapTables {
  cd* m2m = nullptr;  // 4 child quadrants
  cd* l2l = nullptr;  // 4 child quadrants
  cd* m2l = nullptr;  // 49 offsets (dy+3)*7 + (dx+3), |dx|,|dy| <= 3
};

std::vector<double> binomials(int n) {
  std::vector<double> C((size_t)(n + 1) * (n + 1), 0.0);
  for (int l = 0; l <= n; ++l) {
    C[(size_t)l * (n + 1)] = 1.0;
    for (int m = 1; m <= l; ++m)
      C[(size_t)l * (n + 1) + m] = C[(size_t)(l - 1) * (n + 1) + m - 1] +
                                   (m <= l - 1 ? C[(size_t)(l - 1) * (n + 1) + m] : 0.0);
  }
  return C;
}

std::complex<double> cpowi(std::complex<double> z, int n) {
  std::complex<double> r = 1.0;
  for (int i = 0; i < n; ++i) r *= z;
  return r;
}

void make_tables(int p, std::vector<cd>& h) {
  const int P = p + 1, N2 = P * P;
  h.assign((size_t)(4 + 4 + 49) * N2, cd{0, 0});
  auto C = binomials(2 * p + 2);
  auto bin = [&](int l, int m) { return C[(size_t)l * (2 * p + 3) + m]; };
  auto set = [&](cd* T, int i, int j, std::complex<double> v) { T[(size_t)j * P + i] = cd{v.real(), v.imag()}; };
  using Z = std::complex<double>;
  for (int q = 0; q < 4; ++q) {
    // child centre - parent centre, in units of the child box size
    Z w(q & 1 ? 0.5 : -0.5, q & 2 ? 0.5 : -0.5);
    // M2M (laprouts2d.c:666-766): r1 = child size, r2 = 2 r1
    cd* T = h.data() + (size_t)q * N2;
    Z a = 1.0 / w, b = w / 2.0;
    set(T, 0, 0, 1.0);
    for (int i = 1; i <= p; ++i) {
      Z bi = cpowi(b, i);
      set(T, i, 0, -bi / (double)i);
      for (int j = 1; j <= i; ++j) set(T, i, j, bi * bin(i - 1, j - 1) * cpowi(a, j));
    }
    // L2L (laprouts2d.c:777-867): r1 = parent size = 2 r2
    T = h.data() + (size_t)(4 + q) * N2;
    a = w / 2.0;
    b = 1.0 / w;
    for (int i = 0; i <= p; ++i)
      for (int j = i; j <= p; ++j) set(T, i, j, cpowi(b, i) * bin(j, i) * cpowi(a, j));
  }
  // M2L (laprouts2d.c:886-990), r1 = r2 = box size; log(bs) is added in-kernel
  for (int dy = -3; dy <= 3; ++dy)
    for (int dx = -3; dx <= 3; ++dx) {
      if (std::abs(dx) <= 1 && std::abs(dy) <= 1) continue;
      cd* T = h.data() + (size_t)(8 + (dy + 3) * 7 + (dx + 3)) * N2;
      Z w((double)dx, (double)dy);
      Z a = -1.0 / w, b = 1.0 / w;
      set(T, 0, 0, std::log(std::abs(w)));
      for (int j = 1; j <= p; ++j) set(T, 0, j, cpowi(a, j));
      for (int i = 1; i <= p; ++i) {
        Z bi = cpowi(b, i);
        set(T, i, 0, -bi / (double)i);
        for (int j = 1; j <= p; ++j) set(T, i, j, bi * bin(i + j - 1, j - 1) * cpowi(a, j));
      }
    }
}

// device-resident tables, built once per (device, order)
int get_tables(Ctx& c, int p, LapTables& t) {
  static std::mutex mu;
  static std::map<std::pair<int, int>, cd*> cache;
  std::lock_guard<std::mutex> g(mu);
  auto key = std::make_pair(c.device, p);
  auto it = cache.find(key);
  cd* d = nullptr;
  const size_t N2 = (size_t)(p + 1) * (p + 1);
  if (it == cache.end()) {
    std::vector<cd> h;
    make_tables(p, h);
    if (hipMalloc(&d, sizeof(cd) * h.size()) != hipSuccess) return IER_ALLOC;
    if (hipMemcpy(d, h.data(), sizeof(cd) * h.size(), hipMemcpyHostToDevice) != hipSuccess)
      return IER_HIP;
    cache[key] = d;
  } else {
    d = it->second;
  }
  t.m2m = d;
  t.l2l = d + 4 * N2;
  t.m2l = d + 8 * N2;
  return IER_OK;
}

// ---------------------------------------------------------------------
// form multipole (leaf sources) / local (list-4 sources)
// Wave per box, sources staged in LDS in chunks of CH; lane n owns
// coefficient n (looping when p >= 64).
// ---------------------------------------------------------------------
template <bool MP>
__global__ void __launch_bounds__(WAVE) k_form(const int* work, DTree D, const double2* pts,
                                               const double2* chg, const double2* dip, int N,
                                               int nd, int p, int CH, double2* out) {
  extern __shared__ double2 sm[];
  const int b = work[blockIdx.x];
  const int lane = threadIdx.x;
  const int P = p + 1;
  double2* Wc = sm;                    // [P][CH]
  double2* Wd = sm + (size_t)P * CH;   // [P][CH]
  const double2 ctr = D.ctr[b];
  const double r = D.bs[D.level[b]];
  // source boxes: the box itself (multipole) or its list-4 entries (local)
  int nsrcbox = MP ? 1 : D.o4[b + 1] - D.o4[b];
  for (int sb = 0; sb < nsrcbox; ++sb) {
    int jb = MP ? b : D.l4[D.o4[b] + sb];
    int2 rg = D.srng[jb];
    for (int c0 = rg.x; c0 < rg.y; c0 += CH) {
      int cnt = min(CH, rg.y - c0);
      for (int j = lane; j < cnt; j += WAVE) {
        double2 s = pts[c0 + j];
        cd z = mk(s.x - ctr.x, s.y - ctr.y);
        if (MP) {
          // -(z/r)^n / n for charges, z^(n-1)/r^n for dipoles (laprouts2d.c:41-182)
          cd zr = z / r, pc = mk(-1.0, 0.0), pd = mk(1.0 / r, 0.0);
          Wc[j] = make_double2(1.0, 0.0);
          Wd[j] = make_double2(0.0, 0.0);
          for (int n = 1; n <= p; ++n) {
            pc = pc * zr;
            cd v = pc / (double)n;
            Wc[(size_t)n * CH + j] = make_double2(v.x, v.y);
            Wd[(size_t)n * CH + j] = make_double2(pd.x, pd.y);
            pd = pd * zr;
          }
        } else {
          // log|z0|, -(r/z0)^n / n ; -r^n / z0^(n+1)  (laprouts2d.c:358-500)
          cd t = r * inv(z), pc = mk(-1.0, 0.0), pd = -inv(z);
          Wc[j] = make_double2(0.5 * log(norm2(z)), 0.0);
          Wd[j] = make_double2(pd.x, pd.y);
          for (int n = 1; n <= p; ++n) {
            pc = pc * t;
            cd v = pc / (double)n;
            pd = pd * t;
            Wc[(size_t)n * CH + j] = make_double2(v.x, v.y);
            Wd[(size_t)n * CH + j] = make_double2(pd.x, pd.y);
          }
        }
      }
      __syncthreads();
      for (int d = 0; d < nd; ++d) {
        for (int n = lane; n <= p; n += WAVE) {
          cd acc = mk(0, 0);
          for (int j = 0; j < cnt; ++j) {
            if (chg) {
              double2 cc = chg[(size_t)d * N + c0 + j], w = Wc[(size_t)n * CH + j];
              acc += mk(cc.x, cc.y) * mk(w.x, w.y);
            }
            if (dip && (n > 0 || !MP)) {
              double2 qq = dip[(size_t)d * N + c0 + j], w = Wd[(size_t)n * CH + j];
              acc += mk(qq.x, qq.y) * mk(w.x, w.y);
            }
          }
          double2* o = out + ((size_t)b * nd + d) * P + n;
          *o = make_double2(o->x + acc.x, o->y + acc.y);
        }
      }
      __syncthreads();
    }
  }
}

// ---------------------------------------------------------------------
// evaluation at points: wave per box, lane per point
// ---------------------------------------------------------------------
__device__ inline void addc(double2* o, cd v) { *o = make_double2(o->x + v.x, o->y + v.y); }

// list 3: multipoles of finer far boxes at this box's points (laprouts2d.c:193-344)
__global__ void __launch_bounds__(WAVE) k_eval_mp(const int* work, DTree D, const int2* rng,
                                                  const double2* pts, int N, int nd, int p,
                                                  int pg, const double2* mp, double2* pot,
                                                  double2* grad, double2* hess) {
  const int b = work[blockIdx.x];
  const int P = p + 1;
  int2 rg = rng[b];
  for (int i = rg.x + (int)threadIdx.x; i < rg.y; i += WAVE) {
    double2 x = pts[i];
    for (int k = D.o3[b]; k < D.o3[b + 1]; ++k) {
      int s = D.l3[k];
      double2 c = D.ctr[s];
      double r = D.bs[D.level[s]];
      cd z = mk(x.x - c.x, x.y - c.y);
      cd zi = inv(z), t = r * zi;
      double lz = 0.5 * log(norm2(z));
      for (int d = 0; d < nd; ++d) {
        const double2* m = mp + ((size_t)s * nd + d) * P;
        cd m0 = mk(m[0].x, m[0].y);
        cd s1 = mk(0, 0), s2 = mk(0, 0), s3 = mk(0, 0), tn = mk(1, 0);
        for (int n = 1; n <= p; ++n) {
          tn = tn * t;
          cd v = mk(m[n].x, m[n].y) * tn;
          s1 += v;
          s2 += v * (double)n;
          s3 += v * (double)(n * (n + 1));
        }
        addc(pot + (size_t)d * N + i, m0 * lz + s1);
        if (pg >= 2) addc(grad + (size_t)d * N + i, (m0 - s2) * zi);
        if (pg >= 3) addc(hess + (size_t)d * N + i, (s3 - m0) * (zi * zi));
      }
    }
  }
}

// leaf local expansions at the leaf's points (laprouts2d.c:508-642)
__global__ void __launch_bounds__(WAVE) k_eval_lo(const int* work, DTree D, const int2* rng,
                                                  const double2* pts, int N, int nd, int p,
                                                  int pg, const double2* lo, double2* pot,
                                                  double2* grad, double2* hess) {
  const int b = work[blockIdx.x];
  const int P = p + 1;
  int2 rg = rng[b];
  const double2 c = D.ctr[b];
  const double r = D.bs[D.level[b]], ri = 1.0 / r;
  for (int i = rg.x + (int)threadIdx.x; i < rg.y; i += WAVE) {
    double2 x = pts[i];
    cd z = mk((x.x - c.x) / r, (x.y - c.y) / r);
    for (int d = 0; d < nd; ++d) {
      const double2* L = lo + ((size_t)b * nd + d) * P;
      // Horner from the top: S = sum L_n z^n, G = sum n L_n z^(n-1), H = sum n(n-1) L_n z^(n-2)
      cd S = mk(0, 0), G = mk(0, 0), H = mk(0, 0);
      for (int n = p; n >= 0; --n) {
        H = H * z + G * 2.0;
        G = G * z + S;
        S = S * z + mk(L[n].x, L[n].y);
      }
      addc(pot + (size_t)d * N + i, S);
      if (pg >= 2) addc(grad + (size_t)d * N + i, G * ri);
      if (pg >= 3) addc(hess + (size_t)d * N + i, H * (ri * ri));
    }
  }
}

// ---------------------------------------------------------------------
// list-1 near field.  Wave per target box; list-1 sources are staged in LDS
// (coalesced loads), lanes are (target, source slice) pairs so small leaves
// still fill the wave, slices are reduced through LDS at the end.
// MODE: skip test 0: rr <= th^2, 1: rr < th^2 (cauchykernels2d.c:10-14).
// ---------------------------------------------------------------------
constexpr int P2P_CAP = 256;

template <bool C, bool Q, int PG, int MODE>
__global__ void __launch_bounds__(WAVE) k_p2p(const int* work, DTree D, const double2* src,
                                              const double2* chg, const double2* dip, int NS,
                                              const int2* trng, const double2* tpts, int NT,
                                              int nd, double th2, double2* pot, double2* grad,
                                              double2* hess) {
  __shared__ double2 sx[P2P_CAP];
  __shared__ double2 sc[2][P2P_CAP];
  __shared__ double2 sq[2][P2P_CAP];
  __shared__ int boff[14], bsrc[14];
  const int b = work[blockIdx.x];
  const int lane = threadIdx.x;
  const int n1 = D.o1[b + 1] - D.o1[b];
  if (lane == 0) {
    int acc = 0;
    for (int k = 0; k < n1; ++k) {
      int jb = D.l1[D.o1[b] + k];
      int2 r = D.srng[jb];
      boff[k] = acc;
      bsrc[k] = r.x;
      acc += r.y - r.x;
    }
    boff[n1] = acc;
  }
  __syncthreads();
  const int S = boff[n1];
  const int2 tr = trng[b];
  for (int t0 = tr.x; t0 < tr.y; t0 += WAVE) {
    const int Tc = min(WAVE, tr.y - t0);
    const int K = WAVE / Tc;
    const int t = lane / K, k = lane % K;
    const bool act = t < Tc;
    const double2 x = act ? tpts[t0 + t] : make_double2(0, 0);
    for (int d0 = 0; d0 < nd; d0 += 2) {
      const int nch = min(2, nd - d0);
      cd ap[2] = {mk(0, 0), mk(0, 0)}, ag[2] = {mk(0, 0), mk(0, 0)}, ah[2] = {mk(0, 0), mk(0, 0)};
      for (int g0 = 0; g0 < S; g0 += P2P_CAP) {
        const int cnt = min(P2P_CAP, S - g0);
        for (int m = lane; m < cnt; m += WAVE) {
          int g = g0 + m, kb = 0;
          while (boff[kb + 1] <= g) ++kb;
          int si = bsrc[kb] + (g - boff[kb]);
          sx[m] = src[si];
          for (int e = 0; e < nch; ++e) {
            if (C) sc[e][m] = chg[(size_t)(d0 + e) * NS + si];
            if (Q) sq[e][m] = dip[(size_t)(d0 + e) * NS + si];
          }
        }
        __syncthreads();
        if (act) {
          for (int m = k; m < cnt; m += K) {
            double2 s = sx[m];
            double dx = x.x - s.x, dy = x.y - s.y;
            double rr = dx * dx + dy * dy;
            if (MODE == 0 ? (rr <= th2) : (rr < th2)) continue;
            double rinv = 1.0 / rr;
            cd zi = mk(dx * rinv, -dy * rinv);
            cd z2 = -(zi * zi);
            double lr = C ? 0.5 * log(rr) : 0.0;
            for (int e = 0; e < nch; ++e) {
              if (C) {
                cd cc = mk(sc[e][m].x, sc[e][m].y);
                ap[e] += cc * lr;
                if (PG >= 2) ag[e] += cc * zi;
                if (PG >= 3) ah[e] += cc * z2;
              }
              if (Q) {
                cd qq = mk(sq[e][m].x, sq[e][m].y);
                ap[e] += qq * zi;
                if (PG >= 2) ag[e] += qq * z2;
                if (PG >= 3) ah[e] += qq * ((z2 * zi) * -2.0);
              }
            }
          }
        }
        __syncthreads();
      }
      // reduce the K slices of each target through LDS (reusing sc/sq)
      for (int e = 0; e < nch; ++e) {
        double2* red = &sc[0][0];
        red[lane] = make_double2(ap[e].x, ap[e].y);
        red[WAVE + lane] = make_double2(ag[e].x, ag[e].y);
        red[2 * WAVE + lane] = make_double2(ah[e].x, ah[e].y);
        __syncthreads();
        if (act && k == 0) {
          cd P = mk(0, 0), G = mk(0, 0), H = mk(0, 0);
          for (int u = 0; u < K; ++u) {
            int l = t * K + u;
            P += mk(red[l].x, red[l].y);
            G += mk(red[WAVE + l].x, red[WAVE + l].y);
            H += mk(red[2 * WAVE + l].x, red[2 * WAVE + l].y);
          }
          size_t o = (size_t)(d0 + e) * NT + t0 + t;
          addc(pot + o, P);
          if (PG >= 2) addc(grad + o, G);
          if (PG >= 3) addc(hess + o, H);
        }
        __syncthreads();
      }
    }
  }
}

template <bool C, bool Q, int PG>
void launch_p2p_mode(int mode, int nbx, hipStream_t st, const int* work, const DTree& D,
                     const double2* src, const double2* chg, const double2* dip, int NS,
                     const int2* trng, const double2* tpts, int NT, int nd, double th2,
                     double2* pot, double2* grad, double2* hess) {
  if (mode == 0)
    k_p2p<C, Q, PG, 0><<<nbx, WAVE, 0, st>>>(work, D, src, chg, dip, NS, trng, tpts, NT, nd, th2, pot, grad, hess);
  else
    k_p2p<C, Q, PG, 1><<<nbx, WAVE, 0, st>>>(work, D, src, chg, dip, NS, trng, tpts, NT, nd, th2, pot, grad, hess);
}

template <bool C, bool Q>
void launch_p2p_pg(int pg, int mode, int nbx, hipStream_t st, const int* work, const DTree& D,
                   const double2* src, const double2* chg, const double2* dip, int NS,
                   const int2* trng, const double2* tpts, int NT, int nd, double th2,
                   double2* pot, double2* grad, double2* hess) {
  if (pg == 1) launch_p2p_mode<C, Q, 1>(mode, nbx, st, work, D, src, chg, dip, NS, trng, tpts, NT, nd, th2, pot, grad, hess);
  if (pg == 2) launch_p2p_mode<C, Q, 2>(mode, nbx, st, work, D, src, chg, dip, NS, trng, tpts, NT, nd, th2, pot, grad, hess);
  if (pg == 3) launch_p2p_mode<C, Q, 3>(mode, nbx, st, work, D, src, chg, dip, NS, trng, tpts, NT, nd, th2, pot, grad, hess);
}

}  // namespace

int lap_run(Ctx& c, Scratch& S, const SortedPoints& Pt, const Plan& plan, const LapArgs& a) {
  const DTree& D = plan.D;
  const int nd = a.nd, L = plan.L, nb = plan.nb;
  const int p = l2dterms(a.eps);
  if (p > 63) return IER_P_TOO_LARGE;
  const int P = p + 1;
  LapTables tab;
  int e = get_tables(c, p, tab);
  if (e) return e;
  const size_t per = (size_t)nd * P;
  double2* mp = S.get<double2>(per * nb);
  double2* lo = S.get<double2>(per * nb);
  if (S.err) return S.err;
  HIP_TRY(hipMemsetAsync(mp, 0, sizeof(double2) * per * nb, c.st));
  HIP_TRY(hipMemsetAsync(lo, 0, sizeof(double2) * per * nb, c.st));
  const double2* T2 = reinterpret_cast<const double2*>(tab.m2m);
  const double2* TL = reinterpret_cast<const double2*>(tab.l2l);
  const double2* TM = reinterpret_cast<const double2*>(tab.m2l);
  const int* W = plan.dwork;
  const double2* chg = a.ifcharge ? a.chg : nullptr;
  const double2* dip = a.ifdipole ? a.dip : nullptr;
  if (!chg && !dip) return IER_OK;  // zero strengths: outputs stay zero

  int CH = 64;
  while (CH > 8 && (size_t)2 * P * CH * sizeof(double2) > 40960) CH /= 2;
  size_t shm = (size_t)2 * P * CH * sizeof(double2);
  StepTimer* tm = step_timer();
  {
    SolveStats& st = last_stats();
    for (int l = 2; l <= L; ++l) st.m2l_pairs += plan.m2l_pairs[l];
    st.m2l_flops = st.m2l_pairs * 8.0 * P * P;
    st.pmax = p;
  }
  tm->mark(c.st);
  // step 1: leaf multipoles
  if (plan.formmp.n)
    k_form<true><<<plan.formmp.n, WAVE, shm, c.st>>>(W + plan.formmp.off, D, Pt.src, chg, dip,
                                                    Pt.ns, nd, p, CH, mp);
  tm->mark(c.st);
  // step 3: upward pass
  int maxw = 1;
  for (int l = 0; l <= L; ++l)
    maxw = std::max({maxw, plan.m2m.at(l).n, plan.m2l.at(l).n, plan.l2l_kids.at(l).n});
  int* map = S.get<int>((size_t)maxw * 49);
  if (S.err) return S.err;
  for (int l = L - 1; l >= 1; --l) {
    const Span sp = plan.m2m.at(l);
    if (!sp.n) continue;
    k_map_m2m<<<nblk(sp.n, 128), 128, 0, c.st>>>(W + sp.off, sp.n, D.child, D.srng, map);
    HIP_TRY(launch_xlate(c.st, S, W + sp.off, sp.n, map, 4, 4, T2, P, P, mp, per, mp, per, nd, 0.0));
  }
  tm->mark(c.st);
  // step 2: list 4
  if (plan.list4.n)
    k_form<false><<<plan.list4.n, WAVE, shm, c.st>>>(W + plan.list4.off, D, Pt.src, chg, dip,
                                                    Pt.ns, nd, p, CH, lo);
  tm->mark(c.st);
  // step 4: M2L level by level (log of the box size enters entry (0,0) only)
  for (int l = 2; l <= L; ++l) {
    const Span sp = plan.m2l.at(l);
    if (!sp.n) continue;
    k_map_m2l<<<nblk(sp.n, 128), 128, 0, c.st>>>(W + sp.off, sp.n, D.ctr, D.o2, D.l2, plan.bs[l], map);
    HIP_TRY(launch_m2l(c.st, S, W + sp.off, sp.n, map, TM, P, mp, per, lo, per, nd,
                       std::log(plan.bs[l])));
  }
  tm->mark(c.st);
  // step 5: L2L into the children of boxes that hold targets
  for (int l = 1; l <= L - 1; ++l) {
    const Span sp = plan.l2l_kids.at(l);
    if (!sp.n) continue;
    k_map_l2l<<<nblk(sp.n, 128), 128, 0, c.st>>>(W + sp.off, sp.n, D.parent, D.child, map);
    HIP_TRY(launch_xlate(c.st, S, W + sp.off, sp.n, map, 4, 1, TL, P, P, lo, per, lo, per, nd, 0.0));
  }
  // steps 6 + 7: list 3 and local expansions at sources / targets
  struct Set { int pg; Span l3, lv, nn; const int2* rng; const double2* pts; int N; double2 *po, *gr, *he; };
  Set sets[2] = {{a.ifpgh, plan.list3_s.all(), plan.leaves_s.all(), plan.p2p_s.all(), D.srng, Pt.src, Pt.ns, a.pot, a.grad, a.hess},
                 {a.ifpghtarg, plan.list3_t.all(), plan.leaves_t.all(), plan.p2p_t.all(), D.trng, Pt.trg, Pt.nt, a.pott, a.gradt, a.hesst}};
  // reference skip tests: charges (with or without dipoles) use rr vs thresh^2,
  // '<=' for potential only and '<' otherwise; dipole-only routines compare |z|
  const double thresh = Pt.bs0 * 0x1p-51;
  tm->mark(c.st);
  for (auto& s : sets)
    if (s.pg > 0 && s.N > 0 && s.l3.n)
      k_eval_mp<<<s.l3.n, WAVE, 0, c.st>>>(W + s.l3.off, D, s.rng, s.pts, s.N, nd, p, s.pg, mp, s.po, s.gr, s.he);
  tm->mark(c.st);
  for (auto& s : sets)
    if (s.pg > 0 && s.N > 0 && s.lv.n)
      k_eval_lo<<<s.lv.n, WAVE, 0, c.st>>>(W + s.lv.off, D, s.rng, s.pts, s.N, nd, p, s.pg, lo, s.po, s.gr, s.he);
  tm->mark(c.st);
  for (auto& s : sets) {
    if (s.pg <= 0 || s.N == 0) continue;
    if (a.ifnear && s.nn.n) {
      int mode = s.pg == 1 ? 0 : 1;
      double th2 = thresh * thresh;
      if (chg && dip)
        launch_p2p_pg<true, true>(s.pg, mode, s.nn.n, c.st, W + s.nn.off, D, Pt.src, chg, dip, Pt.ns, s.rng, s.pts, s.N, nd, th2, s.po, s.gr, s.he);
      else if (chg)
        launch_p2p_pg<true, false>(s.pg, mode, s.nn.n, c.st, W + s.nn.off, D, Pt.src, chg, dip, Pt.ns, s.rng, s.pts, s.N, nd, th2, s.po, s.gr, s.he);
      else
        launch_p2p_pg<false, true>(s.pg, mode, s.nn.n, c.st, W + s.nn.off, D, Pt.src, chg, dip, Pt.ns, s.rng, s.pts, s.N, nd, th2, s.po, s.gr, s.he);
    }
  }
  tm->mark(c.st);
  HIP_TRY(hipGetLastError());
  return IER_OK;
}

}  // namespace fmm

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
  float2* m2l32 = nullptr;  // the M2L operators rounded to FP32
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
  static std::map<std::pair<int, int>, std::pair<cd*, float2*>> cache;
  std::lock_guard<std::mutex> g(mu);
  auto key = std::make_pair(c.device, p);
  auto it = cache.find(key);
  cd* d = nullptr;
  float2* d32 = nullptr;
  const size_t N2 = (size_t)(p + 1) * (p + 1);
  if (it == cache.end()) {
    std::vector<cd> h;
    make_tables(p, h);
    std::vector<float2> h32(49 * N2);
    for (size_t e = 0; e < h32.size(); ++e)
      h32[e] = make_float2((float)h[8 * N2 + e].x, (float)h[8 * N2 + e].y);
    if (hipMalloc(&d, sizeof(cd) * h.size()) != hipSuccess) return IER_ALLOC;
    if (hipMalloc(&d32, sizeof(float2) * h32.size()) != hipSuccess) {
      (void)hipFree(d);
      return IER_ALLOC;
    }
    if (hipMemcpy(d, h.data(), sizeof(cd) * h.size(), hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(d32, h32.data(), sizeof(float2) * h32.size(), hipMemcpyHostToDevice) != hipSuccess)
      return IER_HIP;
    cache[key] = {d, d32};
  } else {
    d = it->second.first;
    d32 = it->second.second;
  }
  t.m2m = d;
  t.l2l = d + 4 * N2;
  t.m2l = d + 8 * N2;
  t.m2l32 = d32;
  return IER_OK;
}

// ---------------------------------------------------------------------
// Per-point kernels work in 16-lane groups, four boxes per wave: lane u of a
// group takes points u, u + 16, ... of its box.  Expansion coefficients of
// a box are read by every lane of the group (one cache line per request).
// ---------------------------------------------------------------------
__device__ inline void addc(double2* o, cd v) { *o = make_double2(o->x + v.x, o->y + v.y); }
__device__ inline cd ldc(const double2* p) { const double2 v = *p; return mk(v.x, v.y); }

// final value of sorted point i, channel dd, to the caller's arrays (LapOut)
template <int PG>
__device__ inline void lap_final(const LapOut& u, int dd, size_t i, cd P, cd G, cd H) {
  const size_t t = (size_t)u.perm[i];
  const size_t nd = (size_t)u.nd;
  if (u.kind == 0) {
    static_cast<double2*>(u.pot)[t * nd + dd] = make_double2(P.x, P.y);
    if (PG >= 2) static_cast<double2*>(u.grad)[t * nd + dd] = make_double2(G.x, G.y);
    if (PG >= 3) static_cast<double2*>(u.hess)[t * nd + dd] = make_double2(H.x, H.y);
  } else if (u.kind == 1) {  // lfmm2d_ndiv.c:171-243
    const size_t d = dd >> 1, h = dd & 1;
    double* p = static_cast<double*>(u.pot);
    p[2 * (t * nd + d) + h] = P.x;
    if (PG >= 2) {
      double* g = static_cast<double*>(u.grad);
      g[2 * ((t * 2 + 0) * nd + d) + h] = G.x;
      g[2 * ((t * 2 + 1) * nd + d) + h] = -G.y;
    }
    if (PG >= 3) {
      double* q = static_cast<double*>(u.hess);
      q[2 * ((t * 3 + 0) * nd + d) + h] = H.x;
      q[2 * ((t * 3 + 1) * nd + d) + h] = -H.y;
      q[2 * ((t * 3 + 2) * nd + d) + h] = -H.x;
    }
  } else {  // rfmm2d_ndiv.c:156-214
    static_cast<double*>(u.pot)[t * nd + dd] = P.x;
    if (PG >= 2) {
      double* g = static_cast<double*>(u.grad);
      g[(t * 2 + 0) * nd + dd] = G.x;
      g[(t * 2 + 1) * nd + dd] = -G.y;
    }
    if (PG >= 3) {
      double* q = static_cast<double*>(u.hess);
      q[(t * 3 + 0) * nd + dd] = H.x;
      q[(t * 3 + 1) * nd + dd] = -H.y;
      q[(t * 3 + 2) * nd + dd] = -H.x;
    }
  }
}


// Leaf multipoles (MP) from the leaf's sources, or local expansions from
// the sources of the list-4 boxes (!MP).  Four boxes per wave, a 16-lane
// group per box, lane u on point u of each chunk of 16 sources: the lane
// forms its point's terms of 16 coefficients at a time into a 16 x 17 LDS
// tile, and lane u then sums coefficient u's 16 terms in point order (p <= 63):
//   multipole (laprouts2d.c:41-182):  mp_0 += c,
//     mp_n += -c (z/r)^n / n + d z^(n-1) / r^n
//   local (laprouts2d.c:358-500):     lo_0 += c log|z| - d / z,
//     lo_n += -c (r/z)^n / n - d (r/z)^n / z
// with z = source - centre and r the box size.
constexpr int FT_PAD = 17;  // tile row length (16 terms + 1: no bank conflicts)

__device__ inline void lap_wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

template <bool MP>
__global__ void __launch_bounds__(WAVE) k_form(const int* work, int nwork, DTree D,
                                               const double2* pts, const double2* chg,
                                               const double2* dip, int N, int nd, int p,
                                               double2* out) {
  __shared__ cd tile[4 * 16 * FT_PAD];
  const int grp = (int)threadIdx.x >> 4, u = (int)threadIdx.x & 15;
  const int w = blockIdx.x * 4 + grp;
  if (w >= nwork) return;  // whole groups leave; only wave-level syncs below
  cd* T = tile + grp * 16 * FT_PAD;
  const int b = work[w];
  const int P = p + 1;
  const double2 ctr = D.ctr[b];
  const double r = D.bsb[b], ri = 1.0 / r;
  const int nsb = MP ? 1 : D.o4[b + 1] - D.o4[b];
  const int l4 = MP ? 0 : D.o4[b];
  for (int d = 0; d < nd; ++d) {
    cd acc[4] = {mk(0, 0), mk(0, 0), mk(0, 0), mk(0, 0)};
    for (int sb = 0; sb < nsb; ++sb) {
      const int jb = MP ? b : D.l4[l4 + sb];
      const int2 rg = D.srng[jb];
      for (int c0 = rg.x; c0 < rg.y; c0 += 16) {
        const int j = c0 + u;
        const bool ok = j < rg.y;
        const double2 s = pts[ok ? j : rg.x];
        const cd c = ok && chg ? ldc(&chg[(size_t)d * N + j]) : mk(0, 0);
        const cd q = ok && dip ? ldc(&dip[(size_t)d * N + j]) : mk(0, 0);
        const cd z = mk(s.x - ctr.x, s.y - ctr.y);
        // running powers: MP pw = (z/r)^n, qw = d z^(n-1) / r^n;
        // local pw = (r/z)^n, qw = -d (r/z)^n / z
        cd step, qw, term0;
        if (MP) {
          step = z * ri;
          qw = q * ri;
          term0 = c;
        } else {
          const cd zi = inv(z);
          step = r * zi;
          qw = -(q * zi);
          term0 = c * (0.5 * log_pos(norm2(z))) + qw;
        }
        cd pw = mk(1, 0);
#pragma unroll
        for (int hi = 0; hi < 4; ++hi) {
          if (16 * hi > p) break;
          const int nfill = min(16, P - 16 * hi);
          for (int lo = 0; lo < nfill; ++lo) {
            const int n = 16 * hi + lo;
            cd t;
            if (n == 0) {
              t = term0;
            } else {
              pw = pw * step;
              const double inv_n = 1.0 / n;  // uniform over the wave
              if (MP) {
                t = -(c * pw) * inv_n + qw;
                qw = qw * step;
              } else {
                qw = qw * step;
                t = -(c * pw) * inv_n + qw;
              }
            }
            T[lo * FT_PAD + u] = t;
          }
          lap_wave_sync();
          if (u < nfill) {
            cd sum = mk(0, 0);
            for (int k = 0; k < 16; ++k) sum += T[u * FT_PAD + k];
            acc[hi] += sum;
          }
          lap_wave_sync();
        }
      }
    }
    double2* o = out + ((size_t)b * nd + d) * P;
#pragma unroll
    for (int hi = 0; hi < 4; ++hi) {
      const int n = 16 * hi + u;
      if (n <= p) addc(o + n, acc[hi]);
    }
  }
}

// list 3: multipoles of finer far boxes at this box's points
// (laprouts2d.c:193-344): pot += m_0 log|z| + sum m_n t^n, t = r/z;
// grad = (m_0 - sum n m_n t^n) / z; hess = (sum n(n+1) m_n t^n - m_0) / z^2.
// NC channels at a time share each box's z, 1/z, log|z| and powers t^n (the
// log kernel's complex densities always come as channel pairs); each
// channel's sums are formed exactly as alone.
template <int PG, int NC>
__global__ void __launch_bounds__(WAVE) k_eval_mp(const int* work, int nwork, DTree D,
                                                  const int2* rng, const double2* pts, int N,
                                                  int nd, int p, const double2* mp, double2* pot,
                                                  double2* grad, double2* hess) {
  const int u = threadIdx.x & 15;
  const int w = blockIdx.x * 4 + (threadIdx.x >> 4);
  if (w >= nwork) return;
  const int b = work[w];
  const int P = p + 1;
  const int2 rg = rng[b];
  const int k0 = D.o3[b], k1 = D.o3[b + 1];
  for (int i = rg.x + u; i < rg.y; i += 16) {
    const double2 x = pts[i];
    for (int d0 = 0; d0 < nd; d0 += NC) {
      cd sp[NC], sg[NC], sh[NC];
#pragma unroll
      for (int e = 0; e < NC; ++e) sp[e] = sg[e] = sh[e] = mk(0, 0);
      for (int k = k0; k < k1; ++k) {
        const int s = D.l3[k];
        const double2 c = D.ctr[s];
        const double r = D.bsb[s];
        const cd z = mk(x.x - c.x, x.y - c.y);
        const cd zi = inv(z), t = r * zi;
        const double lg = 0.5 * log_pos(norm2(z));
        const double2* m = mp + ((size_t)s * nd + d0) * P;
        cd m0[NC], s1[NC], s2[NC], s3[NC];
#pragma unroll
        for (int e = 0; e < NC; ++e) {
          m0[e] = ldc(m + (size_t)e * P);
          s1[e] = s2[e] = s3[e] = mk(0, 0);
        }
        cd tn = mk(1, 0);
        for (int n = 1; n <= p; ++n) {
          tn = tn * t;
#pragma unroll
          for (int e = 0; e < NC; ++e) {
            const cd v = ldc(m + (size_t)e * P + n) * tn;
            s1[e] += v;
            if (PG >= 2) s2[e] += v * (double)n;
            if (PG >= 3) s3[e] += v * (double)(n * (n + 1));
          }
        }
#pragma unroll
        for (int e = 0; e < NC; ++e) {
          sp[e] += m0[e] * lg + s1[e];
          if (PG >= 2) sg[e] += (m0[e] - s2[e]) * zi;
          if (PG >= 3) sh[e] += (s3[e] - m0[e]) * (zi * zi);
        }
      }
#pragma unroll
      for (int e = 0; e < NC; ++e) {
        const int d = d0 + e;
        addc(pot + (size_t)d * N + i, sp[e]);
        if (PG >= 2) addc(grad + (size_t)d * N + i, sg[e]);
        if (PG >= 3) addc(hess + (size_t)d * N + i, sh[e]);
      }
    }
  }
}

// leaf local expansions at the leaf's points (laprouts2d.c:508-642), Horner
// from the top: S = sum L_n w^n, G = sum n L_n w^(n-1), H = sum n(n-1) L_n
// w^(n-2), w = (x - c)/r; grad = G / r, hess = H / r^2
template <int PG>
__global__ void __launch_bounds__(WAVE) k_eval_lo(const int* work, int nwork, DTree D,
                                                  const int2* rng, const double2* pts, int N,
                                                  int nd, int p, const double2* lo, double2* pot,
                                                  double2* grad, double2* hess) {
  const int u = threadIdx.x & 15;
  const int w = blockIdx.x * 4 + (threadIdx.x >> 4);
  if (w >= nwork) return;
  const int b = work[w];
  const int P = p + 1;
  const int2 rg = rng[b];
  const double2 c = D.ctr[b];
  const double r = D.bsb[b], ri = 1.0 / r;
  for (int i = rg.x + u; i < rg.y; i += 16) {
    const double2 x = pts[i];
    const cd z = mk((x.x - c.x) * ri, (x.y - c.y) * ri);
    for (int d = 0; d < nd; ++d) {
      const double2* L = lo + ((size_t)b * nd + d) * P;
      cd S = mk(0, 0), G = mk(0, 0), H = mk(0, 0);
      for (int n = p; n >= 0; --n) {
        if (PG >= 3) H = H * z + G * 2.0;
        if (PG >= 2) G = G * z + S;
        S = S * z + ldc(L + n);
      }
      addc(pot + (size_t)d * N + i, S);
      if (PG >= 2) addc(grad + (size_t)d * N + i, G * ri);
      if (PG >= 3) addc(hess + (size_t)d * N + i, H * (ri * ri));
    }
  }
}

template <int PG>
void launch_eval(bool mpk, int nw, hipStream_t st, const int* work, const DTree& D,
                 const int2* rng, const double2* pts, int N, int nd, int p, const double2* ex,
                 double2* pot, double2* grad, double2* hess) {
  const int nbx = (nw + 3) / 4;
  if (mpk) {
    if (nd % 2 == 0)
      k_eval_mp<PG, 2><<<nbx, WAVE, 0, st>>>(work, nw, D, rng, pts, N, nd, p, ex, pot, grad, hess);
    else
      k_eval_mp<PG, 1><<<nbx, WAVE, 0, st>>>(work, nw, D, rng, pts, N, nd, p, ex, pot, grad, hess);
  } else {
    k_eval_lo<PG><<<nbx, WAVE, 0, st>>>(work, nw, D, rng, pts, N, nd, p, ex, pot, grad, hess);
  }
}

// ---------------------------------------------------------------------
// list-1 near field.  Wave per target box, one channel at a time; list-1
// sources are staged in LDS LCAP at a time (coalesced loads), lanes are
// (target, source slice) pairs so small leaves still fill the wave, and the
// slices are reduced through LDS at the end.
// MODE: skip test 0: rr <= th^2, 1: rr < th^2 (cauchykernels2d.c:10-14).
// ---------------------------------------------------------------------
constexpr int LCAP = 128;

// source boxes of list 1 as a concatenated range: boff[k] = first slot of
// box k, bsrc[k] = its first sorted source; returns the total
__device__ inline int list1_ranges(const DTree& D, int b, int* boff, int* bsrc) {
  const int n1 = D.o1[b + 1] - D.o1[b];
  if (threadIdx.x == 0) {
    int acc = 0;
    for (int k = 0; k < n1; ++k) {
      const int2 r = D.srng[D.l1[D.o1[b] + k]];
      boff[k] = acc;
      bsrc[k] = r.x;
      acc += r.y - r.x;
    }
    boff[n1] = acc;
  }
  __syncthreads();
  return boff[n1];
}
__device__ inline int list1_source(const int* boff, const int* bsrc, int g) {
  int kb = 0;
  while (boff[kb + 1] <= g) ++kb;
  return bsrc[kb] + (g - boff[kb]);
}

// NC channels at a time (2 for the log kernel's channel pairs): each pair's
// skip test, 1/z, -1/z^2 and log|z| serve every channel, whose sums are
// formed exactly as alone.
template <bool C, bool Q, int PG, int MODE, int NC>
__global__ void __launch_bounds__(WAVE) k_p2p(const int* work, DTree D, const double2* src,
                                              const double2* chg, const double2* dip, int NS,
                                              const int2* trng, const double2* tpts, int NT,
                                              int nd, double th2, double2* pot, double2* grad,
                                              double2* hess, LapOut out) {
  // staging (x, charges, dipoles of NC channels) and the slice reduction
  // share one buffer
  __shared__ double2 lds[(1 + 2 * NC) * LCAP];
  double2* sx = lds;
  double2* sc = lds + LCAP;             // [NC][LCAP]
  double2* sq = lds + (1 + NC) * LCAP;  // [NC][LCAP]
  double2* red = lds;
  static_assert(3 * NC * WAVE <= (1 + 2 * NC) * LCAP, "reduction buffer");
  __shared__ int boff[14], bsrc[14];
  const int b = work[blockIdx.x];
  const int lane = threadIdx.x;
  const int S = list1_ranges(D, b, boff, bsrc);
  const int2 tr = trng[b];
  for (int t0 = tr.x; t0 < tr.y; t0 += WAVE) {
    const int Tc = min(WAVE, tr.y - t0);
    const int K = WAVE / Tc;
    const int t = lane / K, k = lane % K;
    const bool act = t < Tc;
    const double2 x = act ? tpts[t0 + t] : make_double2(0, 0);
    for (int d0 = 0; d0 < nd; d0 += NC) {
      cd ap[NC], ag[NC], ah[NC];
#pragma unroll
      for (int e = 0; e < NC; ++e) ap[e] = ag[e] = ah[e] = mk(0, 0);
      for (int g0 = 0; g0 < S; g0 += LCAP) {
        const int cnt = min(LCAP, S - g0);
        for (int m = lane; m < cnt; m += WAVE) {
          const int si = list1_source(boff, bsrc, g0 + m);
          sx[m] = src[si];
#pragma unroll
          for (int e = 0; e < NC; ++e) {
            if (C) sc[e * LCAP + m] = chg[(size_t)(d0 + e) * NS + si];
            if (Q) sq[e * LCAP + m] = dip[(size_t)(d0 + e) * NS + si];
          }
        }
        __syncthreads();
        if (act) {
          for (int m = k; m < cnt; m += K) {
            const double2 s = sx[m];
            const double dx = x.x - s.x, dy = x.y - s.y;
            const double rr = dx * dx + dy * dy;
            if (MODE == 0 ? (rr <= th2) : (rr < th2)) continue;
            const double rinv = 1.0 / rr;
            const cd zi = mk(dx * rinv, -dy * rinv);
            const cd z2 = -(zi * zi);
            const double lg = C ? 0.5 * log_tab(rr) : 0.0;  // rr > th2 > 0
            const cd z3 = (z2 * zi) * -2.0;
#pragma unroll
            for (int e = 0; e < NC; ++e) {
              if (C) {
                const cd cc = mk(sc[e * LCAP + m].x, sc[e * LCAP + m].y);
                ap[e] += cc * lg;
                if (PG >= 2) ag[e] += cc * zi;
                if (PG >= 3) ah[e] += cc * z2;
              }
              if (Q) {
                const cd qq = mk(sq[e * LCAP + m].x, sq[e * LCAP + m].y);
                ap[e] += qq * zi;
                if (PG >= 2) ag[e] += qq * z2;
                if (PG >= 3) ah[e] += qq * z3;
              }
            }
          }
        }
        __syncthreads();
      }
#pragma unroll
      for (int e = 0; e < NC; ++e) {
        red[(3 * e + 0) * WAVE + lane] = make_double2(ap[e].x, ap[e].y);
        if (PG >= 2) red[(3 * e + 1) * WAVE + lane] = make_double2(ag[e].x, ag[e].y);
        if (PG >= 3) red[(3 * e + 2) * WAVE + lane] = make_double2(ah[e].x, ah[e].y);
      }
      __syncthreads();
      if (act && k == 0) {
#pragma unroll
        for (int e = 0; e < NC; ++e) {
          cd P = mk(0, 0), G = mk(0, 0), H = mk(0, 0);
          for (int u = 0; u < K; ++u) {
            const int l = t * K + u;
            P += ldc(&red[(3 * e + 0) * WAVE + l]);
            if (PG >= 2) G += ldc(&red[(3 * e + 1) * WAVE + l]);
            if (PG >= 3) H += ldc(&red[(3 * e + 2) * WAVE + l]);
          }
          const int d = d0 + e;
          const size_t o = (size_t)d * NT + t0 + t;
          if (out.pot) {
            lap_final<PG>(out, d, t0 + t, ldc(pot + o) + P, PG >= 2 ? ldc(grad + o) + G : G,
                          PG >= 3 ? ldc(hess + o) + H : H);
          } else {
            addc(pot + o, P);
            if (PG >= 2) addc(grad + o, G);
            if (PG >= 3) addc(hess + o, H);
          }
        }
      }
      __syncthreads();
    }
  }
}

// Mixed-precision list-1 kernel (potential and gradient).  Coordinate
// differences and the skip test are FP64, as in k_p2p; the pair terms are
// then formed in FP32 from z scaled by sc = 2^-e (2^e near the root box
// size, so |z| sc stays far from the FP32 range limits), and summed in FP32
// per staged chunk of at most LCAP sources before the chunk sums, rescaled,
// join FP64 accumulators.  Terms of each power of sc are summed apart:
//   log|z| = log|z sc| - log sc,  1/z = sc / (z sc),  1/z^2 = sc^2 / (z sc)^2.
struct cf {
  float x, y;
};
__device__ inline cf cfm(cf a, cf b) { return {a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x}; }
__device__ inline cf cfa(cf a, cf b) { return {a.x + b.x, a.y + b.y}; }

template <bool C, bool Q, int PG, int MODE>
__global__ void __launch_bounds__(WAVE) k_p2p32(const int* work, DTree D, const double2* src,
                                                const double2* chg, const double2* dip, int NS,
                                                const int2* trng, const double2* tpts, int NT,
                                                int nd, double th2, double sc, double2* pot,
                                                double2* grad, LapOut out) {
  // staging: x (FP64), charge and dipole strengths (FP32); the slice
  // reduction reuses the x buffer
  __shared__ double2 lds[2 * LCAP];
  double2* sx = lds;
  float2* sc32 = reinterpret_cast<float2*>(lds + LCAP);
  float2* sq32 = sc32 + LCAP;
  double2* red = lds;
  static_assert(2 * WAVE <= LCAP, "reduction buffer");
  __shared__ int boff[14], bsrc[14];
  const int b = work[blockIdx.x];
  const int lane = threadIdx.x;
  const int S = list1_ranges(D, b, boff, bsrc);
  const int2 tr = trng[b];
  const float lsc = (float)-log(sc);
  const double sc2 = sc * sc;
  for (int t0 = tr.x; t0 < tr.y; t0 += WAVE) {
    const int Tc = min(WAVE, tr.y - t0);
    const int K = WAVE / Tc;
    const int t = lane / K, k = lane % K;
    const bool act = t < Tc;
    const double2 x = act ? tpts[t0 + t] : make_double2(0, 0);
    for (int d = 0; d < nd; ++d) {
      cd ap = mk(0, 0), ag = mk(0, 0);
      for (int g0 = 0; g0 < S; g0 += LCAP) {
        const int cnt = min(LCAP, S - g0);
        for (int m = lane; m < cnt; m += WAVE) {
          const int si = list1_source(boff, bsrc, g0 + m);
          sx[m] = src[si];
          if (C) {
            const double2 v = chg[(size_t)d * NS + si];
            sc32[m] = make_float2((float)v.x, (float)v.y);
          }
          if (Q) {
            const double2 v = dip[(size_t)d * NS + si];
            sq32[m] = make_float2((float)v.x, (float)v.y);
          }
        }
        __syncthreads();
        if (act) {
          // chunk sums: p0 (scale-free) and p1 (x sc) of the potential,
          // g1 (x sc) and g2 (x sc^2) of the gradient
          cf p0 = {0, 0}, p1 = {0, 0}, g1 = {0, 0}, g2 = {0, 0};
          for (int m = k; m < cnt; m += K) {
            const double2 s = sx[m];
            const double dx = x.x - s.x, dy = x.y - s.y;
            const double rr = dx * dx + dy * dy;
            if (MODE == 0 ? (rr <= th2) : (rr < th2)) continue;
            const float fx = (float)(dx * sc), fy = (float)(dy * sc);
            const float frr = fx * fx + fy * fy;
            const float ri = __builtin_amdgcn_rcpf(frr);
            const cf zi = {fx * ri, -fy * ri};
            if (C) {
              const float lr = 0.5f * 0.693147180559945309f * __builtin_amdgcn_logf(frr) + lsc;
              const float2 cc = sc32[m];
              p0 = cfa(p0, {cc.x * lr, cc.y * lr});
              if (PG >= 2) g1 = cfa(g1, cfm({cc.x, cc.y}, zi));
            }
            if (Q) {
              const float2 qq = sq32[m];
              p1 = cfa(p1, cfm({qq.x, qq.y}, zi));
              if (PG >= 2) {
                const cf z2 = {zi.y * zi.y - zi.x * zi.x, -2.0f * zi.x * zi.y};  // -1/z^2
                g2 = cfa(g2, cfm({qq.x, qq.y}, z2));
              }
            }
          }
          ap += mk((double)p0.x + (double)p1.x * sc, (double)p0.y + (double)p1.y * sc);
          if (PG >= 2)
            ag += mk((double)g1.x * sc + (double)g2.x * sc2, (double)g1.y * sc + (double)g2.y * sc2);
        }
        __syncthreads();
      }
      red[lane] = make_double2(ap.x, ap.y);
      if (PG >= 2) red[WAVE + lane] = make_double2(ag.x, ag.y);
      __syncthreads();
      if (act && k == 0) {
        cd P = mk(0, 0), G = mk(0, 0);
        for (int u = 0; u < K; ++u) {
          const int l = t * K + u;
          P += mk(red[l].x, red[l].y);
          if (PG >= 2) G += mk(red[WAVE + l].x, red[WAVE + l].y);
        }
        const size_t o = (size_t)d * NT + t0 + t;
        if (out.pot) {
          lap_final<PG>(out, d, t0 + t, ldc(pot + o) + P, PG >= 2 ? ldc(grad + o) + G : G, G);
        } else {
          addc(pot + o, P);
          if (PG >= 2) addc(grad + o, G);
        }
      }
      __syncthreads();
    }
  }
}

template <bool C, bool Q>
void launch_p2p32(int pg, int mode, int nbx, hipStream_t st, const int* work, const DTree& D,
                  const double2* src, const double2* chg, const double2* dip, int NS,
                  const int2* trng, const double2* tpts, int NT, int nd, double th2, double sc,
                  double2* pot, double2* grad, const LapOut& out) {
#define FMM_P2P32(PGV, MV)                                                                       \
  k_p2p32<C, Q, PGV, MV><<<nbx, WAVE, 0, st>>>(work, D, src, chg, dip, NS, trng, tpts, NT, nd, th2, \
                                               sc, pot, grad, out)
  if (pg == 1) {
    if (mode == 0) FMM_P2P32(1, 0);
    else FMM_P2P32(1, 1);
  } else {
    if (mode == 0) FMM_P2P32(2, 0);
    else FMM_P2P32(2, 1);
  }
#undef FMM_P2P32
}

template <bool C, bool Q, int PG>
void launch_p2p_mode(int mode, int nbx, hipStream_t st, const int* work, const DTree& D,
                     const double2* src, const double2* chg, const double2* dip, int NS,
                     const int2* trng, const double2* tpts, int NT, int nd, double th2,
                     double2* pot, double2* grad, double2* hess, const LapOut& out) {
#define FMM_P2P(MV, NCV)                                                                 \
  k_p2p<C, Q, PG, MV, NCV><<<nbx, WAVE, 0, st>>>(work, D, src, chg, dip, NS, trng, tpts, NT, nd, \
                                                 th2, pot, grad, hess, out)
  if (nd % 2 == 0) {
    if (mode == 0) FMM_P2P(0, 2);
    else FMM_P2P(1, 2);
  } else {
    if (mode == 0) FMM_P2P(0, 1);
    else FMM_P2P(1, 1);
  }
#undef FMM_P2P
}

template <bool C, bool Q>
void launch_p2p_pg(int pg, int mode, int nbx, hipStream_t st, const int* work, const DTree& D,
                   const double2* src, const double2* chg, const double2* dip, int NS,
                   const int2* trng, const double2* tpts, int NT, int nd, double th2,
                   double2* pot, double2* grad, double2* hess, const LapOut& out) {
  if (pg == 1) launch_p2p_mode<C, Q, 1>(mode, nbx, st, work, D, src, chg, dip, NS, trng, tpts, NT, nd, th2, pot, grad, hess, out);
  if (pg == 2) launch_p2p_mode<C, Q, 2>(mode, nbx, st, work, D, src, chg, dip, NS, trng, tpts, NT, nd, th2, pot, grad, hess, out);
  if (pg == 3) launch_p2p_mode<C, Q, 3>(mode, nbx, st, work, D, src, chg, dip, NS, trng, tpts, NT, nd, th2, pot, grad, hess, out);
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
  fill_async(mp, 0u, 4 * per * nb, c.st);
  fill_async(lo, 0u, 4 * per * nb, c.st);
  const double2* T2 = reinterpret_cast<const double2*>(tab.m2m);
  const double2* TL = reinterpret_cast<const double2*>(tab.l2l);
  const double2* TM = reinterpret_cast<const double2*>(tab.m2l);
  const int* W = plan.dwork;
  const bool forest = plan.forest.M >= 0;
  const double2* chg = a.ifcharge ? a.chg : nullptr;
  const double2* dip = a.ifdipole ? a.dip : nullptr;
  if (!chg && !dip) return IER_OK;  // zero strengths: outputs stay zero

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
    k_form<true><<<nblk(plan.formmp.n, 4), WAVE, 0, c.st>>>(W + plan.formmp.off, plan.formmp.n, D,
                                                           Pt.src, chg, dip, Pt.ns, nd, p, mp);
  tm->mark(c.st);
  // step 3: upward pass
  // the M2M and L2L slot maps of every level, one launch each
  int* map_m2m = S.get<int>((size_t)std::max(plan.m2m.n, 1) * 4);
  int* map_l2l = S.get<int>((size_t)std::max(plan.l2l_kids.n, 1) * 4);
  if (S.err) return S.err;
  if (plan.m2m.n)
    k_map_m2m<<<nblk(plan.m2m.n, 128), 128, 0, c.st>>>(W + plan.m2m.off, plan.m2m.n, D.child,
                                                        D.srng, map_m2m);
  if (plan.l2l_kids.n)
    k_map_l2l<<<nblk(plan.l2l_kids.n, 128), 128, 0, c.st>>>(W + plan.l2l_kids.off,
                                                             plan.l2l_kids.n, D.parent, D.child,
                                                             map_l2l);
  for (int l = L - 1; l >= 1; --l) {
    const Span sp = plan.m2m.at(l);
    if (!sp.n) continue;
    const int* map = map_m2m + (size_t)(sp.off - plan.m2m.off) * 4;
    HIP_TRY(launch_xlate(c.st, S, W + sp.off, sp.n, map, 4, 4, T2, P, P, mp, per, mp, per, nd, 0.0));
  }
  tm->mark(c.st);
  // step 2: list 4
  if (plan.list4.n)
    k_form<false><<<nblk(plan.list4.n, 4), WAVE, 0, c.st>>>(W + plan.list4.off, plan.list4.n, D,
                                                           Pt.src, chg, dip, Pt.ns, nd, p, lo);
  tm->mark(c.st);
  // step 4: M2L level by level (log of the box size enters entry (0,0) only);
  // the slot maps of every level in one launch
  int* map2 = S.get<int>((size_t)std::max(plan.m2l.n, 1) * 49);
  if (S.err) return S.err;
  if (plan.m2l.n) {
    fill_async(map2, 0xffffffffu, 49 * (size_t)plan.m2l.n, c.st);
    k_map_m2l_all<<<nblk(32L * plan.m2l.n, 256), 256, 0, c.st>>>(W + plan.m2l.off, plan.m2l.n, D,
                                                                map2);
  }
  std::vector<M2LLevel> m2l_levels;
  for (int l = 2; l <= L; ++l) {
    const Span sp = plan.m2l.at(l);
    if (!sp.n) continue;
    const int* ml = map2 + (size_t)(sp.off - plan.m2l.off) * 49;
    // (a forest's boxes of one level differ in size: each target's own)
    m2l_levels.push_back({W + sp.off, sp.n, ml, TM, P, std::log(plan.bs[l]),
                          a.mixed ? tab.m2l32 : nullptr, forest ? D.bsb : nullptr});
  }
  if (!m2l_levels.empty()) HIP_TRY(launch_m2l_levels(c.st, m2l_levels, mp, per, lo, per, nd, &S));
  tm->mark(c.st);
  // step 5: L2L into the children of boxes that hold targets
  for (int l = 1; l <= L - 1; ++l) {
    const Span sp = plan.l2l_kids.at(l);
    if (!sp.n) continue;
    const int* map = map_l2l + (size_t)(sp.off - plan.l2l_kids.off) * 4;
    HIP_TRY(launch_xlate(c.st, S, W + sp.off, sp.n, map, 4, 1, TL, P, P, lo, per, lo, per, nd, 0.0));
  }
  // steps 6 + 7: list 3 and local expansions at sources / targets
  struct Set {
    int pg; Span l3, lv, nn; const int2* rng; const double2* pts; int N; double2 *po, *gr, *he;
    const LapOut* out; bool* fused;
  };
  Set sets[2] = {{a.ifpgh, plan.list3_s.all(), plan.leaves_s.all(), plan.p2p_s.all(), D.srng, Pt.src, Pt.ns, a.pot, a.grad, a.hess, &a.out_s, &a.fused_s},
                 {a.ifpghtarg, plan.list3_t.all(), plan.leaves_t.all(), plan.p2p_t.all(), D.trng, Pt.trg, Pt.nt, a.pott, a.gradt, a.hesst, &a.out_t, &a.fused_t}};
  a.fused_s = a.fused_t = false;
  // reference skip tests: charges (with or without dipoles) use rr vs thresh^2,
  // '<=' for potential only and '<' otherwise; dipole-only routines compare |z|
  const double thresh = Pt.bs0 * 0x1p-51;
  tm->mark(c.st);
  auto eval = [&](bool mpk, const Set& s, const Span& sp, const double2* ex) {
    if (s.pg == 1) launch_eval<1>(mpk, sp.n, c.st, W + sp.off, D, s.rng, s.pts, s.N, nd, p, ex, s.po, s.gr, s.he);
    if (s.pg == 2) launch_eval<2>(mpk, sp.n, c.st, W + sp.off, D, s.rng, s.pts, s.N, nd, p, ex, s.po, s.gr, s.he);
    if (s.pg == 3) launch_eval<3>(mpk, sp.n, c.st, W + sp.off, D, s.rng, s.pts, s.N, nd, p, ex, s.po, s.gr, s.he);
  };
  for (auto& s : sets)
    if (s.pg > 0 && s.N > 0 && s.l3.n) eval(true, s, s.l3, mp);
  tm->mark(c.st);
  for (auto& s : sets)
    if (s.pg > 0 && s.N > 0 && s.lv.n) eval(false, s, s.lv, lo);
  tm->mark(c.st);
  for (auto& s : sets) {
    if (s.pg <= 0 || s.N == 0) continue;
    if (a.ifnear && s.nn.n) {
      int mode = s.pg == 1 ? 0 : 1;
      double th2 = thresh * thresh;
      if (a.mixed && s.pg <= 2) {
        // FP32 pair terms (the Hessian's 1/z^3 can leave the FP32 range,
        // so pg = 3 stays FP64)
        const double sc = Pt.bs0 > 0.0 ? std::ldexp(1.0, -std::ilogb(Pt.bs0)) : 1.0;
        *s.fused = s.out->pot != nullptr;
        if (chg && dip)
          launch_p2p32<true, true>(s.pg, mode, s.nn.n, c.st, W + s.nn.off, D, Pt.src, chg, dip, Pt.ns, s.rng, s.pts, s.N, nd, th2, sc, s.po, s.gr, *s.out);
        else if (chg)
          launch_p2p32<true, false>(s.pg, mode, s.nn.n, c.st, W + s.nn.off, D, Pt.src, chg, dip, Pt.ns, s.rng, s.pts, s.N, nd, th2, sc, s.po, s.gr, *s.out);
        else
          launch_p2p32<false, true>(s.pg, mode, s.nn.n, c.st, W + s.nn.off, D, Pt.src, chg, dip, Pt.ns, s.rng, s.pts, s.N, nd, th2, sc, s.po, s.gr, *s.out);
        continue;
      }
      *s.fused = s.out->pot != nullptr;
      if (chg && dip)
        launch_p2p_pg<true, true>(s.pg, mode, s.nn.n, c.st, W + s.nn.off, D, Pt.src, chg, dip, Pt.ns, s.rng, s.pts, s.N, nd, th2, s.po, s.gr, s.he, *s.out);
      else if (chg)
        launch_p2p_pg<true, false>(s.pg, mode, s.nn.n, c.st, W + s.nn.off, D, Pt.src, chg, dip, Pt.ns, s.rng, s.pts, s.N, nd, th2, s.po, s.gr, s.he, *s.out);
      else
        launch_p2p_pg<false, true>(s.pg, mode, s.nn.n, c.st, W + s.nn.off, D, Pt.src, chg, dip, Pt.ns, s.rng, s.pts, s.N, nd, th2, s.po, s.gr, s.he, *s.out);
    }
  }
  tm->mark(c.st);
  HIP_TRY(hipGetLastError());
  return IER_OK;
}

}  // namespace fmm

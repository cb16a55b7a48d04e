// helm.hip -- Helmholtz FMM kernels (FP64) and their launch sequence,
// low-frequency path (boxes of at most 16 wavelengths).
//
// Conventions are those of c_translation/src/helmrouts2d.c: expansions of
// order p over n = -p..p, multipoles scaled as M_n / rs^|n| and locals as
// L_n rs^|n| with rs = min(|zk bs / 2 pi|, 1) per level (hfmm2d.c:1768-1776);
// a field is evaluated as (i/4) sum_n e_n v_n(k r) e^{i n theta} with
// v = scaled H (multipole) or scaled J (local).  Expansions live box-major:
// [box][channel][n + p], box stride nd*(2*pmax+1).
#include <algorithm>
#include <cmath>
#include <complex>
#include <map>
#include <mutex>
#include <tuple>
#include <type_traits>
#include <vector>

#include "engine.h"
#include "special.h"
#include "translate.h"

namespace fmm {

namespace {
constexpr int WAVE = 64;
inline int nblk(int n, int b) { return (n + b - 1) / b; }
__device__ inline void addc(double2* o, cd v) { *o = make_double2(o->x + v.x, o->y + v.y); }
__device__ inline cd ld(const double2* p) { return cd{p->x, p->y}; }

struct HLev {   // per-level parameters on the device
  int p;
  double rs;
};

// orders LDS writes and reads among the lanes of one wavefront
__device__ inline void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}


// ---------------------------------------------------------------------
// formation: multipoles of leaves (scaled J) and locals from list 4
// (scaled H); charge part helmrouts2d.c:43-64, dipole part :67-105
// ---------------------------------------------------------------------
// Bessel tables of one point, kept out of line: their constants then stay
// inside the call instead of being held in registers across the kernel
template <int N>
__device__ __noinline__ void hvals_real_point(int nmax, double x, double rs, cd* out, int stride) {
  hvals_real<N>(nmax, x, rs, out, stride);
}
__device__ __noinline__ void jvals_real_point(int nmax, double x, double rs, double* out,
                                              int stride) {
  jvals_real(nmax, x, rs, out, stride);
}

// NR > 0: real positive wavenumber (for list 4 also the length of the
// H0/H1 series, valid over the launch's k r): J tables are real (VT =
// double) and every Bessel value comes from real arithmetic; NR = 0:
// complex path.
template <bool MP, int NR>
__global__ void __launch_bounds__(WAVE)
k_hform(const int* work, DTree D, const HLev* lev, cd zk, const double2* pts, const double2* chg,
        const double2* dip, const double* dv, int N, int nd, int CH, size_t per, double2* out) {
  using VT = typename std::conditional<MP && (NR > 0), double, cd>::type;
  extern __shared__ double2 sm[];
  const int b = work[blockIdx.x];
  const int lane = threadIdx.x;
  const int L = D.level[b];
  const int p = lev[L].p;
  const double rs = lev[L].rs;
  const int Q = p + 2;                   // v_0 .. v_{p+1} (J) / .. v_{p+2} (H)
  cd* Z = reinterpret_cast<cd*>(sm);     // [Q+1][CH] powers of e^{-i theta}
  cd* F = Z + (size_t)(Q + 1) * CH;      // [6][CH] per-point factors of one channel
  VT* V = reinterpret_cast<VT*>(F + (size_t)6 * CH);  // [Q+1][CH]
  const double2 ctr = D.ctr[b];
  const double rsd = MP ? rs : 1.0 / rs;  // dipole scale (locals use 1/rs)
  const int nsb = MP ? 1 : D.o4[b + 1] - D.o4[b];
  for (int sb = 0; sb < nsb; ++sb) {
    int jb = MP ? b : D.l4[D.o4[b] + sb];
    int2 rg = D.srng[jb];
    for (int c0 = rg.x; c0 < rg.y; c0 += CH) {
      const int cnt = min(CH, rg.y - c0);
      for (int j = lane; j < cnt; j += WAVE) {
        double2 s = pts[c0 + j];
        double dx = s.x - ctr.x, dy = s.y - ctr.y;
        double r = sqrt(dx * dx + dy * dy);
        // e^{i theta} = (dx, dy) / r, theta = 0 at the centre
        const double ri = r > 0.0 ? 1.0 / r : 0.0;
        const double ct = r > 0.0 ? dx * ri : 1.0, st = dy * ri;
        if constexpr (NR > 0) {
          const double x = zk.x * r;
          if constexpr (MP) jvals_real_point(p + 1, x, rs, reinterpret_cast<double*>(V) + j, CH);
          else hvals_real_point<(NR > 0 ? NR : 1)>(p + 2, x, rs, reinterpret_cast<cd*>(V) + j, CH);
        } else {
          cd z = zk * r;
          if (MP) jvals(p + 1, z, rs, reinterpret_cast<cd*>(V) + j, CH);
          else hvals(p + 2, z, rs, reinterpret_cast<cd*>(V) + j, CH);
        }
        cd zm = cd{ct, -st}, pw = mk(1, 0);   // e^{-i theta}
        for (int n = 0; n <= Q; ++n) {
          Z[(size_t)n * CH + j] = pw;
          pw = pw * zm;
        }
      }
      __syncthreads();
      for (int d = 0; d < nd; ++d) {
        // per-point factors of this channel: charge, and the dipole terms of
        // orders m > 0 (TA, TB), m < 0 (TC, TD) and m = 0 (K0), so that the
        // coefficient loop below is the same for every lane
        for (int j = lane; j < cnt; j += WAVE) {
          const size_t g = (size_t)d * N + c0 + j;
          F[j] = chg ? ld(chg + g) : mk(0, 0);
          if (dip) {
            const cd qf = ld(dip + g) * (zk * -0.5);
            const double v1 = dv[g], v2 = dv[(size_t)nd * N + g];
            const cd zm = Z[(size_t)CH + j], zi = conj(zm);
            F[CH + j] = ((zi / rsd) * mk(-v1, v2)) * qf;
            F[2 * CH + j] = ((zm * rsd) * mk(v1, v2)) * qf;
            F[3 * CH + j] = ((zm / rsd) * mk(-v1, -v2)) * qf;
            F[4 * CH + j] = ((zi * rsd) * mk(v1, -v2)) * qf;
            F[5 * CH + j] = (qf * rsd) * ((zm + zi) * v1 + mul_i(zm - zi) * v2);
          }
        }
        __syncthreads();
        for (int m = lane - p; m <= p; m += WAVE) {
          const int n = m < 0 ? -m : m;
          const bool neg = m < 0;
          const int ia = m == 0 ? 1 : n - 1, ib = m == 0 ? 1 : n + 1;
          const cd* FA = F + (m > 0 ? CH : neg ? 3 * CH : 5 * CH);
          const cd* FB = F + (m > 0 ? 2 * CH : 4 * CH);
          const double sb = m == 0 ? 0.0 : 1.0;  // m = 0 has one dipole term
          cd acc = mk(0, 0);
          for (int j = 0; j < cnt; ++j) {
            const cd zn = Z[(size_t)n * CH + j];
            const cd ph = neg ? ((n & 1) ? -conj(zn) : conj(zn)) : zn;  // (-zinv)^n for m < 0
            cd t = F[j] * V[(size_t)n * CH + j];
            if (dip) t += FA[j] * V[(size_t)ia * CH + j] + FB[j] * (V[(size_t)ib * CH + j] * sb);
            acc += t * ph;
          }
          addc(out + (size_t)b * per + (size_t)d * (2 * p + 1) + (m + p), acc);
        }
        __syncthreads();  // F is rewritten for the next channel
      }
    }
  }
}

// ---------------------------------------------------------------------
// evaluation of an expansion and its derivatives at a point.
// Derivative expansions (helmrouts2d.c:187-287) are formed coefficient by
// coefficient: with z1 = zk/(2 rs'), z2 = zk rs'/2,
//   d/dx: e'_k = z2 e_{k+1} - z1 e_{k-1} (k>0), z2 (e_1 - e_-1) (k=0),
//               z1 e_{k+1} - z2 e_{k-1} (k<0);
//   d/dy: e'_k = i (z2 e_{k+1} + z1 e_{k-1}), i z2 (e_1 + e_-1), i (z1 e_{k+1} + z2 e_{k-1}).
// ---------------------------------------------------------------------
struct Ex {
  const double2* e;
  int p;
  cd z1, z2;
  __device__ cd c(int k) const { return (k < -p || k > p) ? mk(0, 0) : ld(e + k + p); }
  __device__ cd dx(int k) const {  // valid for |k| <= p+1
    if (k > 0) return z2 * c(k + 1) - z1 * c(k - 1);
    if (k < 0) return z1 * c(k + 1) - z2 * c(k - 1);
    return z2 * (c(1) - c(-1));
  }
  __device__ cd dy(int k) const {
    if (k > 0) return mul_i(z2 * c(k + 1) + z1 * c(k - 1));
    if (k < 0) return mul_i(z1 * c(k + 1) + z2 * c(k - 1));
    return mul_i(z2 * (c(1) + c(-1)));
  }
  template <bool Y>
  __device__ cd d1(int k) const { return Y ? dy(k) : dx(k); }
  // second derivatives: apply the x or y rule to the first-derivative expansion
  template <bool Y1, bool Y2>
  __device__ cd d2(int k) const {
    if (k > 0) {
      cd a = d1<Y1>(k + 1), b = d1<Y1>(k - 1);
      return Y2 ? mul_i(z2 * a + z1 * b) : z2 * a - z1 * b;
    }
    if (k < 0) {
      cd a = d1<Y1>(k + 1), b = d1<Y1>(k - 1);
      return Y2 ? mul_i(z1 * a + z2 * b) : z1 * a - z2 * b;
    }
    cd a = d1<Y1>(1), b = d1<Y1>(-1);
    return Y2 ? mul_i(z2 * (a + b)) : z2 * (a - b);
  }
};

struct EvalAcc {
  cd P, Gx, Gy, Hxx, Hxy, Hyy;
};
struct NearAcc {
  cd Pc, Pd, Gx, Gy, Hxx, Hxy, Hyy;  // Pc carries i/4, the rest i k/4
};

// contribution of order n (with value v = scaled H or J at the point,
// complex or, for J of a real argument, real; pos = e^{i n theta},
// neg = (-e^{-i theta})^n)
template <int PG, class EA, class V>
__device__ inline void eval_term(const EA& E, int n, V v, cd pos, cd neg, EvalAcc& A) {
  if (n == 0) {
    A.P += v * E.c(0);
    if (PG >= 2) { A.Gx += v * E.dx(0); A.Gy += v * E.dy(0); }
    if (PG >= 3) {
      A.Hxx += v * E.template d2<false, false>(0);
      A.Hxy += v * E.template d2<false, true>(0);
      A.Hyy += v * E.template d2<true, true>(0);
    }
    return;
  }
  const int p = E.p;
  // the order's value times its two phases, shared by every quantity
  const cd wp = v * pos, wn = v * neg;
  if (n <= p) A.P += E.c(n) * wp + E.c(-n) * wn;
  if (PG >= 2 && n <= p + 1) {
    A.Gx += E.dx(n) * wp + E.dx(-n) * wn;
    A.Gy += E.dy(n) * wp + E.dy(-n) * wn;
  }
  if (PG >= 3 && n <= p + 2) {
    A.Hxx += E.template d2<false, false>(n) * wp + E.template d2<false, false>(-n) * wn;
    A.Hxy += E.template d2<false, true>(n) * wp + E.template d2<false, true>(-n) * wn;
    A.Hyy += E.template d2<true, true>(n) * wp + E.template d2<true, true>(-n) * wn;
  }
}

// multipole at a point: v_n = H_n rs^n by upward recurrence, n = 0..p+2
template <int PG, class EA>
__device__ void eval_mp_at(const EA& E, cd zk, double rs, double r, cd e1, EvalAcc& A) {
  cd z = zk * r;
  if (cabs_(z) < 1e-200) return;
  cd h0, h1;
  hank01(z, h0, h1);
  cd pos = mk(1, 0), neg = mk(1, 0);
  cd zi = inv(z) * rs;
  double s2 = rs * rs;
  cd vm = h0, v = h1 * rs;
  eval_term<PG>(E, 0, h0, pos, neg, A);
  const int top = E.p + (PG >= 3 ? 2 : PG >= 2 ? 1 : 0);
  for (int n = 1; n <= top; ++n) {
    pos = pos * e1;
    neg = (n & 1) ? -conj(pos) : conj(pos);  // (-e^{-i theta})^n
    eval_term<PG>(E, n, v, pos, neg, A);
    cd vn = zi * v * (2.0 * n) - vm * s2;
    vm = v;
    v = vn;
  }
}

// local at a point: v_n = J_n / rs^n by Miller's downward recurrence with
// the generating-function normalisation applied at the end
template <int PG, class EA>
__device__ void eval_lo_at(const EA& E, cd zk, double rs, double r, cd e1, EvalAcc& A) {
  cd z = zk * r;
  const int top_n = E.p + (PG >= 3 ? 2 : PG >= 2 ? 1 : 0);
  if (cabs_(z) < 1e-200) {  // J_0 = 1, others 0
    eval_term<PG>(E, 0, mk(1, 0), mk(1, 0), mk(1, 0), A);
    return;
  }
  int top = miller_top(top_n, z);
  cd zr = inv(z) * rs;
  double r2 = rs * rs;
  cd u = z.y >= 0 ? mk(0, -1) : mk(0, 1), ur = u * rs;
  // phases at n = top_n downward
  cd pos = cpow_n(e1, top_n);
  cd ie1 = conj(e1);
  EvalAcc B = {mk(0, 0), mk(0, 0), mk(0, 0), mk(0, 0), mk(0, 0), mk(0, 0)};
  cd gp = mk(0, 0), g = mk(1, 0), H = mk(0, 0);
  for (int n = top - 1; n >= 1; --n) {
    if (n <= top_n) {
      cd neg = (n & 1) ? -conj(pos) : conj(pos);
      eval_term<PG>(E, n, g, pos, neg, B);
      pos = pos * ie1;
    }
    H = ur * (g + H);
    cd gm = zr * g * (2.0 * n) - gp * r2;
    gp = g;
    g = gm;
    if (norm2(g) > 1e40) {
      g = g * 1e-40; gp = gp * 1e-40; H = H * 1e-40;
      B.P = B.P * 1e-40; B.Gx = B.Gx * 1e-40; B.Gy = B.Gy * 1e-40;
      B.Hxx = B.Hxx * 1e-40; B.Hxy = B.Hxy * 1e-40; B.Hyy = B.Hyy * 1e-40;
    }
  }
  eval_term<PG>(E, 0, g, mk(1, 0), mk(1, 0), B);
  cd sc = inv((g + H * 2.0) * cexp_(-(u * z)));
  A.P += B.P * sc; A.Gx += B.Gx * sc; A.Gy += B.Gy * sc;
  A.Hxx += B.Hxx * sc; A.Hxy += B.Hxy * sc; A.Hyy += B.Hyy * sc;
}

// Real positive wavenumber, x = k r: the same two evaluations with real
// Bessel arithmetic.  Multipole: H0, H1 from hank_series_real<NR> (the
// launch guarantees x <= JY_XMAX for NR) and the upward recurrence with real
// multipliers.  Local: J_n / rs^n by Miller's downward recurrence, normalised
// with the Neumann sum J_0 + 2 sum_k J_2k = 1 (Abramowitz & Stegun 9.1.46).
struct H01 {
  cd h0, h1;
};
// out of line, so the series constants are not held in registers across the
// caller's loops
template <int N>
__device__ __noinline__ H01 hank_real_point(double x) {
  H01 h;
  hank_series_real<N>(x, h.h0, h.h1);
  return h;
}

// ---------------------------------------------------------------------
// Formation for a real positive wavenumber with every lane busy: four
// boxes per wave, a 16-lane group per box, lane u on point u of each chunk
// of 16 source points.  A lane streams its point's Bessel values in
// registers -- H_n rs^n by the upward recurrence for locals, J_n / rs^n
// for multipoles by Miller's downward recurrence run twice (once for the
// normalisation, once to emit the values, which equal jvals_real's bit for
// bit) -- and forms its terms of the coefficients in the order they become
// available.  A tile in LDS (FG_ROWS coefficients of 16 terms) then hands
// the 16 terms of each coefficient to one lane, which sums them in point
// order and adds the sum to the expansion.  The per-point factors (charge; dipole terms of orders
// m > 0, m < 0 and m = 0) and the term of coefficient m,
//   t_m = (F0 V_n + FA V_{n-1} + FB V_{n+1}) e^{-i m theta},  n = |m|,
// are those of k_hform (helmrouts2d.c:294-391, 513-612).  Needs p <= 63.
// ---------------------------------------------------------------------
constexpr int FG_PAD = 17;  // tile row length (16 terms + 1: no bank conflicts)
// tile rows: coefficients summed per flush.  8 rows keep the tile at 8.5 KB
// per wave, so LDS admits four waves per SIMD (16 rows: 17 KB, two)
constexpr int FG_ROWS = 8;

template <bool MP, int NR>
__global__ void __launch_bounds__(WAVE)
k_hform_g(const int* work, int nwork, DTree D, const HLev* lev, double k, const double2* pts,
          const double2* chg, const double2* dip, const double* dv, int N, int nd, size_t per,
          double2* out) {
  __shared__ cd tile[4 * FG_ROWS * FG_PAD];
  const int grp = (int)threadIdx.x >> 4, u = (int)threadIdx.x & 15;
  const int w = blockIdx.x * 4 + grp;
  if (w >= nwork) return;  // the whole group leaves; only wave-level syncs below
  cd* T = tile + grp * FG_ROWS * FG_PAD;
  const int b = work[w];
  const int L = D.level[b];
  const int p = lev[L].p;
  const double rs = lev[L].rs;
  const double rsd = MP ? rs : 1.0 / rs;
  const double2 ctr = D.ctr[b];
  const int nsb = MP ? 1 : D.o4[b + 1] - D.o4[b];
  const int ncoef = 2 * p + 1;
  // coefficient of emission e: locals ascending 0, +1, -1, +2, ...;
  // multipoles descending +p, -p, +(p-1), ..., 0
  auto m_of = [&](int e) {
    if (MP) return e == 2 * p ? 0 : ((e & 1) ? -(p - e / 2) : p - e / 2);
    return e == 0 ? 0 : ((e & 1) ? (e + 1) / 2 : -(e / 2));
  };
  for (int d = 0; d < nd; ++d) {
    double2* o = out + (size_t)b * per + (size_t)d * ncoef + p;  // o[m], |m| <= p
    // source ranges of 16 boxes at a time, one box per lane (no chain of
    // dependent loads per box), handed out by shuffles
    int2 mrg = make_int2(0, 0);
    auto load_rng = [&](int sb0) {
      if (sb0 + u < nsb) mrg = D.srng[MP ? b : D.l4[D.o4[b] + sb0 + u]];
    };
    load_rng(0);
    for (int sb = 0; sb < nsb; ++sb) {
      if (sb > 0 && (sb & 15) == 0) load_rng(sb);
      const int2 rg = make_int2(__shfl(mrg.x, sb & 15, 16), __shfl(mrg.y, sb & 15, 16));
      for (int c0 = rg.x; c0 < rg.y; c0 += 16) {
        const int j = c0 + u;
        const bool ok = j < rg.y;
        const int jj = ok ? j : rg.x;
        const double2 s = pts[jj];
        const double dx = s.x - ctr.x, dy = s.y - ctr.y;
        const double r = sqrt(dx * dx + dy * dy);
        const double ri = r > 0.0 ? 1.0 / r : 0.0;
        const double ct = r > 0.0 ? dx * ri : 1.0, st = dy * ri;
        const cd zm = cd{ct, -st}, zi = cd{ct, st};  // e^{-i theta}, e^{i theta}
        // per-point factors of this channel (zero for lanes past the chunk)
        const size_t g = (size_t)d * N + jj;
        cd F0 = mk(0, 0), FAp = mk(0, 0), FBp = mk(0, 0), FAm = mk(0, 0), FBm = mk(0, 0),
           F0d = mk(0, 0);
        if (ok && chg) F0 = ld(chg + g);
        if (ok && dip) {
          const cd qf = ld(dip + g) * (-0.5 * k);
          const double v1 = dv[g], v2 = dv[(size_t)nd * N + g];
          FAp = ((zi / rsd) * mk(-v1, v2)) * qf;
          FBp = ((zm * rsd) * mk(v1, v2)) * qf;
          FAm = ((zm / rsd) * mk(-v1, -v2)) * qf;
          FBm = ((zi * rsd) * mk(v1, -v2)) * qf;
          F0d = (qf * rsd) * ((zm + zi) * v1 + mul_i(zm - zi) * v2);
        }
        int e = 0;  // emissions so far (uniform over the group)
        // lane u < FG_ROWS adds coefficient base + u of each group of FG_ROWS
        // emissions; its current value is loaded one group ahead, so the
        // load's latency hides behind the group's emissions (each coefficient
        // has one writer, this lane, and is written once per group)
        double2 pre = u < FG_ROWS && u < ncoef ? o[m_of(u)] : make_double2(0.0, 0.0);
        auto emit = [&](cd term) {
          T[(e % FG_ROWS) * FG_PAD + u] = term;
          ++e;
          if (e % FG_ROWS == 0 || e == ncoef) {
            const int base = (e - 1) / FG_ROWS * FG_ROWS, nfill = e - base;
            wave_sync();
            if (u < nfill) {
              cd sum = mk(0, 0);
#pragma unroll 4
              for (int q = 0; q < 16; ++q) sum += T[u * FG_PAD + q];
              o[m_of(base + u)] = make_double2(pre.x + sum.x, pre.y + sum.y);
            }
            if (u < FG_ROWS && base + FG_ROWS + u < ncoef) pre = o[m_of(base + FG_ROWS + u)];
            wave_sync();
          }
        };
        // terms of orders +n and -n from V_{n-1}, V_n, V_{n+1} and Z_n = e^{-i n theta}
        auto pair = [&](auto vm, auto v0, auto vp, cd zn, int n) {
          const cd tp = F0 * v0 + FAp * vm + FBp * vp;
          const cd tm = F0 * v0 + FAm * vm + FBm * vp;
          const cd zc = (n & 1) ? -conj(zn) : conj(zn);  // (-e^{i theta})^n
          emit(tp * zn);
          emit(tm * zc);
        };
        const double x = k * r;
        if constexpr (!MP) {
          // H_n rs^n upward (hvals_real); every list-4 source is off centre
          const H01 hh = hank_real_point<(NR > 0 ? NR : 1)>(x);
          const double zr = rs / x, s2 = rs * rs;
          cd vm = hh.h0, v0 = hh.h1 * rs;  // V_0, V_1
          emit(F0 * vm + F0d * v0);        // m = 0: one dipole term
          cd zn = mk(1, 0);
          for (int n = 1; n <= p; ++n) {
            const cd vp = v0 * (zr * (2.0 * n)) - vm * s2;  // V_{n+1}
            zn = zn * zm;
            pair(vm, v0, vp, zn, n);
            vm = v0;
            v0 = vp;
          }
        } else {
          // J_n / rs^n, n = 0..p+1 (jvals_real), emitted downward.  Each
          // lane's recurrence starts at its own top; the loop index is the
          // group's, so every lane emits at the same iterations
          const int nmax = p + 1;
          cd zn = mk(1, 0);
          for (int n = 1; n <= p; ++n) zn = zn * zm;  // Z_p
          const bool centre = x < 1e-200;             // J_0 = 1, the rest 0
          const double xi = centre ? 0.0 : 1.0 / x;
          int top = nmax + 1;
          if (!centre) {
            top = nmax + 10000;
            double fn = 1.0, fm = 0.0;
            for (int i = nmax; i <= nmax + 10000; ++i) {
              const double fp = xi * fn * (2.0 * i) - fm;
              if (fp * fp > 1e40) {
                top = i + 1;
                break;
              }
              fm = fn;
              fn = fp;
            }
          }
          int topg = top;
#pragma unroll
          for (int q = 8; q >= 1; q >>= 1) topg = max(topg, __shfl_xor(topg, q, 16));
          const double zr = xi * rs, r2 = rs * rs;
          // pass 1: normalisation and the number of rescalings
          double gp = 0.0, g1 = 1.0, H = 0.0;
          int R = 0;
          if (!centre) {
            for (int n = top - 1; n >= 1; --n) {
              H = rs * ((n & 1) ? H : g1 + H);
              const double gm = zr * g1 * (2.0 * n) - gp * r2;
              gp = g1;
              g1 = gm;
              if (g1 * g1 > 1e40) {
                g1 *= 1e-40; gp *= 1e-40; H *= 1e-40;
                ++R;
              }
            }
          }
          const double sc = centre ? 1.0 : 1.0 / (g1 + 2.0 * H);
          // pass 2: the same recurrence; J_n = g_n (1e-40)^(rescalings at
          // or below n) sc
          double g = 1.0;
          gp = 0.0;
          int seen = 0;
          double va = 0.0, vb = 0.0;  // J_{n+1}, J_{n+2} of the next emission
          for (int n = topg - 1; n >= 1; --n) {
            if (n <= nmax) {  // n < top here: the lane's recurrence has started
              double v = 0.0;
              if (!centre) {
                v = g;
                for (int q = seen; q < R; ++q) v *= 1e-40;
                v *= sc;
              }
              if (n <= p - 1) {  // J_n completes order n + 1
                pair(v, va, vb, zn, n + 1);
                zn = zn * zi;
              }
              vb = va;
              va = v;
            }
            if (!centre && n <= top - 1) {
              const double gm = zr * g * (2.0 * n) - gp * r2;
              gp = g;
              g = gm;
              if (g * g > 1e40) {
                g *= 1e-40; gp *= 1e-40;
                ++seen;
              }
            }
          }
          const double j0 = centre ? 1.0 : g * sc;
          pair(j0, va, vb, zn, 1);    // order 1: J_0, J_1, J_2
          emit(F0 * j0 + F0d * va);   // order 0: J_0 and J_1
        }
      }
    }
  }
}

template <int PG, int NR, class EA>
__device__ void eval_mp_at_real(const EA& E, double k, double rs, double r, cd e1,
                                EvalAcc& A) {
  const double x = k * r;
  if (x < 1e-200) return;
  const H01 hh = hank_real_point<NR>(x);
  const cd h0 = hh.h0, h1 = hh.h1;
  cd pos = mk(1, 0), neg = mk(1, 0);
  const double zi = rs / x, s2 = rs * rs;
  cd vm = h0, v = h1 * rs;
  eval_term<PG>(E, 0, h0, pos, neg, A);
  const int top = E.p + (PG >= 3 ? 2 : PG >= 2 ? 1 : 0);
  for (int n = 1; n <= top; ++n) {
    pos = pos * e1;
    neg = (n & 1) ? -conj(pos) : conj(pos);  // (-e^{-i theta})^n
    eval_term<PG>(E, n, v, pos, neg, A);
    cd vn = v * (zi * (2.0 * n)) - vm * s2;
    vm = v;
    v = vn;
  }
}

__device__ inline int miller_top_real(int nterms, double x) {
  const double xi = 1.0 / x;
  double fn = 1.0, fm = 0.0;
  for (int i = nterms; i <= nterms + 10000; ++i) {
    const double fp = xi * fn * (2.0 * i) - fm;
    if (fp * fp > 1e40) return i + 1;
    fm = fn;
    fn = fp;
  }
  return nterms + 10000;
}

template <int PG, class EA>
__device__ void eval_lo_at_real(const EA& E, double k, double rs, double r, cd e1,
                                EvalAcc& A) {
  const double x = k * r;
  const int top_n = E.p + (PG >= 3 ? 2 : PG >= 2 ? 1 : 0);
  if (x < 1e-200) {  // J_0 = 1, others 0
    eval_term<PG>(E, 0, 1.0, mk(1, 0), mk(1, 0), A);
    return;
  }
  const int top = miller_top_real(top_n, x);
  const double zr = rs / x, r2 = rs * rs;
  cd pos = cpow_n(e1, top_n);
  const cd ie1 = conj(e1);
  EvalAcc B = {mk(0, 0), mk(0, 0), mk(0, 0), mk(0, 0), mk(0, 0), mk(0, 0)};
  double gp = 0.0, g = 1.0, H = 0.0;  // H = sum over even n >= 2 of rs^n g_n (Horner)
  for (int n = top - 1; n >= 1; --n) {
    if (n <= top_n) {
      cd neg = (n & 1) ? -conj(pos) : conj(pos);
      eval_term<PG>(E, n, g, pos, neg, B);
      pos = pos * ie1;
    }
    H = rs * ((n & 1) ? H : g + H);
    const double gm = zr * g * (2.0 * n) - gp * r2;
    gp = g;
    g = gm;
    if (g * g > 1e40) {
      g *= 1e-40; gp *= 1e-40; H *= 1e-40;
      B.P = B.P * 1e-40; B.Gx = B.Gx * 1e-40; B.Gy = B.Gy * 1e-40;
      B.Hxx = B.Hxx * 1e-40; B.Hxy = B.Hxy * 1e-40; B.Hyy = B.Hyy * 1e-40;
    }
  }
  eval_term<PG>(E, 0, g, mk(1, 0), mk(1, 0), B);
  const double sc = 1.0 / (g + 2.0 * H);
  A.P += B.P * sc; A.Gx += B.Gx * sc; A.Gy += B.Gy * sc;
  A.Hxx += B.Hxx * sc; A.Hxy += B.Hxy * sc; A.Hyy += B.Hyy * sc;
}

// An expansion staged in LDS per order n = 0 .. top for the staged
// evaluation: with pos = e^{i n theta} = C + i S and the negative order's
// phase (-e^{-i theta})^n = (-1)^n conj(pos), a quantity Q (the expansion,
// or its x or y derivative expansion) contributes v (Q(n) pos + Q(-n)
// (-1)^n conj(pos)) = v (s C + d S) at order n, with
//   s = Q(n) + (-1)^n Q(-n),  d = i (Q(n) - (-1)^n Q(-n))   (s = Q(0), d = 0
// at n = 0).  Entries (s, d) of the expansion, then of the x and y
// derivatives (PG >= 2), NQ = 2 or 6 per order.
template <int PG>
struct SdTerm {
  static constexpr int NQ = PG >= 2 ? 6 : 2;
  // v (s C + d S) for each staged quantity of order n
  template <class V>
  __device__ static void add(const cd* e, V v, double C, double S, EvalAcc& A) {
    A.P += v * (e[0] * C + e[1] * S);
    if (PG >= 2) {
      A.Gx += v * (e[2] * C + e[3] * S);
      A.Gy += v * (e[4] * C + e[5] * S);
    }
  }
};

// multipole at a point from the staged form, real positive k: H_n rs^n by
// the upward recurrence from hank_series_real<NR> values
template <int PG, int NR>
__device__ void eval_mp_sd_real(const cd* E, int top, double k, double rs, double r, cd e1,
                                EvalAcc& A) {
  using T = SdTerm<PG>;
  const double x = k * r;
  if (x < 1e-200) return;
  const H01 hh = hank_real_point<NR>(x);
  const double zi = rs / x, s2 = rs * rs;
  cd vm = hh.h0, v = hh.h1 * rs, pos = mk(1, 0);
  T::add(E, hh.h0, 1.0, 0.0, A);
  for (int n = 1; n <= top; ++n) {
    pos = pos * e1;
    T::add(E + T::NQ * n, v, pos.x, pos.y, A);
    const cd vn = v * (zi * (2.0 * n)) - vm * s2;
    vm = v;
    v = vn;
  }
}

// multipole at a point from the staged form, any k
template <int PG>
__device__ void eval_mp_sd(const cd* E, int top, cd zk, double rs, double r, cd e1, EvalAcc& A) {
  using T = SdTerm<PG>;
  const cd z = zk * r;
  if (cabs_(z) < 1e-200) return;
  cd h0, h1;
  hank01(z, h0, h1);
  const cd zi = inv(z) * rs;
  const double s2 = rs * rs;
  cd vm = h0, v = h1 * rs, pos = mk(1, 0);
  T::add(E, h0, 1.0, 0.0, A);
  for (int n = 1; n <= top; ++n) {
    pos = pos * e1;
    T::add(E + T::NQ * n, v, pos.x, pos.y, A);
    const cd vn = zi * v * (2.0 * n) - vm * s2;
    vm = v;
    v = vn;
  }
}

// local at a point from the staged form, real positive k: J_n / rs^n by
// Miller's downward recurrence, normalised with J_0 + 2 sum_k J_2k = 1
template <int PG>
__device__ void eval_lo_sd_real(const cd* E, int top_n, double k, double rs, double r, cd e1,
                                EvalAcc& A) {
  using T = SdTerm<PG>;
  const double x = k * r;
  if (x < 1e-200) {  // J_0 = 1, others 0
    T::add(E, 1.0, 1.0, 0.0, A);
    return;
  }
  const int top = miller_top_real(top_n, x);
  const double zr = rs / x, r2 = rs * rs;
  cd pos = cpow_n(e1, top_n);
  const cd ie1 = conj(e1);
  EvalAcc B = {mk(0, 0), mk(0, 0), mk(0, 0), mk(0, 0), mk(0, 0), mk(0, 0)};
  double gp = 0.0, g = 1.0, H = 0.0;  // H = sum over even n >= 2 of rs^n g_n (Horner)
  for (int n = top - 1; n >= 1; --n) {
    if (n <= top_n) {
      T::add(E + T::NQ * n, g, pos.x, pos.y, B);
      pos = pos * ie1;
    }
    H = rs * ((n & 1) ? H : g + H);
    const double gm = zr * g * (2.0 * n) - gp * r2;
    gp = g;
    g = gm;
    if (g * g > 1e40) {
      g *= 1e-40; gp *= 1e-40; H *= 1e-40;
      B.P = B.P * 1e-40; B.Gx = B.Gx * 1e-40; B.Gy = B.Gy * 1e-40;
    }
  }
  T::add(E, g, 1.0, 0.0, B);
  const double sc = 1.0 / (g + 2.0 * H);
  A.P += B.P * sc; A.Gx += B.Gx * sc; A.Gy += B.Gy * sc;
}

// local at a point from the staged form, any k (generating-function
// normalisation, as eval_lo_at)
template <int PG>
__device__ void eval_lo_sd(const cd* E, int top_n, cd zk, double rs, double r, cd e1,
                           EvalAcc& A) {
  using T = SdTerm<PG>;
  const cd z = zk * r;
  if (cabs_(z) < 1e-200) {  // J_0 = 1, others 0
    T::add(E, 1.0, 1.0, 0.0, A);
    return;
  }
  const int top = miller_top(top_n, z);
  const cd zr = inv(z) * rs;
  const double r2 = rs * rs;
  const cd u = z.y >= 0 ? mk(0, -1) : mk(0, 1), ur = u * rs;
  cd pos = cpow_n(e1, top_n);
  const cd ie1 = conj(e1);
  EvalAcc B = {mk(0, 0), mk(0, 0), mk(0, 0), mk(0, 0), mk(0, 0), mk(0, 0)};
  cd gp = mk(0, 0), g = mk(1, 0), H = mk(0, 0);
  for (int n = top - 1; n >= 1; --n) {
    if (n <= top_n) {
      T::add(E + T::NQ * n, g, pos.x, pos.y, B);
      pos = pos * ie1;
    }
    H = ur * (g + H);
    const cd gm = zr * g * (2.0 * n) - gp * r2;
    gp = g;
    g = gm;
    if (norm2(g) > 1e40) {
      g = g * 1e-40; gp = gp * 1e-40; H = H * 1e-40;
      B.P = B.P * 1e-40; B.Gx = B.Gx * 1e-40; B.Gy = B.Gy * 1e-40;
    }
  }
  T::add(E, g, 1.0, 0.0, B);
  const cd sc = inv((g + H * 2.0) * cexp_(-(u * z)));
  A.P += B.P * sc; A.Gx += B.Gx * sc; A.Gy += B.Gy * sc;
}

// Four boxes per wave, 16 lanes per box, lane per point; for each source
// expansion the box's 16 lanes stage it in LDS in the (s, d) form of
// SdTerm (the derivative expansions computed once per box instead of once
// per point and order), then evaluate.  The work list is ordered by
// decreasing cost within each level, so the four boxes of a wave cost about
// the same.  PG <= 2.  Per group, W >= NQ (pmax + 2) entries of staged
// form and then, for list 3, HE_RAW (>= 2 pmax + 1) raw coefficients.
//
// Latency (list 3): the source boxes' ids, levels, orders and centres are loaded
// once per 16 boxes, one box per lane, and handed out by shuffles; the raw
// coefficients of box k + 1 are loaded into registers (HE_PRE per lane,
// orders up to 8 HE_PRE - 1; any beyond load at staging) while box k is
// staged and evaluated, so no dependent chain of loads stalls the group
// between boxes.
constexpr int HE_PRE = 4;

template <bool MPK, int PG, int NR>
__global__ void __launch_bounds__(WAVE)
k_heval_s(const int* work, int nwork, DTree D, const HLev* lev, cd zk, const int2* rng,
          const double2* pts, int N, int nd, size_t per, const double2* ex, int W, int HE_RAW,
          double2* pot, double2* grad) {
  using T = SdTerm<PG>;
  extern __shared__ double2 shx[];
  const int grp = (int)threadIdx.x >> 4, gl = (int)threadIdx.x & 15;
  const int w = blockIdx.x * 4 + grp;
  if (w >= nwork) return;  // the whole group leaves; no block barriers below
  cd* E = reinterpret_cast<cd*>(shx) + (size_t)grp * (W + HE_RAW);
  cd* Raw = E + W;
  const int b = work[w];
  const int2 rg = rng[b];
  const cd i4 = mk(0, 0.25);
  const int nsrc = MPK ? D.o3[b + 1] - D.o3[b] : 1;
  const int o3b = MPK ? D.o3[b] : 0;
  for (int i0 = rg.x; i0 < rg.y; i0 += 16) {
    const int i = i0 + gl;
    const bool act = i < rg.y;
    const double2 x = act ? pts[i] : make_double2(0.0, 0.0);
    for (int d = 0; d < nd; ++d) {
      EvalAcc A = {mk(0, 0), mk(0, 0), mk(0, 0), mk(0, 0), mk(0, 0), mk(0, 0)};
      // box k0 + gl of the current 16: id, order, scale, centre
      int ms = 0, mp = 0;
      double mrs = 1.0;
      double2 mc = make_double2(0.0, 0.0);
      auto load_meta = [&](int k0) {
        const int kk = k0 + gl;
        if (kk < nsrc) {
          ms = MPK ? D.l3[o3b + kk] : b;
          const HLev lv = lev[D.level[ms]];
          mp = lv.p;
          mrs = lv.rs;
          mc = D.ctr[ms];
        }
      };
      // raw coefficients -p..p of a box, lane gl holding gl + 16 j
      double2 pre[HE_PRE];
      auto load_raw = [&](int s, int p) {
        const double2* e = ex + (size_t)s * per + (size_t)d * (2 * p + 1);
#pragma unroll
        for (int j = 0; j < HE_PRE; ++j) {
          const int q = gl + 16 * j;
          pre[j] = q < 2 * p + 1 ? e[q] : make_double2(0.0, 0.0);
        }
      };
      if constexpr (MPK) {
        load_meta(0);
        load_raw(__shfl(ms, 0, 16), __shfl(mp, 0, 16));
      }
      for (int k = 0; k < nsrc; ++k) {
        // the box's id, order, scale and centre: list 3 from the lanes' loads,
        // the local expansion (one box, the group's own) directly
        int s, p;
        double rs;
        double2 cc;
        if constexpr (MPK) {
          const int kl = k & 15;
          s = __shfl(ms, kl, 16);
          p = __shfl(mp, kl, 16);
          rs = __shfl(mrs, kl, 16);
          cc = make_double2(__shfl(mc.x, kl, 16), __shfl(mc.y, kl, 16));
        } else {
          s = b;
          const HLev lv = lev[D.level[b]];
          p = lv.p;
          rs = lv.rs;
          cc = D.ctr[b];
        }
        const int top = p + (PG >= 2 ? 1 : 0);
        wave_sync();  // the group's previous readers of Raw and E are done
        const double2* e = ex + (size_t)s * per + (size_t)d * (2 * p + 1);
        if constexpr (MPK) {
#pragma unroll
          for (int j = 0; j < HE_PRE; ++j) {
            const int q = gl + 16 * j;
            if (q < 2 * p + 1) Raw[q] = ld(&pre[j]);
          }
          for (int q = 16 * HE_PRE + gl; q < 2 * p + 1; q += 16) Raw[q] = ld(e + q);
          // box k + 1: its loads are in flight while box k is staged and
          // evaluated
          if (k + 1 < nsrc) {
            if (((k + 1) & 15) == 0) load_meta(k + 1);
            const int k1 = (k + 1) & 15;
            load_raw(__shfl(ms, k1, 16), __shfl(mp, k1, 16));
          }
          wave_sync();
        }
        const double rsd = MPK ? rs : 1.0 / rs;
        const cd z1 = zk * (0.5 / rsd), z2 = zk * (0.5 * rsd);
        // coefficient of order q: list 3 from the raw copy in LDS, the local
        // expansion from memory
        auto co = [&](int q) {
          if (q < -p || q > p) return mk(0, 0);
          if constexpr (MPK) return Raw[p + q];
          else return ld(e + p + q);
        };
        // the derivative expansions at order q (from orders q - 1 and q + 1)
        auto grads = [&](int q, cd& gx, cd& gy) {
          const cd a = co(q + 1), bm = co(q - 1);
          if (q > 0) { gx = z2 * a - z1 * bm; gy = mul_i(z2 * a + z1 * bm); }
          else if (q < 0) { gx = z1 * a - z2 * bm; gy = mul_i(z1 * a + z2 * bm); }
          else { gx = z2 * (a - bm); gy = mul_i(z2 * (a + bm)); }
        };
        for (int n = gl; n <= top; n += 16) {
          cd* o = E + T::NQ * n;
          if (n == 0) {
            o[0] = co(0);
            o[1] = mk(0, 0);
            if (PG >= 2) {
              cd gx, gy;
              grads(0, gx, gy);
              o[2] = gx; o[3] = mk(0, 0);
              o[4] = gy; o[5] = mk(0, 0);
            }
          } else {
            const double sg = (n & 1) ? -1.0 : 1.0;
            const cd cp = co(n), cm = co(-n) * sg;
            o[0] = cp + cm;
            o[1] = mul_i(cp - cm);
            if (PG >= 2) {
              cd gxp, gyp, gxm, gym;
              grads(n, gxp, gyp);
              grads(-n, gxm, gym);
              gxm = gxm * sg;
              gym = gym * sg;
              o[2] = gxp + gxm; o[3] = mul_i(gxp - gxm);
              o[4] = gyp + gym; o[5] = mul_i(gyp - gym);
            }
          }
        }
        wave_sync();
        if (act) {
          const double ddx = x.x - cc.x, ddy = x.y - cc.y;
          const double r = sqrt(ddx * ddx + ddy * ddy);
          const double ri = r > 0.0 ? 1.0 / r : 0.0;
          const cd e1 = r > 0.0 ? mk(ddx * ri, ddy * ri) : mk(1, 0);  // e^{i theta}
          if (NR > 0) {
            if (MPK) eval_mp_sd_real<PG, (NR > 0 ? NR : 1)>(E, top, zk.x, rs, r, e1, A);
            else eval_lo_sd_real<PG>(E, top, zk.x, rs, r, e1, A);
          } else {
            if (MPK) eval_mp_sd<PG>(E, top, zk, rs, r, e1, A);
            else eval_lo_sd<PG>(E, top, zk, rs, r, e1, A);
          }
        }
      }
      if (act) {
        addc(pot + (size_t)d * N + i, A.P * i4);
        if (PG >= 2) {
          addc(grad + (size_t)d * N + i, A.Gx * i4);
          addc(grad + (size_t)(nd + d) * N + i, A.Gy * i4);
        }
      }
    }
  }
}

// Four boxes per wave, 16 lanes per box, lane per point (leaves hold about
// ndiv points); MPK: list-3 multipoles, else the box's own local expansion.
// NR > 0: real positive wavenumber (for MPK also the series length of
// hank_series_real, valid over the launch's k r); NR = 0: complex path.
template <bool MPK, int PG, int NR>
__global__ void __launch_bounds__(WAVE)
k_heval(const int* work, int nwork, DTree D, const HLev* lev, cd zk, const int2* rng,
        const double2* pts, int N, int nd, size_t per, const double2* ex, double2* pot,
        double2* grad, double2* hess) {
  const int w = blockIdx.x * 4 + ((int)threadIdx.x >> 4);
  if (w >= nwork) return;
  const int b = work[w];
  const int2 rg = rng[b];
  const cd i4 = mk(0, 0.25);
  for (int i = rg.x + ((int)threadIdx.x & 15); i < rg.y; i += 16) {
    double2 x = pts[i];
    const int ns = MPK ? D.o3[b + 1] - D.o3[b] : 1;
    for (int d = 0; d < nd; ++d) {
      EvalAcc A = {mk(0, 0), mk(0, 0), mk(0, 0), mk(0, 0), mk(0, 0), mk(0, 0)};
      for (int k = 0; k < ns; ++k) {
        int s = MPK ? D.l3[D.o3[b] + k] : b;
        int L = D.level[s];
        HLev lv = lev[L];
        double2 c = D.ctr[s];
        double ddx = x.x - c.x, ddy = x.y - c.y;
        double r = sqrt(ddx * ddx + ddy * ddy);
        const double ri = r > 0.0 ? 1.0 / r : 0.0;
        const cd e1 = r > 0.0 ? mk(ddx * ri, ddy * ri) : mk(1, 0);  // e^{i theta}
        double rsd = MPK ? lv.rs : 1.0 / lv.rs;
        Ex E{ex + (size_t)s * per + (size_t)d * (2 * lv.p + 1), lv.p, zk * (0.5 / rsd), zk * (0.5 * rsd)};
        if (NR > 0) {
          if (MPK) eval_mp_at_real<PG, (NR > 0 ? NR : 1)>(E, zk.x, lv.rs, r, e1, A);
          else eval_lo_at_real<PG>(E, zk.x, lv.rs, r, e1, A);
        } else {
          if (MPK) eval_mp_at<PG>(E, zk, lv.rs, r, e1, A);
          else eval_lo_at<PG>(E, zk, lv.rs, r, e1, A);
        }
      }
      addc(pot + (size_t)d * N + i, A.P * i4);
      if (PG >= 2) {
        addc(grad + (size_t)d * N + i, A.Gx * i4);
        addc(grad + (size_t)(nd + d) * N + i, A.Gy * i4);
      }
      if (PG >= 3) {
        addc(hess + (size_t)d * N + i, A.Hxx * i4);
        addc(hess + (size_t)(nd + d) * N + i, A.Hxy * i4);
        addc(hess + (size_t)(2 * nd + d) * N + i, A.Hyy * i4);
      }
    }
  }
}

// ---------------------------------------------------------------------
// list-1 near field (helmkernels2d.c:21-520); same wave layout as the
// Laplace kernel.  CP: charge-only potential, which skips on |z| <= thresh
// (all other routines on <).
// ---------------------------------------------------------------------
constexpr int HCAP = 128;  // sources staged per pass

// NR > 0: real positive wavenumber and every k r of the launch at most
// JY_XMAX for the series length NR, so H0/H1 come from hank_series_real<NR>;
// NR = 0: any wavenumber, complex evaluator
template <bool C, bool Q, int PG, int NR>
__global__ void __launch_bounds__(WAVE)
k_hp2p(const int* work, DTree D, cd zk, const double2* src, const double2* chg,
       const double2* dip, const double* dvv, int NS, const int2* trng, const double2* tpts,
       int NT, int nd, double thresh, double2* pot, double2* grad, double2* hess,
       const int* perm, double2* opot, double2* ograd, double2* ohess) {
  // staging (x, charge, dipole strength, dipole vector) and, after each
  // target group, the slice reduction share one LDS buffer
  __shared__ double2 lds[4 * HCAP];
  double2* sx = lds;
  double2* sc = lds + HCAP;
  double2* sq = lds + 2 * HCAP;
  double2* sv = lds + 3 * HCAP;
  double2* red = lds;
  static_assert(7 * WAVE <= 4 * HCAP, "reduction buffer");
  __shared__ int boff[14], bsrc[14];
  const int b = work[blockIdx.x];
  const int lane = threadIdx.x;
  const int n1 = D.o1[b + 1] - D.o1[b];
  if (lane == 0) {
    int acc = 0;
    for (int k = 0; k < n1; ++k) {
      int2 r = D.srng[D.l1[D.o1[b] + k]];
      boff[k] = acc;
      bsrc[k] = r.x;
      acc += r.y - r.x;
    }
    boff[n1] = acc;
  }
  __syncthreads();
  const int S = boff[n1];
  const int2 tr = trng[b];
  const cd ik4 = zk * mk(0, 0.25);
  const bool cp_mode = C && !Q && PG == 1;
  const double izk = NR > 0 ? 1.0 / zk.x : 0.0;
  const double kk4 = zk.x * zk.x * 0.25;  // w = (k r)^2 / 4 = kk4 r^2
  // real k: k/2, the threshold on k r / 2 (a pair skips on k r < thresh,
  // i.e. k r / 2 < thresh / 2, both sides scaled exactly), and the factors
  // of H1 / r below
  const double hk = 0.5 * zk.x, hth = 0.5 * thresh;
  const double gc1 = -(2.0 / kPi) * izk, gc2 = -zk.x / kPi;
  for (int t0 = tr.x; t0 < tr.y; t0 += WAVE) {
    const int Tc = min(WAVE, tr.y - t0);
    const int K = WAVE / Tc;
    const int t = lane / K, k = lane % K;
    const bool act = t < Tc;
    const double2 x = act ? tpts[t0 + t] : make_double2(0, 0);
    for (int d = 0; d < nd; ++d) {
      NearAcc A = {mk(0, 0), mk(0, 0), mk(0, 0), mk(0, 0), mk(0, 0), mk(0, 0), mk(0, 0)};
      for (int g0 = 0; g0 < S; g0 += HCAP) {
        const int cnt = min(HCAP, S - g0);
        for (int m = lane; m < cnt; m += WAVE) {
          int g = g0 + m, kb = 0;
          while (boff[kb + 1] <= g) ++kb;
          int si = bsrc[kb] + (g - boff[kb]);
          sx[m] = src[si];
          if (C) sc[m] = chg[(size_t)d * NS + si];
          if (Q) {
            sq[m] = dip[(size_t)d * NS + si];
            sv[m] = make_double2(dvv[(size_t)d * NS + si], dvv[(size_t)(nd + d) * NS + si]);
          }
        }
        __syncthreads();
        if (act && NR > 0 && PG <= 2) {
          // real k, branch-free: a pair under the threshold (a coincident one
          // included, whose r^2 is raised to the smallest normal double) runs
          // the series at its own tiny but finite arguments with H0 and
          // g = H1 / r set to 0, which zeroes every term it adds.
          // With k H2 = 2 g - k H0, the charge and dipole gradients are
          // -a (dx, dy) + b v, a = g c + q k H2 (d.v) / r^2, b = q g, and the
          // dipole potential is b (d.v).
          for (int m = k; m < cnt; m += K) {
            const double dx = x.x - sx[m].x, dy = x.y - sx[m].y;
            const double rr = fmax(dx * dx + dy * dy, 0x1p-1022);
            const double ri = rsqrt_pos(rr);
            const double hx = hk * (rr * ri);  // k r / 2
            const bool skip = cp_mode ? hx <= hth : hx < hth;
            // H0 = P + i (2/pi) (lg P + Q) and, from J1 = -P' k r / 2 and
            // Y1 = -(2/pi) [P / (k r) + (k r / 2)(lg P' + Q')] (A&S 9.1.11,
            // 9.1.13), g = -(k/2) P' - i (2/pi) [P / (k r^2) + (k/2)(lg P' + Q')],
            // lg = ln(k r / 2) + gamma, P and Q in w = (k r / 2)^2
            double p, dp, q, dq;
            jy_pq<(NR > 0 ? NR : 1)>(kk4 * rr, p, dp, q, dq);
            const double lg = log_tab(hx) + kEuler;
            const double ri2 = ri * ri;
            cd h0 = cd{p, (lg * p + q) * (2.0 / kPi)};
            cd g = cd{dp * -hk, fma(p * ri2, gc1, fma(lg, dp, dq) * gc2)};
            if (skip) {
              h0 = cd{0.0, 0.0};
              g = cd{0.0, 0.0};
            }
            cd a = mk(0, 0);
            if (C) {
              const double2 cv = sc[m];
              const cd c = cd{cv.x, cv.y};
              A.Pc += h0 * c;
              if (PG >= 2) a = g * c;
            }
            if (Q) {
              const double2 qv = sq[m];
              const cd q = cd{qv.x, qv.y};
              const double v1 = sv[m].x, v2 = sv[m].y;
              const double dv = dx * v1 + dy * v2;
              const cd bq = q * g;
              A.Pd += bq * dv;
              if (PG >= 2) {
                const cd kh2 = g * 2.0 - h0 * zk.x;
                a += (q * kh2) * (dv * ri2);
                A.Gx += bq * v1;
                A.Gy += bq * v2;
              }
            }
            if (PG >= 2) {
              A.Gx -= a * dx;
              A.Gy -= a * dy;
            }
          }
        } else if (act) {
          for (int m = k; m < cnt; m += K) {
            double dx = x.x - sx[m].x, dy = x.y - sx[m].y;
            double rr = dx * dx + dy * dy, r, ri;
            cd z, h0, h1;
            double az;
            if (NR > 0) {
              // r and 1/r from one reciprocal square root (r = 0 for a
              // coincident pair, which the threshold then skips)
              ri = rr > 0.0 ? rsqrt(rr) : 0.0;
              r = rr * ri;
              az = zk.x * r;
              z = cd{az, 0.0};
            } else {
              r = sqrt(rr);
              ri = 1.0 / r;
              z = zk * r;
              az = cabs_(z);
            }
            if (cp_mode ? az <= thresh : az < thresh) continue;
            const double azi = ri * izk;  // 1/(k r) for a real wavenumber
            if (NR > 0) hank_series_real<(NR > 0 ? NR : 1)>(az, azi, h0, h1);
            else hank01(z, h0, h1);
            if (C) {
              cd c = ld(&sc[m]);
              A.Pc += h0 * c;
              if (PG >= 2) {
                cd cd_ = -(h1 * c * ri);
                A.Gx += cd_ * dx;
                A.Gy += cd_ * dy;
              }
              if (PG >= 3) {
                cd f = c * (ri / rr), h2z = -(z * h0) + h1 * 2.0;
                A.Hxx += f * (h2z * (dx * dx) - h1 * rr);
                A.Hxy += f * (h2z * (dx * dy));
                A.Hyy += f * (h2z * (dy * dy) - h1 * rr);
              }
            }
            if (Q) {
              cd q = ld(&sq[m]);
              double v1 = sv[m].x, v2 = sv[m].y;
              const double dv = dx * v1 + dy * v2;
              A.Pd += (h1 * (ri * dv)) * q;
              if (PG >= 2) {
                // grad of H1(kr) (x-y).v / r:  -k [H2 (dv / r^2) (dx, dy)
                //   - (H1 / (k r)) v],  H2 = 2 H1 / (k r) - H0
                const cd zi = NR > 0 ? cd{azi, 0.0} : inv(z);
                const cd w = h1 * zi;
                const cd h2 = w * 2.0 - h0;
                const cd u = h2 * (dv * ri * ri);
                const cd qk = NR > 0 ? cd{-q.x * zk.x, -q.y * zk.x} : -(q * zk);
                A.Gx += qk * (u * dx - w * v1);
                A.Gy += qk * (u * dy - w * v2);
                if (PG >= 3) {
                  double ct = dx * ri, st = dy * ri;
                  cd h3 = h2 * zi * 4.0 - h1;
                  cd a = (h1 * 0.75 - h3 * (0.5 * (ct * ct - 0.5 - st * st))) * ct;
                  cd bb = (h1 * 0.25 - h3 * (0.5 * (1.5 * ct * ct - 0.5 * st * st))) * st;
                  cd c2 = (h1 * 0.25 - h3 * (0.5 * (1.5 * st * st - 0.5 * ct * ct))) * ct;
                  cd e = (h1 * 0.75 - h3 * (0.5 * (st * st - 0.5 - ct * ct))) * st;
                  cd qk2 = -(q * zk * zk);
                  A.Hxx += qk2 * (a * v1 + bb * v2);
                  A.Hxy += qk2 * (bb * v1 + c2 * v2);
                  A.Hyy += qk2 * (c2 * v1 + e * v2);
                }
              }
            }
          }
        }
        __syncthreads();
      }
      // slice reduction; charge potential carries i/4, every other sum i k/4
      cd vals[7] = {A.Pc, A.Pd, A.Gx, A.Gy, A.Hxx, A.Hxy, A.Hyy};
      const int nv = PG >= 3 ? 7 : PG >= 2 ? 4 : 2;
      for (int v = 0; v < nv; ++v) red[v * WAVE + lane] = make_double2(vals[v].x, vals[v].y);
      __syncthreads();
      if (act && k == 0) {
        cd s[7];
        for (int v = 0; v < nv; ++v) {
          s[v] = mk(0, 0);
          for (int u = 0; u < K; ++u) s[v] += ld(&red[v * WAVE + t * K + u]);
        }
        const size_t o = t0 + t;
        if (opot) {
          // final value straight to the input-order output (point perm[o])
          const size_t u = (size_t)perm[o];
          auto put = [](double2* y, const double2* x, cd v) {
            *y = make_double2(x->x + v.x, x->y + v.y);
          };
          put(opot + u * nd + d, pot + (size_t)d * NT + o, s[0] * mk(0, 0.25) + s[1] * ik4);
          if (PG >= 2) {
            put(ograd + (u * 2 + 0) * nd + d, grad + (size_t)d * NT + o, s[2] * ik4);
            put(ograd + (u * 2 + 1) * nd + d, grad + (size_t)(nd + d) * NT + o, s[3] * ik4);
          }
          if (PG >= 3) {
            put(ohess + (u * 3 + 0) * nd + d, hess + (size_t)d * NT + o, s[4] * ik4);
            put(ohess + (u * 3 + 1) * nd + d, hess + (size_t)(nd + d) * NT + o, s[5] * ik4);
            put(ohess + (u * 3 + 2) * nd + d, hess + (size_t)(2 * nd + d) * NT + o, s[6] * ik4);
          }
        } else {
          addc(pot + (size_t)d * NT + o, s[0] * mk(0, 0.25) + s[1] * ik4);
          if (PG >= 2) {
            addc(grad + (size_t)d * NT + o, s[2] * ik4);
            addc(grad + (size_t)(nd + d) * NT + o, s[3] * ik4);
          }
          if (PG >= 3) {
            addc(hess + (size_t)d * NT + o, s[4] * ik4);
            addc(hess + (size_t)(nd + d) * NT + o, s[5] * ik4);
            addc(hess + (size_t)(2 * nd + d) * NT + o, s[6] * ik4);
          }
        }
      }
      __syncthreads();
    }
  }
}

// ---------------------------------------------------------------------
// translation operators per level (helmrouts2d.c:736-1042):
//   out_i = sum_j in_j B(i-j) F(i,j),  B(m) = v_|m| e^{-i m theta} (-1)^{m<0 ? m : 0}
// with theta = arg(c2 - c1) - pi, v = scaled J (shifts) or H (M2L) at
// zk |c2 - c1|, and F the scale factors of the scaled conventions:
//   M2M  rs1^(|j|+|i-j|-|i|) (rs1/rs2)^|i|
//   L2L  rs2^|i| rs1^(|i-j|-|j|)
//   M2L  rs1^(|j|-|i-j|) rs2^|i|
// One block per (level, slot); T stored [j][i].
// ---------------------------------------------------------------------
struct OpJob {
  int kind;          // 0 M2M, 1 L2L, 2 M2L
  int p1, p2;
  double rs1, rs2;
  double dx, dy;     // c2 - c1
  size_t off;        // into the table buffer (double2 units)
};

__global__ void __launch_bounds__(256) k_hops(const OpJob* jobs, cd zk, double2* T) {
  extern __shared__ double2 sv[];
  const OpJob J = jobs[blockIdx.x];
  const int P = J.p1 + J.p2;
  cd* v = reinterpret_cast<cd*>(sv);  // P + 4 entries
  double r = sqrt(J.dx * J.dx + J.dy * J.dy);
  double th = atan2(J.dy, J.dx) - kPi;
  if (threadIdx.x == 0) {
    cd z = zk * r;
    if (J.kind == 2) hvals(P + 1, z, J.rs1, v, 1);
    else jvals(P + 3, z, J.rs1, v, 1);
  }
  __syncthreads();
  const int nin = 2 * J.p1 + 1, nout = 2 * J.p2 + 1;
  const double l1 = log(J.rs1), l2 = log(J.rs2);
  for (int e = threadIdx.x; e < nin * nout; e += blockDim.x) {
    int jj = e / nout, ii = e - jj * nout;
    int j = jj - J.p1, i = ii - J.p2;
    int m = i - j, am = abs(m), ai = abs(i), aj = abs(j);
    double lf;
    if (J.kind == 0) lf = (aj + am - ai) * l1 + ai * (l1 - l2);
    else if (J.kind == 1) lf = ai * l2 + (am - aj) * l1;
    else lf = (aj - am) * l1 + ai * l2;
    double ang = -m * th;
    cd ph = mk(cos(ang), sin(ang));
    if (m < 0 && (am & 1)) ph = -ph;
    cd va = v[am];
    double mag = cabs_(va);
    cd t = mag > 0 ? (va * ph) * (exp(lf + log(mag)) / mag) : mk(0, 0);
    T[J.off + e] = make_double2(t.x, t.y);
  }
}

// xmax: bound on k r over the launch's pairs (0 when zk is not real and
// positive)
template <bool C, bool Q>
void launch_hp2p(int pg, int nbx, hipStream_t st, const int* work, const DTree& D, cd zk,
                 const double2* src, const double2* chg, const double2* dip, const double* dv,
                 int NS, const int2* trng, const double2* tpts, int NT, int nd, double th,
                 double xmax, double2* pot, double2* grad, double2* hess, const int* perm,
                 double2* opot, double2* ograd, double2* ohess) {
  const int nr = jy_nterms(xmax);
#define FMM_HP2P(PGV, NRV)                                                                  \
  k_hp2p<C, Q, PGV, NRV><<<nbx, WAVE, 0, st>>>(work, D, zk, src, chg, dip, dv, NS, trng, tpts, \
                                               NT, nd, th, pot, grad, hess, perm, opot, ograd, \
                                               ohess)
#define FMM_HP2P_NR(PGV)                                 \
  do {                                                   \
    if (nr == JY_NTERMS0) FMM_HP2P(PGV, JY_NTERMS0);     \
    else if (nr == JY_NTERMS1) FMM_HP2P(PGV, JY_NTERMS1); \
    else if (nr == JY_NTERMS2) FMM_HP2P(PGV, JY_NTERMS2); \
    else if (nr == JY_NTERMS3) FMM_HP2P(PGV, JY_NTERMS3); \
    else FMM_HP2P(PGV, 0);                               \
  } while (0)
  if (pg == 1) FMM_HP2P_NR(1);
  if (pg == 2) FMM_HP2P_NR(2);
  if (pg == 3) FMM_HP2P_NR(3);
#undef FMM_HP2P_NR
#undef FMM_HP2P
}

// xmax: bound on k r over the launch (multipoles), 0 when zk is not real
// and positive
template <bool MPK>
void launch_heval(int pg, int nw, hipStream_t st, const int* work, const DTree& D,
                  const HLev* lev, cd zk, const int2* rng, const double2* pts, int N, int nd,
                  size_t per, const double2* ex, double xmax, int pmax, double2* pot,
                  double2* grad, double2* hess) {
  const bool rz = zk.y == 0.0 && zk.x > 0.0;
  int nr = 0;
  if (rz) {
    if (!MPK) nr = 1;
    else nr = jy_nterms(xmax);
  }
  const int nbx = (nw + 3) / 4;
  // staged expansions (value and gradient, NQ entries per order) while
  // they fit comfortably in LDS
  const int W = (pg >= 2 ? 6 : 2) * (pmax + 2);
  // raw coefficients of a list-3 box, 2 pmax + 1 rounded up (the local
  // expansion stages from memory)
  const int RAW = MPK ? 2 * pmax + 2 : 0;
  const bool staged = pg <= 2 && pmax <= 77;
  const size_t shs = (size_t)4 * (W + RAW) * sizeof(double2);
#define FMM_HEVAL(PGV, NRV)                                                                      \
  do {                                                                                           \
    if (staged && PGV <= 2)                                                                      \
      k_heval_s<MPK, (PGV <= 2 ? PGV : 2), NRV><<<nbx, WAVE, shs, st>>>(                         \
          work, nw, D, lev, zk, rng, pts, N, nd, per, ex, W, RAW, pot, grad);                    \
    else                                                                                         \
      k_heval<MPK, PGV, NRV><<<nbx, WAVE, 0, st>>>(work, nw, D, lev, zk, rng, pts, N, nd, per,   \
                                                   ex, pot, grad, hess);                         \
  } while (0)
#define FMM_HEVAL_NR(PGV)                                                \
  do {                                                                   \
    if constexpr (!MPK) {                                                \
      if (nr) FMM_HEVAL(PGV, 1);                                         \
      else FMM_HEVAL(PGV, 0);                                            \
    } else {                                                             \
      if (nr == JY_NTERMS0) FMM_HEVAL(PGV, JY_NTERMS0);                  \
      else if (nr == JY_NTERMS1) FMM_HEVAL(PGV, JY_NTERMS1);             \
      else if (nr == JY_NTERMS2) FMM_HEVAL(PGV, JY_NTERMS2);             \
      else if (nr == JY_NTERMS3) FMM_HEVAL(PGV, JY_NTERMS3);             \
      else FMM_HEVAL(PGV, 0);                                            \
    }                                                                    \
  } while (0)
  if (pg == 1) FMM_HEVAL_NR(1);
  if (pg == 2) FMM_HEVAL_NR(2);
  if (pg == 3) FMM_HEVAL_NR(3);
#undef FMM_HEVAL_NR
#undef FMM_HEVAL
}

}  // namespace

// ---------------------------------------------------------------------
// host side
// ---------------------------------------------------------------------
namespace {

// expansion order per box size (h2dterms.c:22-78), memoised: the rule needs
// 50000-term recurrences and depends only on (zk, box size, eps)
int h2dterms_cached(double bs, std::complex<double> zk, double eps, int* ier) {
  static std::mutex mu;
  static std::map<std::tuple<double, double, double, double>, std::pair<int, int>> memo;
  auto key = std::make_tuple(bs, zk.real(), zk.imag(), eps);
  {
    std::lock_guard<std::mutex> g(mu);
    auto it = memo.find(key);
    if (it != memo.end()) {
      *ier = it->second.second;
      return it->second.first;
    }
  }
  const int ntmax = 50000;
  double rs = 1.0;
  cd zb = cd{zk.real() * bs, zk.imag() * bs};
  if (cabs_(zb) < 2 * kPi) rs = cabs_(zb);
  auto ab = [](cd a) { return cabs_(a); };
  // The rule scans j = 2..ntmax and stops at the first small term, almost
  // always within a few hundred orders.  Orders 0..nlim are evaluated with
  // the ntmax-term recurrences' own starting points (bit for bit the values
  // of the full evaluation), widening nlim only when the scan gets past it.
  int nt = 10001, e = 13;
  for (int nlim : {256, 4096, ntmax}) {
    std::vector<cd> hf(nlim + 2), jf(nlim + 2);
    hvals(nlim, zb * 1.5, rs, hf.data(), 1);
    jvals_part(ntmax, nlim, zb * (std::sqrt(2.0) / 2.0), rs, jf.data(), 1);
    double x0 = ab(jf[0] * hf[0]) + ab(jf[1] * hf[1]);
    bool found = false;
    for (int j = 2; j <= nlim; ++j) {
      double x = (ab(jf[j] * hf[j]) + ab(jf[j - 1] * hf[j - 1])) * ab(hf[0]);
      if (x < eps * x0) {
        nt = j + 1;
        e = 0;
        found = true;
        break;
      }
    }
    if (found) break;
  }
  std::lock_guard<std::mutex> g(mu);
  memo[key] = {nt, e};
  *ier = e;
  return nt;
}

}  // namespace

int h2dterms_host(double bs, std::complex<double> zk, double eps, int* ier) {
  return h2dterms_cached(bs, zk, eps, ier);
}

// Consecutive levels (from lfrom) of a level-sorted work list, grouped by
// the kind of H0/H1 evaluation each level's bound on k r allows: a short
// real series (up to 11 terms), the 18-term series, or the general
// evaluator (all one group when rz is false): f(span, xmax) per group, xmax
// the group's largest bound, which sets its series length.  A tree whose
// leaves span many sizes -- adaptive trees, forests of problems of many
// sizes -- keeps short series on its small leaves and the general
// evaluator on the large ones only.  (Splitting by every series length cost
// more in launch tails than 8 instead of 9 terms saved: 1.19 -> 1.26 ms
// for the near field of 1M + 1M uniform points.)
template <class XF, class F>
void by_tier(const WorkList& wl, int lfrom, int L, bool rz, XF xmax, F f) {
  int l = std::max(lfrom, 0);
  auto tier = [&](int q) {
    if (!rz) return 0;
    const int nr = jy_nterms(xmax(q));
    return nr == 0 ? 0 : nr <= JY_NTERMS2 ? 1 : 2;
  };
  while (l <= L) {
    if (!wl.at(l).n) {
      ++l;
      continue;
    }
    const int nr = tier(l);
    double x = rz ? xmax(l) : 0.0;
    int l2 = l + 1;
    while (l2 <= L && (!wl.at(l2).n || tier(l2) == nr)) {
      if (wl.at(l2).n && rz) x = std::max(x, xmax(l2));
      ++l2;
    }
    f(Span{wl.off + wl.lev[l], wl.lev[l2] - wl.lev[l]}, x);
    l = l2;
  }
}

int helm_run(Ctx& c, Scratch& S, const SortedPoints& Pt, const Plan& plan, const HelmArgs& a) {
  const DTree& D = plan.D;
  const int nd = a.nd, L = plan.L, nb = plan.nb;
  const std::vector<double>& bs = plan.bs;
  const cd zk = cd{a.zk.real(), a.zk.imag()};
  const double zi = a.zk.imag();
  const double t_setup = wall();
  // per-level order and scale (hfmm2d.c:1768-1779)
  std::vector<HLev> lev(L + 1);
  std::vector<char> gate(L + 2, 0);
  // the first level with expansions: 2, or in a forest 2 below the highest
  // problem root (coarser levels are shared cells above the roots)
  const int lf = plan.lfirst;
  int pmax = 0;
  for (int l = 0; l <= L; ++l) {
    std::complex<double> zt = a.zk * bs[l] / (2.0 * kPi);
    lev[l].rs = std::min(std::abs(zt), 1.0);
    int e = 0;
    // (in a forest, the shared cells above the roots need no order)
    lev[l].p = l >= lf || plan.forest.M < 0 ? h2dterms_cached(bs[l], a.zk, a.eps, &e) : 0;
    gate[l] = zi * bs[l] < 40.0;
    // levels 0 and 1 take part in no M2L, list 3 or list 4 interaction, so
    // their expansions stay zero and are neither formed nor stored
    if (gate[l] && l >= lf) pmax = std::max(pmax, lev[l].p);
  }
  if (pmax > 2000) return IER_P_TOO_LARGE;
  // Levels whose boxes exceed 16 wavelengths translate in the diagonal form
  // (hfmm2d.c:946-960: M2M into them, M2L at them and L2L out of them;
  // levels 0 and 1 form no expansion that any step reads).
  std::vector<char> hf(L + 2, 0);
  for (int l = lf; l <= L; ++l)
    hf[l] = gate[l] && bs[l] * a.zk.real() / (2 * kPi) > 16.0;
  // the largest order the per-point steps meet: formation, list 4 and local
  // evaluation at their own level, list 3 at the level below (the orders of
  // the diagonal-form levels are far above those of the leaves)
  int pmax_pt = 0;
  for (int l = 0; l <= L; ++l) {
    if (!gate[l] || l < lf) continue;
    if (plan.formmp.at(l).n || plan.list4.at(l).n || plan.leaves_s.at(l).n ||
        plan.leaves_t.at(l).n)
      pmax_pt = std::max(pmax_pt, lev[l].p);
  }
  for (int l = 0; l < L; ++l)
    if (gate[l + 1] && l + 1 >= lf && (plan.list3_s.at(l).n || plan.list3_t.at(l).n))
      pmax_pt = std::max(pmax_pt, lev[l + 1].p);
  const double2* chg = a.ifcharge ? a.chg : nullptr;
  const double2* dip = a.ifdipole ? a.dip : nullptr;
  if (!chg && !dip) return IER_OK;
  const size_t per = (size_t)nd * (2 * pmax + 1);
  HLev* dlev = S.get<HLev>(L + 1);
  double2* mp = S.get<double2>(per * nb);
  double2* lo = S.get<double2>(per * nb);
  if (S.err) return S.err;
  if (int e = upload_small(dlev, lev.data(), sizeof(HLev) * (L + 1), c.st)) return e;
  fill_async(mp, 0u, 4 * per * nb, c.st);
  fill_async(lo, 0u, 4 * per * nb, c.st);

  // ---- operator tables
  std::vector<OpJob> jobs;
  size_t toff = 0;
  struct Tab { size_t m2m[4], l2l[4], m2l[49]; };
  std::vector<Tab> tab(L + 1);
  int vmax = 0;
  for (int l = 0; l <= L; ++l) {
    if (!gate[l]) continue;
    if (l >= lf && l <= L - 1 && !hf[l]) {
      double h = bs[l + 1] / 2;
      size_t sz1 = (size_t)(2 * lev[l + 1].p + 1) * (2 * lev[l].p + 1);
      size_t base_m = toff, base_l = toff + 4 * sz1;
      toff += 8 * sz1;
      for (int q = 0; q < 4; ++q) {
        double sx = (q & 1) ? h : -h, sy = (q & 2) ? h : -h;
        // M2M child (l+1) -> parent (l): c2 - c1 = parent - child
        jobs.push_back({0, lev[l + 1].p, lev[l].p, lev[l + 1].rs, lev[l].rs, -sx, -sy, base_m + q * sz1});
        // L2L parent (l) -> child (l+1): same matrix size, transposed roles
        jobs.push_back({1, lev[l].p, lev[l + 1].p, lev[l].rs, lev[l + 1].rs, sx, sy, base_l + q * sz1});
      }
      tab[l].m2m[0] = base_m;
      tab[l].l2l[0] = base_l;
      vmax = std::max(vmax, lev[l].p + lev[l + 1].p + 4);
    }
    if (l >= lf && !hf[l]) {
      size_t sz = (size_t)(2 * lev[l].p + 1) * (2 * lev[l].p + 1);
      size_t base = toff;
      toff += 49 * sz;  // near slots stay unused (no list-2 pair maps to them)
      tab[l].m2l[0] = base;
      for (int s = 0; s < 49; ++s) {
        int dx = s % 7 - 3, dy = s / 7 - 3;
        if (std::abs(dx) <= 1 && std::abs(dy) <= 1) continue;
        // source at (dx, dy) bs from the target: c2 - c1 = target - source
        jobs.push_back({2, lev[l].p, lev[l].p, lev[l].rs, lev[l].rs, -dx * bs[l], -dy * bs[l], base + s * sz});
      }
      vmax = std::max(vmax, 2 * lev[l].p + 4);
    }
  }
  double2* dT = S.get<double2>(std::max<size_t>(toff, 1));
  OpJob* djobs = S.get<OpJob>(std::max<size_t>(jobs.size(), 1));
  if (S.err) return S.err;
  if (!jobs.empty()) {
    if (int e = upload_small(djobs, jobs.data(), sizeof(OpJob) * jobs.size(), c.st)) return e;
    k_hops<<<(int)jobs.size(), 256, sizeof(double2) * (vmax + 8), c.st>>>(djobs, zk, dT);
  }
  // FP32 copy of the operators for the mixed-precision M2L
  float2* dT32 = nullptr;
  if (a.mixed && toff) {
    dT32 = S.get<float2>(toff);
    if (S.err) return S.err;
    k_to_f32<<<(unsigned)((toff + 255) / 256), 256, 0, c.st>>>(dT, dT32, toff);
  }
  // M2M/L2L slot tables are the 4 quadrant matrices in a row; M2L slots are
  // addressed through a per-level array of 49 offsets
  const int* W = plan.dwork;
  const double* dv = a.dipvec;
  // formation: points staged CH at a time, Z (complex) and V (J real for a
  // real wavenumber, else complex) for orders 0..p+2; 16 points keep the
  // footprint near 12-16 KB so several waves share a CU
  // (fewer points per stage for the high orders of leaves above a few
  // wavelengths, so that a block stays within 64 KB of LDS)
  const bool rz = zk.y == 0.0 && zk.x > 0.0;
  auto shm_for = [&](int ch, size_t vsz) {
    return (size_t)(pmax_pt + 3) * ch * (sizeof(double2) + vsz) + (size_t)6 * ch * sizeof(double2);
  };
  int CH = 16;
  while (CH > 1 && shm_for(CH, sizeof(double2)) > 65536) CH /= 2;
  if (shm_for(CH, sizeof(double2)) > 65536) return IER_P_TOO_LARGE;
  const size_t shm = shm_for(CH, sizeof(double2));
  const size_t shm_mp = shm_for(CH, rz ? sizeof(double) : sizeof(double2));

  // levels with gate[l] false (complex zk, boxes far above the decay length)
  // form no expansions; gate[l] only switches on as boxes shrink, so the
  // gated work is a level suffix of every (level-sorted) work list
  int lg = 0;
  while (lg <= L && !gate[lg]) ++lg;
  for (int l = lg; l <= L; ++l)
    if (!gate[l]) return IER_ARGS;
  const Span s_form = plan.formmp.from(lg), s_l4 = plan.list4.from(lg);
  const Span s_l3s = plan.list3_s.from(lg - 1), s_l3t = plan.list3_t.from(lg - 1);
  const Span s_lvs = plan.leaves_s.from(lg), s_lvt = plan.leaves_t.from(lg);
  auto wptr = [&](const Span& sp) { return W + sp.off; };

  // k r over list 4: a source of a coarse leaf to the centre of a target of
  // size s is at most 3.5 s away per coordinate, r <= 4.95 s
  auto xmax4 = [&](int l) { return zk.x * 4.95 * bs[l]; };
  StepTimer* tm = step_timer();
  {
    SolveStats& st = last_stats();
    st.host_s[6] = wall() - t_setup;
    for (int l = lf; l <= L; ++l) {
      if (!gate[l]) continue;
      double np = plan.m2l_pairs[l];
      double nc = 2 * lev[l].p + 1;
      st.m2l_pairs += np;
      // a diagonal-form pair is one complex multiply-add per signature point
      st.m2l_flops += hf[l] ? np * 8.0 * next235(4.0 * lev[l].p + 1) : np * 8.0 * nc * nc;
    }
    st.pmax = pmax;
  }
  tm->mark(c.st);
  // step 1: leaf multipoles
  // all-lanes formation for a real wavenumber up to order 63
  const bool fg = rz && pmax_pt <= 63;
  if (s_form.n) {
    if (fg)
      k_hform_g<true, 1><<<nblk(s_form.n, 4), WAVE, 0, c.st>>>(wptr(s_form), s_form.n, D, dlev,
                                                              zk.x, Pt.src, chg, dip, dv, Pt.ns,
                                                              nd, per, mp);
    else
      (rz ? k_hform<true, 1> : k_hform<true, 0>)<<<s_form.n, WAVE, shm_mp, c.st>>>(
          wptr(s_form), D, dlev, zk, Pt.src, chg, dip, dv, Pt.ns, nd, CH, per, mp);
  }
  tm->mark(c.st);
  // step 3: M2M
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
  for (int l = L - 1; l >= lf; --l) {
    const Span sp = plan.m2m.at(l);
    if (!sp.n || !gate[l]) continue;
    const int* map = map_m2m + (size_t)(sp.off - plan.m2m.off) * 4;
    if (hf[l]) {  // hfmm2d.c:993-1051
      HfJob j;
      j.kind = 0;
      j.pin = lev[l + 1].p;
      j.pout = lev[l].p;
      j.nsig = next235(2.0 * (lev[l].p + lev[l + 1].p) + 1);
      j.work = W + sp.off;
      j.nwork = sp.n;
      j.map = map;
      j.nslot = 4;
      j.in_first = plan.loff[l + 1];
      j.in_count = plan.loff[l + 2] - plan.loff[l + 1];
      j.bs = bs[l];
      const int e = hf_translate(c.st, S, a.zk, j, mp, per, mp, per, nd);
      if (e) return e;
      continue;
    }
    HIP_TRY(launch_xlate(c.st, S, W + sp.off, sp.n, map, 4, 4, dT + tab[l].m2m[0], 2 * lev[l + 1].p + 1,
                             2 * lev[l].p + 1, mp, per, mp, per, nd, 0.0));
  }
  tm->mark(c.st);
  // step 2: list 4, by groups of levels of one series length
  if (s_l4.n) by_tier(plan.list4, lg, L, rz, xmax4, [&](const Span& sp, double xm) {
    const int nr4 = rz ? jy_nterms(xm) : 0;
    if (fg && nr4 == JY_NTERMS0)
      k_hform_g<false, JY_NTERMS0><<<nblk(sp.n, 4), WAVE, 0, c.st>>>(
          wptr(sp), sp.n, D, dlev, zk.x, Pt.src, chg, dip, dv, Pt.ns, nd, per, lo);
    else if (fg && nr4 == JY_NTERMS1)
      k_hform_g<false, JY_NTERMS1><<<nblk(sp.n, 4), WAVE, 0, c.st>>>(
          wptr(sp), sp.n, D, dlev, zk.x, Pt.src, chg, dip, dv, Pt.ns, nd, per, lo);
    else if (fg && nr4 == JY_NTERMS2)
      k_hform_g<false, JY_NTERMS2><<<nblk(sp.n, 4), WAVE, 0, c.st>>>(
          wptr(sp), sp.n, D, dlev, zk.x, Pt.src, chg, dip, dv, Pt.ns, nd, per, lo);
    else if (fg && nr4 == JY_NTERMS3)
      k_hform_g<false, JY_NTERMS3><<<nblk(sp.n, 4), WAVE, 0, c.st>>>(
          wptr(sp), sp.n, D, dlev, zk.x, Pt.src, chg, dip, dv, Pt.ns, nd, per, lo);
    else
      (nr4 == JY_NTERMS0   ? k_hform<false, JY_NTERMS0>
       : nr4 == JY_NTERMS1 ? k_hform<false, JY_NTERMS1>
       : nr4 == JY_NTERMS2 ? k_hform<false, JY_NTERMS2>
       : nr4 == JY_NTERMS3 ? k_hform<false, JY_NTERMS3>
                           : k_hform<false, 0>)<<<sp.n, WAVE, shm, c.st>>>(
          wptr(sp), D, dlev, zk, Pt.src, chg, dip, dv, Pt.ns, nd, CH, per, lo);
  });
  tm->mark(c.st);
  // step 4: M2L (49 slot matrices per level, contiguous); the slot maps of
  // every level in one launch
  int* map2 = S.get<int>((size_t)std::max(plan.m2l.n, 1) * 49);
  if (S.err) return S.err;
  if (plan.m2l.n) {
    fill_async(map2, 0xffffffffu, 49 * (size_t)plan.m2l.n, c.st);
    k_map_m2l_all<<<nblk(32L * plan.m2l.n, 256), 256, 0, c.st>>>(W + plan.m2l.off, plan.m2l.n, D,
                                                                map2);
  }
  std::vector<M2LLevel> m2l_levels;
  for (int l = lf; l <= L; ++l) {
    const Span sp = plan.m2l.at(l);
    if (!sp.n || !gate[l]) continue;
    const int* ml = map2 + (size_t)(sp.off - plan.m2l.off) * 49;
    if (hf[l]) {  // hfmm2d.c:1057-1160, the boxlam > 16 branch
      HfJob j;
      j.kind = 1;
      j.pin = j.pout = lev[l].p;
      j.nsig = next235(4.0 * lev[l].p + 1);
      j.work = W + sp.off;
      j.nwork = sp.n;
      j.map = ml;
      j.nslot = 49;
      j.in_first = plan.loff[l];
      j.in_count = plan.loff[l + 1] - plan.loff[l];
      j.bs = bs[l];
      const int e = hf_translate(c.st, S, a.zk, j, mp, per, lo, per, nd);
      if (e) return e;
      continue;
    }
    m2l_levels.push_back({W + sp.off, sp.n, ml, dT + tab[l].m2l[0], 2 * lev[l].p + 1, 0.0,
                          dT32 ? dT32 + tab[l].m2l[0] : nullptr, nullptr, 1});
  }
  if (!m2l_levels.empty()) HIP_TRY(launch_m2l_levels(c.st, m2l_levels, mp, per, lo, per, nd, &S));
  tm->mark(c.st);
  // step 5: L2L
  for (int l = lf; l <= L - 1; ++l) {
    const Span sp = plan.l2l_kids.at(l);
    if (!sp.n || !gate[l]) continue;
    const int* map = map_l2l + (size_t)(sp.off - plan.l2l_kids.off) * 4;
    if (hf[l]) {  // hfmm2d.c:1165-1254, the boxlam > 16 branch
      HfJob j;
      j.kind = 2;
      j.pin = lev[l].p;
      j.pout = lev[l + 1].p;
      j.nsig = next235(2.0 * (lev[l].p + lev[l + 1].p) + 1);
      j.work = W + sp.off;
      j.nwork = sp.n;
      j.map = map;
      j.nslot = 4;
      j.in_first = plan.loff[l];
      j.in_count = plan.loff[l + 1] - plan.loff[l];
      j.bs = bs[l];
      const int e = hf_translate(c.st, S, a.zk, j, lo, per, lo, per, nd);
      if (e) return e;
      continue;
    }
    HIP_TRY(launch_xlate(c.st, S, W + sp.off, sp.n, map, 4, 1, dT + tab[l].l2l[0], 2 * lev[l].p + 1,
                             2 * lev[l + 1].p + 1, lo, per, lo, per, nd, 0.0));
  }
  // steps 6-8 at sources and targets
  struct Set {
    int pg; Span l3, lv, nn; const int2* rng; const double2* pts; int N; double2 *po, *gr, *he;
    const int* perm; double2 *uo, *ug, *uh; bool* fused;
    const WorkList *l3w, *nnw;
  };
  Set sets[2] = {{a.ifpgh, s_l3s, s_lvs, plan.p2p_s.all(), D.srng, Pt.src, Pt.ns, a.pot, a.grad,
                  a.hess, Pt.perm_s, a.upot, a.ugrad, a.uhess, &a.fused_s, &plan.list3_s,
                  &plan.p2p_s},
                 {a.ifpghtarg, s_l3t, s_lvt, plan.p2p_t.all(), D.trng, Pt.trg, Pt.nt, a.pott,
                  a.gradt, a.hesst, Pt.perm_t, a.upott, a.ugradt, a.uhesst, &a.fused_t,
                  &plan.list3_t, &plan.p2p_t}};
  a.fused_s = a.fused_t = false;
  const double thresh = std::abs(a.zk) * Pt.bs0 * 0x1p-51;
  // k r over list 1, per level of the target leaves, from the tree's largest
  // list-1 box-pair distance there (for leaves of one size s, 2 sqrt(2) s;
  // with a neighbour of size 2s, up to 4.25 s)
  // (a level with no list-1 pair gets the shortest series, which it never uses)
  auto xmax1 = [&](int l) { return zk.x * std::max(plan.p2p_rmax_lev[l], 1e-300); };
  // k r over list 3: a point of a leaf of size s to the centre of a child
  // of a colleague is at most 1.75 s away per coordinate, r <= 2.48 s
  auto xmax3 = [&](int l) { return zk.x * 2.48 * bs[l]; };
  tm->mark(c.st);
  for (auto& s : sets)
    if (s.pg > 0 && s.N > 0 && s.l3.n)
      by_tier(*s.l3w, lg - 1, L, rz, xmax3, [&](const Span& sp, double xm) {
        launch_heval<true>(s.pg, sp.n, c.st, wptr(sp), D, dlev, zk, s.rng, s.pts, s.N, nd, per,
                           mp, xm, pmax_pt, s.po, s.gr, s.he);
      });
  tm->mark(c.st);
  for (auto& s : sets)
    if (s.pg > 0 && s.N > 0 && s.lv.n)
      launch_heval<false>(s.pg, s.lv.n, c.st, wptr(s.lv), D, dlev, zk, s.rng, s.pts, s.N, nd, per, lo, 0.0, pmax_pt, s.po, s.gr, s.he);
  tm->mark(c.st);
  for (auto& s : sets) {
    if (s.pg <= 0 || s.N == 0 || !a.ifnear || !s.nn.n) continue;
    by_tier(*s.nnw, 0, L, rz, xmax1, [&](const Span& sp, double xm) {
      if (chg && dip)
        launch_hp2p<true, true>(s.pg, sp.n, c.st, W + sp.off, D, zk, Pt.src, chg, dip, dv, Pt.ns, s.rng, s.pts, s.N, nd, thresh, xm, s.po, s.gr, s.he, s.perm, s.uo, s.ug, s.uh);
      else if (chg)
        launch_hp2p<true, false>(s.pg, sp.n, c.st, W + sp.off, D, zk, Pt.src, chg, dip, dv, Pt.ns, s.rng, s.pts, s.N, nd, thresh, xm, s.po, s.gr, s.he, s.perm, s.uo, s.ug, s.uh);
      else
        launch_hp2p<false, true>(s.pg, sp.n, c.st, W + sp.off, D, zk, Pt.src, chg, dip, dv, Pt.ns, s.rng, s.pts, s.N, nd, thresh, xm, s.po, s.gr, s.he, s.perm, s.uo, s.ug, s.uh);
    });
    *s.fused = s.uo != nullptr;
  }
  tm->mark(c.st);
  HIP_TRY(hipGetLastError());
  return IER_OK;
}

}  // namespace fmm

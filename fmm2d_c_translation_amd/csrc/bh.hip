// Biharmonic FMM (bhfmm2d_, c_translation/src/bhfmm2d.c:315-1282) and the
// Stokes FMM built on it (stfmm2d_, stfmm2d.c:40-249), on the device tree
// and translation engine of the other kernels; biharmonic direct sums
// (bh2d_direct*_, bhkernels2d.c:30-398).
//
// A biharmonic expansion holds five coefficient sets k = 1..5 of orders
// n = 0..p per channel (bhrouts2d.c:16-31).  Here a box's expansions are
// 5 nd rows of P = p + 1 complex values, row (k - 1) nd + d, and sets 2 and
// 3 are stored conjugated: psi = (m1, conj m2, conj m3, m4, m5).  In that
// form every translation of bhrouts2d.c:745-1128 is one P x P operator V
// applied to all five rows, plus a coupling of row 3 into row 2:
//   psi_k' = V psi_k (k = 1..5),  psi_2' += (-/+ conj zd) V psi_3
// (minus for M2M and M2L, plus for L2L; zd the shift of the centre), since
// their set-2/3 blocks are conj(V) and -/+ zd conj(V).  V is the Laplace
// operator of the same shift (with its log column: orders 0 of sets 1-3 of
// a multipole are zero, so that column only acts on sets 4 and 5), so the
// slot-GEMM kernels of translate.h run the 5 nd rows as channels, and the
// coupling as a second pass over row 3.  The operators depend on the box
// size through zd and log|zd|, so they are built per level, on the device,
// at every solve.
//
// With z = target - source and w = 1/z, the kernel is (bhkernels2d.c):
//   charges (c1, c2):  vel += 2 c1 log|z| + c2 conj(w) z
//                      g1 += c1 w, g2 += c2 conj(w), g3 += c1 conj(w) - c2 conj(w^2) z
//   dipoles (d1, d2, d3): vel += d1 w + d3 conj(w) + d2 conj(w^2) z
//                      g1 -= d1 w^2, g2 += d2 conj(w^2),
//                      g3 -= d3 conj(w^2) + 2 d2 conj(w^3) z
// skipping pairs with |z| <= thresh.  Charges are (nd,2,n) complex,
// dipoles (nd,3,n), vel (nd,n) and grad (nd,3,n).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <type_traits>
#include <vector>

#include "engine.h"
#include "special.h"
#include "translate.h"

namespace fmm {
namespace {

constexpr int BW = 64;      // threads of the per-box kernels
constexpr int BCAP = 128;   // near-field sources staged per pass

inline int nblkb(long n, int b) { return (int)((n + b - 1) / b); }

__device__ inline cd ldz(const double2* p) { const double2 v = *p; return cd{v.x, v.y}; }
__device__ inline void addz(double2* o, cd v) { *o = make_double2(o->x + v.x, o->y + v.y); }
__device__ inline cd mul_i(cd a) { return cd{-a.y, a.x}; }

// ---------------------------------------------------------------------
// operators: per level l, slots 0-3 M2M (children at l + 1 into the
// parent at l), 4-7 L2L (parent at l into children at l + 1), 8-56 M2L at
// level l (slot (dy+3)*7 + (dx+3), source minus target in box units);
// T[j][i], j input order, i output order (translate.h's layout); V the
// five-row operator, U the coupling of row 3 into row 2.
// Closed forms of bh2dmpmp / bh2dlocloc / bh2dmploc (bhrouts2d.c:745-1128):
// with the powers a_j and b_i below and K the binomial weight,
// v = b_i K a_j, and the log entries of sets 4, 5 at order 0 in.
// ---------------------------------------------------------------------
constexpr int BH_SLOTS = 57;
__global__ void k_bh_ops(int L, int P, const double* bs, const double* binom, int bw, double2* V,
                         double2* U) {
  const long e = (long)blockIdx.x * blockDim.x + threadIdx.x;
  const long total = (long)(L + 1) * BH_SLOTS * P * P;
  if (e >= total) return;
  const int j = (int)(e % P), i = (int)((e / P) % P);
  const int s = (int)((e / ((long)P * P)) % BH_SLOTS), l = (int)(e / ((long)P * P * BH_SLOTS));
  auto C = [&](int a, int b) { return binom[(long)a * bw + b]; };
  cd v = mk(0, 0), cpl = mk(0, 0), l4 = mk(0, 0);
  bool ok = true, logcol = false;
  if (s < 8) {
    const int q = s & 3;
    if (l + 1 > L) ok = false;  // no children below the deepest level
    const double bc = ok ? bs[l + 1] : 1.0;
    const cd wq = mk(q & 1 ? 0.5 : -0.5, q & 2 ? 0.5 : -0.5);  // child - parent, child units
    const cd zd = wq * bc;
    if (s < 4) {
      // M2M (bhrouts2d.c:860-995): a_j = (rs1/zd)^j, b_i = (zd/rs2)^i,
      // K = C(i-1, j-1) for 1 <= j <= i; sets 4, 5: 1 at (0 <- 0), -b_i / i
      // at (i <- 0); coupling -conj(zd)
      const cd a = cpow_n(inv(wq), j), b = cpow_n(wq * 0.5, i);
      const double K = (j >= 1 && j <= i) ? C(i - 1, j - 1) : 0.0;
      v = b * a * K;
      cpl = -conj(zd);
      logcol = true;
      if (j == 0) l4 = i == 0 ? mk(1, 0) : b * (-1.0 / i);
    } else {
      // L2L (bhrouts2d.c:745-858): a_j = (zd/rs1)^j, b_i = (rs2/zd)^i,
      // K = C(j, i) for j >= i; coupling +conj(zd)
      const cd a = cpow_n(wq * 0.5, j), b = cpow_n(inv(wq), i);
      const double K = j >= i ? C(j, i) : 0.0;
      v = b * a * K;
      cpl = conj(zd);
    }
  } else {
    const int m = s - 8, dx = m % 7 - 3, dy = m / 7 - 3;
    if (abs(dx) <= 1 && abs(dy) <= 1) ok = false;
    // M2L (bhrouts2d.c:997-1128): zd = source - target, a_j = (-rs/zd)^j,
    // b_i = (rs/zd)^i, K = C(i+j-1, i) for j >= 1; sets 4, 5: log|zd| at
    // (0 <- 0), -b_i / i at (i <- 0); coupling -conj(zd)
    const cd w = mk((double)dx, (double)dy);
    const cd zd = w * bs[l];
    const cd a = cpow_n(-inv(w), j), b = cpow_n(inv(w), i);
    const double K = j >= 1 ? C(i + j - 1, i) : 0.0;
    v = b * a * K;
    cpl = -conj(zd);
    logcol = true;
    if (j == 0) l4 = i == 0 ? mk(log(cabs_(zd)), 0) : b * (-1.0 / i);
  }
  cd t = logcol && j == 0 ? l4 : v, u = cpl * v;
  if (!ok) t = u = mk(0, 0);
  const long o = ((long)l * BH_SLOTS + s) * P * P + (long)j * P + i;
  V[o] = make_double2(t.x, t.y);
  U[o] = make_double2(u.x, u.y);
}

// ---------------------------------------------------------------------
// expansions from points (bh2dformmp*, bh2dformta*; bhrouts2d.c:354-743):
// a wave per box, lane n computes order n of the five sets over the box's
// sources (multipoles) or its list-4 sources (local expansions), in order.
// ---------------------------------------------------------------------
template <bool MP, bool C, bool Q>
__global__ void __launch_bounds__(BW) k_bh_form(const int* work, DTree D, const double2* src,
                                                const double2* chg, const double2* dip, int nd,
                                                int p, double2* ex) {
  const int b = work[blockIdx.x];
  const int n = threadIdx.x;
  const int P = p + 1;
  const double r = D.bs[D.level[b]], ri = 1.0 / r;
  const double2 c = D.ctr[b];
  const int nbox = MP ? 1 : D.o4[b + 1] - D.o4[b];
  for (int d = 0; d < nd; ++d) {
    cd A[5] = {mk(0, 0), mk(0, 0), mk(0, 0), mk(0, 0), mk(0, 0)};
    for (int k = 0; k < nbox; ++k) {
      const int jb = MP ? b : D.l4[D.o4[b] + k];
      const int2 rg = D.srng[jb];
      for (int s = rg.x; s < rg.y; ++s) {
        if (n > p) continue;
        const cd dz = mk(src[s].x - c.x, src[s].y - c.y);
        const cd c1 = C ? ldz(chg + ((size_t)s * 2 + 0) * nd + d) : mk(0, 0);
        const cd c2 = C ? ldz(chg + ((size_t)s * 2 + 1) * nd + d) : mk(0, 0);
        const cd d1 = Q ? ldz(dip + ((size_t)s * 3 + 0) * nd + d) : mk(0, 0);
        const cd d2 = Q ? ldz(dip + ((size_t)s * 3 + 1) * nd + d) : mk(0, 0);
        const cd d3 = Q ? ldz(dip + ((size_t)s * 3 + 2) * nd + d) : mk(0, 0);
        const double re2 = 2.0 * c1.x, im2 = 2.0 * c1.y;
        if (MP) {
          const cd w = dz * ri;
          if (C && n == 0) {
            A[3] += mk(re2, 0);
            A[4] += mk(im2, 0);
          }
          if (cabs_(w) <= 1.0e-16) {  // source at the centre
            if (C && n == 1) A[2] += c2 * ri;
            if (Q && n == 1) {
              A[0] += d1 * ri;
              A[1] += d3 * ri;
            }
            if (Q && n == 2) A[2] += d2 * (ri * ri);
            continue;
          }
          if (n == 0) continue;
          const cd wn = cpow_n(w, n), cwn = conj(wn);
          const cd ct = inv(conj(dz));
          if (C) {
            A[3] -= wn * (re2 / n);
            A[4] -= wn * (im2 / n);
            const cd a = c2 * ct;
            A[1] -= a * cwn * dz;
            A[2] += a * cwn;
          }
          if (Q) {
            A[0] += d1 * wn * inv(dz);
            const cd bq = d2 * (ct * ct) * (double)(n - 1);
            A[1] -= bq * cwn * dz;
            A[2] += bq * cwn;
            A[1] += d3 * ct * cwn;
          }
        } else {
          const cd t = inv(dz), ct = conj(t);
          const cd qn = cpow_n(t * r, n), cqn = conj(qn);
          if (C) {
            if (n == 0) {
              const double lg = log(cabs_(dz));
              A[3] += mk(re2 * lg, 0);
              A[4] += mk(im2 * lg, 0);
            } else {
              A[3] -= qn * (re2 / n);
              A[4] -= qn * (im2 / n);
            }
            A[1] += c2 * cqn * ct * inv(t);
            A[2] -= c2 * cqn * ct;
          }
          if (Q) {
            A[0] -= d1 * qn * t;
            A[1] -= d3 * cqn * ct;
            const cd e2 = d2 * cqn * (ct * ct) * (double)(n + 1);
            A[1] -= e2 * inv(t);
            A[2] += e2;
          }
        }
      }
    }
    if (n <= p) {
      A[1] = conj(A[1]);
      A[2] = conj(A[2]);
      double2* o = ex + ((size_t)b * 5 * nd + d) * P + n;
      for (int k = 0; k < 5; ++k) addz(o + (size_t)k * nd * P, A[k]);
    }
  }
}

// ---------------------------------------------------------------------
// expansions at points (bh2dmpeval*, bh2dtaeval*; bhrouts2d.c:37-352):
// 16-lane groups, four boxes per wave, lane u of a group takes points
// u, u + 16, ... of its box; MP: the multipoles of the box's list 3,
// else its own local expansion.  pot (nd, n), grad (nd, 3, n), sorted.
// ---------------------------------------------------------------------
template <bool MP, int PG>
__global__ void __launch_bounds__(BW) k_bh_eval(const int* work, int nwork, DTree D,
                                                const int2* rng, const double2* pts, int nd,
                                                int p, const double2* ex, double2* pot,
                                                double2* grad) {
  const int w = blockIdx.x * 4 + (int)(threadIdx.x >> 4), u = threadIdx.x & 15;
  if (w >= nwork) return;
  const int b = work[w];
  const int P = p + 1, R = nd * P;  // R: distance between the sets of a box
  const int2 rg = rng[b];
  const int nbox = MP ? D.o3[b + 1] - D.o3[b] : 1;
  for (int t = rg.x + u; t < rg.y; t += 16) {
    const double2 x = pts[t];
    for (int d = 0; d < nd; ++d) {
      cd v = mk(0, 0), g1 = mk(0, 0), g2 = mk(0, 0), g3 = mk(0, 0);
      for (int k = 0; k < nbox; ++k) {
        const int jb = MP ? D.l3[D.o3[b] + k] : b;
        const double r = D.bs[D.level[jb]], ri = 1.0 / r;
        const double2 c = D.ctr[jb];
        const double2* E = ex + ((size_t)jb * 5 * nd + d) * P;
        const cd dz = mk(x.x - c.x, x.y - c.y);
        if (MP) {
          const cd uu = inv(dz) * r;
          const cd A0 = ldz(E + 3 * R) + mul_i(ldz(E + 4 * R));
          v += A0 * log(cabs_(dz));
          if (PG >= 2) {
            g1 += A0 * uu * (0.5 * ri);
            g3 += A0 * conj(uu) * (0.5 * ri);
          }
          cd ui = mk(1, 0);
          for (int i = 1; i <= p; ++i) {
            ui = ui * uu;
            const cd m1 = ldz(E + i), m2 = conj(ldz(E + R + i)), m3 = conj(ldz(E + 2 * R + i));
            const cd m4 = ldz(E + 3 * R + i), m5 = ldz(E + 4 * R + i);
            const cd cu = conj(ui);
            const cd r4 = m4 * ui, r5 = m5 * ui;
            v += m1 * ui + m2 * cu + m3 * cu * dz + mk(r4.x, r5.x);
            if (PG >= 2) {
              const cd u1 = ui * uu, cu1 = conj(u1);
              const double f = i * ri;
              g1 -= (m1 + m4 * 0.5 + mul_i(m5) * 0.5) * u1 * f;
              g2 += m3 * cu;
              g3 -= (m2 + dz * m3) * cu1 * f;
              g3 -= conj(m4 * u1 * 0.5) * f;
              g3 += conj(mul_i(m5) * u1 * 0.5) * f;
            }
          }
        } else {
          const cd vv = dz * ri;
          cd vi = mk(1, 0), vim = mk(0, 0);
          for (int i = 0; i <= p; ++i) {
            const cd m1 = ldz(E + i), m2 = conj(ldz(E + R + i)), m3 = conj(ldz(E + 2 * R + i));
            const cd m4 = ldz(E + 3 * R + i), m5 = ldz(E + 4 * R + i);
            const cd cv = conj(vi);
            const cd r4 = m4 * vi, r5 = m5 * vi;
            v += m1 * vi + m2 * cv + m3 * cv * dz + mk(r4.x, r5.x);
            if (PG >= 2) {
              if (i == 0) {
                g2 += m3;
              } else {
                const double f = i * ri;
                const cd cvm = conj(vim);
                g1 += (m1 + m4 * 0.5 + mul_i(m5) * 0.5) * vim * f;
                g2 += m3 * cv;
                g3 += (m2 + m3 * dz) * cvm * f;
                g3 += conj(m4 * vim * 0.5) * f;
                g3 -= conj(mul_i(m5) * vim * 0.5) * f;
              }
            }
            vim = vi;
            vi = vi * vv;
          }
        }
      }
      addz(pot + (size_t)t * nd + d, v);
      if (PG >= 2) {
        addz(grad + ((size_t)t * 3 + 0) * nd + d, g1);
        addz(grad + ((size_t)t * 3 + 1) * nd + d, g2);
        addz(grad + ((size_t)t * 3 + 2) * nd + d, g3);
      }
    }
  }
}

// one source-target pair of the kernel above
template <bool C, bool Q, int PG>
__device__ __forceinline__ void bh_pair(cd z, cd c1, cd c2, cd d1, cd d2, cd d3, cd& v, cd& g1,
                                        cd& g2, cd& g3) {
  const cd w = inv(z), cw = conj(w);
  if (C) {
    v += c1 * (2.0 * log(cabs_(z))) + c2 * cw * z;
    if (PG >= 2) {
      g1 += c1 * w;
      g2 += c2 * cw;
      g3 += c1 * cw - c2 * (cw * cw) * z;
    }
  }
  if (Q) {
    const cd cw2 = cw * cw;
    v += d1 * w + d3 * cw + d2 * cw2 * z;
    if (PG >= 2) {
      g1 -= d1 * (w * w);
      g2 += d2 * cw2;
      g3 -= d3 * cw2 + d2 * (cw2 * cw) * z * 2.0;
    }
  }
}

// ---------------------------------------------------------------------
// list-1 near field: a wave per target leaf, lanes = (target, source
// slice) with K = 64 / Tc slices, list-1 sources staged in LDS, the K
// partial sums of a target added in slice order.
// ---------------------------------------------------------------------
template <bool C, bool Q, int PG>
__global__ void __launch_bounds__(BW) k_bh_p2p(const int* work, DTree D, const double2* src,
                                               const double2* chg, const double2* dip,
                                               const int2* trng, const double2* tpts, int nd,
                                               double thresh, double2* pot, double2* grad) {
  __shared__ double2 sx[BCAP], sc[2][BCAP], sq[3][BCAP];
  __shared__ double2 red[4][BW];
  __shared__ int boff[14], bsrc[14];
  const int b = work[blockIdx.x];
  const int lane = threadIdx.x;
  const int n1 = D.o1[b + 1] - D.o1[b];
  if (lane == 0) {
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
  const int S = boff[n1];
  const int2 tr = trng[b];
  for (int t0 = tr.x; t0 < tr.y; t0 += BW) {
    const int Tc = min(BW, tr.y - t0), K = BW / Tc;
    const int t = lane / K, k = lane % K;
    const bool act = t < Tc;
    const double2 x = act ? tpts[t0 + t] : make_double2(0, 0);
    for (int d = 0; d < nd; ++d) {
      cd v = mk(0, 0), g1 = mk(0, 0), g2 = mk(0, 0), g3 = mk(0, 0);
      for (int g0 = 0; g0 < S; g0 += BCAP) {
        const int cnt = min(BCAP, S - g0);
        for (int m = lane; m < cnt; m += BW) {
          const int g = g0 + m;
          int kb = 0;
          while (boff[kb + 1] <= g) ++kb;
          const int si = bsrc[kb] + (g - boff[kb]);
          sx[m] = src[si];
          if (C) {
            sc[0][m] = chg[((size_t)si * 2 + 0) * nd + d];
            sc[1][m] = chg[((size_t)si * 2 + 1) * nd + d];
          }
          if (Q)
            for (int e = 0; e < 3; ++e) sq[e][m] = dip[((size_t)si * 3 + e) * nd + d];
        }
        __syncthreads();
        if (act)
          for (int m = k; m < cnt; m += K) {
            const cd z = mk(x.x - sx[m].x, x.y - sx[m].y);
            if (cabs_(z) <= thresh) continue;
            bh_pair<C, Q, PG>(z, C ? ldz(&sc[0][m]) : mk(0, 0), C ? ldz(&sc[1][m]) : mk(0, 0),
                              Q ? ldz(&sq[0][m]) : mk(0, 0), Q ? ldz(&sq[1][m]) : mk(0, 0),
                              Q ? ldz(&sq[2][m]) : mk(0, 0), v, g1, g2, g3);
          }
        __syncthreads();
      }
      red[0][lane] = make_double2(v.x, v.y);
      red[1][lane] = make_double2(g1.x, g1.y);
      red[2][lane] = make_double2(g2.x, g2.y);
      red[3][lane] = make_double2(g3.x, g3.y);
      __syncthreads();
      if (act && k == 0) {
        const size_t o = t0 + t;
        const int nv = PG >= 2 ? 4 : 1;
        for (int q = 0; q < nv; ++q) {
          cd s = mk(0, 0);
          for (int u = 0; u < K; ++u) s += ldz(&red[q][t * K + u]);
          if (q == 0) addz(pot + o * nd + d, s);
          else addz(grad + (o * 3 + (q - 1)) * nd + d, s);
        }
      }
      __syncthreads();
    }
  }
}

// rows of `row` complex values per point, between input and sorted order
__global__ void k_bh_gather(const double2* in, double2* out, const int* perm, int n, int row) {
  const long e = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= (long)n * row) return;
  const int i = (int)(e / row), c = (int)(e % row);
  out[e] = in[(size_t)perm[i] * row + c];
}
__global__ void k_bh_scatter(const double2* in, double2* out, const int* perm, int n, int row) {
  const long e = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= (long)n * row) return;
  const int i = (int)(e / row), c = (int)(e % row);
  out[(size_t)perm[i] * row + c] = in[e];
}

template <bool C, bool Q>
void launch_bh_p2p(int pg, int nbx, hipStream_t st, const int* work, const DTree& D,
                   const double2* src, const double2* chg, const double2* dip, const int2* trng,
                   const double2* tpts, int nd, double th, double2* pot, double2* grad) {
  if (pg == 1)
    k_bh_p2p<C, Q, 1><<<nbx, BW, 0, st>>>(work, D, src, chg, dip, trng, tpts, nd, th, pot, grad);
  else
    k_bh_p2p<C, Q, 2><<<nbx, BW, 0, st>>>(work, D, src, chg, dip, trng, tpts, nd, th, pot, grad);
}

template <bool MP>
void launch_bh_eval(int pg, int nw, hipStream_t st, const int* work, const DTree& D,
                    const int2* rng, const double2* pts, int nd, int p, const double2* ex,
                    double2* pot, double2* grad) {
  const int nbx = (nw + 3) / 4;
  if (pg == 1)
    k_bh_eval<MP, 1><<<nbx, BW, 0, st>>>(work, nw, D, rng, pts, nd, p, ex, pot, grad);
  else
    k_bh_eval<MP, 2><<<nbx, BW, 0, st>>>(work, nw, D, rng, pts, nd, p, ex, pot, grad);
}

template <bool MP>
void launch_bh_form(bool c, bool q, int nw, hipStream_t st, const int* work, const DTree& D,
                    const double2* src, const double2* chg, const double2* dip, int nd, int p,
                    double2* ex) {
  if (c && q) k_bh_form<MP, true, true><<<nw, BW, 0, st>>>(work, D, src, chg, dip, nd, p, ex);
  else if (c) k_bh_form<MP, true, false><<<nw, BW, 0, st>>>(work, D, src, chg, dip, nd, p, ex);
  else k_bh_form<MP, false, true><<<nw, BW, 0, st>>>(work, D, src, chg, dip, nd, p, ex);
}

}  // namespace

// leaf size of bhndiv2d (bhndiv2d.c:17-50)
int bh_ndiv(double eps, int ns, int nt) {
  if (eps >= 0.5e-0) return 3;
  if (eps >= 0.5e-1) return 5;
  if (eps >= 0.5e-2) return 8;
  if (eps >= 0.5e-3) return 10;
  if (eps >= 0.5e-6) return 15;
  if (eps >= 0.5e-9) return 20;
  if (eps >= 0.5e-12) return 25;
  if (eps >= 0.5e-15) return 45;
  return ns + nt;
}

// The eight steps of bhfmm2dmain (bhfmm2d.c:315-920) on sorted data:
// chg (n,2,nd) and dip (n,3,nd) point-major, outputs (n,nd) / (n,3,nd)
// sorted, incremented.  Orders: bh2dterms = the l2dterms rule.
struct BhArgs {
  int nd = 1;
  double eps = 1e-6;
  int ifcharge = 0, ifdipole = 0, ifpgh = 0, ifpghtarg = 0;
  const double2 *chg = nullptr, *dip = nullptr;
  double2 *pot = nullptr, *grad = nullptr, *pott = nullptr, *gradt = nullptr;
};
int bh_run(Ctx& c, Scratch& S, const SortedPoints& Pt, const Plan& plan, const BhArgs& a) {
  const DTree& D = plan.D;
  const int nd = a.nd, L = plan.L, nb = plan.nb;
  const int p = l2dterms(a.eps);
  if (p > 63) return IER_P_TOO_LARGE;
  const int P = p + 1;
  const bool C = a.ifcharge == 1, Q = a.ifdipole == 1;
  if (!C && !Q) return IER_OK;
  // operators of every level
  const int bw = 2 * P + 1;
  std::vector<double> hb((size_t)bw * bw, 0.0);
  for (int l = 0; l < bw; ++l) {
    hb[(size_t)l * bw] = 1.0;
    for (int m = 1; m <= l; ++m)
      hb[(size_t)l * bw + m] = hb[(size_t)(l - 1) * bw + m - 1] + (m < l ? hb[(size_t)(l - 1) * bw + m] : 0.0);
  }
  double* dbin = S.get<double>(hb.size());
  double* dbs = S.get<double>(L + 2);
  const size_t per_op = (size_t)P * P;
  double2* V = S.get<double2>((size_t)(L + 1) * BH_SLOTS * per_op);
  double2* U = S.get<double2>((size_t)(L + 1) * BH_SLOTS * per_op);
  // a box's expansions: 5 nd rows of P; rows 2 nd .. 3 nd - 1 (set 3) couple
  // into rows nd .. 2 nd - 1 (set 2)
  const size_t per = (size_t)5 * nd * P, r2 = (size_t)nd * P, r3 = 2 * r2;
  const int rows = 5 * nd;
  double2* mp = S.get<double2>(per * nb);
  double2* lo = S.get<double2>(per * nb);
  if (S.err) return S.err;
  std::vector<double> hbs(plan.bs.begin(), plan.bs.end());
  hbs.push_back(hbs.back() * 0.5);
  double* pin = static_cast<double*>(pinned(sizeof(double) * (hb.size() + hbs.size()), 7));
  if (!pin) return IER_ALLOC;
  std::memcpy(pin, hb.data(), sizeof(double) * hb.size());
  std::memcpy(pin + hb.size(), hbs.data(), sizeof(double) * hbs.size());
  HIP_TRY(hipMemcpyAsync(dbin, pin, sizeof(double) * hb.size(), hipMemcpyHostToDevice, c.st));
  HIP_TRY(hipMemcpyAsync(dbs, pin + hb.size(), sizeof(double) * hbs.size(), hipMemcpyHostToDevice, c.st));
  {
    const long tot = (long)(L + 1) * BH_SLOTS * P * P;
    k_bh_ops<<<nblkb(tot, 256), 256, 0, c.st>>>(L, P, dbs, dbin, bw, V, U);
  }
  HIP_TRY(hipMemsetAsync(mp, 0, sizeof(double2) * per * nb, c.st));
  HIP_TRY(hipMemsetAsync(lo, 0, sizeof(double2) * per * nb, c.st));
  const int* W = plan.dwork;
  auto opv = [&](int l, int s) { return V + ((size_t)l * BH_SLOTS + s) * per_op; };
  auto opu = [&](int l, int s) { return U + ((size_t)l * BH_SLOTS + s) * per_op; };
  StepTimer* tm = step_timer();
  {
    SolveStats& st = last_stats();
    for (int l = 2; l <= L; ++l) st.m2l_pairs += plan.m2l_pairs[l];
    st.m2l_flops = st.m2l_pairs * 8.0 * P * P * (rows + nd);
    st.pmax = p;
  }
  tm->mark(c.st);
  // step 1: multipoles of the leaves at levels >= 2
  if (plan.formmp.n)
    launch_bh_form<true>(C, Q, plan.formmp.n, c.st, W + plan.formmp.off, D, Pt.src, a.chg, a.dip,
                         nd, p, mp);
  tm->mark(c.st);
  // step 3: M2M, finest level first
  int* map_m2m = S.get<int>((size_t)std::max(plan.m2m.n, 1) * 4);
  int* map_l2l = S.get<int>((size_t)std::max(plan.l2l_kids.n, 1) * 4);
  int* map2 = S.get<int>((size_t)std::max(plan.m2l.n, 1) * 49);
  if (S.err) return S.err;
  if (plan.m2m.n)
    k_map_m2m<<<nblkb(plan.m2m.n, 128), 128, 0, c.st>>>(W + plan.m2m.off, plan.m2m.n, D.child,
                                                         D.srng, map_m2m);
  if (plan.l2l_kids.n)
    k_map_l2l<<<nblkb(plan.l2l_kids.n, 128), 128, 0, c.st>>>(
        W + plan.l2l_kids.off, plan.l2l_kids.n, D.parent, D.child, map_l2l);
  for (int l = L - 1; l >= 1; --l) {
    const Span sp = plan.m2m.at(l);
    if (!sp.n) continue;
    const int* map = map_m2m + (size_t)(sp.off - plan.m2m.off) * 4;
    HIP_TRY(launch_xlate(c.st, S, W + sp.off, sp.n, map, 4, 4, opv(l, 0), P, P, mp, per, mp, per,
                         rows, 0.0));
    HIP_TRY(launch_xlate(c.st, S, W + sp.off, sp.n, map, 4, 4, opu(l, 0), P, P, mp + r3, per,
                         mp + r2, per, nd, 0.0));
  }
  tm->mark(c.st);
  // step 2: local expansions from list 4
  if (plan.list4.n)
    launch_bh_form<false>(C, Q, plan.list4.n, c.st, W + plan.list4.off, D, Pt.src, a.chg, a.dip,
                          nd, p, lo);
  tm->mark(c.st);
  // step 4: M2L, every level
  if (plan.m2l.n) {
    HIP_TRY(hipMemsetAsync(map2, 0xff, sizeof(int) * 49 * (size_t)plan.m2l.n, c.st));
    k_map_m2l_all<<<nblkb(32L * plan.m2l.n, 256), 256, 0, c.st>>>(W + plan.m2l.off, plan.m2l.n,
                                                                  D, map2);
  }
  std::vector<M2LLevel> m2l_v, m2l_u;
  for (int l = 2; l <= L; ++l) {
    const Span sp = plan.m2l.at(l);
    if (!sp.n) continue;
    const int* map = map2 + (size_t)(sp.off - plan.m2l.off) * 49;
    m2l_v.push_back({W + sp.off, sp.n, map, opv(l, 8), P, 0.0, nullptr});
    m2l_u.push_back({W + sp.off, sp.n, map, opu(l, 8), P, 0.0, nullptr});
  }
  if (!m2l_v.empty()) {
    HIP_TRY(launch_m2l_levels(c.st, m2l_v, mp, per, lo, per, rows));
    HIP_TRY(launch_m2l_levels(c.st, m2l_u, mp + r3, per, lo + r2, per, nd));
  }
  tm->mark(c.st);
  // step 5: L2L into the children
  for (int l = 1; l <= L - 1; ++l) {
    const Span sp = plan.l2l_kids.at(l);
    if (!sp.n) continue;
    const int* map = map_l2l + (size_t)(sp.off - plan.l2l_kids.off) * 4;
    HIP_TRY(launch_xlate(c.st, S, W + sp.off, sp.n, map, 4, 1, opv(l, 4), P, P, lo, per, lo, per,
                         rows, 0.0));
    HIP_TRY(launch_xlate(c.st, S, W + sp.off, sp.n, map, 4, 1, opu(l, 4), P, P, lo + r3, per,
                         lo + r2, per, nd, 0.0));
  }
  // steps 6-8 at sources and targets; the reference evaluates pg 1 and 2
  // only (bhfmm2d.c:780-920), so pg 3 adds nothing
  struct Set { int pg; Span l3, lv, nn; const int2* rng; const double2* pts; double2 *po, *gr; };
  Set sets[2] = {{a.ifpgh, plan.list3_s.all(), plan.leaves_s.all(), plan.p2p_s.all(), D.srng,
                  Pt.src, a.pot, a.grad},
                 {a.ifpghtarg, plan.list3_t.all(), plan.leaves_t.all(), plan.p2p_t.all(), D.trng,
                  Pt.trg, a.pott, a.gradt}};
  tm->mark(c.st);
  for (auto& s : sets)
    if ((s.pg == 1 || s.pg == 2) && s.l3.n)
      launch_bh_eval<true>(s.pg, s.l3.n, c.st, W + s.l3.off, D, s.rng, s.pts, nd, p, mp, s.po, s.gr);
  tm->mark(c.st);
  for (auto& s : sets)
    if ((s.pg == 1 || s.pg == 2) && s.lv.n)
      launch_bh_eval<false>(s.pg, s.lv.n, c.st, W + s.lv.off, D, s.rng, s.pts, nd, p, lo, s.po, s.gr);
  tm->mark(c.st);
  const double thresh = Pt.bs0 * 0x1p-51;
  for (auto& s : sets) {
    if (!(s.pg == 1 || s.pg == 2) || !s.nn.n) continue;
    const int* w = W + s.nn.off;
    if (C && Q) launch_bh_p2p<true, true>(s.pg, s.nn.n, c.st, w, D, Pt.src, a.chg, a.dip, s.rng, s.pts, nd, thresh, s.po, s.gr);
    else if (C) launch_bh_p2p<true, false>(s.pg, s.nn.n, c.st, w, D, Pt.src, a.chg, a.dip, s.rng, s.pts, nd, thresh, s.po, s.gr);
    else launch_bh_p2p<false, true>(s.pg, s.nn.n, c.st, w, D, Pt.src, a.chg, a.dip, s.rng, s.pts, nd, thresh, s.po, s.gr);
  }
  tm->mark(c.st);
  HIP_TRY(hipGetLastError());
  return IER_OK;
}

// bhfmm2d_ on host or device arrays (bhfmm2d.c:924-1282): tree from
// bhndiv2d, strengths gathered to sorted order, the eight steps, outputs
// scattered back; pg 3 writes zeros to pot, grad and hess, as the
// reference does.
struct BhReq {
  int nd = 1;
  double eps = 1e-6;
  int ns = 0, nt = 0;
  const double *src = nullptr, *trg = nullptr;
  int ifc = 0, ifd = 0;
  const double2 *chg = nullptr, *dip = nullptr;
  int pg = 0, pgt = 0;
  double2 *pot = nullptr, *grad = nullptr, *hess = nullptr;
  double2 *pott = nullptr, *gradt = nullptr, *hesst = nullptr;
  bool device = false;  // every array in device memory
};
int run_bh(Ctx& c, Scratch& S, const BhReq& r) {
  if (r.nd < 1 || r.ns < 1 || r.nt < 0 || r.pg < 0 || r.pg > 3 || r.pgt < 0 || r.pgt > 3)
    return IER_ARGS;
  const int nd = r.nd, ns = r.ns, nt = r.nt;
  const double *src = r.src, *trg = r.trg;
  const double2 *chg = r.chg, *dip = r.dip;
  auto up = [&](const auto* h, size_t n, auto** d) -> int {
    using T = std::remove_const_t<std::remove_pointer_t<decltype(h)>>;
    T* p = S.get<T>(std::max<size_t>(n, 1));
    if (S.err) return S.err;
    if (n) HIP_TRY(hipMemcpyAsync(p, h, sizeof(T) * n, hipMemcpyHostToDevice, c.st));
    *d = p;
    return IER_OK;
  };
  int e = IER_OK;
  if (!r.device) {
    if ((e = up(r.src, 2 * (size_t)ns, &src))) return e;
    if ((e = up(r.trg, 2 * (size_t)nt, &trg))) return e;
    if (r.ifc == 1 && (e = up(r.chg, 2 * (size_t)nd * ns, &chg))) return e;
    if (r.ifd == 1 && (e = up(r.dip, 3 * (size_t)nd * ns, &dip))) return e;
  }
  SortedPoints Pt;
  Plan plan;
  e = build_plan(c, S, src, ns, trg, nt, 0, bh_ndiv(r.eps, ns, nt), Pt, plan, true, r.pg, r.pgt);
  if (e) return e;
  BhArgs a;
  a.nd = nd;
  a.eps = r.eps;
  a.ifcharge = r.ifc;
  a.ifdipole = r.ifd;
  a.ifpgh = r.pg;
  a.ifpghtarg = r.pgt;
  double2* cs = r.ifc == 1 ? S.get<double2>(2 * (size_t)nd * ns) : nullptr;
  double2* qs = r.ifd == 1 ? S.get<double2>(3 * (size_t)nd * ns) : nullptr;
  const size_t np_s = (size_t)nd * ns, np_t = (size_t)nd * std::max(nt, 1);
  double2* ps = S.get<double2>(np_s);
  double2* gs = S.get<double2>(3 * np_s);
  double2* pt = S.get<double2>(np_t);
  double2* gt = S.get<double2>(3 * np_t);
  if (S.err) return S.err;
  if (cs) k_bh_gather<<<nblkb(2L * nd * ns, 256), 256, 0, c.st>>>(chg, cs, Pt.perm_s, ns, 2 * nd);
  if (qs) k_bh_gather<<<nblkb(3L * nd * ns, 256), 256, 0, c.st>>>(dip, qs, Pt.perm_s, ns, 3 * nd);
  HIP_TRY(hipMemsetAsync(ps, 0, sizeof(double2) * np_s, c.st));
  HIP_TRY(hipMemsetAsync(gs, 0, sizeof(double2) * 3 * np_s, c.st));
  HIP_TRY(hipMemsetAsync(pt, 0, sizeof(double2) * np_t, c.st));
  HIP_TRY(hipMemsetAsync(gt, 0, sizeof(double2) * 3 * np_t, c.st));
  a.chg = cs;
  a.dip = qs;
  a.pot = ps;
  a.grad = gs;
  a.pott = pt;
  a.gradt = gt;
  e = bh_run(c, S, Pt, plan, a);
  if (e) return e;
  // outputs in input order: pg 1 pot; pg 2 pot, grad; pg 3 pot, grad, hess
  // (zeros: nothing is evaluated)
  struct Out { int pg, n; const int* perm; double2 *sp, *sg, *up, *ug, *uh; };
  Out outs[2] = {{r.pg, ns, Pt.perm_s, ps, gs, r.pot, r.grad, r.hess},
                 {r.pgt, nt, Pt.perm_t, pt, gt, r.pott, r.gradt, r.hesst}};
  for (const Out& o : outs) {
    if (o.pg <= 0 || o.n == 0) continue;
    const size_t np = (size_t)nd * o.n;
    double2 *dp = o.up, *dg = o.ug, *dh = o.uh;
    if (!r.device) {
      dp = S.get<double2>(np);
      dg = o.pg >= 2 ? S.get<double2>(3 * np) : nullptr;
      dh = o.pg >= 3 ? S.get<double2>(3 * np) : nullptr;
      if (S.err) return S.err;
    }
    k_bh_scatter<<<nblkb((long)np, 256), 256, 0, c.st>>>(o.sp, dp, o.perm, o.n, nd);
    if (o.pg >= 2) k_bh_scatter<<<nblkb(3L * np, 256), 256, 0, c.st>>>(o.sg, dg, o.perm, o.n, 3 * nd);
    if (o.pg >= 3) HIP_TRY(hipMemsetAsync(dh, 0, sizeof(double2) * 3 * np, c.st));
    if (!r.device) {
      HIP_TRY(hipMemcpyAsync(o.up, dp, sizeof(double2) * np, hipMemcpyDeviceToHost, c.st));
      if (o.pg >= 2) HIP_TRY(hipMemcpyAsync(o.ug, dg, sizeof(double2) * 3 * np, hipMemcpyDeviceToHost, c.st));
      if (o.pg >= 3) HIP_TRY(hipMemcpyAsync(o.uh, dh, sizeof(double2) * 3 * np, hipMemcpyDeviceToHost, c.st));
    }
  }
  HIP_TRY(hipGetLastError());
  HIP_TRY(hipStreamSynchronize(c.st));
  return sort_status(plan);
}

namespace {

// ---------------------------------------------------------------------
// Stokes on the biharmonic FMM (stfmm2d.c:40-249): per source
//   charge 1 = (-i f1 + f2) / 4, charge 2 = conj(charge 1)
//   z1 = -(-m2 + i m1) / 2, z2 = -n2 + i n1 (stresslet mu = m, nu = n)
//   dip 1 = -i z1 z2, dip 2 = -conj(dip 1), dip 3 = i (z1 conj(z2) + conj(z1) z2)
// zsum = -sum of charge 1 over the sources; then with V = vel + zsum
// (+ charge 1 at a source itself): u = (Im V, -Re V), p = -4 Im g1,
// grad(1,1) = Im g3 = -grad(2,2), grad(2,1) = Re(2 g1 - g3),
// grad(1,2) = -Re(2 g1 + g3).
// ---------------------------------------------------------------------
__global__ void k_st_prep(const double* stok, const double* strs, const double* strv, int ns,
                          int nd, int ifst, int ifsr, double2* chg, double2* dip) {
  const long e = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= (long)ns * nd) return;
  const int i = (int)(e / nd), d = (int)(e % nd);
  cd c1 = mk(0, 0), q1 = mk(0, 0), q2 = mk(0, 0), q3 = mk(0, 0);
  if (ifst == 1) {
    const double f1 = stok[((size_t)i * 2 + 0) * nd + d], f2 = stok[((size_t)i * 2 + 1) * nd + d];
    c1 = mk(f2, -f1) * 0.25;
  }
  if (ifsr == 1) {
    const double m1 = strs[((size_t)i * 2 + 0) * nd + d], m2 = strs[((size_t)i * 2 + 1) * nd + d];
    const double n1 = strv[((size_t)i * 2 + 0) * nd + d], n2 = strv[((size_t)i * 2 + 1) * nd + d];
    const cd z1 = mk(m2, -m1) * 0.5, z2 = mk(-n2, n1);
    q1 = -mul_i(z1 * z2);
    q2 = -conj(q1);
    q3 = mul_i(z1 * conj(z2) + conj(z1) * z2);
  }
  chg[((size_t)i * 2 + 0) * nd + d] = make_double2(c1.x, c1.y);
  chg[((size_t)i * 2 + 1) * nd + d] = make_double2(c1.x, -c1.y);
  dip[((size_t)i * 3 + 0) * nd + d] = make_double2(q1.x, q1.y);
  dip[((size_t)i * 3 + 1) * nd + d] = make_double2(q2.x, q2.y);
  dip[((size_t)i * 3 + 2) * nd + d] = make_double2(q3.x, q3.y);
}

// zsum[d] = -sum over sources of charge 1, deterministic: block (g, d)
// sums sources 256 (g + k ZB) + lane, k = 0, 1, ..., by lanes and a fixed
// tree into part[d][g]; one block of ZB = 256 threads per channel then
// adds the ZB partial sums the same way.
constexpr int ZB = 256;
__device__ inline void block_sum(double2* sh, double sx, double sy) {
  sh[threadIdx.x] = make_double2(sx, sy);
  __syncthreads();
  for (int h = 128; h > 0; h >>= 1) {
    if ((int)threadIdx.x < h)
      sh[threadIdx.x] = make_double2(sh[threadIdx.x].x + sh[threadIdx.x + h].x,
                                     sh[threadIdx.x].y + sh[threadIdx.x + h].y);
    __syncthreads();
  }
}
__global__ void __launch_bounds__(256) k_st_zpart(const double2* chg, int ns, int nd, double2* part) {
  __shared__ double2 sh[256];
  const int g = blockIdx.x, d = blockIdx.y;
  double sx = 0, sy = 0;
  for (long i = (long)g * 256 + threadIdx.x; i < ns; i += (long)ZB * 256) {
    const double2 v = chg[((size_t)i * 2 + 0) * nd + d];
    sx += v.x;
    sy += v.y;
  }
  block_sum(sh, sx, sy);
  if (threadIdx.x == 0) part[(size_t)d * ZB + g] = sh[0];
}
__global__ void __launch_bounds__(256) k_st_zsum(const double2* part, double2* zsum) {
  __shared__ double2 sh[256];
  const int d = blockIdx.x;
  const double2 v = part[(size_t)d * ZB + threadIdx.x];
  block_sum(sh, v.x, v.y);
  if (threadIdx.x == 0) zsum[d] = make_double2(-sh[0].x, -sh[0].y);
}

// velocity, pressure and velocity gradient from the biharmonic outputs
__global__ void k_st_unpack(const double2* vel, const double2* g, const double2* zsum,
                            const double2* chg, int n, int nd, int flag, double* pot, double* pre,
                            double* grad) {
  const long e = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= (long)n * nd) return;
  const int i = (int)(e / nd), d = (int)(e % nd);
  cd V = ldz(vel + e) + ldz(zsum + d);
  if (chg) V += ldz(chg + ((size_t)i * 2 + 0) * nd + d);
  pot[((size_t)i * 2 + 0) * nd + d] = V.y;
  pot[((size_t)i * 2 + 1) * nd + d] = -V.x;
  if (flag >= 2) {
    const cd g1 = ldz(g + ((size_t)i * 3 + 0) * nd + d), g3 = ldz(g + ((size_t)i * 3 + 2) * nd + d);
    pre[(size_t)i * nd + d] = -4.0 * g1.y;
    if (flag >= 3) {
      double* G = grad + (size_t)i * 4 * nd + d;  // (nd, 2, 2): (a, b) at ((b*2 + a) nd)
      G[(0 * 2 + 0) * nd] = g3.y;
      G[(1 * 2 + 1) * nd] = -g3.y;
      G[(0 * 2 + 1) * nd] = 2.0 * g1.x - g3.x;
      G[(1 * 2 + 0) * nd] = -(2.0 * g1.x + g3.x);
    }
  }
}

// biharmonic direct sums: grid over targets, sources in LDS tiles
template <bool C, bool Q, int PG>
__global__ void __launch_bounds__(256) k_bh_direct(const double2* src, int ns, const double2* chg,
                                                   const double2* dip, const double2* trg, int nt,
                                                   int nd, double th, double2* vel, double2* grad) {
  __shared__ double2 sx[256];
  const int j = blockIdx.x * 256 + threadIdx.x;
  const bool act = j < nt;
  const double2 x = act ? trg[j] : make_double2(0, 0);
  for (int d = 0; d < nd; ++d) {
    cd v = mk(0, 0), g1 = mk(0, 0), g2 = mk(0, 0), g3 = mk(0, 0);
    for (int s0 = 0; s0 < ns; s0 += 256) {
      const int cnt = min(256, ns - s0);
      __syncthreads();
      if ((int)threadIdx.x < cnt) sx[threadIdx.x] = src[s0 + threadIdx.x];
      __syncthreads();
      if (act)
        for (int m = 0; m < cnt; ++m) {
          const cd z = mk(x.x - sx[m].x, x.y - sx[m].y);
          if (cabs_(z) <= th) continue;
          const size_t s = (size_t)(s0 + m);
          bh_pair<C, Q, PG>(z, C ? ldz(chg + (s * 2 + 0) * nd + d) : mk(0, 0),
                            C ? ldz(chg + (s * 2 + 1) * nd + d) : mk(0, 0),
                            Q ? ldz(dip + (s * 3 + 0) * nd + d) : mk(0, 0),
                            Q ? ldz(dip + (s * 3 + 1) * nd + d) : mk(0, 0),
                            Q ? ldz(dip + (s * 3 + 2) * nd + d) : mk(0, 0), v, g1, g2, g3);
        }
    }
    if (act) {
      addz(vel + (size_t)j * nd + d, v);
      if (PG >= 2) {
        addz(grad + ((size_t)j * 3 + 0) * nd + d, g1);
        addz(grad + ((size_t)j * 3 + 1) * nd + d, g2);
        addz(grad + ((size_t)j * 3 + 2) * nd + d, g3);
      }
    }
  }
}

template <bool C, bool Q, int PG>
int bh_direct(int nd, const double* sources, int ns, const double2* charges, const double2* dip,
              const double* targ, int nt, double2* vel, double2* grad, double th) {
  if (nd < 1 || ns < 0 || nt < 0) return IER_ARGS;
  if (ns == 0 || nt == 0) return IER_OK;
  Ctx c;
  int e = get_ctx(c);
  if (e) return e;
  Scratch S(c.st);
  double2* s = S.get<double2>(ns);
  double2* t = S.get<double2>(nt);
  double2* cs = C ? S.get<double2>(2 * (size_t)nd * ns) : nullptr;
  double2* qs = Q ? S.get<double2>(3 * (size_t)nd * ns) : nullptr;
  double2* v = S.get<double2>((size_t)nd * nt);
  double2* g = PG >= 2 ? S.get<double2>(3 * (size_t)nd * nt) : nullptr;
  if (S.err) return S.err;
  HIP_TRY(hipMemcpyAsync(s, sources, 16 * (size_t)ns, hipMemcpyHostToDevice, c.st));
  HIP_TRY(hipMemcpyAsync(t, targ, 16 * (size_t)nt, hipMemcpyHostToDevice, c.st));
  if (C) HIP_TRY(hipMemcpyAsync(cs, charges, 32 * (size_t)nd * ns, hipMemcpyHostToDevice, c.st));
  if (Q) HIP_TRY(hipMemcpyAsync(qs, dip, 48 * (size_t)nd * ns, hipMemcpyHostToDevice, c.st));
  HIP_TRY(hipMemcpyAsync(v, vel, 16 * (size_t)nd * nt, hipMemcpyHostToDevice, c.st));
  if (g) HIP_TRY(hipMemcpyAsync(g, grad, 48 * (size_t)nd * nt, hipMemcpyHostToDevice, c.st));
  k_bh_direct<C, Q, PG><<<nblkb(nt, 256), 256, 0, c.st>>>(s, ns, cs, qs, t, nt, nd, th, v, g);
  HIP_TRY(hipGetLastError());
  HIP_TRY(hipMemcpyAsync(vel, v, 16 * (size_t)nd * nt, hipMemcpyDeviceToHost, c.st));
  if (g) HIP_TRY(hipMemcpyAsync(grad, g, 48 * (size_t)nd * nt, hipMemcpyDeviceToHost, c.st));
  HIP_TRY(hipStreamSynchronize(c.st));
  return IER_OK;
}

int guarded_bh(const BhReq& r) {
  try {
    Ctx c;
    int e = get_ctx(c);
    if (e) return e;
    Scratch S(c.st);
    return run_bh(c, S, r);
  } catch (const std::bad_alloc&) {
    return IER_ALLOC;
  } catch (...) {
    return IER_HIP;
  }
}

// stfmm2d_: device prep, zsum, the biharmonic FMM on device arrays, unpack
int run_st(int nd, double eps, int ns, const double* source, int ifst, const double* stoklet,
           int ifsr, const double* strslet, const double* strsvec, int ifpp, double* pot,
           double* pre, double* grad, int nt, const double* targ, int ifppt, double* pott,
           double* pret, double* gradt) {
  if (nd < 1 || ns < 1 || nt < 0 || ifpp < 0 || ifpp > 3 || ifppt < 0 || ifppt > 3) return IER_ARGS;
  Ctx c;
  int e = get_ctx(c);
  if (e) return e;
  Scratch S(c.st);
  const size_t n2 = 2 * (size_t)nd * ns;
  double2* src = S.get<double2>(ns);
  double2* trg = S.get<double2>(std::max(nt, 1));
  double* f = S.get<double>(n2);
  double* m = S.get<double>(n2);
  double* nv = S.get<double>(n2);
  double2* chg = S.get<double2>(n2);
  double2* dip = S.get<double2>(3 * (size_t)nd * ns);
  double2* zsum = S.get<double2>(nd);
  double2* zpart = S.get<double2>((size_t)ZB * nd);
  const size_t ps = (size_t)nd * ns, pt = (size_t)nd * std::max(nt, 1);
  double2* vs = S.get<double2>(ps);
  double2* gs = S.get<double2>(3 * ps);
  double2* vt = S.get<double2>(pt);
  double2* gt = S.get<double2>(3 * pt);
  double* ups = S.get<double>(2 * ps);
  double* pps = S.get<double>(ps);
  double* ugs = S.get<double>(4 * ps);
  double* upt = S.get<double>(2 * pt);
  double* ppt = S.get<double>(pt);
  double* ugt = S.get<double>(4 * pt);
  if (S.err) return S.err;
  HIP_TRY(hipMemcpyAsync(src, source, 16 * (size_t)ns, hipMemcpyHostToDevice, c.st));
  if (nt) HIP_TRY(hipMemcpyAsync(trg, targ, 16 * (size_t)nt, hipMemcpyHostToDevice, c.st));
  if (ifst == 1) HIP_TRY(hipMemcpyAsync(f, stoklet, 8 * n2, hipMemcpyHostToDevice, c.st));
  if (ifsr == 1) {
    HIP_TRY(hipMemcpyAsync(m, strslet, 8 * n2, hipMemcpyHostToDevice, c.st));
    HIP_TRY(hipMemcpyAsync(nv, strsvec, 8 * n2, hipMemcpyHostToDevice, c.st));
  }
  k_st_prep<<<nblkb((long)ns * nd, 256), 256, 0, c.st>>>(f, m, nv, ns, nd, ifst, ifsr, chg, dip);
  k_st_zpart<<<dim3(ZB, nd), 256, 0, c.st>>>(chg, ns, nd, zpart);
  k_st_zsum<<<nd, 256, 0, c.st>>>(zpart, zsum);
  BhReq r;
  r.nd = nd;
  r.eps = eps;
  r.ns = ns;
  r.nt = nt;
  r.src = reinterpret_cast<const double*>(src);
  r.trg = reinterpret_cast<const double*>(trg);
  r.ifc = ifst == 1 ? 1 : 0;
  r.ifd = ifsr == 1 ? 1 : 0;
  r.chg = chg;
  r.dip = dip;
  r.pg = ifpp >= 2 ? 2 : ifpp;
  r.pgt = ifppt >= 2 ? 2 : ifppt;
  r.pot = vs;
  r.grad = gs;
  r.pott = vt;
  r.gradt = gt;
  r.device = true;
  HIP_TRY(hipMemsetAsync(vs, 0, sizeof(double2) * ps, c.st));
  HIP_TRY(hipMemsetAsync(gs, 0, sizeof(double2) * 3 * ps, c.st));
  HIP_TRY(hipMemsetAsync(vt, 0, sizeof(double2) * pt, c.st));
  HIP_TRY(hipMemsetAsync(gt, 0, sizeof(double2) * 3 * pt, c.st));
  e = run_bh(c, S, r);
  if (e) return e;
  if (ifpp >= 1) {
    k_st_unpack<<<nblkb((long)ps, 256), 256, 0, c.st>>>(vs, gs, zsum, r.ifc ? chg : nullptr, ns,
                                                         nd, ifpp, ups, pps, ugs);
    HIP_TRY(hipMemcpyAsync(pot, ups, 16 * ps, hipMemcpyDeviceToHost, c.st));
    if (ifpp >= 2) HIP_TRY(hipMemcpyAsync(pre, pps, 8 * ps, hipMemcpyDeviceToHost, c.st));
    if (ifpp >= 3) HIP_TRY(hipMemcpyAsync(grad, ugs, 32 * ps, hipMemcpyDeviceToHost, c.st));
  }
  if (ifppt >= 1 && nt > 0) {
    const size_t q = (size_t)nd * nt;
    k_st_unpack<<<nblkb((long)q, 256), 256, 0, c.st>>>(vt, gt, zsum, nullptr, nt, nd, ifppt, upt,
                                                        ppt, ugt);
    HIP_TRY(hipMemcpyAsync(pott, upt, 16 * q, hipMemcpyDeviceToHost, c.st));
    if (ifppt >= 2) HIP_TRY(hipMemcpyAsync(pret, ppt, 8 * q, hipMemcpyDeviceToHost, c.st));
    if (ifppt >= 3) HIP_TRY(hipMemcpyAsync(gradt, ugt, 32 * q, hipMemcpyDeviceToHost, c.st));
  }
  HIP_TRY(hipGetLastError());
  HIP_TRY(hipStreamSynchronize(c.st));
  return IER_OK;
}

}  // namespace
}  // namespace fmm

extern "C" {
void bhfmm2d_(const fmm2d_int* nd, const double* eps, const fmm2d_int* ns, const double* sources,
              const fmm2d_int* ifcharge, const fmm2d_cplx* charge, const fmm2d_int* ifdipole,
              const fmm2d_cplx* dip, fmm2d_int* iper, const fmm2d_int* ifpgh, fmm2d_cplx* pot,
              fmm2d_cplx* grad, fmm2d_cplx* hess, const fmm2d_int* nt, const double* targ,
              const fmm2d_int* ifpghtarg, fmm2d_cplx* pottarg, fmm2d_cplx* gradtarg,
              fmm2d_cplx* hesstarg, fmm2d_int* ier) {
  if (iper) *iper = 0;
  fmm::BhReq r;
  r.nd = *nd;
  r.eps = *eps;
  r.ns = *ns;
  r.nt = *nt;
  r.src = sources;
  r.trg = targ;
  r.ifc = *ifcharge;
  r.ifd = *ifdipole;
  r.chg = reinterpret_cast<const double2*>(charge);
  r.dip = reinterpret_cast<const double2*>(dip);
  r.pg = *ifpgh;
  r.pgt = *ifpghtarg;
  r.pot = reinterpret_cast<double2*>(pot);
  r.grad = reinterpret_cast<double2*>(grad);
  r.hess = reinterpret_cast<double2*>(hess);
  r.pott = reinterpret_cast<double2*>(pottarg);
  r.gradt = reinterpret_cast<double2*>(gradtarg);
  r.hesst = reinterpret_cast<double2*>(hesstarg);
  *ier = fmm::guarded_bh(r);
}

void stfmm2d_(const fmm2d_int* nd, const double* eps, const fmm2d_int* nsource,
              const double* source, const fmm2d_int* ifstoklet, const double* stoklet,
              const fmm2d_int* ifstrslet, const double* strslet, const double* strsvec,
              const fmm2d_int* ifppreg, double* pot, double* pre, double* grad,
              const fmm2d_int* ntarg, const double* targ, const fmm2d_int* ifppregtarg,
              double* pottarg, double* pretarg, double* gradtarg, fmm2d_int* ier) {
  try {
    *ier = fmm::run_st(*nd, *eps, *nsource, source, *ifstoklet, stoklet, *ifstrslet, strslet,
                       strsvec, *ifppreg, pot, pre, grad, *ntarg, targ, *ifppregtarg, pottarg,
                       pretarg, gradtarg);
  } catch (const std::bad_alloc&) {
    *ier = fmm::IER_ALLOC;
  } catch (...) {
    *ier = fmm::IER_HIP;
  }
}

void bh2dterms_(const double* eps, fmm2d_int* nterms, fmm2d_int* ier) {
  *ier = 0;
  *nterms = fmm::l2dterms(*eps);
}

void bhndiv2d_(const double* eps, const fmm2d_int* ns, const fmm2d_int* nt, const fmm2d_int*,
               const fmm2d_int*, const fmm2d_int*, const fmm2d_int*, fmm2d_int* ndiv,
               fmm2d_int* idivflag) {
  *idivflag = 0;
  *ndiv = fmm::bh_ndiv(*eps, *ns, *nt);
}

// bh2d_direct{c,d,cd}{p,g}_ (bhkernels2d.c:30-398): charges (nd,2,ns),
// dipoles (nd,3,ns), vel (nd,nt), grad (nd,3,nt), incremented
void bh2d_directcp_(const fmm2d_int* nd, const double* sources, const fmm2d_int* ns,
                    const fmm2d_cplx* charges, const double* targ, const fmm2d_int* nt,
                    fmm2d_cplx* vel, const double* thresh) {
  fmm::bh_direct<true, false, 1>(*nd, sources, *ns, reinterpret_cast<const double2*>(charges),
                                 nullptr, targ, *nt, reinterpret_cast<double2*>(vel), nullptr,
                                 *thresh);
}
void bh2d_directcg_(const fmm2d_int* nd, const double* sources, const fmm2d_int* ns,
                    const fmm2d_cplx* charges, const double* targ, const fmm2d_int* nt,
                    fmm2d_cplx* vel, fmm2d_cplx* grad, const double* thresh) {
  fmm::bh_direct<true, false, 2>(*nd, sources, *ns, reinterpret_cast<const double2*>(charges),
                                 nullptr, targ, *nt, reinterpret_cast<double2*>(vel),
                                 reinterpret_cast<double2*>(grad), *thresh);
}
void bh2d_directdp_(const fmm2d_int* nd, const double* sources, const fmm2d_int* ns,
                    const fmm2d_cplx* dippar, const double* targ, const fmm2d_int* nt,
                    fmm2d_cplx* vel, const double* thresh) {
  fmm::bh_direct<false, true, 1>(*nd, sources, *ns, nullptr,
                                 reinterpret_cast<const double2*>(dippar), targ, *nt,
                                 reinterpret_cast<double2*>(vel), nullptr, *thresh);
}
void bh2d_directdg_(const fmm2d_int* nd, const double* sources, const fmm2d_int* ns,
                    const fmm2d_cplx* dippar, const double* targ, const fmm2d_int* nt,
                    fmm2d_cplx* vel, fmm2d_cplx* grad, const double* thresh) {
  fmm::bh_direct<false, true, 2>(*nd, sources, *ns, nullptr,
                                 reinterpret_cast<const double2*>(dippar), targ, *nt,
                                 reinterpret_cast<double2*>(vel),
                                 reinterpret_cast<double2*>(grad), *thresh);
}
void bh2d_directcdp_(const fmm2d_int* nd, const double* sources, const fmm2d_int* ns,
                     const fmm2d_cplx* charges, const fmm2d_cplx* dippar, const double* targ,
                     const fmm2d_int* nt, fmm2d_cplx* vel, const double* thresh) {
  fmm::bh_direct<true, true, 1>(*nd, sources, *ns, reinterpret_cast<const double2*>(charges),
                                reinterpret_cast<const double2*>(dippar), targ, *nt,
                                reinterpret_cast<double2*>(vel), nullptr, *thresh);
}
void bh2d_directcdg_(const fmm2d_int* nd, const double* sources, const fmm2d_int* ns,
                     const fmm2d_cplx* charges, const fmm2d_cplx* dippar, const double* targ,
                     const fmm2d_int* nt, fmm2d_cplx* vel, fmm2d_cplx* grad,
                     const double* thresh) {
  fmm::bh_direct<true, true, 2>(*nd, sources, *ns, reinterpret_cast<const double2*>(charges),
                                reinterpret_cast<const double2*>(dippar), targ, *nt,
                                reinterpret_cast<double2*>(vel),
                                reinterpret_cast<double2*>(grad), *thresh);
}
}

// translate.h -- dense expansion translations as small batched GEMMs.
//
// Every translation in the FMM (M2M, M2L, L2L, Laplace or Helmholtz) is
//     Y[b] += sum_s  T_s  X[src(b, s)]
// with a handful of operator "slots" s that depend only on the relative
// position of the two boxes (4 child quadrants, 49 offsets of list 2).
// Blocks own a tile of TB target boxes and walk the slots: the slot's
// operator is staged once in LDS and applied to the whole tile, so each
// operator is read from HBM/L2 once per tile instead of once per box pair.
// Operators are stored T[s][j][i] (j = input coefficient, i = output).
#pragma once

#include "common.h"
#include "engine.h"

#include <algorithm>
#include <vector>

namespace fmm {
// every translation unit gets its own copy of these kernels and launcher
namespace {

constexpr int XT_THREADS = 256;

// map[w * nslot + s]: source box of work item w in slot s, or -1.
// Y/X are [box][channel][coef] with per-box strides ystride/xstride
// (in double2 units), channel stride nout/nin.
// lap_log: Laplace M2L adds x_0 * log(bs) to output 0 per source (the only
// level-dependent entry of the otherwise scale-free operator).
template <int TB>
__global__ void __launch_bounds__(XT_THREADS)
k_translate(const int* __restrict__ work, int nwork, const int* __restrict__ map, int nslot,
            const double2* __restrict__ T, int nin, int nout, const double2* __restrict__ X,
            size_t xstride, double2* __restrict__ Y, size_t ystride, int nd, double lap_log) {
  extern __shared__ double2 lds[];
  double2* sT = lds;                           // nin * nout
  double2* sX = lds + (size_t)nin * nout;      // TB * nin
  __shared__ int ssrc[TB];
  __shared__ int sany;
  const int w0 = blockIdx.x * TB;
  const int nb = min(TB, nwork - w0);
  const int npair = nb * nout;
  constexpr int MAXP = (TB * 128 + XT_THREADS - 1) / XT_THREADS;  // nout <= 128
  for (int d = 0; d < nd; ++d) {
    cd acc[MAXP];
#pragma unroll
    for (int k = 0; k < MAXP; ++k) acc[k] = mk(0, 0);
    for (int s = 0; s < nslot; ++s) {
      if (threadIdx.x == 0) sany = 0;
      __syncthreads();
      if ((int)threadIdx.x < TB) {
        int src = (int)threadIdx.x < nb ? map[(size_t)(w0 + threadIdx.x) * nslot + s] : -1;
        ssrc[threadIdx.x] = src;
        if (src >= 0) atomicOr(&sany, 1);
      }
      __syncthreads();
      if (!sany) continue;  // block-uniform
      for (int e = threadIdx.x; e < nin * nout; e += XT_THREADS) sT[e] = T[(size_t)s * nin * nout + e];
      for (int e = threadIdx.x; e < TB * nin; e += XT_THREADS) {
        int b = e / nin, j = e - b * nin;
        int src = ssrc[b];
        sX[e] = src >= 0 ? X[(size_t)src * xstride + (size_t)d * nin + j] : make_double2(0, 0);
      }
      __syncthreads();
#pragma unroll
      for (int k = 0; k < MAXP; ++k) {
        int q = threadIdx.x + k * XT_THREADS;
        if (q < npair) {
          int b = q / nout, i = q - b * nout;
          if (ssrc[b] >= 0) {
            const double2* xb = sX + b * nin;
            double ar = 0, ai = 0;
            for (int j = 0; j < nin; ++j) {
              double2 t = sT[j * nout + i], x = xb[j];
              ar += t.x * x.x - t.y * x.y;
              ai += t.x * x.y + t.y * x.x;
            }
            if (lap_log != 0.0 && i == 0) {
              ar += xb[0].x * lap_log;
              ai += xb[0].y * lap_log;
            }
            acc[k] += mk(ar, ai);
          }
        }
      }
      __syncthreads();
    }
#pragma unroll
    for (int k = 0; k < MAXP; ++k) {
      int q = threadIdx.x + k * XT_THREADS;
      if (q < npair) {
        int b = q / nout, i = q - b * nout;
        double2* y = Y + (size_t)work[w0 + b] * ystride + (size_t)d * nout + i;
        *y = make_double2(y->x + acc[k].x, y->y + acc[k].y);
      }
    }
  }
}

// list-2 slot of a box pair: (dy+3)*7 + (dx+3) from centre offsets
__device__ inline int m2l_slot(double2 cs, double2 ct, double bs) {
  int dx = (int)rint((cs.x - ct.x) / bs), dy = (int)rint((cs.y - ct.y) / bs);
  if (dx < -3 || dx > 3 || dy < -3 || dy > 3) return -1;
  return (dy + 3) * 7 + (dx + 3);
}

// maps for the translation kinds
// M2L slot maps of a whole M2L work list (every level) in one launch: row e
// of map (49 entries, filled with -1 beforehand) gets, at each list-2
// offset of work item e whose box holds sources, that box.  One thread per
// (work item, list-2 entry), so the entry, source-range and centre loads of
// one item run in parallel; list 2 holds at most 27 boxes.
__global__ void k_map_m2l_all(const int* work, int nwork, DTree D, int* map) {
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  const int e = (int)(t >> 5), k = (int)(t & 31);
  if (e >= nwork) return;
  const int b = work[e], k0 = D.o2[b];
  if (k >= D.o2[b + 1] - k0) return;
  const int j = D.l2[k0 + k];
  if (D.srng[j].y <= D.srng[j].x) return;
  const int s = m2l_slot(D.ctr[j], D.ctr[b], D.bsb[b]);
  if (s >= 0) map[(size_t)e * 49 + s] = j;
}

// M2M: slot q = child quadrant, children with sources only
__global__ void k_map_m2m(const int* work, int nwork, const int4* child, const int2* srng,
                          int* map) {
  int w = blockIdx.x * blockDim.x + threadIdx.x;
  if (w >= nwork) return;
  int4 c = child[work[w]];
  int k[4] = {c.x, c.y, c.z, c.w};
  for (int q = 0; q < 4; ++q) {
    int s = k[q];
    map[(size_t)w * 4 + q] = (s >= 0 && srng[s].y > srng[s].x) ? s : -1;
  }
}

// L2L: work items are children; slot = the child's quadrant in its parent
__global__ void k_map_l2l(const int* work, int nwork, const int* parent, const int4* child,
                          int* map) {
  int w = blockIdx.x * blockDim.x + threadIdx.x;
  if (w >= nwork) return;
  int c = work[w], p = parent[c];
  int4 k = child[p];
  map[(size_t)w * 4 + 0] = k.x == c ? p : -1;
  map[(size_t)w * 4 + 1] = k.y == c ? p : -1;
  map[(size_t)w * 4 + 2] = k.z == c ? p : -1;
  map[(size_t)w * 4 + 3] = k.w == c ? p : -1;
}

// Operators too large for LDS (coarse Helmholtz levels): one block per
// target, operator read from global memory.
__global__ void __launch_bounds__(XT_THREADS)
k_translate_direct(const int* __restrict__ work, const int* __restrict__ map, int nslot,
                   const double2* __restrict__ T, int nin, int nout, const double2* __restrict__ X,
                   size_t xstride, double2* __restrict__ Y, size_t ystride, int nd, double lap_log) {
  const int w = blockIdx.x, b = work[w];
  for (int d = 0; d < nd; ++d)
    for (int i = threadIdx.x; i < nout; i += XT_THREADS) {
      double ar = 0, ai = 0;
      for (int s = 0; s < nslot; ++s) {
        int src = map[(size_t)w * nslot + s];
        if (src < 0) continue;
        const double2* x = X + (size_t)src * xstride + (size_t)d * nin;
        const double2* Ts = T + (size_t)s * nin * nout;
        for (int j = 0; j < nin; ++j) {
          double2 t = Ts[(size_t)j * nout + i], v = x[j];
          ar += t.x * v.x - t.y * v.y;
          ai += t.x * v.y + t.y * v.x;
        }
        if (lap_log != 0.0 && i == 0) {
          ar += x[0].x * lap_log;
          ai += x[0].y * lap_log;
        }
      }
      double2* y = Y + (size_t)b * ystride + (size_t)d * nout + i;
      *y = make_double2(y->x + ar, y->y + ai);
    }
}

// ---------------------------------------------------------------------
// Translations on the FP64 matrix cores.  For one level,
//   Y[work[w]] += sum_s T_s X[map[w][s]],
// T_s the nin x nout complex operator of slot s (list-2 offset or child
// quadrant), stored T[s][j][i] (j input, i output).  Complex products as
// four real 16x16x4 MFMAs: Yr += Tr Xr + Ti (-Xi), Yi += Tr Xi + Ti Xr.
// A wave owns 16 work items (the N columns) and MW row tiles of 16 outputs;
// lane l holds A[l%16][l/16] = T_s[k0 + l/16][16t + l%16] and
// B[l/16][l%16] = X[k0 + l/16] of column l%16, and accumulates
// D[4v + l/16][l%16] (v = 0..3) -- the v_mfma_f64_16x16x4f64 operand layout.
// A slot is skipped when none of the wave's 16 columns has a source there,
// so tiles of work items that share a child quadrant waste no work: the
// wave stages its 16 map rows in LDS and takes the set of used slots from
// one ballot.  The operands of step k0 + 4 are loaded before the matrix
// operations of step k0, and the last step of a slot loads the next used
// slot's first operands.  nslot <= 64.  Grid:
// x = column tiles (4 per block), y = row-tile groups, z = slot groups
// (slots s = z, z + G, ...).  With G > 1 each slot group writes its partial
// sums to part[(g nd + d) nwork + w][i] and k_xlate_reduce adds them up in
// a fixed order; with G = 1 the wave adds into Y directly.  lap_log as in
// k_translate (nin == nout).
// ---------------------------------------------------------------------
typedef double mfma_d4 __attribute__((ext_vector_type(4)));

inline hipError_t launch_translate(hipStream_t st, const int* work, int nwork, const int* map,
                                   int nslot, const double2* T, int nin, int nout,
                                   const double2* X, size_t xs, double2* Y, size_t ys, int nd,
                                   double lap_log);

// One block of k_xlate_mfma: column-tile block bx, row-tile group by, slot
// group g.
template <int MW>
__device__ __forceinline__ void xlate_block(
    int bx, int by, int g, const int* __restrict__ work, int nwork, const int* __restrict__ map,
    int nslot, const double2* __restrict__ T, int nin, int nout, const double2* __restrict__ X,
    size_t xstride, double2* __restrict__ Y, size_t ystride, int nd, double lap_log, int G,
    double2* __restrict__ part, const double* __restrict__ lap_bsb = nullptr) {
  __shared__ int smap[4][16 * 64];
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const int n = lane & 15, kq = lane >> 4;
  const int tile = bx * 4 + wv;
  const int w = tile * 16 + n;
  const bool wok = w < nwork;
  const int rt0 = by * MW;
  int* sm = smap[wv];
  // per-target log(box size) when box sizes differ within a level (forests)
  if (lap_bsb && wok) lap_log = log(lap_bsb[work[w]]);
  {
    const size_t base = (size_t)tile * 16 * nslot, end = (size_t)nwork * nslot;
    for (int e = lane; e < 16 * nslot; e += 64) sm[e] = base + e < end ? map[base + e] : -1;
  }
  __syncthreads();
  bool used = false;
  if (lane < nslot && lane % G == g)
    for (int c = 0; c < 16; ++c) used |= sm[c * nslot + lane] >= 0;
  const uint64_t act = __ballot(used);
  for (int d = 0; d < nd; ++d) {
    mfma_d4 k1[MW], k2[MW], k3[MW];
#pragma unroll
    for (int t = 0; t < MW; ++t) {
      k1[t] = mfma_d4{0.0, 0.0, 0.0, 0.0};
      k2[t] = mfma_d4{0.0, 0.0, 0.0, 0.0};
      k3[t] = mfma_d4{0.0, 0.0, 0.0, 0.0};
    }
    double s0r = 0.0, s0i = 0.0;  // sum of x_0 over the sources (lanes kq == 0)
    int irow[MW];
    bool rok[MW];
#pragma unroll
    for (int t = 0; t < MW; ++t) {
      const int i = 16 * (rt0 + t) + n;
      rok[t] = i < nout;
      irow[t] = rok[t] ? i : nout - 1;
    }
    const int j0 = kq < nin ? kq : nin - 1;
    uint64_t rem = act;
    int s = rem ? __builtin_ctzll(rem) : 0;
    int src = sm[n * nslot + s];
    const double2* xs = X + (size_t)(src >= 0 ? src : 0) * xstride + (size_t)d * nin;
    const double2* Ts = T + (size_t)s * nin * nout;
    double2 xl = xs[j0];
    double2 al[MW];
#pragma unroll
    for (int t = 0; t < MW; ++t) al[t] = Ts[(size_t)j0 * nout + irow[t]];
    while (rem) {
      rem &= rem - 1;
      const bool sok = src >= 0;
      if (sok && kq == 0) {
        const double2 x0 = xs[0];
        s0r += x0.x;
        s0i += x0.y;
      }
      // the next used slot (its first operands load during this slot's last step)
      const int sn = rem ? __builtin_ctzll(rem) : s;
      const int srcn = sm[n * nslot + sn];
      const double2* xsn = X + (size_t)(srcn >= 0 ? srcn : 0) * xstride + (size_t)d * nin;
      const double2* Tsn = T + (size_t)sn * nin * nout;
      // Branch-free body: indexes are clamped into range and out-of-range
      // operands zeroed by selects, so the accumulators stay in one register
      // class across the loop.
      for (int k0 = 0; k0 < nin; k0 += 4) {
        const bool jok = k0 + kq < nin;
        const bool xok = sok && jok;
        const double2 x = make_double2(xok ? xl.x : 0.0, xok ? xl.y : 0.0);
        const double b1 = x.x, b2 = x.y - x.x, b3 = x.x + x.y;
        double ar[MW], ai[MW], as[MW];
#pragma unroll
        for (int t = 0; t < MW; ++t) {
          const bool aok = jok && rok[t];
          ar[t] = aok ? al[t].x : 0.0;
          ai[t] = aok ? al[t].y : 0.0;
          as[t] = ar[t] + ai[t];
        }
        if (k0 + 4 < nin) {
          const int jn = k0 + 4 + kq < nin ? k0 + 4 + kq : nin - 1;
          xl = xs[jn];
#pragma unroll
          for (int t = 0; t < MW; ++t) al[t] = Ts[(size_t)jn * nout + irow[t]];
        } else {
          xl = xsn[j0];
#pragma unroll
          for (int t = 0; t < MW; ++t) al[t] = Tsn[(size_t)j0 * nout + irow[t]];
        }
        __builtin_amdgcn_sched_barrier(0);  // keep the next step's loads ahead
#pragma unroll
        for (int t = 0; t < MW; ++t) {
          k1[t] = __builtin_amdgcn_mfma_f64_16x16x4f64(as[t], b1, k1[t], 0, 0, 0);
          k2[t] = __builtin_amdgcn_mfma_f64_16x16x4f64(ar[t], b2, k2[t], 0, 0, 0);
          k3[t] = __builtin_amdgcn_mfma_f64_16x16x4f64(ai[t], b3, k3[t], 0, 0, 0);
        }
      }
      s = sn;
      src = srcn;
      xs = xsn;
      Ts = Tsn;
    }
    if (!wok) continue;
    double2* y = G == 1 ? Y + (size_t)work[w] * ystride + (size_t)d * nout
                        : part + ((size_t)(g * nd + d) * nwork + w) * nout;
#pragma unroll
    for (int t = 0; t < MW; ++t)
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const int i = 16 * (rt0 + t) + 4 * v + kq;
        if (i >= nout) continue;
        double re = k1[t][v] - k3[t][v], im = k1[t][v] + k2[t][v];
        if (i == 0 && lap_log != 0.0) {
          re += s0r * lap_log;
          im += s0i * lap_log;
        }
        if (G == 1) y[i] = make_double2(y[i].x + re, y[i].y + im);
        else y[i] = make_double2(re, im);
      }
  }
}

template <int MW>
__global__ void __launch_bounds__(256)
k_xlate_mfma(const int* __restrict__ work, int nwork, const int* __restrict__ map, int nslot,
             const double2* __restrict__ T, int nin, int nout, const double2* __restrict__ X,
             size_t xstride, double2* __restrict__ Y, size_t ystride, int nd, double lap_log,
             int G, double2* __restrict__ part) {
  xlate_block<MW>(blockIdx.x, blockIdx.y, blockIdx.z, work, nwork, map, nslot, T, nin, nout, X,
                  xstride, Y, ystride, nd, lap_log, G, part);
}

// Several M2L levels in one launch (they read and write disjoint expansions):
// entry j -- a level, or one of its row-tile groups -- owns the blocks
// [block0, next entry's block0), a wave per 16 targets with MW row tiles
// from row tile by * MW, and every slot.
struct XLevel {
  const int* work;
  int nwork;
  const int* map;
  const double2* T;
  int nin, nout;
  double lap_log;
  const double* lap_bsb;  // per-box sizes for log(box size), or null (lap_log)
  int block0;
  int by;
  int G;           // slot groups: blocks block0 + G b + g, partial sums in part
  double2* part;   // G > 1: the level's partial sums (k_xlate_reduce adds them)
};
constexpr int kMaxXLevels = 30;
struct XLevels {
  XLevel lv[kMaxXLevels];
  int n;
};
template <int MW>
__global__ void __launch_bounds__(256)
k_xlate_mfma_levels(XLevels L, int nslot, const double2* __restrict__ X, size_t xstride,
                    double2* __restrict__ Y, size_t ystride, int nd) {
  int j = 0;
  while (j + 1 < L.n && (int)blockIdx.x >= L.lv[j + 1].block0) ++j;
  const XLevel& v = L.lv[j];
  const int local = blockIdx.x - v.block0;
  xlate_block<MW>(local / v.G, v.by, local % v.G, v.work, v.nwork, v.map, nslot, v.T, v.nin,
                  v.nout, X, xstride, Y, ystride, nd, v.lap_log, v.G, v.part, v.lap_bsb);
}

// Y[work[w]] += sum over slot groups of the partials, group 0 first
__global__ void k_xlate_reduce(const int* __restrict__ work, int nwork, int nout, int nd, int G,
                               const double2* __restrict__ part, double2* __restrict__ Y,
                               size_t ystride) {
  const size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  const size_t per = (size_t)nwork * nout;
  if (e >= per * nd) return;
  const int d = (int)(e / per);
  const size_t r = e - (size_t)d * per;
  const int w = (int)(r / nout), i = (int)(r - (size_t)w * nout);
  double re = 0.0, im = 0.0;
  for (int g = 0; g < G; ++g) {
    const double2 v = part[((size_t)(g * nd + d) * nwork + w) * nout + i];
    re += v.x;
    im += v.y;
  }
  double2* y = Y + (size_t)work[w] * ystride + (size_t)d * nout + i;
  *y = make_double2(y->x + re, y->y + im);
}

// ---------------------------------------------------------------------
// FP32 variant of k_xlate_mfma for the mixed-precision M2L: the operator is
// stored in FP32, the input expansion is rounded to FP32 as it is loaded,
// each slot's contraction (at most nin products per output) runs on the
// FP32 matrix cores, and the per-slot results are summed in FP64 before
// they are added into the FP64 output.  v_mfma_f32_16x16x4f32 operand
// layout: A[l%16][l/16], B[l/16][l%16], D[4(l/16) + v][l%16].
// ---------------------------------------------------------------------
typedef float mfma_f4 __attribute__((ext_vector_type(4)));

// One block of k_xlate_mfma32 (column-tile block bx, row-tile group by,
// slot group g); the wave's map rows are staged in LDS and the used slots
// taken from one ballot, as in xlate_block.
template <int MW>
__device__ __forceinline__ void xlate32_block(
    int bx, int by, int g, const int* __restrict__ work, int nwork, const int* __restrict__ map,
    int nslot, const float2* __restrict__ T, int nin, int nout, const double2* __restrict__ X,
    size_t xstride, double2* __restrict__ Y, size_t ystride, int nd, double lap_log, int G,
    double2* __restrict__ part, const double* __restrict__ lap_bsb = nullptr) {
  __shared__ int smap[4][16 * 64];
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const int n = lane & 15, kq = lane >> 4;
  const int tile = bx * 4 + wv;
  const int w = tile * 16 + n;
  const bool wok = w < nwork;
  const int rt0 = by * MW;
  int* sm = smap[wv];
  // per-target log(box size) when box sizes differ within a level (forests)
  if (lap_bsb && wok) lap_log = log(lap_bsb[work[w]]);
  {
    const size_t base = (size_t)tile * 16 * nslot, end = (size_t)nwork * nslot;
    for (int e = lane; e < 16 * nslot; e += 64) sm[e] = base + e < end ? map[base + e] : -1;
  }
  __syncthreads();
  bool used = false;
  if (lane < nslot && lane % G == g)
    for (int c = 0; c < 16; ++c) used |= sm[c * nslot + lane] >= 0;
  const uint64_t act = __ballot(used);
  for (int d = 0; d < nd; ++d) {
    double yr[MW][4], yi[MW][4];
#pragma unroll
    for (int t = 0; t < MW; ++t)
#pragma unroll
      for (int v = 0; v < 4; ++v) yr[t][v] = yi[t][v] = 0.0;
    double s0r = 0.0, s0i = 0.0;  // sum of x_0 over the sources (lanes kq == 0)
    for (uint64_t rem = act; rem; rem &= rem - 1) {
      const int s = __builtin_ctzll(rem);
      const int src = sm[n * nslot + s];
      const bool sok = src >= 0;
      const double2* xs = X + (size_t)(sok ? src : 0) * xstride + (size_t)d * nin;
      const float2* Ts = T + (size_t)s * nin * nout;
      if (sok && kq == 0) {
        const double2 x0 = xs[0];
        s0r += x0.x;
        s0i += x0.y;
      }
      mfma_f4 cr[MW], ci[MW];
#pragma unroll
      for (int t = 0; t < MW; ++t) {
        cr[t] = mfma_f4{0.0f, 0.0f, 0.0f, 0.0f};
        ci[t] = mfma_f4{0.0f, 0.0f, 0.0f, 0.0f};
      }
      int irow[MW];
      bool rok[MW];
#pragma unroll
      for (int t = 0; t < MW; ++t) {
        const int i = 16 * (rt0 + t) + n;
        rok[t] = i < nout;
        irow[t] = rok[t] ? i : nout - 1;
      }
      double2 xl = xs[kq < nin ? kq : nin - 1];
      float2 al[MW];
#pragma unroll
      for (int t = 0; t < MW; ++t) al[t] = Ts[(size_t)(kq < nin ? kq : nin - 1) * nout + irow[t]];
      for (int k0 = 0; k0 < nin; k0 += 4) {
        const bool jok = k0 + kq < nin;
        const bool xok = sok && jok;
        const float xr = xok ? (float)xl.x : 0.0f, xi = xok ? (float)xl.y : 0.0f;
        const float xin = -xi;
        float ar[MW], ai[MW];
#pragma unroll
        for (int t = 0; t < MW; ++t) {
          const bool aok = jok && rok[t];
          ar[t] = aok ? al[t].x : 0.0f;
          ai[t] = aok ? al[t].y : 0.0f;
        }
        const int jn = k0 + 4 + kq < nin ? k0 + 4 + kq : nin - 1;
        xl = xs[jn];
#pragma unroll
        for (int t = 0; t < MW; ++t) al[t] = Ts[(size_t)jn * nout + irow[t]];
        __builtin_amdgcn_sched_barrier(0);  // keep the next step's loads ahead
#pragma unroll
        for (int t = 0; t < MW; ++t) {
          cr[t] = __builtin_amdgcn_mfma_f32_16x16x4f32(ar[t], xr, cr[t], 0, 0, 0);
          ci[t] = __builtin_amdgcn_mfma_f32_16x16x4f32(ar[t], xi, ci[t], 0, 0, 0);
        }
#pragma unroll
        for (int t = 0; t < MW; ++t) {
          cr[t] = __builtin_amdgcn_mfma_f32_16x16x4f32(ai[t], xin, cr[t], 0, 0, 0);
          ci[t] = __builtin_amdgcn_mfma_f32_16x16x4f32(ai[t], xr, ci[t], 0, 0, 0);
        }
      }
#pragma unroll
      for (int t = 0; t < MW; ++t)
#pragma unroll
        for (int v = 0; v < 4; ++v) {
          yr[t][v] += (double)cr[t][v];
          yi[t][v] += (double)ci[t][v];
        }
    }
    if (!wok) continue;
    double2* y = G == 1 ? Y + (size_t)work[w] * ystride + (size_t)d * nout
                        : part + ((size_t)(g * nd + d) * nwork + w) * nout;
#pragma unroll
    for (int t = 0; t < MW; ++t)
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const int i = 16 * (rt0 + t) + 4 * kq + v;
        if (i >= nout) continue;
        double re = yr[t][v], im = yi[t][v];
        if (i == 0 && lap_log != 0.0) {
          // x_0 sums sit in the kq == 0 lanes, which own output row 0
          re += s0r * lap_log;
          im += s0i * lap_log;
        }
        if (G == 1) y[i] = make_double2(y[i].x + re, y[i].y + im);
        else y[i] = make_double2(re, im);
      }
  }
}

template <int MW>
__global__ void __launch_bounds__(256)
k_xlate_mfma32(const int* __restrict__ work, int nwork, const int* __restrict__ map, int nslot,
               const float2* __restrict__ T, int nin, int nout, const double2* __restrict__ X,
               size_t xstride, double2* __restrict__ Y, size_t ystride, int nd, double lap_log,
               int G, double2* __restrict__ part) {
  xlate32_block<MW>(blockIdx.x, blockIdx.y, blockIdx.z, work, nwork, map, nslot, T, nin, nout, X,
                    xstride, Y, ystride, nd, lap_log, G, part);
}

// Several M2L levels in one FP32 launch (as k_xlate_mfma_levels).
struct XLevel32 {
  const int* work;
  int nwork;
  const int* map;
  const float2* T;
  int nin, nout;
  double lap_log;
  const double* lap_bsb;
  int block0;
  int by;
  int G;
  double2* part;
};
struct XLevels32 {
  XLevel32 lv[kMaxXLevels];
  int n;
};
template <int MW>
__global__ void __launch_bounds__(256)
k_xlate_mfma32_levels(XLevels32 L, int nslot, const double2* __restrict__ X, size_t xstride,
                      double2* __restrict__ Y, size_t ystride, int nd) {
  int j = 0;
  while (j + 1 < L.n && (int)blockIdx.x >= L.lv[j + 1].block0) ++j;
  const XLevel32& v = L.lv[j];
  const int local = blockIdx.x - v.block0;
  xlate32_block<MW>(local / v.G, v.by, local % v.G, v.work, v.nwork, v.map, nslot, v.T, v.nin,
                    v.nout, X, xstride, Y, ystride, nd, v.lap_log, v.G, v.part, v.lap_bsb);
}

// Translation through k_xlate_mfma when the output order fits eight row
// tiles (nout <= 128), else k_translate.  Rows and slots are split over waves to keep about 4096
// waves in flight whatever the level's size; maxg bounds the slot groups
// (the number of slots a work item can use: 27 for M2L, 4 for M2M, 1 for
// L2L).  S provides the partial-sum buffer.
inline hipError_t launch_xlate(hipStream_t st, Scratch& S, const int* work, int nwork,
                               const int* map, int nslot, int maxg, const double2* T, int nin,
                               int nout, const double2* X, size_t xs, double2* Y, size_t ys,
                               int nd, double lap_log) {
  if (nwork <= 0) return hipSuccess;
  const int MT = (nout + 15) / 16;
  if (MT > 8 || nslot > 64)
    return launch_translate(st, work, nwork, map, nslot, T, nin, nout, X, xs, Y, ys, nd, lap_log);
  const int ntile = (nwork + 15) / 16, nb = (ntile + 3) / 4;
  const int kWaves = xlate_waves();
  int mw = MT <= 4 ? MT : (MT + 1) / 2, G = 1;
  if (ntile * MT < kWaves) {  // small level: one row tile per wave and split the slots
    mw = 1;
    const int per_g = ntile * MT;
    G = std::min(std::min(maxg, nslot), std::max(1, (kWaves + per_g - 1) / per_g));
  } else if (ntile < kWaves) {
    mw = MT % 2 == 0 ? 2 : 1;  // no wave on padding rows only
  }
  double2* part = nullptr;
  if (G > 1) {
    part = S.get<double2>((size_t)G * nd * nwork * nout);
    if (!part) return hipErrorOutOfMemory;
  }
  const dim3 grid(nb, (MT + mw - 1) / mw, G);
#define FMM_XL(MWV)                                                                       \
  k_xlate_mfma<MWV><<<grid, 256, 0, st>>>(work, nwork, map, nslot, T, nin, nout, X, xs, Y, ys, \
                                          nd, lap_log, G, part)
  switch (mw) {
    case 1: FMM_XL(1); break;
    case 2: FMM_XL(2); break;
    case 3: FMM_XL(3); break;
    default: FMM_XL(4); break;
  }
#undef FMM_XL
  if (G > 1) {
    const size_t ne = (size_t)nd * nwork * nout;
    k_xlate_reduce<<<(unsigned)((ne + 255) / 256), 256, 0, st>>>(work, nwork, nout, nd, G, part, Y,
                                                                ys);
  }
  return hipGetLastError();
}


// M2L of several levels.  Levels whose order fits eight MFMA row tiles
// (P <= 128; beyond four, two row groups) are grouped by row tiles per wave,
// FP64 and FP32 (mixed precision) apart, and each group runs as one
// k_xlate_mfma_levels launch, so the small levels share the GPU with the
// large ones instead of running alone; larger orders go through
// launch_translate level by level.
struct M2LLevel {
  const int* work;
  int nwork;
  const int* map;       // 49 slots per work item
  const double2* T;     // the level's 49 operators, P x P each
  int P;
  double lap_log;
  const float2* T32;    // the same in FP32 (mixed precision), or null
  // Laplace: per-box sizes whose log replaces lap_log (a forest's levels
  // hold boxes of different sizes), or null
  const double* lap_bsb = nullptr;
  int family = 0;  // for the tile choice: 0 Laplace and biharmonic, 1 Helmholtz
};
// With S, a launch group whose waves would not fill the GPU once (small
// trees: a few long waves would set the time) splits every level's slots
// over G groups of blocks, whose partial sums k_xlate_reduce then adds,
// group 0 first.
inline hipError_t launch_m2l_levels(hipStream_t st, const std::vector<M2LLevel>& lv,
                                    const double2* X, size_t xs, double2* Y, size_t ys, int nd,
                                    Scratch* S = nullptr) {
  constexpr int kFill = 2048;  // waves that keep every SIMD busy twice over
  // row tiles per wave: all of them up to 4, else two groups of ceil(MT/2)
  // (or the autotuner's choice, m2l_tiles)
  for (int mw = 1; mw <= 4; ++mw) {
    XLevels L;
    XLevels32 L32;
    L.n = L32.n = 0;
    int nb = 0, nb32 = 0;
    // waves of this launch group without a slot split
    long waves = 0;
    for (const M2LLevel& v : lv) {
      const int MT = (v.P + 15) / 16;
      if (v.nwork <= 0 || MT > 8 || m2l_tiles(MT, v.T32 != nullptr, v.family) != mw) continue;
      waves += 4L * ((v.nwork + 15) / 16 + 3) / 4 * ((MT + mw - 1) / mw);
    }
    const int G = S && waves > 0 && waves < kFill ? (int)std::min<long>(8, (kFill + waves - 1) / waves)
                                                  : 1;
    struct Red {
      const int* work;
      int nwork, nout;
      double2* part;
    };
    std::vector<Red> reds;
    auto flush = [&]() {
      if (L.n) switch (mw) {
        case 1: k_xlate_mfma_levels<1><<<nb, 256, 0, st>>>(L, 49, X, xs, Y, ys, nd); break;
        case 2: k_xlate_mfma_levels<2><<<nb, 256, 0, st>>>(L, 49, X, xs, Y, ys, nd); break;
        case 3: k_xlate_mfma_levels<3><<<nb, 256, 0, st>>>(L, 49, X, xs, Y, ys, nd); break;
        default: k_xlate_mfma_levels<4><<<nb, 256, 0, st>>>(L, 49, X, xs, Y, ys, nd); break;
      }
      if (L32.n) switch (mw) {
        case 1: k_xlate_mfma32_levels<1><<<nb32, 256, 0, st>>>(L32, 49, X, xs, Y, ys, nd); break;
        case 2: k_xlate_mfma32_levels<2><<<nb32, 256, 0, st>>>(L32, 49, X, xs, Y, ys, nd); break;
        case 3: k_xlate_mfma32_levels<3><<<nb32, 256, 0, st>>>(L32, 49, X, xs, Y, ys, nd); break;
        default: k_xlate_mfma32_levels<4><<<nb32, 256, 0, st>>>(L32, 49, X, xs, Y, ys, nd); break;
      }
      L.n = L32.n = 0;
      nb = nb32 = 0;
    };
    for (const M2LLevel& v : lv) {
      const int MT = (v.P + 15) / 16;
      const bool f32 = v.T32 != nullptr;
      if (v.nwork <= 0 || MT > 8 || m2l_tiles(MT, f32, v.family) != mw) continue;
      last_stats().m2l_mt_mask |= uint64_t(1) << ((v.family * 2 + (f32 ? 1 : 0)) * 9 + MT);
      const int nbl = ((v.nwork + 15) / 16 + 3) / 4 * G;
      double2* part = nullptr;
      if (G > 1) {  // shared by the level's row groups (disjoint rows)
        part = S->get<double2>((size_t)G * nd * v.nwork * v.P);
        if (!part) return hipErrorOutOfMemory;
        reds.push_back(Red{v.work, v.nwork, v.P, part});
      }
      for (int by = 0; by * mw < MT; ++by) {
        if (L.n == kMaxXLevels || L32.n == kMaxXLevels) flush();
        if (v.T32) {
          L32.lv[L32.n++] = XLevel32{v.work, v.nwork, v.map, v.T32, v.P, v.P, v.lap_log,
                                     v.lap_bsb, nb32, by, G, part};
          nb32 += nbl;
        } else {
          L.lv[L.n++] = XLevel{v.work, v.nwork, v.map, v.T, v.P, v.P, v.lap_log, v.lap_bsb, nb, by,
                               G, part};
          nb += nbl;
        }
      }
    }
    flush();
    for (const Red& r : reds) {
      const size_t ne = (size_t)nd * r.nwork * r.nout;
      k_xlate_reduce<<<(unsigned)((ne + 255) / 256), 256, 0, st>>>(r.work, r.nwork, r.nout, nd, G,
                                                                  r.part, Y, ys);
    }
  }
  for (const M2LLevel& v : lv) {
    if (v.nwork <= 0 || (v.P + 15) / 16 <= 8) continue;
    if (v.lap_bsb) return hipErrorInvalidValue;  // per-box sizes: MFMA kernels only
    const hipError_t e = launch_translate(st, v.work, v.nwork, v.map, 49, v.T, v.P, v.P, X, xs, Y,
                                          ys, nd, v.lap_log);
    if (e != hipSuccess) return e;
  }
  return hipGetLastError();
}

// FP64 -> FP32 copy of operator tables.
__global__ void k_to_f32(const double2* __restrict__ in, float2* __restrict__ out, size_t n) {
  const size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e < n) out[e] = make_float2((float)in[e].x, (float)in[e].y);
}

constexpr int XT_TB = 16;
constexpr size_t XT_LDS_MAX = 150 * 1024;

// Y[work] += sum_s T_s X[map]; picks the LDS-tiled kernel when the operator
// fits, the direct one otherwise.
inline hipError_t launch_translate(hipStream_t st, const int* work, int nwork, const int* map,
                                   int nslot, const double2* T, int nin, int nout,
                                   const double2* X, size_t xs, double2* Y, size_t ys, int nd,
                                   double lap_log) {
  if (nwork <= 0) return hipSuccess;
  size_t lds = ((size_t)nin * nout + (size_t)XT_TB * nin) * sizeof(double2);
  if (nout <= 128 && lds <= XT_LDS_MAX) {
    static bool init = false;
    if (!init) {
      hipError_t e = hipFuncSetAttribute((const void*)k_translate<XT_TB>,
                                         hipFuncAttributeMaxDynamicSharedMemorySize,
                                         (int)XT_LDS_MAX);
      if (e != hipSuccess) return e;
      init = true;
    }
    k_translate<XT_TB><<<(nwork + XT_TB - 1) / XT_TB, XT_THREADS, lds, st>>>(
        work, nwork, map, nslot, T, nin, nout, X, xs, Y, ys, nd, lap_log);
  } else {
    k_translate_direct<<<nwork, XT_THREADS, 0, st>>>(work, map, nslot, T, nin, nout, X, xs, Y, ys,
                                                     nd, lap_log);
  }
  return hipGetLastError();
}

}  // namespace
}  // namespace fmm

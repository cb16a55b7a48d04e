// dtree.hip -- the adaptive, level-restricted quadtree, its interaction lists
// and the per-step work lists, built on the device from sorted Morton keys.
//
// A box is (level l, Morton prefix key of 2l bits); box ids are level-major
// with Morton order inside a level, so every per-level range of boxes is a
// contiguous id range and every work list compacted in id order is sorted
// by level.  The box set is the reference's:
//   1. adaptive refinement (c_translation/src/pts_tree2d.c:530-610): a box
//      is split while its point count (the length of its key range in the
//      sorted key array) exceeds ndiv; all split boxes of all levels are
//      found in one pass over the sorted keys, each by its first point;
//   2. level restriction (pts_tree2d.c:767-960 with tree_routs2d.c:197-318):
//      the reference's flag passes split exactly the leaves that the 2:1
//      rule forces ("every same-level neighbour of a split box exists"),
//      so the result is the closure of the adaptive tree under that rule;
//      here the closure is taken directly, finest level first, on the sorted
//      key sets of split boxes: for each split box r at level k, the parents
//      of r's in-domain neighbours join the split set of level k-1;
//   3. the boxes of level l+1 are the four children of each split box of
//      level l.
// Centres use the reference's recurrence c + (+-1) bs/2 (tree_routs2d.c:81-84)
// along the key's digits, so they are bit-identical to the reference's.
// Colleagues and lists 1-4 (tree_routs2d.c:140-185, 369-510) are computed in
// integer cell coordinates; the reference's 1.05-scaled centre-distance tests
// separate exactly the same cell offsets.
#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <cstring>

#include "engine.h"
#include "rsort.h"
#include "translate.h"

namespace fmm {
namespace {

inline int nblk(long n, int b) { return (int)((n + b - 1) / b); }

FMM_HD uint32_t compact_bits(uint64_t v) {
  v &= 0x5555555555555555ull;
  v = (v | (v >> 1)) & 0x3333333333333333ull;
  v = (v | (v >> 2)) & 0x0f0f0f0f0f0f0f0full;
  v = (v | (v >> 4)) & 0x00ff00ff00ff00ffull;
  v = (v | (v >> 8)) & 0x0000ffff0000ffffull;
  v = (v | (v >> 16)) & 0x00000000ffffffffull;
  return (uint32_t)v;
}
FMM_HD uint64_t spread_bits(uint32_t x) {
  uint64_t v = x;
  v = (v | (v << 16)) & 0x0000ffff0000ffffull;
  v = (v | (v << 8)) & 0x00ff00ff00ff00ffull;
  v = (v | (v << 4)) & 0x0f0f0f0f0f0f0f0full;
  v = (v | (v << 2)) & 0x3333333333333333ull;
  v = (v | (v << 1)) & 0x5555555555555555ull;
  return v;
}
// digit bit 0 is the x half, bit 1 the y half (common.h quadrant)
FMM_HD void key_cell(uint64_t key, int& ix, int& iy) {
  ix = (int)compact_bits(key);
  iy = (int)compact_bits(key >> 1);
}
FMM_HD uint64_t cell_key(int ix, int iy) {
  return spread_bits((uint32_t)ix) | (spread_bits((uint32_t)iy) << 1);
}

__device__ inline int lbound(const uint64_t* a, int n, uint64_t v) {
  int lo = 0, hi = n;
  while (lo < hi) {
    int m = (lo + hi) >> 1;
    if (a[m] < v) lo = m + 1;
    else hi = m;
  }
  return lo;
}

// points of box (key, level) in a sorted key array
__device__ inline int2 key_range(const uint64_t* k, int n, uint64_t key, int level) {
  int a = lbound(k, n, box_key_lo(key, level));
  int b = lbound(k, n, box_key_lo(key + 1, level));
  return make_int2(a, b);
}

// Deepest split level: points i .. i+ndiv of the sorted keys share the
// level-c box, c = the number of leading key digits the two ends agree on,
// so that box holds more than ndiv points and is split; no deeper box is.
__global__ void k_depth(const uint64_t* k, int n, int ndiv, int* depth) {
  int c = -1;
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i + ndiv < n; i += gridDim.x * blockDim.x) {
    const uint64_t x = k[i] ^ k[i + ndiv];
    c = max(c, x == 0 ? kMaxDepth
                      : min(kMaxDepth, (__clzll((long long)x) - (64 - 2 * kMaxDepth)) >> 1));
  }
  // the block's maximum in depth[blockIdx.x] (the host takes the maximum)
  __shared__ int wmax[4];
  for (int o = 32; o >= 1; o >>= 1) c = max(c, __shfl_xor(c, o));
  if ((threadIdx.x & 63) == 0) wmax[threadIdx.x >> 6] = c;
  __syncthreads();
  if (threadIdx.x == 0) {
    int m = wmax[0];
    for (int k = 1; k < (int)(blockDim.x >> 6); ++k) m = max(m, wmax[k]);
    depth[blockIdx.x] = m;
  }
}

// leading base-4 digits (levels) two keys share, at most kMaxDepth
__device__ inline int common_levels(uint64_t a, uint64_t b) {
  const uint64_t x = a ^ b;
  return x == 0 ? kMaxDepth : min(kMaxDepth, (__clzll((long long)x) - (64 - 2 * kMaxDepth)) >> 1);
}

// Adaptive refinement in one pass over the sorted keys.  Point i starts its
// level-l box when it shares fewer than l digits with point i-1, and that
// box holds more than ndiv points -- and is split -- when point i+ndiv
// shares l digits with point i.  So the split boxes first met at point i
// are those of levels a_i < l <= b_i: every split box of every level is
// found once, by its first point, in Morton order within its level.  (A box
// with more than ndiv points has ancestors with more, all split, so it is
// a box of the tree.)  Levels are capped at the deepest split level Ls.
__device__ inline int2 split_span(const uint64_t* k, int n, int ndiv, int Ls, int i) {
  const int a = i == 0 ? -1 : common_levels(k[i - 1], k[i]);
  const int b = i + ndiv < n ? min(common_levels(k[i], k[i + ndiv]), Ls) : -1;
  return make_int2(a + 1, b);  // levels a+1 .. b
}
// cnt[0..n]: split boxes first met at each point (cnt[n] = 0, so an
// exclusive scan over n+1 entries ends with the total)
__global__ void k_split_count(const uint64_t* k, int n, int ndiv, int Ls, int* cnt) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i > n) return;
  if (i == n) {
    cnt[n] = 0;
    return;
  }
  const int2 r = split_span(k, n, ndiv, Ls, i);
  cnt[i] = max(0, r.y - r.x + 1);
}
// the split boxes as (level << 2L) | Morton prefix, written from *base + off[i]
// (cap bounds the output: at most n / (ndiv + 1) split boxes per level)
__global__ void k_split_emit(const uint64_t* k, int n, int ndiv, int Ls, int L, const int* off,
                             const int* base, long cap, uint64_t* out) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int2 r = split_span(k, n, ndiv, Ls, i);
  long o = off[i] + (base ? *base : 0);
  for (int l = r.x; l <= r.y && o < cap; ++l)
    out[o++] = ((uint64_t)l << (2 * L)) | (k[i] >> (2 * (kMaxDepth - l)));
}
// used = the split keys emitted (the totals of the two scans, capped at
// cap).  With nsplit null, [used, cap) is padded with v, which sorts after
// every split key (for a unique pass over all cap keys); else
// *nsplit = used, and the sort takes that count from the device.
__global__ void k_pad_tail(uint64_t* a, const int* tot_s, const int* tot_t, long cap, uint64_t v,
                           int* used, int* nsplit) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  const long u = min(cap, (long)(tot_s ? *tot_s : 0) + (tot_t ? *tot_t : 0));
  if (i == 0) {
    *used = (int)u;
    if (nsplit) *nsplit = (int)u;
  }
  if (!nsplit && i < cap && i >= u) a[i] = v;
}

// Level-restriction closure in one pass.  A split box r of level k forces
// the parents of its in-domain neighbours to be split at level k-1, whose
// neighbours' parents must then be split at k-2, and so on: at each coarser
// level the forced boxes form a rectangle of at most 3 x 3 cells.  Written
// as (level << 2L | key) for a tree of depth L, 9 slots per level below k
// and above mlo (short rectangles repeat their first cell); the adaptive
// split boxes themselves come first, already level-tagged.  The domain is
// the unit square (mlo = 0), or in a forest the split box's problem root
// (cell (rx, ry) of level m >= mlo): there only boxes at least two levels
// below the root force anything, and only down to the level below the
// root, as the problem's own tree would; unused slots repeat the split key.
__global__ void k_closure_cand(const uint64_t* R, const int* roff, const long* cbase, int L,
                               ForestDev F, int mlo, uint64_t* out) {
  const int kLevShift = 2 * L;
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= roff[L + 1]) return;
  const int k = (int)(R[r] >> kLevShift);  // R: (level << 2L) | key, sorted
  const uint64_t key = R[r] & ((uint64_t(1) << kLevShift) - 1);
  out[r] = R[r];
  if (k < mlo + 2) return;
  uint64_t* o = out + cbase[k] + (long)(r - roff[k]) * 9 * (k - 1 - mlo);
  int m = 0, rx = 0, ry = 0;
  if (F.M >= 0) {
    ForestSlot s;
    if (!forest_slot(F, key, k, s) || k - s.m < 2) {
      for (long q = 0; q < 9L * (k - 1 - mlo); ++q) o[q] = R[r];
      return;
    }
    m = s.m;
    rx = s.rx;
    ry = s.ry;
  }
  int ix, iy;
  key_cell(key, ix, iy);
  // the domain's cells at level k: [x0, x1] x [y0, y1]
  int x0 = rx << (k - m), x1 = ((rx + 1) << (k - m)) - 1;
  int y0 = ry << (k - m), y1 = ((ry + 1) << (k - m)) - 1;
  int ax = max(ix - 1, x0), bx = min(ix + 1, x1);
  int ay = max(iy - 1, y0), by = min(iy + 1, y1);
  for (int kk = k - 1; kk >= mlo + 1; --kk) {
    if (kk <= m) {  // at or above the root: nothing forced
      for (int q = 0; q < 9; ++q) o[q] = R[r];
      o += 9;
      continue;
    }
    ax >>= 1; bx >>= 1; ay >>= 1; by >>= 1;
    const uint64_t lev = (uint64_t)kk << kLevShift;
    int n = 0;
    for (int y = ay; y <= by && y <= ay + 2; ++y)
      for (int x = ax; x <= bx && x <= ax + 2; ++x) o[n++] = lev | cell_key(x, y);
    for (; n < 9; ++n) o[n] = lev | cell_key(ax, ay);
    o += 9;
    x0 = rx << (kk - m); x1 = ((rx + 1) << (kk - m)) - 1;
    y0 = ry << (kk - m); y1 = ((ry + 1) << (kk - m)) - 1;
    ax = max(ax - 1, x0); bx = min(bx + 1, x1);
    ay = max(ay - 1, y0); by = min(by + 1, y1);
  }
}

// per-level offsets of the sorted (level, key) split set, then the keys
// without their level bits
__global__ void k_split_levels(const uint64_t* U, const int* n, int L, int* off) {
  int l = blockIdx.x * blockDim.x + threadIdx.x;
  if (l > L + 1) return;
  const uint64_t v = (uint64_t)l << (2 * L);
  int lo = 0, hi = *n;
  while (lo < hi) {
    int m = (lo + hi) >> 1;
    if (U[m] < v) lo = m + 1;
    else hi = m;
  }
  off[l] = lo;
}
__global__ void k_strip_level(uint64_t* U, const int* n, int L) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < *n) U[i] &= (uint64_t(1) << (2 * L)) - 1;
}

struct Levels {
  int L;
  int loff[kMaxDepth + 2];  // first box id of level l; loff[L+1] = nb
  int roff[kMaxDepth + 2];  // first split box of level l in the split-key array
  double bs[kMaxDepth + 1];
  double c0x, c0y;
};


__device__ inline int level_of(const Levels& lv, int b) {
  int l = 0;
  while (l < lv.L && b >= lv.loff[l + 1]) ++l;
  return l;
}

__global__ void k_box_keys(Levels lv, const uint64_t* R, int nb, uint64_t* key) {
  int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= nb) return;
  int l = level_of(lv, b), i = b - lv.loff[l];
  key[b] = l == 0 ? 0 : 4 * R[lv.roff[l - 1] + (i >> 2)] + (uint64_t)(i & 3);
}

// per box: level, centre, children (and their parent link)
__global__ void k_boxes(Levels lv, const uint64_t* R, const uint64_t* key, int nb, int* level,
                        double2* ctr, int* nchild, int4* child, int* parent, double* bs_out,
                        double* bsb) {
  int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b == 0)
    for (int l = 0; l <= lv.L; ++l) bs_out[l] = lv.bs[l];
  if (b >= nb) return;
  const int l = level_of(lv, b);
  const uint64_t k = key[b];
  level[b] = l;
  bsb[b] = lv.bs[l];
  double cx = lv.c0x, cy = lv.c0y;
  for (int j = 1; j <= l; ++j) {
    int q = (int)(k >> (2 * (l - j))) & 3;
    cx = child_coord(cx, q & 1, lv.bs[j]);
    cy = child_coord(cy, q & 2, lv.bs[j]);
  }
  ctr[b] = make_double2(cx, cy);
  const int nr = lv.roff[l + 1] - lv.roff[l];
  const uint64_t* Rl = R + lv.roff[l];
  int rank = lbound(Rl, nr, k);
  if (rank < nr && Rl[rank] == k) {
    int c0 = lv.loff[l + 1] + 4 * rank;
    nchild[b] = 4;
    child[b] = make_int4(c0, c0 + 1, c0 + 2, c0 + 3);
    for (int q = 0; q < 4; ++q) parent[c0 + q] = b;
  } else {
    nchild[b] = 0;
    child[b] = make_int4(-1, -1, -1, -1);
  }
  if (b == 0) parent[0] = -1;
}

// The binary searches of every box, one per thread (blockIdx.y = search):
// 0-3 the ends of its source and target ranges, 4-12 the colleague in slot
// s = (dy+1)*3 + (dx+1) (-1 none).  Independent searches in separate
// threads overlap their load latencies.
constexpr int kBoxSearches = 13;
__global__ void k_box_search(Levels lv, const uint64_t* key, int nb, const uint64_t* ks, int ns,
                             const uint64_t* kt, int nt, int* srng, int* trng, int* nbr) {
  const int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= nb) return;
  const int job = blockIdx.y;
  const int l = level_of(lv, b);
  const uint64_t k = key[b];
  if (job < 4) {
    const uint64_t* a = job < 2 ? ks : kt;
    const int n = job < 2 ? ns : nt;
    const int r = lbound(a, n, box_key_lo(k + (job & 1), l));
    (job < 2 ? srng : trng)[2 * b + (job & 1)] = r;
    return;
  }
  const int s = job - 4;
  int ix, iy;
  key_cell(k, ix, iy);
  const int side = 1 << l, first = lv.loff[l], n = lv.loff[l + 1] - first;
  const int jx = ix + s % 3 - 1, jy = iy + s / 3 - 1;
  int id = -1;
  if (jx >= 0 && jy >= 0 && jx < side && jy < side) {
    const uint64_t kk = cell_key(jx, jy);
    const int p = lbound(key + first, n, kk);
    if (p < n && key[first + p] == kk) id = first + p;
  }
  nbr[9 * b + s] = id;
}

struct Int4Sum {
  __device__ int4 operator()(const int4& a, const int4& b) const {
    return make_int4(a.x + b.x, a.y + b.y, a.z + b.z, a.w + b.w);
  }
};

struct ListOut {
  int *o1, *o2, *o3, *o4, *l1, *l2, *l3, *l4;
  long cap1, cap2, cap3, cap4;  // entries each list array holds
  // per box: bit (dy + 3) 7 + (dx + 3) for each list-2 box at cell offset
  // (dx, dy) that holds sources (the M2L grouping key's mask), or null
  uint64_t* m2l_mask;
};

// Lists of box b (tree_routs2d.c:369-510):
//   list 2: children of the parent's colleagues two or more cells away;
//   list 1 (leaf b): leaf colleagues (b included), children of split
//           colleagues touching b, and -- for any b, as the reference adds
//           them from the coarse side -- leaf colleagues of the parent that
//           touch b;
//   list 3 (leaf b): children of split colleagues not touching b;
//   list 4: leaf colleagues of the parent not touching b.
// FILL = false counts into cnt[b]; FILL = true writes at the scanned offsets
// (and the list-2 source mask when out.m2l_mask is set; srng: source ranges).
template <bool FILL>
__global__ void k_lists(const uint64_t* key, const int* level, const int* nchild,
                        const int4* child, const int* parent, const int* nbr, const int2* srng,
                        int nb, int4* cnt, const int4* off, ListOut out) {
  int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b > nb) return;
  if (b == nb) {
    if (FILL) {
      int4 o = off[nb];
      out.o1[nb] = o.x; out.o2[nb] = o.y; out.o3[nb] = o.z; out.o4[nb] = o.w;
    } else {
      cnt[nb] = make_int4(0, 0, 0, 0);
    }
    return;
  }
  int4 o = FILL ? off[b] : make_int4(0, 0, 0, 0);
  if (FILL) {
    out.o1[b] = o.x; out.o2[b] = o.y; out.o3[b] = o.z; out.o4[b] = o.w;
  }
  int c1 = 0, c2 = 0, c3 = 0, c4 = 0;
  uint64_t mask = 0;
  // an entry past the array's capacity is counted but not written; the
  // host then finds the total above the capacity and rejects the tree
  auto put = [&](int* lst, long cap, int base, int& c, int v) {
    if (FILL && (long)base + c < cap) lst[base + c] = v;
    ++c;
  };
  const int l = level[b];
  const bool leaf = nchild[b] == 0;
  int ix, iy;
  key_cell(key[b], ix, iy);
  if (l == 0) {
    if (leaf) put(out.l1, out.cap1, o.x, c1, 0);
  } else {
    const int p = parent[b];
    for (int s = 0; s < 9; ++s) {
      int P = nbr[9 * p + s];
      if (P < 0) continue;
      if (nchild[P]) {
        int4 ch = child[P];
        int kk[4] = {ch.x, ch.y, ch.z, ch.w};
        for (int q = 0; q < 4; ++q) {
          int kx, ky;
          key_cell(key[kk[q]], kx, ky);
          if (abs(kx - ix) >= 2 || abs(ky - iy) >= 2) {
            put(out.l2, out.cap2, o.y, c2, kk[q]);
            if (FILL && out.m2l_mask) {
              const int2 r = srng[kk[q]];
              if (r.y > r.x) mask |= 1ull << ((ky - iy + 3) * 7 + (kx - ix + 3));
            }
          }
        }
      } else if (s != 4) {
        int px, py;
        key_cell(key[P], px, py);
        bool touch = ix >= 2 * px - 1 && ix <= 2 * px + 2 && iy >= 2 * py - 1 && iy <= 2 * py + 2;
        if (touch) put(out.l1, out.cap1, o.x, c1, P);
        else put(out.l4, out.cap4, o.w, c4, P);
      }
    }
    if (leaf) {
      for (int s = 0; s < 9; ++s) {
        int J = nbr[9 * b + s];
        if (J < 0) continue;
        if (!nchild[J]) {
          put(out.l1, out.cap1, o.x, c1, J);
          continue;
        }
        int4 ch = child[J];
        int kk[4] = {ch.x, ch.y, ch.z, ch.w};
        for (int q = 0; q < 4; ++q) {
          int kx, ky;
          key_cell(key[kk[q]], kx, ky);
          bool touch = kx >= 2 * ix - 1 && kx <= 2 * ix + 2 && ky >= 2 * iy - 1 && ky <= 2 * iy + 2;
          if (touch) put(out.l1, out.cap1, o.x, c1, kk[q]);
          else put(out.l3, out.cap3, o.z, c3, kk[q]);
        }
      }
    }
  }
  if (!FILL) cnt[b] = make_int4(c1, c2, c3, c4);
  if (FILL && out.m2l_mask) out.m2l_mask[b] = mask;
}

// ---- work lists: one ordered multi-list compaction over the boxes
enum { WL_FORMMP, WL_M2M, WL_M2L, WL_L2L, WL_LIST4, WL_LIST3_S, WL_LIST3_T, WL_LEAVES_S,
       WL_LEAVES_T, WL_P2P_S, WL_P2P_T, WL_N };
constexpr int WLB = 256;  // boxes per block (4 waves)

struct WlArgs {
  const int* level;
  const int* nchild;
  const int2* srng;
  const int2* trng;
  const int4* cnt;  // list lengths n1..n4
  int nb, pg, pgt;
  const double2* ctr;  // box centres
  const double* bs;    // box size per level
  const uint64_t* key;
  ForestDev F;
};

// In a forest, a box's level in its own problem's tree takes the place of
// its level: expansions (formation, M2M, M2L, L2L, list 4) start two
// levels below the root as they start at level 2 alone, and the near field
// runs over leaves only (above the roots, where no level restriction
// applies, a split box can have a coarse empty leaf in its list 1; below
// them, as alone, level restriction leaves a split box no list-1 entry).
__device__ inline unsigned wl_mask(const WlArgs& a, int b) {
  if (b >= a.nb) return 0;
  int l = a.level[b];
  const bool fo = a.F.M >= 0;
  if (fo) l = forest_plev(a.F, a.key[b], l);
  bool leaf = a.nchild[b] == 0;
  int2 s = a.srng[b], t = a.trng[b];
  int ns = s.y - s.x, nt = t.y - t.x;
  int np = (a.pgt > 0 ? nt : 0) + (a.pg > 0 ? ns : 0);
  int4 n = a.cnt[b];
  unsigned m = 0;
  m |= (unsigned)(l >= 2 && leaf && ns > 0) << WL_FORMMP;
  m |= (unsigned)(!leaf && ns > 0 && (!fo || l >= 2)) << WL_M2M;
  m |= (unsigned)(l >= 2 && np > 0 && n.y > 0) << WL_M2L;
  m |= (unsigned)(np > 0 && !leaf && (!fo || l >= 2)) << WL_L2L;
  m |= (unsigned)(l >= 2 && np > 0 && n.w > 0) << WL_LIST4;
  m |= (unsigned)(n.z > 0 && ns > 0) << WL_LIST3_S;
  m |= (unsigned)(n.z > 0 && nt > 0) << WL_LIST3_T;
  m |= (unsigned)(leaf && ns > 0) << WL_LEAVES_S;
  m |= (unsigned)(leaf && nt > 0) << WL_LEAVES_T;
  m |= (unsigned)(n.x > 0 && ns > 0 && (!fo || leaf)) << WL_P2P_S;
  m |= (unsigned)(n.x > 0 && nt > 0 && (!fo || leaf)) << WL_P2P_T;
  return m;
}

// per-block list counts; also the solve's work counts (list-1 pairs, list-2
// pairs per level) and, in pairs[4] (as the bits of a non-negative double,
// whose order is that of the integers), the square of the largest
// point-to-point distance over list 1: per axis |centre difference| + the
// two half sizes, over leaves with points and list-1 boxes with sources
__global__ void __launch_bounds__(WLB)
k_wl_count(WlArgs a, const int* o1, const int* l1, long cap1, int* bcnt, int nblocks,
           double* pairs, unsigned long long* m2l_lev) {
  __shared__ int sc[WL_N];
  __shared__ unsigned long long sm2l[kMaxDepth + 1];
  __shared__ unsigned long long sr2[kMaxDepth + 1];  // per level: squared bound, as bits
  if (threadIdx.x < WL_N) sc[threadIdx.x] = 0;
  if (threadIdx.x <= kMaxDepth) sm2l[threadIdx.x] = sr2[threadIdx.x] = 0;
  __syncthreads();
  int b = blockIdx.x * WLB + threadIdx.x;
  unsigned m = wl_mask(a, b);
  const int lane = threadIdx.x & 63;
  for (int j = 0; j < WL_N; ++j) {
    unsigned long long bal = __ballot((m >> j) & 1);
    if (lane == 0 && bal) atomicAdd(&sc[j], __popcll(bal));
  }
  // pairs[0], [1]: list-1 pairs at source and at target leaves; [2], [3]:
  // list-1 source reads (sum over list 1 of the source counts) of the
  // leaves holding sources and of those holding targets
  double acc[4] = {0, 0, 0, 0};
  double r2max = 0.0;
  if (b < a.nb) {
    if ((m >> WL_M2L) & 1) atomicAdd(&sm2l[a.level[b]], (unsigned long long)a.cnt[b].y);
    if (a.cnt[b].x > 0) {
      const int nsb = a.srng[b].y - a.srng[b].x, ntb = a.trng[b].y - a.trng[b].x;
      const double2 cb = a.ctr[b];
      const double hb = 0.5 * a.bs[a.level[b]];
      double srcs = 0;
      // entries past the array's capacity were not written (the host
      // rejects such a tree after this launch)
      const int kend = (int)min((long)o1[b + 1], cap1);
      for (int k = o1[b]; k < kend; ++k) {
        const int j = l1[k];
        int2 r = a.srng[j];
        srcs += r.y - r.x;
        if (r.y > r.x && nsb + ntb > 0) {
          const double2 cj = a.ctr[j];
          const double h = hb + 0.5 * a.bs[a.level[j]];
          const double ex = fabs(cb.x - cj.x) + h, ey = fabs(cb.y - cj.y) + h;
          r2max = fmax(r2max, ex * ex + ey * ey);
        }
      }
      acc[0] = srcs * nsb;
      acc[1] = srcs * ntb;
      acc[2] = nsb > 0 ? srcs : 0.0;
      acc[3] = ntb > 0 ? srcs : 0.0;
    }
  }
  __shared__ double sp[4][WLB / 64];
  for (int q = 0; q < 4; ++q) {
    double v = acc[q];
    for (int w = 32; w >= 1; w >>= 1) v += __shfl_xor(v, w);
    if (lane == 0) sp[q][threadIdx.x >> 6] = v;
  }
  __syncthreads();
  if (threadIdx.x < 4) {  // one atomic per block and sum
    double t = 0;
    for (int w = 0; w < WLB / 64; ++w) t += sp[threadIdx.x][w];
    if (t != 0) atomicAdd(&pairs[threadIdx.x], t);
  }
  // per level of the box (in LDS, then one atomic per level and block into
  // pairs[5 + l]), and over all levels (pairs[4])
  if (r2max > 0.0)
    atomicMax(&sr2[a.level[b]], (unsigned long long)__double_as_longlong(r2max));
  for (int w = 32; w >= 1; w >>= 1) r2max = fmax(r2max, __shfl_xor(r2max, w));
  if (lane == 0 && r2max > 0.0)
    atomicMax(reinterpret_cast<unsigned long long*>(pairs + 4),
              (unsigned long long)__double_as_longlong(r2max));
  if (threadIdx.x < WL_N) bcnt[threadIdx.x * nblocks + blockIdx.x] = sc[threadIdx.x];
  if (threadIdx.x <= kMaxDepth && sm2l[threadIdx.x])
    atomicAdd(&m2l_lev[threadIdx.x], sm2l[threadIdx.x]);
  __syncthreads();  // every lane's per-level bound is in sr2
  if (threadIdx.x <= kMaxDepth && sr2[threadIdx.x])
    atomicMax(reinterpret_cast<unsigned long long*>(pairs + 5 + threadIdx.x), sr2[threadIdx.x]);
}

// exclusive scan of the per-block counts of every list (one block, wave j
// scans list j); info[j] = list base in the output, info[WL_N + 1 + j] =
// list length
__global__ void __launch_bounds__(64 * WL_N) k_wl_scan(int* bcnt, int nblocks, int* info) {
  __shared__ int tot[WL_N];
  const int j = threadIdx.x >> 6, lane = threadIdx.x & 63;
  int* c = bcnt + (size_t)j * nblocks;
  int run = 0;
  for (int i0 = 0; i0 < nblocks; i0 += 64) {
    const int i = i0 + lane;
    const int v = i < nblocks ? c[i] : 0;
    int inc = v;
    for (int o = 1; o < 64; o <<= 1) {
      const int y = __shfl_up(inc, o);
      if (lane >= o) inc += y;
    }
    if (i < nblocks) c[i] = run + inc - v;
    run += __shfl(inc, 63);
  }
  if (lane == 0) tot[j] = run;
  __syncthreads();
  if (lane == 0) {
    int base = 0;
    for (int q = 0; q < j; ++q) base += tot[q];
    info[j] = base;
    info[WL_N + 1 + j] = run;
    if (j == WL_N - 1) info[WL_N] = base + run;  // start of the list-2 L2L children
  }
}

__global__ void __launch_bounds__(WLB)
k_wl_fill(WlArgs a, const int* boff, int nblocks, const int* info, int* out) {
  __shared__ int wc[WLB / 64][WL_N];
  int b = blockIdx.x * WLB + threadIdx.x;
  unsigned m = wl_mask(a, b);
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  unsigned long long below = (1ull << lane) - 1;
  int rank[WL_N];
  for (int j = 0; j < WL_N; ++j) {
    unsigned long long bal = __ballot((m >> j) & 1);
    rank[j] = __popcll(bal & below);
    if (lane == 0) wc[wv][j] = __popcll(bal);
  }
  __syncthreads();
  for (int j = 0; j < WL_N; ++j) {
    if (!((m >> j) & 1)) continue;
    int pos = info[j] + boff[(size_t)j * nblocks + blockIdx.x] + rank[j];
    for (int w = 0; w < wv; ++w) pos += wc[w][j];
    out[pos] = b;
  }
}

// children of the boxes of the L2L list, after all lists: for the parents
// of level l (entries lev[l] .. lev[l+1] of the L2L list) the children take
// positions 4 lev[l] .. 4 lev[l+1], grouped by quadrant, so the targets of
// a translation tile share their one operator
__global__ void k_wl_kids(const int* info, const int* lev, int W, const int4* child, int* out) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  int n = info[WL_N + 1 + WL_L2L];
  if (i >= n) return;
  const int* lv = lev + WL_L2L * W;
  int l = 0;
  while (l + 1 < W && lv[l + 1] <= i) ++l;
  const int a = lv[l], nl = lv[l + 1] - a;
  int4 c = child[out[info[WL_L2L] + i]];
  int* k = out + info[WL_N] + 4 * a + (i - a);
  k[0] = c.x; k[nl] = c.y; k[2 * nl] = c.z; k[3 * nl] = c.w;
}

// lev[j*(L+2) + l] = first entry of list j whose box is at level >= l
__global__ void k_wl_levels(Levels lv, const int* info, const int* out, int* lev) {
  int t = blockIdx.x * blockDim.x + threadIdx.x;
  const int W = lv.L + 2;
  if (t >= WL_N * W) return;
  int j = t / W, l = t - j * W;
  const int* v = out + info[j];
  int n = info[WL_N + 1 + j];
  int key = lv.loff[l];
  int lo = 0, hi = n;
  while (lo < hi) {
    int m = (lo + hi) >> 1;
    if (v[m] < key) lo = m + 1;
    else hi = m;
  }
  lev[t] = lo;
}

// M2L grouping key of a target: its level (5 bits) above a 19-bit hash
// (Fibonacci hashing) of the 49-bit mask of the list-2 offsets (slots) whose
// source box holds sources (k_lists' m2l_mask).  Equal masks get equal keys;
// a rare collision only puts two masks in one group.
constexpr int kM2lKeyBits = 5 + 19;
__global__ void k_m2l_keys(Levels lv, const uint64_t* m2l_mask, const int* list, int n,
                           uint32_t* key, int* val) {
  int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= n) return;
  const int b = list[e], l = level_of(lv, b);
  const uint64_t mask = m2l_mask[b];
  key[e] = ((uint32_t)l << 19) | (uint32_t)((mask * 0x9e3779b97f4a7c15ull) >> (64 - 19));
  val[e] = b;
}

// Keys that order the adjacent work lists list3_s, list3_t, leaves_s and
// leaves_t (one sort over the four) by list, level, then decreasing cost: the
// 16-point chunks of the box's points times, for list 3, its list-3 count
// (capped at 255).  The evaluation kernels run four boxes per wave, so boxes
// of equal cost side by side keep the wave's groups equally busy, and the
// heaviest waves of a level start first.
struct CostLists {
  int end[4];  // end of list j, relative to the first
};
__global__ void k_cost_keys(const int* work, int n, CostLists cl, const int* level,
                            const int2* srng, const int2* trng, const int* o3, uint32_t* key,
                            int* val) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int j = (i >= cl.end[0]) + (i >= cl.end[1]) + (i >= cl.end[2]);
  const int b = work[i];
  const int2 r = (j & 1) ? trng[b] : srng[b];
  const int cost = ((r.y - r.x + 15) >> 4) * (j < 2 ? o3[b + 1] - o3[b] : 1);
  key[i] = ((uint32_t)j << 13) | ((uint32_t)level[b] << 8) | (uint32_t)(255 - min(cost, 255));
  val[i] = b;
}

// FMM2DHIP_DEBUG=1 in the environment prints the tree's shape per solve
bool debug_enabled() {
  static const bool on = [] {
    const char* v = std::getenv("FMM2DHIP_DEBUG");
    return v && *v && *v != '0';
  }();
  return on;
}

// Reads n ints back; the stream is then synchronised, so a sort whose
// look-back wait gave up is known here, before its output is used as
// indices: that is reported as IER_HIP.
int sync_read(Ctx& c, const int* d, int n, int* h, const uint32_t* sort_err) {
  int* pin = static_cast<int*>(pinned(sizeof(int) * n, 4));
  if (!pin) return IER_ALLOC;
  HIP_TRY(hipMemcpyAsync(pin, d, sizeof(int) * n, hipMemcpyDeviceToHost, c.st));
  HIP_TRY(hipStreamSynchronize(c.st));
  std::memcpy(h, pin, sizeof(int) * n);
  return sort_err && *sort_err ? IER_HIP : IER_OK;
}

}  // namespace

int build_dtree(Ctx& c, Scratch& S, const uint64_t* ks, int ns, const uint64_t* kt, int nt,
                const double* geo, int idivflag, int ndiv, int ifpgh, int ifpghtarg,
                int sorted_levels, Plan& plan, double* host_s, const ForestDev* forest,
                int forest_mlo) {
  double t0 = wall();
  // ---- 0. depth of the adaptive tree (one read-back of per-block maxima)
  const bool depth_s = idivflag != 1 && ns > ndiv;
  const bool depth_t = (idivflag == 1 || idivflag == 2) && nt > ndiv;
  const int nbs = depth_s ? std::min(1024, nblk(ns - ndiv, 256)) : 0;
  const int nbt = depth_t ? std::min(1024, nblk(nt - ndiv, 256)) : 0;
  int* d_depth = S.get<int>(std::max(1, nbs + nbt));  // per block: deepest split level, -1 none
  if (S.err) return S.err;
  if (nbs) k_depth<<<nbs, 256, 0, c.st>>>(ks, ns, ndiv, d_depth);
  if (nbt) k_depth<<<nbt, 256, 0, c.st>>>(kt, nt, ndiv, d_depth + nbs);
  int Ls = -1;
  int e = IER_OK;
  if (nbs + nbt) {
    std::vector<int> bm(nbs + nbt);
    e = sync_read(c, d_depth, nbs + nbt, bm.data(), plan.sort_err);
    if (e) return e;
    for (int m : bm) {
      if (m < -1 || m > kMaxDepth) return IER_HIP;  // read-back out of range
      Ls = std::max(Ls, m);
    }
  } else {
    // nothing splits: the copies of the root box still have to land
    HIP_TRY(hipStreamSynchronize(c.st));
    if (*plan.sort_err) return IER_HIP;
  }
  if (idivflag < 0 || idivflag > 2) Ls = -1;
  // exact while every box that splits lies above the sorted levels: a box
  // at level sorted_levels holding more than ndiv points shows up as a
  // split at that level or below, and then the keys need a full sort
  if (Ls >= sorted_levels && sorted_levels < kMaxDepth) return IER_RESORT;
  // The keys resolve kMaxDepth levels: refinement stops there and the
  // deepest leaves keep more than ndiv points, as the reference's leaves do
  // when it stops at nlmax (pts_tree2d.c:191; nlmax = 51, cfmm2d.c:1120).
  // Coincident points (more than ndiv of them) reach this.
  if (Ls >= kMaxDepth) Ls = kMaxDepth - 1;
  const int L = Ls + 1;

  // ---- 1. adaptive refinement in one pass over the points (split_span):
  // per-point counts, a scan, the level-tagged split keys, one sort by
  // (level, key) -- padded to a capacity, so no read-back -- and the
  // per-level offsets.  With idivflag 2 both point sets contribute and
  // duplicates are removed.
  const bool use_s = idivflag == 0 || idivflag == 2, use_t = idivflag == 1 || idivflag == 2;
  const long cap = std::max<long>(
      1, (long)(Ls + 1) * ((use_s ? ns / (ndiv + 1) : 0) + (use_t ? nt / (ndiv + 1) : 0)));
  // level above 2L key bits, in as many bits as L + 1 needs: the pad key's
  // level field (all ones) then exceeds every level the splits can have
  const int kbits = 2 * L + (32 - __builtin_clz((unsigned)(L + 1)));
  const uint64_t kPad = (uint64_t(1) << kbits) - 1;
  uint64_t* Rs = S.get<uint64_t>(cap);   // split keys, sorted by (level, key)
  int* d_roff = S.get<int>(L + 2);
  int* d_nsplit = S.get<int>(2);
  if (S.err) return S.err;
  {
    uint64_t* Re = S.get<uint64_t>(cap);  // emitted, unsorted
    int* cs = use_s ? S.get<int>((size_t)ns + 1) : nullptr;
    int* os = use_s ? S.get<int>((size_t)ns + 1) : nullptr;
    int* ct = use_t ? S.get<int>((size_t)nt + 1) : nullptr;
    int* ot = use_t ? S.get<int>((size_t)nt + 1) : nullptr;
    if (S.err) return S.err;
    size_t tbs = 0, tbt = 0, tbu = 0;
    if (use_s) HIP_TRY(hipcub::DeviceScan::ExclusiveSum(nullptr, tbs, cs, os, ns + 1, c.st));
    if (use_t) HIP_TRY(hipcub::DeviceScan::ExclusiveSum(nullptr, tbt, ct, ot, nt + 1, c.st));
    if (idivflag == 2)
      HIP_TRY(hipcub::DeviceSelect::Unique(nullptr, tbu, Re, Rs, d_nsplit, (int)cap, c.st));
    void* tmp = S.get<char>(std::max({tbs, tbt, tbu, size_t(1)}));
    if (S.err) return S.err;
    int* d_used = d_nsplit + 1;  // emitted keys
    if (Ls >= 0) {
      if (use_s) {
        k_split_count<<<nblk(ns + 1, 256), 256, 0, c.st>>>(ks, ns, ndiv, Ls, cs);
        HIP_TRY(hipcub::DeviceScan::ExclusiveSum(tmp, tbs, cs, os, ns + 1, c.st));
        k_split_emit<<<nblk(ns, 256), 256, 0, c.st>>>(ks, ns, ndiv, Ls, L, os, nullptr, cap, Re);
      }
      if (use_t) {
        k_split_count<<<nblk(nt + 1, 256), 256, 0, c.st>>>(kt, nt, ndiv, Ls, ct);
        HIP_TRY(hipcub::DeviceScan::ExclusiveSum(tmp, tbt, ct, ot, nt + 1, c.st));
        k_split_emit<<<nblk(nt, 256), 256, 0, c.st>>>(kt, nt, ndiv, Ls, L, ot,
                                                      use_s ? os + ns : nullptr, cap, Re);
      }
    }
    k_pad_tail<<<idivflag == 2 ? (unsigned)((cap + 255) / 256) : 1u, 256, 0, c.st>>>(
        Re, Ls >= 0 && use_s ? os + ns : nullptr, Ls >= 0 && use_t ? ot + nt : nullptr, cap, kPad,
        d_used, idivflag == 2 ? nullptr : d_nsplit);
    if (idivflag == 2) {
      uint64_t* Rt = S.get<uint64_t>(cap);
      if (S.err) return S.err;
      e = rsort::radix_sort(c, S, Re, Rt, nullptr, nullptr, (int)cap, 0, kbits, plan.sort_err);
      if (e) return e;
      HIP_TRY(hipcub::DeviceSelect::Unique(tmp, tbu, Rt, Rs, d_nsplit, (int)cap, c.st));
    } else {
      // the emitted keys only (their count on the device): the capacity is
      // a bound several times the usual count
      e = rsort::radix_sort(c, S, Re, Rs, nullptr, nullptr, (int)cap, 0, kbits, plan.sort_err,
                            d_nsplit);
      if (e) return e;
    }
    k_split_levels<<<1, 64, 0, c.st>>>(Rs, d_nsplit, L, d_roff);
  }
  std::vector<int> nR(L + 1, 0);
  {
    std::vector<int> ro(L + 2);
    e = sync_read(c, d_roff, L + 2, ro.data(), plan.sort_err);
    if (e) return e;
    for (int l = 0; l <= L; ++l) nR[l] = ro[l + 1] - ro[l];
    // counts read back are checked before they size anything: at most cap
    // split boxes in all, none negative
    if (ro[0] != 0 || ro[L + 1] > cap) return IER_HIP;
    for (int l = 0; l <= L; ++l)
      if (nR[l] < 0) return IER_HIP;
  }
  if (Ls >= 0 && nR[Ls] == 0) return IER_ARGS;  // the depth and the refinement disagree
  double t1 = wall();

  // ---- 2. level restriction: closure of the split sets in one sort
  // (in a forest nothing is forced at or above the highest root, level mlo)
  const int mlo = forest && forest->M >= 0 ? forest_mlo : 0;
  uint64_t* Rall = nullptr;
  {
    std::vector<int> roff(L + 2, 0);
    std::vector<long> cbase(L + 1, 0);
    long nc = 0;
    for (int l = 0; l <= L; ++l) roff[l + 1] = roff[l] + nR[l];
    nc = roff[L + 1];
    for (int l = 0; l <= L; ++l) {
      cbase[l] = nc;
      if (l >= mlo + 2) nc += 9L * (l - 1 - mlo) * nR[l];
    }
    long* d_cbase = S.get<long>(L + 1);
    uint64_t* cand = S.get<uint64_t>(std::max<long>(nc, 1));
    uint64_t* sorted = S.get<uint64_t>(std::max<long>(nc, 1));
    Rall = S.get<uint64_t>(std::max<long>(nc, 1));
    int* d_nu = S.get<int>(1);
    int* d_loff = S.get<int>(L + 2);
    if (S.err) return S.err;
    long* hcb = static_cast<long*>(pinned(sizeof(long) * (L + 1), 5));
    if (!hcb) return IER_ALLOC;
    std::memcpy(hcb, cbase.data(), sizeof(long) * (L + 1));
    HIP_TRY(hipMemcpyAsync(d_cbase, hcb, sizeof(long) * (L + 1), hipMemcpyHostToDevice, c.st));
    // cand[0 .. roff[L]): the split keys (level-tagged), then the forced boxes
    if (roff[L] > 0)
      k_closure_cand<<<nblk(roff[L], 256), 256, 0, c.st>>>(
          Rs, d_roff, d_cbase, L, forest ? *forest : ForestDev{}, mlo, cand);
    size_t tb2 = 0;
    HIP_TRY(hipcub::DeviceSelect::Unique(nullptr, tb2, sorted, Rall, d_nu, (int)nc, c.st));
    void* tmp = S.get<char>(std::max(tb2, size_t(1)));
    if (S.err) return S.err;
    if (nc > 0) {
      e = rsort::radix_sort(c, S, cand, sorted, nullptr, nullptr, (int)nc, 0, kbits, plan.sort_err);
      if (e) return e;
      HIP_TRY(hipcub::DeviceSelect::Unique(tmp, tb2, sorted, Rall, d_nu, (int)nc, c.st));
    } else {
      HIP_TRY(hipMemsetAsync(d_nu, 0, sizeof(int), c.st));
    }
    k_split_levels<<<1, 64, 0, c.st>>>(Rall, d_nu, L, d_loff);
    k_strip_level<<<nblk(std::max<long>(nc, 1), 256), 256, 0, c.st>>>(Rall, d_nu, L);
    std::vector<int> lo(L + 2);
    e = sync_read(c, d_loff, L + 2, lo.data(), plan.sort_err);
    if (e) return e;
    if (lo[0] != 0 || lo[L + 1] > nc) return IER_HIP;
    for (int l = 0; l <= L; ++l) {
      nR[l] = lo[l + 1] - lo[l];
      // split boxes of level l have split parents at level l - 1, and a box
      // has 4 children: level l holds at most 4 nR[l-1] of them
      if (nR[l] < 0 || (l > 0 && nR[l] > 4L * nR[l - 1])) return IER_HIP;
    }
  }
  double t2 = wall();
  // ---- 3. boxes
  Levels lv{};
  lv.L = L;
  lv.c0x = geo[0];  // valid: the read-backs above synchronised the stream
  lv.c0y = geo[1];
  lv.bs[0] = geo[2];
  for (int l = 1; l <= L; ++l) lv.bs[l] = lv.bs[l - 1] / 2;
  lv.loff[0] = 0;
  lv.roff[0] = 0;
  for (int l = 0; l <= L; ++l) {
    int nl = l == 0 ? 1 : 4 * nR[l - 1];
    lv.loff[l + 1] = lv.loff[l] + nl;
    lv.roff[l + 1] = lv.roff[l] + nR[l];
  }
  const int nb = lv.loff[L + 1];
  uint64_t* key = S.get<uint64_t>(nb);
  int* level = S.get<int>(nb);
  double2* ctr = S.get<double2>(nb);
  int2* srng = S.get<int2>(nb);
  int2* trng = S.get<int2>(nb);
  int* nchild = S.get<int>(nb);
  int4* child = S.get<int4>(nb);
  int* parent = S.get<int>(nb);
  int* nbr = S.get<int>((size_t)9 * nb);
  double* dbs = S.get<double>(L + 1);
  double* bsb = S.get<double>(nb);
  int4* cnt = S.get<int4>(nb + 1);
  int4* off = S.get<int4>(nb + 1);
  int* o = S.get<int>(4 * ((size_t)nb + 1));
  if (S.err) return S.err;
  k_box_keys<<<nblk(nb, 256), 256, 0, c.st>>>(lv, Rall, nb, key);
  k_boxes<<<nblk(nb, 256), 256, 0, c.st>>>(lv, Rall, key, nb, level, ctr, nchild, child, parent,
                                          dbs, bsb);
  k_box_search<<<dim3(nblk(nb, 256), kBoxSearches), 256, 0, c.st>>>(
      lv, key, nb, ks, ns, kt, nt, reinterpret_cast<int*>(srng), reinterpret_cast<int*>(trng), nbr);
  // ---- lists 1-4 (count, scan, fill)
  ListOut lo{o, o + nb + 1, o + 2 * (nb + 1), o + 3 * (nb + 1), nullptr, nullptr, nullptr, nullptr,
             0, 0, 0, 0, nullptr};
  k_lists<false><<<nblk(nb + 1, 256), 256, 0, c.st>>>(key, level, nchild, child, parent, nbr,
                                                     srng, nb, cnt, nullptr, lo);
  size_t tb = 0;
  HIP_TRY(hipcub::DeviceScan::ExclusiveScan(nullptr, tb, cnt, off, Int4Sum(), make_int4(0, 0, 0, 0), nb + 1, c.st));
  void* tmp = S.get<char>(tb);
  if (S.err) return S.err;
  HIP_TRY(hipcub::DeviceScan::ExclusiveScan(tmp, tb, cnt, off, Int4Sum(), make_int4(0, 0, 0, 0), nb + 1, c.st));
  // List lengths per box are bounded (13, 27, 20 and 5 after level
  // restriction); the arrays hold twice that per box, so their totals need
  // no read-back here: they come back with the work-list counts below and
  // a total above its array's capacity rejects the tree there, before any
  // kernel other than the work-list counter (which stops at the capacity)
  // reads the lists.
  const long lim[4] = {26L * nb, 54L * nb, 40L * nb, 10L * nb};
  lo.cap1 = lim[0];
  lo.cap2 = lim[1];
  lo.cap3 = lim[2];
  lo.cap4 = lim[3];
  lo.l1 = S.get<int>(lim[0]);
  lo.l2 = S.get<int>(lim[1]);
  lo.l3 = S.get<int>(lim[2]);
  lo.l4 = S.get<int>(lim[3]);
  lo.m2l_mask = S.get<uint64_t>(nb);
  if (S.err) return S.err;
  k_lists<true><<<nblk(nb + 1, 256), 256, 0, c.st>>>(key, level, nchild, child, parent, nbr,
                                                    srng, nb, nullptr, off, lo);
  // ---- work lists
  const int nbk = nblk(nb, WLB);
  const int W = L + 2;
  const int ninfo = 2 * WL_N + 1;
  // small results read back in one copy: info, level offsets, m2l pairs
  // (as 2 ints each), list-1 pairs, source reads and the squared list-1
  // distance bound, over all levels and per level (5 + kMaxDepth + 1
  // doubles as ints)
  const int nsmall = ninfo + WL_N * W + 2 * (kMaxDepth + 1) + 10 + 2 * (kMaxDepth + 1);
  int* small = S.get<int>(nsmall + 6);  // + the four list totals
  int* bcnt = S.get<int>((size_t)WL_N * nbk);
  if (S.err) return S.err;
  int* info = small;
  int* lev = small + ninfo;
  unsigned long long* m2l_lev = reinterpret_cast<unsigned long long*>(small + ((ninfo + WL_N * W + 1) & ~1));
  double* pairs = reinterpret_cast<double*>(m2l_lev + kMaxDepth + 1);
  HIP_TRY(hipMemsetAsync(small, 0, sizeof(int) * (nsmall + 2), c.st));
  HIP_TRY(hipMemcpyAsync(small + nsmall + 2, off + nb, sizeof(int4), hipMemcpyDeviceToDevice, c.st));
  WlArgs wa{level, nchild, srng, trng, cnt, nb, ifpgh, ifpghtarg, ctr, dbs, key,
            forest ? *forest : ForestDev{}};
  k_wl_count<<<nbk, WLB, 0, c.st>>>(wa, lo.o1, lo.l1, lo.cap1, bcnt, nbk, pairs, m2l_lev);
  k_wl_scan<<<1, 64 * WL_N, 0, c.st>>>(bcnt, nbk, info);
  // every list holds at most nb entries; the L2L children at most 4 nb
  int* work = S.get<int>((size_t)WL_N * nb + 4 * (size_t)nb + 1);
  if (S.err) return S.err;
  k_wl_fill<<<nbk, WLB, 0, c.st>>>(wa, bcnt, nbk, info, work);
  k_wl_levels<<<nblk(WL_N * W, 64), 64, 0, c.st>>>(lv, info, work, lev);
  k_wl_kids<<<nblk(nb, 256), 256, 0, c.st>>>(info, lev, W, child, work);
  HIP_TRY(hipGetLastError());
  std::vector<int> hs(nsmall + 6);
  e = sync_read(c, small, nsmall + 6, hs.data(), plan.sort_err);
  if (e) return e;
  int tot[4];
  for (int k = 0; k < 4; ++k) {
    tot[k] = hs[nsmall + 2 + k];
    if (tot[k] < 0 || tot[k] > lim[k]) return IER_HIP;
  }

  // ---- host view
  plan.nb = nb;
  plan.L = L;
  plan.bs.assign(lv.bs, lv.bs + L + 1);
  plan.loff.assign(lv.loff, lv.loff + L + 2);
  plan.dwork = work;
  WorkList* wl[WL_N] = {&plan.formmp, &plan.m2m, &plan.m2l, &plan.l2l, &plan.list4,
                        &plan.list3_s, &plan.list3_t, &plan.leaves_s, &plan.leaves_t,
                        &plan.p2p_s, &plan.p2p_t};
  for (int j = 0; j < WL_N; ++j) {
    wl[j]->off = hs[j];
    wl[j]->n = hs[WL_N + 1 + j];
    wl[j]->lev.assign(hs.begin() + ninfo + j * W, hs.begin() + ninfo + (j + 1) * W);
    // every list holds at most nb boxes, in level order
    if (wl[j]->n < 0 || wl[j]->n > nb || wl[j]->off < 0 || wl[j]->lev.front() != 0 ||
        wl[j]->lev.back() != wl[j]->n)
      return IER_HIP;
    for (int l = 0; l + 1 < W; ++l)
      if (wl[j]->lev[l] > wl[j]->lev[l + 1]) return IER_HIP;
  }
  plan.l2l_kids.off = hs[WL_N];
  plan.l2l_kids.n = 4 * plan.l2l.n;
  plan.l2l_kids.lev.resize(W);
  for (int l = 0; l < W; ++l) plan.l2l_kids.lev[l] = 4 * plan.l2l.lev[l];
  // M2L targets of a level grouped by the set of list-2 offsets that carry
  // sources (stable, so equal sets stay in Morton order): the 16 targets of
  // an M2L tile then mostly share the slots they use, and a slot is skipped
  // only when no target of the tile has a source there
  if (plan.m2l.n > 0) {
    const int n = plan.m2l.n;
    uint32_t* k0 = S.get<uint32_t>(n);
    uint32_t* k1 = S.get<uint32_t>(n);
    int* v0 = S.get<int>(n);
    if (S.err) return S.err;
    k_m2l_keys<<<nblk(n, 256), 256, 0, c.st>>>(lv, lo.m2l_mask, work + plan.m2l.off, n, k0, v0);
    const int kb = kM2lKeyBits;
    e = rsort::radix_sort(c, S, k0, k1, v0, work + plan.m2l.off, n, 0, kb, plan.sort_err);
    if (e) return e;
    rsort::fallback_unsorted(c.st, plan.sort_err, v0, work + plan.m2l.off, n);
  }
  // list-3 and leaf work in decreasing cost within each level (the four
  // lists are adjacent in the work buffer)
  static_assert(WL_LIST3_T == WL_LIST3_S + 1 && WL_LEAVES_S == WL_LIST3_T + 1 &&
                    WL_LEAVES_T == WL_LEAVES_S + 1, "adjacent lists");
  {
    const int first = plan.list3_s.off;
    CostLists cl;
    const WorkList* q4[4] = {&plan.list3_s, &plan.list3_t, &plan.leaves_s, &plan.leaves_t};
    for (int j = 0; j < 4; ++j) cl.end[j] = q4[j]->off + q4[j]->n - first;
    const int n = cl.end[3];
    if (n > 1) {
      uint32_t* k0 = S.get<uint32_t>(n);
      uint32_t* k1 = S.get<uint32_t>(n);
      int* v0 = S.get<int>(n);
      if (S.err) return S.err;
      k_cost_keys<<<nblk(n, 256), 256, 0, c.st>>>(work + first, n, cl, level, srng, trng, lo.o3,
                                                  k0, v0);
      e = rsort::radix_sort(c, S, k0, k1, v0, work + first, n, 0, 15, plan.sort_err);
      if (e) return e;
      rsort::fallback_unsorted(c.st, plan.sort_err, v0, work + first, n);
    }
  }
  const int m2l_at = (ninfo + WL_N * W + 1) & ~1;
  const unsigned long long* hm = reinterpret_cast<const unsigned long long*>(hs.data() + m2l_at);
  plan.m2l_pairs.assign(hm, hm + L + 1);
  const double* hp = reinterpret_cast<const double*>(hm + kMaxDepth + 1);

  DTree& D = plan.D;
  D.nb = nb;
  D.key = key;
  D.ctr = ctr;
  D.level = level;
  D.bs = dbs;
  D.bsb = bsb;
  D.nchild = nchild;
  D.parent = parent;
  D.child = child;
  D.srng = srng;
  D.trng = trng;
  D.nbr = nbr;
  D.o1 = lo.o1; D.l1 = lo.l1; D.o2 = lo.o2; D.l2 = lo.l2;
  D.o3 = lo.o3; D.l3 = lo.l3; D.o4 = lo.o4; D.l4 = lo.l4;

  if (debug_enabled()) {
    std::fprintf(stderr, "fmm2dhip tree: L=%d nb=%d lists=%d %d %d %d split/level:", L, nb,
                 tot[0], tot[1], tot[2], tot[3]);
    for (int l = 0; l <= L; ++l) std::fprintf(stderr, " %d", nR[l]);
    std::fprintf(stderr, "\nfmm2dhip work lists (off,n):");
    for (int j = 0; j < WL_N; ++j) std::fprintf(stderr, " (%d,%d)", wl[j]->off, wl[j]->n);
    std::fprintf(stderr, " kids (%d,%d)\n", plan.l2l_kids.off, plan.l2l_kids.n);
  }
  SolveStats& st = last_stats();
  st.p2p_pairs_src = hp[0];
  st.p2p_pairs_trg = hp[1];
  st.p2p_reads_src = hp[2];
  st.p2p_reads_trg = hp[3];
  plan.p2p_rmax = hp[4] >= 0.0 && hp[4] < 1e300 ? std::sqrt(hp[4]) : 0.0;
  plan.p2p_rmax_lev.assign(L + 1, 0.0);
  for (int l = 0; l <= L; ++l)
    plan.p2p_rmax_lev[l] = hp[5 + l] >= 0.0 && hp[5 + l] < 1e300 ? std::sqrt(hp[5 + l]) : 0.0;
  plan.forest = forest ? *forest : ForestDev{};
  st.nboxes = nb;
  st.nlevels = L;
  double t3 = wall();
  if (host_s) {
    host_s[0] = t1 - t0;
    host_s[1] = t2 - t1;
    host_s[2] = t3 - t2;
  }
  return IER_OK;
}

}  // namespace fmm

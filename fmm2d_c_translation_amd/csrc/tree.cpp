// tree.cpp -- box-level adaptive quadtree from sorted Morton keys.
//
// Produces the reference's tree bit for bit (box numbering, centres,
// colleagues, lists); the rules implemented here are those of
//   refinement            c_translation/src/pts_tree2d.c:530-610
//   child numbering       c_translation/src/tree_routs2d.c:48-101
//   colleagues            c_translation/src/tree_routs2d.c:140-185
//   2:1 restriction       c_translation/src/pts_tree2d.c:767-960 (+ updateflags,
//                         tree_refine_boxes_flag: tree_routs2d.c:197-318)
//   renumbering           c_translation/src/pts_tree2d.c:966-1081
//   lists 1-4             c_translation/src/tree_routs2d.c:369-510
// but the work is organised around key ranges: the point count of a box is
// the length of its key interval in the sorted key array, so no per-point
// partitioning happens here.
#include "tree.h"

#include <algorithm>
#include <cmath>

#include "common.h"

namespace fmm {
namespace {

int lower(const uint64_t* k, int a, int b, uint64_t v) {
  return (int)(std::lower_bound(k + a, k + b, v) - k);
}

class Builder {
 public:
  Builder(Tree& t, const uint64_t* ks, int ns, const uint64_t* kt, int nt)
      : T(t), ks_(ks), kt_(kt) {
    (void)ns;
    (void)nt;
  }

  int add_box(int lev, uint64_t pre, double x, double y, int par) {
    int id = T.nboxes++;
    T.prefix.push_back(pre);
    T.level.push_back(lev);
    T.parent.push_back(par);
    T.nchild.push_back(0);
    for (int q = 0; q < 4; ++q) T.child.push_back(-1);
    T.cx.push_back(x);
    T.cy.push_back(y);
    T.nnb.push_back(0);
    for (int i = 0; i < 9; ++i) T.nbr.push_back(-1);
    T.s0.push_back(0); T.s1.push_back(0); T.t0.push_back(0); T.t1.push_back(0);
    flag.push_back(0);
    return id;
  }

  // four children of b at level lev+1, point ranges split by key
  void split(int b) {
    int lev = T.level[b];
    double bsc = T.bs[lev + 1];
    uint64_t pre = T.prefix[b];
    int s = T.s0[b], t = T.t0[b];
    T.nchild[b] = 4;
    for (int q = 0; q < 4; ++q) {
      int c = add_box(lev + 1, 4 * pre + q, child_coord(T.cx[b], q & 1, bsc),
                      child_coord(T.cy[b], q & 2, bsc), b);
      T.child[4 * b + q] = c;
      uint64_t hi = box_key_lo(4 * pre + q + 1, lev + 1);
      int se = q == 3 ? T.s1[b] : lower(ks_, s, T.s1[b], hi);
      int te = q == 3 ? T.t1[b] : lower(kt_, t, T.t1[b], hi);
      T.s0[c] = s; T.s1[c] = se; T.t0[c] = t; T.t1[c] = te;
      s = se;
      t = te;
    }
  }

  bool touching(int a, int b, double d) const {
    return std::fabs(T.cx[a] - T.cx[b]) <= d && std::fabs(T.cy[a] - T.cy[b]) <= d;
  }

  // colleagues of b: children of the parent's colleagues within 1.05 bs,
  // visited in (parent-colleague, child) order, i.e. Morton order
  void colleagues(int b) {
    int dad = T.parent[b];
    double d = 1.05 * T.bs[T.level[b]];
    int n = 0;
    for (int i = 0; i < T.nnb[dad]; ++i) {
      int jb = T.nbr[9 * dad + i];
      for (int q = 0; q < 4; ++q) {
        int kb = T.child[4 * jb + q];
        if (kb >= 0 && touching(kb, b, d)) T.nbr[9 * b + n++] = kb;
      }
    }
    for (int i = n; i < 9; ++i) T.nbr[9 * b + i] = -1;
    T.nnb[b] = n;
  }

  void all_colleagues() {
    T.nnb[0] = 1;
    T.nbr[0] = 0;
    for (int i = 1; i < 9; ++i) T.nbr[i] = -1;
    for (int l = 1; l <= T.nlevels; ++l)
      for (int b = T.lfirst[l]; b < T.lfirst[l] + T.lcount[l]; ++b) colleagues(b);
  }

  // refine every box of [first, first+count) whose flag is positive;
  // children of flag-1 boxes start as flag 3, of flag-2 boxes as flag 0
  void refine_flagged(int first, int count) {
    for (int b = first; b < first + count; ++b) {
      if (flag[b] <= 0) continue;
      int f = flag[b];
      split(b);
      for (int q = 0; q < 4; ++q) flag[T.child[4 * b + q]] = (f == 1) ? 3 : 0;
    }
  }

  // renumber: each level becomes (its boxes in current order, then tails)
  void reorg(const std::vector<int>& tfirst, const std::vector<int>& tcount) {
    int n = T.nboxes, L = T.nlevels;
    std::vector<int> order;
    order.reserve(n);
    std::vector<int> nf(L + 1), nc(L + 1);
    for (int l = 0; l <= L; ++l) {
      nf[l] = (int)order.size();
      for (int b = T.lfirst[l]; b < T.lfirst[l] + T.lcount[l]; ++b) order.push_back(b);
      for (int b = tfirst[l]; b < tfirst[l] + tcount[l]; ++b) order.push_back(b);
      nc[l] = (int)order.size() - nf[l];
    }
    std::vector<int> map(n);
    for (int i = 0; i < n; ++i) map[order[i]] = i;
    auto perm = [&](auto& v, int w) {
      auto old = v;
      for (int i = 0; i < n; ++i)
        for (int k = 0; k < w; ++k) v[(size_t)w * i + k] = old[(size_t)w * order[i] + k];
    };
    perm(T.prefix, 1); perm(T.level, 1); perm(T.parent, 1); perm(T.nchild, 1);
    perm(T.child, 4); perm(T.cx, 1); perm(T.cy, 1); perm(T.s0, 1); perm(T.s1, 1);
    perm(T.t0, 1); perm(T.t1, 1); perm(flag, 1);
    for (int i = 0; i < n; ++i) {
      if (T.parent[i] >= 0) T.parent[i] = map[T.parent[i]];
      for (int q = 0; q < 4; ++q)
        if (T.child[4 * i + q] >= 0) T.child[4 * i + q] = map[T.child[4 * i + q]];
    }
    T.lfirst = nf;
    T.lcount = nc;
  }

  void level_restriction() {
    const int L = T.nlevels;
    const std::vector<double>& bs = T.bs;
    auto range = [&](int l, auto&& f) {
      for (int b = T.lfirst[l]; b < T.lfirst[l] + T.lcount[l]; ++b) f(b);
    };
    // leaves next to a box whose children sit two levels below them
    for (int l = L; l >= 2; --l) {
      double d = 1.05 * (bs[l - 1] + bs[l - 2]) / 2.0;
      range(l, [&](int b) {
        int dad = T.parent[b], gd = T.parent[dad];
        for (int i = 0; i < T.nnb[gd]; ++i) {
          int jb = T.nbr[9 * gd + i];
          if (T.nchild[jb] == 0 && flag[jb] == 0 && touching(jb, dad, d)) flag[jb] = 1;
        }
      });
    }
    // leaves next to a flagged box one level below them
    for (int l = L; l >= 1; --l) {
      double d = 1.05 * (bs[l] + bs[l - 1]) / 2.0;
      range(l, [&](int b) {
        if (flag[b] != 1 && flag[b] != 2) return;
        int dad = T.parent[b];
        for (int i = 0; i < T.nnb[dad]; ++i) {
          int jb = T.nbr[9 * dad + i];
          if (T.nchild[jb] == 0 && flag[jb] == 0 && touching(jb, b, d)) flag[jb] = 2;
        }
      });
    }
    std::vector<int> tf(L + 1, 0), tc(L + 1, 0);
    for (int l = 1; l <= L - 2; ++l) {
      tf[l + 1] = T.nboxes;
      refine_flagged(T.lfirst[l], T.lcount[l]);
      tc[l + 1] = T.nboxes - tf[l + 1];
    }
    reorg(tf, tc);
    all_colleagues();
    for (auto& f : flag)
      if (f != 3) f = 0;
    std::fill(tf.begin(), tf.end(), 0);
    std::fill(tc.begin(), tc.end(), 0);
    for (int l = 2; l <= L - 2; ++l) {
      double d = 1.05 * (bs[l] + bs[l + 1]) / 2.0;
      auto update = [&](int first, int count) {
        for (int b = first; b < first + count; ++b) {
          if (flag[b] != 3) continue;
          flag[b] = 0;
          for (int i = 0; i < T.nnb[b] && !flag[b]; ++i) {
            int jb = T.nbr[9 * b + i];
            for (int q = 0; q < 4; ++q) {
              int kb = T.child[4 * jb + q];
              if (kb >= 0 && T.nchild[kb] > 0 && touching(kb, b, d)) {
                flag[b] = 1;
                break;
              }
            }
          }
        }
      };
      update(T.lfirst[l], T.lcount[l]);
      update(tf[l], tc[l]);
      tf[l + 1] = T.nboxes;
      refine_flagged(T.lfirst[l], T.lcount[l]);
      refine_flagged(tf[l], tc[l]);
      tc[l + 1] = T.nboxes - tf[l + 1];
      for (int b = tf[l + 1]; b < tf[l + 1] + tc[l + 1]; ++b) colleagues(b);
    }
    reorg(tf, tc);
    all_colleagues();
  }

  Tree& T;
  std::vector<int> flag;

 private:
  const uint64_t *ks_, *kt_;
};

}  // namespace

int build_tree(const uint64_t* ks, int ns, const uint64_t* kt, int nt, double c0x, double c0y,
               double bs0, int idivflag, int ndiv, Tree& T) {
  T = Tree();
  Builder B(T, ks, ns, kt, nt);
  T.bs.push_back(bs0);
  B.add_box(0, 0, c0x, c0y, -1);
  T.s0[0] = 0; T.s1[0] = ns; T.t0[0] = 0; T.t1[0] = nt;
  T.lfirst.push_back(0);
  T.lcount.push_back(1);
  int lev = 0;
  for (;; ++lev) {
    std::vector<int> ref;
    for (int b = T.lfirst[lev]; b < T.lfirst[lev] + T.lcount[lev]; ++b) {
      int nss = T.s1[b] - T.s0[b], ntt = T.t1[b] - T.t0[b], nn = 0;
      if (idivflag == 0) nn = nss;
      if (idivflag == 1) nn = ntt;
      if (idivflag == 2) nn = std::max(nss, ntt);
      if (nn > ndiv) ref.push_back(b);
    }
    if (ref.empty()) break;
    if (lev + 1 > kMaxDepth) return IER_DEPTH;
    T.bs.push_back(T.bs[lev] / 2);
    T.lfirst.push_back(T.nboxes);
    for (int b : ref) B.split(b);
    T.lcount.push_back(T.nboxes - T.lfirst[lev + 1]);
  }
  T.nlevels = lev;
  B.all_colleagues();
  if (T.nlevels >= 2) B.level_restriction();
  return IER_OK;
}

// Lists in the reference's order: levels coarse to fine, boxes in id order;
// a coarse leaf appends itself to the list1/list4 of finer boxes before those
// boxes add their own entries.
void build_lists(Tree& T) {
  const int n = T.nboxes;
  T.n1.assign(n, 0); T.n2.assign(n, 0); T.n3.assign(n, 0); T.n4.assign(n, 0);
  T.l1.assign((size_t)13 * n, -1); T.l2.assign((size_t)27 * n, -1);
  T.l3.assign((size_t)20 * n, -1); T.l4.assign((size_t)5 * n, -1);
  auto push = [](std::vector<int>& l, std::vector<int>& c, int w, int b, int v) {
    if (c[b] < w) l[(size_t)w * b + c[b]] = v;
    c[b]++;
  };
  if (T.nchild[0] == 0) push(T.l1, T.n1, 13, 0, 0);
  for (int lev = 1; lev <= T.nlevels; ++lev) {
    const double bs = T.bs[lev];
    const double bsn = lev < T.nlevels ? T.bs[lev + 1] : bs / 2;
    for (int b = T.lfirst[lev]; b < T.lfirst[lev] + T.lcount[lev]; ++b) {
      int dad = T.parent[b];
      for (int i = 0; i < T.nnb[dad]; ++i) {
        int jb = T.nbr[9 * dad + i];
        for (int q = 0; q < T.nchild[jb]; ++q) {
          int kb = T.child[4 * jb + q];
          if (std::fabs(T.cx[kb] - T.cx[b]) >= 1.05 * bs || std::fabs(T.cy[kb] - T.cy[b]) >= 1.05 * bs)
            push(T.l2, T.n2, 27, b, kb);
        }
      }
      if (T.nchild[b] != 0) continue;
      for (int i = 0; i < T.nnb[b]; ++i) {
        int jb = T.nbr[9 * b + i];
        if (T.nchild[jb] == 0) {
          push(T.l1, T.n1, 13, b, jb);
          continue;
        }
        double d = 1.05 * (bs + bsn) / 2.0;
        for (int q = 0; q < T.nchild[jb]; ++q) {
          int kb = T.child[4 * jb + q];
          double xd = std::fabs(T.cx[kb] - T.cx[b]), yd = std::fabs(T.cy[kb] - T.cy[b]);
          if (xd < d && yd < d) {
            push(T.l1, T.n1, 13, b, kb);
            push(T.l1, T.n1, 13, kb, b);
          } else {
            push(T.l3, T.n3, 20, b, kb);
            push(T.l4, T.n4, 5, kb, b);
          }
        }
      }
    }
  }
}

}  // namespace fmm

// tree.h -- box-level quadtree built from sorted Morton keys.
//
// The device computes each point's Morton path key (common.h point_key) and
// radix-sorts the keys; this host code turns the sorted key arrays into the
// reference's adaptive, level-restricted quadtree with the reference's box
// numbering, colleagues and interaction lists, using only per-box work
// (binary searches into the key arrays, no per-point work).
#pragma once

#include <cstdint>
#include <vector>

namespace fmm {

struct Tree {
  int nlevels = 0;  // deepest level index
  int nboxes = 0;
  std::vector<double> bs;        // box size per level
  std::vector<int> lfirst;       // first box id of each level (0-based ids)
  std::vector<int> lcount;       // boxes per level; ids of a level are contiguous
  std::vector<uint64_t> prefix;  // Morton prefix of each box at its own level
  std::vector<int> level, parent, nchild, child;  // child[4*b+q] (q = 0..3), -1 none
  std::vector<double> cx, cy;
  std::vector<int> nnb, nbr;     // colleagues, nbr[9*b+i]
  std::vector<int> s0, s1;       // sources of box b: sorted positions [s0, s1)
  std::vector<int> t0, t1;       // targets of box b
  // interaction lists, fixed row widths as in the reference (13/27/20/5)
  std::vector<int> n1, l1, n2, l2, n3, l3, n4, l4;
};

// Builds the tree.  ks/kt are the sorted source/target keys; (c0x, c0y) and
// bs0 are the root centre and size.  Returns 0 or an fmm::IER_* code.
int build_tree(const uint64_t* ks, int ns, const uint64_t* kt, int nt, double c0x,
               double c0y, double bs0, int idivflag, int ndiv, Tree& T);

// Interaction lists 1-4 (needs a finished tree).
void build_lists(Tree& T);

}  // namespace fmm

// Post-order walk program of a rooted binary tree (host code, no HIP).
//
// The four-state post-order kernel evaluates a whole tree per pattern tile
// with a stack of partials. need(v) is the number of stack entries needed to
// evaluate the subtree of v when the internal child with the larger need is
// evaluated first (its result then waits on the stack while the other child
// is evaluated):
//   need(tip) = 0
//   need(v)   = max(need(a), need(b), 1 + min(need(a), need(b)))  both internal
//             = max(1, need(a), need(b))                          otherwise
// which is at most floor(log2(T)) + 1. Each internal node becomes one entry
// of kWalkInts ints, root last:
//   [0] v  [1] left  [2] right
//   [3] slots = (left slot + 1) | (right slot + 1) << 8 | (output slot) << 16,
//       slot + 1 = 0 for a tip child
//   [4] tpos  = (left tip position + 1) | (right tip position + 1) << 8: the
//       row of the tip's codes in the current tip segment (0 = not a tip)
//   [5] seg   = index of the tip segment this entry opens, or -1
//   [6] kinds = kind(left) | kind(right) << 8 with kind 0 = tip, 1 = the
//       result of the previous entry, 2 = an earlier result (stacked: the
//       first-evaluated child of a node whose children are both internal)
// Tips are consumed in walk order; consecutive entries are grouped into
// segments of at most kWalkSegTips tips so a kernel can stage the codes of
// one segment while it works through the previous one. segtab holds
// kWalkSegTips tip ids per segment (-1 padding).
#pragma once

#include <algorithm>
#include <cstdint>
#include <vector>

namespace hx {

constexpr int kWalkInts = 8;
constexpr int kWalkSegTips = 16;

// kids: [2 * nn] children of each node (tips 0..T-1 have -1); postorder:
// [I][3] triples, root last. Returns false if the program is malformed.
inline bool build_walk(int T, const std::vector<int32_t>& kids, const int32_t* postorder,
                       std::vector<int32_t>& walk, std::vector<int32_t>& segtab, int& walk_need) {
  const int I = T - 1, nn = 2 * T - 1;
  const int root = postorder[3 * (I - 1)];
  std::vector<int> need(nn, 0);
  for (int r = 0; r < I; ++r) {
    const int v = postorder[3 * r], a = kids[2 * v], b = kids[2 * v + 1];
    const int na = a >= T ? need[a] : 0, nb = b >= T ? need[b] : 0;
    need[v] = (na > 0 && nb > 0) ? std::max(std::max(na, nb), 1 + std::min(na, nb))
                                 : std::max(1, std::max(na, nb));
  }
  struct Frame {
    int v, base, state;
  };
  walk.clear();
  walk.reserve(kWalkInts * (size_t)I);
  segtab.clear();
  int seg_fill = kWalkSegTips;  // forces a segment at the first entry
  walk_need = 0;
  std::vector<Frame> fs;
  fs.push_back({root, 0, 0});
  while (!fs.empty()) {
    const Frame f = fs.back();
    const int a = kids[2 * f.v], b = kids[2 * f.v + 1];
    int first = -1, second = -1;
    if (a >= T && b >= T) {
      first = need[a] >= need[b] ? a : b;
      second = first == a ? b : a;
    } else if (a >= T) {
      first = a;
    } else if (b >= T) {
      first = b;
    }
    if (f.state == 0) {
      fs.back().state = 1;
      if (first >= 0) {
        fs.push_back({first, f.base, 0});
        continue;
      }
    }
    if (f.state <= 1) {
      fs.back().state = 2;
      if (second >= 0) {
        fs.push_back({second, f.base + 1, 0});
        continue;
      }
    }
    const int sa = a < T ? -1 : (a == first ? f.base : f.base + 1);
    const int sb = b < T ? -1 : (b == first ? f.base : f.base + 1);
    const int ntips = (a < T) + (b < T);
    int seg = -1;
    if (seg_fill + ntips > kWalkSegTips) {
      seg = (int)(segtab.size() / kWalkSegTips);
      segtab.insert(segtab.end(), kWalkSegTips, -1);
      seg_fill = 0;
    }
    const size_t row0 = segtab.size() - kWalkSegTips;
    int pa = -1, pb = -1;
    if (a < T) {
      pa = seg_fill++;
      segtab[row0 + pa] = a;
    }
    if (b < T) {
      pb = seg_fill++;
      segtab[row0 + pb] = b;
    }
    const int prev = walk.empty() ? -1 : walk[walk.size() - kWalkInts];
    const int ka = a < T ? 0 : (a == prev ? 1 : 2);
    const int kb = b < T ? 0 : (b == prev ? 1 : 2);
    walk.push_back(f.v);
    walk.push_back(a);
    walk.push_back(b);
    walk.push_back((sa + 1) | ((sb + 1) << 8) | (f.base << 16));
    walk.push_back((pa + 1) | ((pb + 1) << 8));
    walk.push_back(seg);
    walk.push_back(ka | (kb << 8));
    walk.push_back(0);
    walk_need = std::max(walk_need, f.base + 1 + (second >= 0 ? 1 : 0));
    fs.pop_back();
  }
  return (int)walk.size() == kWalkInts * I && walk[kWalkInts * (I - 1)] == root &&
         walk_need <= 255 && walk_need == need[root] && walk[5] == 0;
}


// Pre-order walk program: one entry per internal node, depth first from the
// root. An entry processes parent p with children a and b given p's adjoint U
// (root: the frequencies): it forms both children's statistics and adjoints.
// One internal child continues the walk (its adjoint stays in registers for
// the next entry); when both are internal the one with the smaller subtree
// continues and the other's adjoint is pushed, to be popped when the walk
// below the continuing child is finished. A push at depth d waits for that
// whole subtree, so shallow depths (long waits) are the ones worth keeping in
// registers. Continuing into the smaller subtree bounds the depth by
// floor(log2 T). kPreInts ints per entry:
//   [0] p  [1] a  [2] b
//   [3] usrc = 0 root, 1 continuation of the previous entry, 2 | d << 8 pop
//       of depth d
//   [4] cont = child (0 a, 1 b) whose adjoint the next entry uses, or -1
//   [5] push = child | d << 8 whose adjoint is pushed at depth d, or -1
//   [6] tips = (a is a tip) | (b is a tip) << 1
//   [7] 0
constexpr int kPreInts = 8;

inline bool build_pre_walk(int T, const std::vector<int32_t>& kids, const int32_t* postorder,
                           std::vector<int32_t>& prog, int& depth) {
  const int I = T - 1, nn = 2 * T - 1;
  const int root = postorder[3 * (I - 1)];
  std::vector<int> size(nn, 0);  // internal nodes in the subtree
  for (int r = 0; r < I; ++r) {
    const int v = postorder[3 * r], a = kids[2 * v], b = kids[2 * v + 1];
    size[v] = 1 + (a >= T ? size[a] : 0) + (b >= T ? size[b] : 0);
  }
  prog.clear();
  prog.reserve((size_t)kPreInts * I);
  depth = 0;
  std::vector<int> pend;  // pushed nodes, deepest last
  int cur = root, usrc = 0;
  for (int guard = 0; guard < I; ++guard) {
    const int a = kids[2 * cur], b = kids[2 * cur + 1];
    const bool ia = a >= T, ib = b >= T;
    int cont = -1, push = -1;
    if (ia && ib) {
      cont = size[a] <= size[b] ? 0 : 1;
      push = (1 - cont) | ((int)pend.size() << 8);
      pend.push_back(cont == 0 ? b : a);
      depth = std::max(depth, (int)pend.size());
    } else if (ia) {
      cont = 0;
    } else if (ib) {
      cont = 1;
    }
    const int32_t e[kPreInts] = {cur, a, b, usrc, cont, push, (a < T ? 1 : 0) | (b < T ? 2 : 0), 0};
    prog.insert(prog.end(), e, e + kPreInts);
    if (cont >= 0) {
      cur = cont == 0 ? a : b;
      usrc = 1;
    } else if (!pend.empty()) {
      cur = pend.back();
      pend.pop_back();
      usrc = 2 | ((int)pend.size() << 8);
    } else {
      break;
    }
  }
  return (int)prog.size() == kPreInts * I && prog[0] == root && pend.empty() && depth <= 255;
}

}  // namespace hx

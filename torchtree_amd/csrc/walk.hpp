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
//   [7] drop  = bit m set when this entry's node is a cherry (both children
//       tips) whose result its parent takes from registers under register
//       mode m (0: no stack depth in registers; 1: depths kWalkRegDepth0 ..
//       kWalkRegDepth0 + kWalkRegDepths - 1 in registers): the partial need
//       not be stored, because the pre-order walk rebuilds it from the two
//       tip codes
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
// Stack depths the post-order walk may keep in registers (register mode 1).
constexpr int kWalkRegDepth0 = 2;
constexpr int kWalkRegDepths = 3;

// kids: [2 * nn] children of each node (tips 0..T-1 have -1); postorder:
// [I][3] triples, root last. drop (optional): per node, the [7] bits of its
// entry. Returns false if the program is malformed.
inline bool build_walk(int T, const std::vector<int32_t>& kids, const int32_t* postorder,
                       std::vector<int32_t>& walk, std::vector<int32_t>& segtab, int& walk_need,
                       std::vector<uint8_t>* drop = nullptr) {
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
  // drop bits: a cherry taken by its parent as the previous result (both
  // modes) or from a register depth (mode 1)
  std::vector<uint8_t> bits(nn, 0);
  std::vector<int> at(nn, -1);
  for (int i = 0; i < (int)walk.size() / kWalkInts; ++i) at[walk[kWalkInts * i]] = i;
  for (int i = 0; i < (int)walk.size() / kWalkInts; ++i) {
    const int32_t* e = walk.data() + (size_t)kWalkInts * i;
    for (int h = 0; h < 2; ++h) {
      const int ch = e[1 + h];
      if (ch < T || kids[2 * ch] >= T || kids[2 * ch + 1] >= T) continue;
      const int kind = (e[6] >> (8 * h)) & 0xff, slot = ((e[3] >> (8 * h)) & 0xff) - 1;
      const bool reg = slot >= kWalkRegDepth0 && slot < kWalkRegDepth0 + kWalkRegDepths;
      bits[ch] = (uint8_t)((kind == 1 ? 1 : 0) | ((kind == 1 || (kind == 2 && reg)) ? 2 : 0));
    }
  }
  for (int i = 0; i < (int)walk.size() / kWalkInts; ++i)
    walk[(size_t)kWalkInts * i + 7] = bits[walk[(size_t)kWalkInts * i]];
  if (drop) *drop = bits;
  return (int)walk.size() == kWalkInts * I && walk[kWalkInts * (I - 1)] == root &&
         walk_need <= 255 && walk_need == need[root] && walk[5] == 0 && bits[root] == 0;
}

// The cherries of a tree (internal nodes whose children are both tips) in
// post-order, as [ncher][2] tip ids, and the row of each node among them
// (-1 for other nodes).
inline void list_cherries(int T, const std::vector<int32_t>& kids, const int32_t* postorder,
                          std::vector<int32_t>& pairs, std::vector<int32_t>& row) {
  const int I = T - 1, nn = 2 * T - 1;
  pairs.clear();
  row.assign(nn, -1);
  for (int r = 0; r < I; ++r) {
    const int v = postorder[3 * r], a = kids[2 * v], b = kids[2 * v + 1];
    if (a < T && b < T) {
      row[v] = (int)(pairs.size() / 2);
      pairs.push_back(a);
      pairs.push_back(b);
    }
  }
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
//   [7] drop = the post-order drop bits (build_walk) of a | those of b << 2:
//       under register mode m, child h's partial was not stored when bit
//       2 h + m is set, and the walk rebuilds it from its tips
//   [8], [9]   code rows of a and b: a tip's id, a cherry's T + its index
//              in list_cherries, otherwise 0
//   [10], [11] a cherry a's two tips (else 0); [12], [13] those of b
//   [14], [15] 0
constexpr int kPreInts = 16;

// drop: per node, the bits build_walk reports (nullptr: none dropped);
// cher_row: list_cherries' rows (nullptr: no cherry rows).
inline bool build_pre_walk(int T, const std::vector<int32_t>& kids, const int32_t* postorder,
                           std::vector<int32_t>& prog, int& depth,
                           const std::vector<uint8_t>* drop = nullptr,
                           const std::vector<int32_t>* cher_row = nullptr) {
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
    auto dbits = [&](int ch) { return (ch >= T && drop) ? (int)((*drop)[ch] & 3) : 0; };
    auto crow = [&](int ch) {
      if (ch < T) return ch;
      return (cher_row && (*cher_row)[ch] >= 0) ? T + (*cher_row)[ch] : 0;
    };
    auto gtip = [&](int ch, int q) {
      return (ch >= T && cher_row && (*cher_row)[ch] >= 0) ? kids[2 * ch + q] : 0;
    };
    const int32_t e[kPreInts] = {cur, a, b, usrc, cont, push, (a < T ? 1 : 0) | (b < T ? 2 : 0),
                                 dbits(a) | (dbits(b) << 2), crow(a), crow(b), gtip(a, 0),
                                 gtip(a, 1), gtip(b, 0), gtip(b, 1), 0, 0};
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

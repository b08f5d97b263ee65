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
// which is at most floor(log2(T)) + 1. Each internal node becomes one int4
// {v, left, right, slots}, slots = (left slot + 1) | (right slot + 1) << 8 |
// (output slot) << 16 with slot + 1 = 0 for a tip; the root comes last.
#pragma once

#include <algorithm>
#include <cstdint>
#include <vector>

namespace hx {

// kids: [2 * nn] children of each node (tips 0..T-1 have -1); postorder:
// [I][3] triples, root last. Returns false if the program is malformed.
inline bool build_walk(int T, const std::vector<int32_t>& kids, const int32_t* postorder,
                       std::vector<int32_t>& walk, int& walk_need) {
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
  walk.reserve(4 * (size_t)I);
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
    walk.push_back(f.v);
    walk.push_back(a);
    walk.push_back(b);
    walk.push_back((sa + 1) | ((sb + 1) << 8) | (f.base << 16));
    walk_need = std::max(walk_need, f.base + 1 + (second >= 0 ? 1 : 0));
    fs.pop_back();
  }
  return (int)walk.size() == 4 * I && walk[4 * (I - 1)] == root && walk_need <= 255 &&
         walk_need == need[root];
}

}  // namespace hx

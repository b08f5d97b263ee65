// Reads T and (T-1) post-order triples on stdin, prints the walk program
// (need and segment count, one "v l r slots tpos seg kinds" line per internal node,
// then one line of 16 tip ids per segment) built by
// hx::build_walk. Host-only test driver for the four-state post-order walk.
#include <cstdio>
#include <vector>

#include "walk.hpp"

int main() {
  int T = 0;
  if (std::scanf("%d", &T) != 1 || T < 2) return 2;
  const int I = T - 1, nn = 2 * T - 1;
  std::vector<int32_t> post(3 * I), kids(2 * nn, -1);
  for (int r = 0; r < I; ++r) {
    if (std::scanf("%d %d %d", &post[3 * r], &post[3 * r + 1], &post[3 * r + 2]) != 3) return 2;
    kids[2 * post[3 * r]] = post[3 * r + 1];
    kids[2 * post[3 * r] + 1] = post[3 * r + 2];
  }
  std::vector<int32_t> walk, segtab;
  int need = 0;
  if (!hx::build_walk(T, kids, post.data(), walk, segtab, need)) return 1;
  const int nseg = (int)(segtab.size() / hx::kWalkSegTips);
  std::printf("%d %d\n", need, nseg);
  for (int i = 0; i < I; ++i) {
    const int32_t* w = walk.data() + hx::kWalkInts * i;
    std::printf("%d %d %d %d %d %d %d %d\n", w[0], w[1], w[2], w[3], w[4], w[5], w[6], w[7]);
  }
  for (int s = 0; s < nseg; ++s) {
    for (int j = 0; j < hx::kWalkSegTips; ++j)
      std::printf("%d%c", segtab[hx::kWalkSegTips * s + j], j + 1 < hx::kWalkSegTips ? ' ' : '\n');
  }
  return 0;
}

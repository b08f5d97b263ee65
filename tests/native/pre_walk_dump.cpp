// Reads T and (T-1) post-order triples on stdin, prints the pre-order walk
// program built by hx::build_pre_walk (with the post-order walk's drop bits
// and the cherry rows, as the plan builds it): its stack depth, then one
// "p a b usrc cont push tips drop row_a row_b a0 a1 b0 b1" line per internal
// node. Host-only test driver for the four-state pre-order walk.
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
  std::vector<int32_t> walk, segtab, cher, cher_row;
  std::vector<uint8_t> drop;
  int need = 0;
  if (!hx::build_walk(T, kids, post.data(), walk, segtab, need, &drop)) return 1;
  hx::list_cherries(T, kids, post.data(), cher, cher_row);
  std::vector<int32_t> prog;
  int depth = 0;
  if (!hx::build_pre_walk(T, kids, post.data(), prog, depth, &drop, &cher_row)) return 1;
  std::printf("%d\n", depth);
  for (int i = 0; i < I; ++i) {
    const int32_t* e = prog.data() + hx::kPreInts * i;
    for (int q = 0; q < 14; ++q) std::printf("%d%c", e[q], q + 1 < 14 ? ' ' : '\n');
  }
  return 0;
}

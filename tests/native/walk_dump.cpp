// Reads T and (T-1) post-order triples on stdin, prints the walk program
// (need, then one "v l r slots" line per internal node) built by
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
  std::vector<int32_t> walk;
  int need = 0;
  if (!hx::build_walk(T, kids, post.data(), walk, need)) return 1;
  std::printf("%d\n", need);
  for (int i = 0; i < I; ++i)
    std::printf("%d %d %d %d\n", walk[4 * i], walk[4 * i + 1], walk[4 * i + 2], walk[4 * i + 3]);
  return 0;
}

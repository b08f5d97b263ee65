// Host-only checks of the library's shape arithmetic (layout.hpp) and walk programs
// (walk.hpp), meant to be built with -fsanitize=address,undefined: workspace regions of
// hx_loglik* and of a job evaluation are ordered, 256-byte aligned and large enough for
// what the kernels address; the tiling covers every pattern; the walk builders succeed on
// random, caterpillar and balanced trees and leave room for the cherry code rows.
// Prints "ok <checks>" and exits 0, or prints the failed check and exits 1.
#include <cstdio>
#include <cstdlib>
#include <vector>

#include "layout.hpp"
#include "walk.hpp"

static long g_checks = 0;
#define CHECK(cond)                                                              \
  do {                                                                           \
    ++g_checks;                                                                  \
    if (!(cond)) {                                                               \
      std::printf("FAILED %s:%d: %s\n", __FILE__, __LINE__, #cond);              \
      std::exit(1);                                                              \
    }                                                                            \
  } while (0)

static uint64_t g_rng = 88172645463325252ull;
static uint32_t rnd(uint32_t n) {  // xorshift64
  g_rng ^= g_rng << 13;
  g_rng ^= g_rng >> 7;
  g_rng ^= g_rng << 17;
  return (uint32_t)(g_rng % n);
}

// postorder triples (v, a, b) of a random rooted binary tree on T tips (random joins)
static std::vector<int32_t> random_tree(int T) {
  std::vector<int> pool(T);
  for (int i = 0; i < T; ++i) pool[i] = i;
  std::vector<int32_t> post;
  int next = T;
  while (pool.size() > 1) {
    const int i = (int)rnd((uint32_t)pool.size());
    const int a = pool[i];
    pool[i] = pool.back();
    pool.pop_back();
    const int j = (int)rnd((uint32_t)pool.size());
    const int b = pool[j];
    pool[j] = next;
    post.push_back(next++);
    post.push_back(a);
    post.push_back(b);
  }
  return post;
}

static std::vector<int32_t> caterpillar(int T) {
  std::vector<int32_t> post;
  int prev = 0;
  for (int t = 1; t < T; ++t) {
    post.push_back(T + t - 1);
    post.push_back(prev);
    post.push_back(t);
    prev = T + t - 1;
  }
  return post;
}

static void check_walks(int T, const std::vector<int32_t>& post) {
  const int nn = 2 * T - 1, I = T - 1;
  std::vector<int32_t> kids(2 * nn, -1);
  for (int r = 0; r < I; ++r) {
    kids[2 * post[3 * r]] = post[3 * r + 1];
    kids[2 * post[3 * r] + 1] = post[3 * r + 2];
  }
  std::vector<int32_t> walk, segtab, cher, row, prew;
  std::vector<uint8_t> drop;
  int need = 0, depth = 0;
  CHECK(hx::build_walk(T, kids, post.data(), walk, segtab, need, &drop));
  CHECK((int)drop.size() == nn);
  hx::list_cherries(T, kids, post.data(), cher, row);
  CHECK(2 * (int)(cher.size() / 2) <= T);  // the 16-bit code rows reserve T / 2 cherries
  CHECK(hx::build_pre_walk(T, kids, post.data(), prew, depth, &drop, &row));
  CHECK((int)prew.size() == hx::kPreInts * I);
  for (int v = 0; v < nn; ++v) {
    if (drop[v]) CHECK(v >= T && row[v] >= 0);  // only cherries are dropped
    CHECK((drop[v] & 1) == 0 || (drop[v] & 2) != 0);  // mode 1 drops what mode 0 drops
  }
  for (int i = 0; i < I; ++i) {
    const int32_t* e = prew.data() + (size_t)hx::kPreInts * i;
    for (int h = 0; h < 2; ++h) {
      const int ch = e[1 + h], r = e[8 + h];
      CHECK(r >= 0 && r < T + (int)(cher.size() / 2));  // a row of the 16-bit code array
      if (ch < T) CHECK(r == ch);
    }
  }
}

static void check_layout(int T, int S, int N, int C, int K, int dtype, int grad) {
  const size_t rb = dtype == HX_F64 ? 8 : 4;
  const hx::Tiling t = hx::make_tiling(N, C, K, S, (int)rb);
  CHECK(t.ntiles >= 1 && (size_t)t.ntiles * t.ppb >= (size_t)N);
  CHECK((size_t)(t.ntiles - 1) * t.ppb < (size_t)N);
  CHECK(t.nchunks >= 1 && (size_t)t.nchunks * t.tiles_per_chunk >= (size_t)t.ntiles);
  CHECK((size_t)(t.nchunks - 1) * t.tiles_per_chunk < (size_t)t.ntiles);
  const hx::WsLayout w = hx::ws_layout_t(T, S, N, C, K, dtype, grad);
  const size_t I = T - 1, B = 2 * (size_t)T - 2;
  const size_t np = S == 4 ? (size_t)t.ntiles * 64 : (size_t)N;
  const size_t offs[] = {w.partials, w.scales, w.rootsums, w.gpart, w.ppad, w.pcodes, w.total};
  for (size_t q = 0; q + 1 < sizeof(offs) / sizeof(offs[0]); ++q) {
    CHECK(offs[q] % 256 == 0);
    CHECK(offs[q] <= offs[q + 1]);
  }
  CHECK(w.tail == w.partials + I * K * np * C * S * rb);
  CHECK(w.tail <= w.scales);
  if (S == 4) CHECK(w.rootsums - w.scales >= (size_t)T * t.ntiles * 64);
  CHECK(w.gpart - w.rootsums >= (size_t)K * t.ntiles * (1 + S + C) * 8);
  if (grad == HX_GRAD_MATS) CHECK(w.ppad - w.gpart >= B * K * (size_t)t.nchunks * C * S * S * rb);
  if (grad == HX_GRAD_TAU) CHECK(w.ppad - w.gpart >= B * K * (size_t)t.nchunks * C * 8);
  if (hx::use_ppad(S, dtype, grad)) CHECK(w.pcodes - w.ppad >= B * K * (size_t)C * 2 * 64 * 64 * rb);
  if (S == 4 && rb == 4) CHECK(w.total - w.pcodes >= (size_t)(T + T / 2) * np * 2);
  CHECK(hx::workspace_bytes_for(T, S, N, C, K, dtype, grad) == w.total);
  if (hx::walk_drop_mode(S, C, N, K, dtype, HX_TIPS_CODES) >= 0) {
    CHECK(S == 4 && rb == 4 && hx::use_pre_walk(S, C, 4, N, K));
    CHECK(t.tiles_per_chunk == hx::kPreWalkW);
  }
  // a resident job's evaluation: one slice of the patterns
  const hx::JobWs j = hx::job_ws_shape(T, S, N, C, K, dtype, grad != 0);
  CHECK(j.mats % 256 == 0 && j.prune % 256 == 0 && j.loglik_tmp % 256 == 0);
  CHECK(j.prune >= (size_t)K * B * C * S * S * rb * (grad ? 2 : 1));
  CHECK(j.prune_bytes == hx::workspace_bytes_for(T, S, N, C, K, dtype, grad ? HX_GRAD_TAU : 0));
  CHECK(j.loglik_tmp >= j.prune + j.prune_bytes);
  CHECK(j.total >= j.loglik_tmp + (size_t)K * 8 + (grad ? 2 * (size_t)K * B * C * 8 : 0));
}

int main() {
  const int taxa[] = {2, 3, 5, 17, 64, 200, 512, 1024, 3000};
  for (int T : taxa) {
    check_walks(T, random_tree(T));
    check_walks(T, caterpillar(T));
  }
  for (int rep = 0; rep < 40; ++rep) {
    const int T = 2 + (int)rnd(2000);
    check_walks(T, random_tree(T));
  }
  const int states[] = {2, 4, 5, 20, 33, 61, 64};
  const int cats[] = {1, 2, 3, 4, 8, 16};
  const int pats[] = {1, 63, 64, 65, 1000, 33000, 100000};
  const int ks[] = {1, 2, 8, 32};
  const int dtypes[] = {HX_F32, HX_F64, HX_F32_BF16X3};
  for (int S : states)
    for (int C : cats) {
      if (S != 4 && 2 * C * (S <= 32 ? 32 : 64) > 512) continue;  // as hx_loglik refuses
      for (int N : pats)
        for (int K : ks)
          for (int d : dtypes)
            for (int g = 0; g < 3; ++g) check_layout(2 + (int)rnd(1024), S, N, C, K, d, g);
    }
  std::printf("ok %ld\n", g_checks);
  return 0;
}

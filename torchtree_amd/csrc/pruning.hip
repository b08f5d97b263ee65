// Felsenstein pruning on gfx950: entry points, workspace layout and the chunk
// reductions shared by the four-state (pruning_dna.hip) and dense-state
// (pruning_wide.hip, pruning_mfma.hip) kernels.
//
// Reference: torchtree/evolution/tree_likelihood.py (calculate_treelikelihood*,
// lines 14-278). The reference walks post_indexing in Python and issues two
// batched matmuls and one multiply per node; here the four-state post-order is
// one launch that walks the whole tree per pattern tile, and the other passes
// cover every node of a tree level (or two levels) for every sample per launch.
//
// Layout in HBM (R = float or double):
//   partials [slot][K][N][C][S]   slot = node - T (S != 4), pattern-major so a
//                                 pattern's C*S values are contiguous;
//            [slot][K][C][np][4]  for S = 4, category-major so a wave reads
//                                 1 KiB runs of one category; np = 64 ntiles
//   scales   [slot][K][N] int32   cumulative power-of-two exponent of the
//                                 subtree below the node
//   rootsums [K][part][1+S+C]     FP64 partial sums at the root per chunk
//                                 (S != 4) or per 64-pattern tile (S = 4)
//   gpart    [B][K][chunk][C][S][S]  per-chunk partial d loglik / d mats, or
//   gtau     [B][K][chunk][C]        per-chunk d loglik / d tau (tau = t_b r_c),
//                                    the contraction of the former with dP/dtau
//
// Rescaling: every node's (C,S)-maximum per (sample, pattern) is brought into
// [0.5, 1) by an exact power of two (frexp/ldexp), the exponent is added to
// the children's cumulative exponents, and the root adds E*ln2 in FP64. This
// is the per-node max rescaling of calculate_treelikelihood_discrete_rescaled
// (tree_likelihood.py:186-221) with the divisor rounded to a power of two, so
// the scaling itself introduces no rounding.
#include "pruning.hpp"

#include <math.h>

#include <type_traits>

namespace hx {

// ------------------------------------------------------------ chunk reductions
// One block per (sample, quantity): thread t sums parts t, t+256, ... in order,
// then a fixed tree over the threads; the result does not depend on K.
// Block (0, 0) also zeroes two runs of words that the dense-state pre-order
// reads past the end of a slab and must find finite: the padding after the
// partials, and the head of the root's first slab (the post-order never
// stores the root's partials).
__global__ __launch_bounds__(256) void k_reduce_root(const double* __restrict__ rootsums, int K,
                                                     int nparts, int S, int C, double* out_l,
                                                     double* out_f, double* out_p, uint32_t* zero0,
                                                     int nzero0, uint32_t* zero1, int nzero1) {
  __shared__ double red[256];
  const int Q = 1 + S + C;
  const int k = blockIdx.x, q = blockIdx.y;
  if (k == 0 && q == 0) {
    for (int i = threadIdx.x; i < nzero0; i += 256) zero0[i] = 0u;
    for (int i = threadIdx.x; i < nzero1; i += 256) zero1[i] = 0u;
  }
  double t = 0.0;
  for (int ch = threadIdx.x; ch < nparts; ch += 256) t += rootsums[((size_t)k * nparts + ch) * Q + q];
  red[threadIdx.x] = t;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if (threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x != 0) return;
  t = red[0];
  if (q == 0) {
    out_l[k] = t;
  } else if (q <= S) {
    if (out_f) out_f[(size_t)k * S + (q - 1)] = t;
  } else {
    if (out_p) out_p[(size_t)k * C + (q - 1 - S)] = t;
  }
}

template <typename R>
__global__ void k_reduce_g(const R* __restrict__ gpart, double* __restrict__ dmats, int K, int B,
                           int nchunks, int CSS) {
  const size_t idx = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  const size_t total = (size_t)K * B * CSS;
  if (idx >= total) return;
  const int q = (int)(idx % CSS);
  const size_t kb = idx / CSS;
  const int b = (int)(kb % B), k = (int)(kb / B);
  const R* src = gpart + (((size_t)b * K + k) * nchunks) * CSS + q;
  // eight loads in flight, then the same sequential sum
  double t = 0.0;
  int ch = 0;
  for (; ch + 8 <= nchunks; ch += 8) {
    R v[8];
#pragma unroll
    for (int u = 0; u < 8; ++u) v[u] = src[(size_t)(ch + u) * CSS];
#pragma unroll
    for (int u = 0; u < 8; ++u) t += (double)v[u];
  }
  for (; ch < nchunks; ++ch) t += (double)src[(size_t)ch * CSS];
  dmats[idx] = t;
}

__global__ void k_reduce_tau(const double* __restrict__ gtau, double* __restrict__ dtau, int K,
                             int B, int nchunks, int C) {
  const size_t idx = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= (size_t)K * B * C) return;
  const int c = (int)(idx % C);
  const size_t kb = idx / C;
  const int b = (int)(kb % B), k = (int)(kb / B);
  const double* src = gtau + (((size_t)b * K + k) * nchunks) * C + c;
  double t = 0.0;
  int ch = 0;
  for (; ch + 8 <= nchunks; ch += 8) {
    double v[8];
#pragma unroll
    for (int u = 0; u < 8; ++u) v[u] = src[(size_t)(ch + u) * C];
#pragma unroll
    for (int u = 0; u < 8; ++u) t += v[u];
  }
  for (; ch < nchunks; ++ch) t += src[(size_t)ch * C];
  dtau[idx] = t;
}

// P and P^T of every (sample, branch, category), zero-padded to 64 x 64, in the order the
// 64 lanes of a dense pre-order wave read them: part 0 holds M = P, part 1 holds M = P^T, and
// M[16 t + n][16 g + 4 w + e] sits at ppad[kbc][part][t][w][lane n + 16 g][e], so that each
// 16-byte load of a wave (fixed t, w) reads 1 KiB of consecutive bytes. One block per matrix.
// With `split` (FP32 only) blocks w = 2 b and 2 b + 1 instead hold the BF16 hi and lo halves
// (hi = bf16(x), lo = bf16(x - hi)) of M[16 t + n][16 g + 8 b + m], m = 0..7, as 8 BF16 values
// each: the two operand registers of a split-precision product, loaded as they are used.
__device__ __forceinline__ uint32_t pad_half(float x, bool lo) {
  const __bf16 hi = (__bf16)x;
  const __bf16 h = lo ? (__bf16)(x - (float)hi) : hi;
  return (uint32_t)__builtin_bit_cast(uint16_t, h);
}
template <typename R>
__global__ __launch_bounds__(256) void k_pad_mats(const R* __restrict__ mats, R* __restrict__ ppad,
                                                  int S, int split) {
  const size_t m = blockIdx.x;
  const R* src = mats + m * S * S;
  R* dst = ppad + m * 2 * 64 * 64;
  auto at = [&](int i, int j, bool tr) -> R {
    return (i < S && j < S) ? (tr ? src[j * S + i] : src[i * S + j]) : R(0);
  };
  for (int e = threadIdx.x; e < 64 * 64; e += 256) {
    // e = ((t 4 + w) 64 + lane) 4 + word
    const int lane = (e >> 2) & 63, t = e >> 10, w = (e >> 8) & 3, wd = e & 3;
    const int i = 16 * t + (lane & 15), g = lane >> 4;
    R p, pt;
    if constexpr (sizeof(R) == 4) {
      if (split) {
        // word wd of block w: elements m = 2 wd, 2 wd + 1 of pack w / 2, half w & 1
        const int j = 16 * g + 8 * (w >> 1) + 2 * wd;
        const bool lo = w & 1;
        p = __builtin_bit_cast(float, pad_half(at(i, j, false), lo) |
                                          (pad_half(at(i, j + 1, false), lo) << 16));
        pt = __builtin_bit_cast(float, pad_half(at(i, j, true), lo) |
                                           (pad_half(at(i, j + 1, true), lo) << 16));
        dst[e] = p;
        dst[64 * 64 + e] = pt;
        continue;
      }
    }
    const int j = 16 * g + 4 * w + wd;
    p = at(i, j, false);
    pt = at(i, j, true);
    dst[e] = p;
    dst[64 * 64 + e] = pt;
  }
}

// ------------------------------------------------------------------- host side
// (tiling and workspace layout: layout.hpp)
static WsLayout ws_layout(const hx_plan_s* plan, int S, int N, int C, int K, int dtype,
                          int grad_mode) {
  return ws_layout_t(plan->T, S, N, C, K, dtype, grad_mode);
}

template <typename R>
static int launch_dense(bool post, const PassArgs<R>& a, const int32_t* nodes, int width, int tips,
                        bool root, int S, hipStream_t st) {
  if (use_mfma(a.C, S, std::is_same<R, float>::value ? HX_F32 : HX_F64))
    return launch_mfma<R>(post, a, nodes, width, tips, root, S, st);
  return post ? launch_post_wide<R>(a, nodes, width, tips, root, S, st)
              : launch_pre_wide<R>(a, nodes, width, tips, root, S, st);
}

struct CallArgs {
  hx_plan_t plan;
  int S, N, C, K, dtype, tip_kind;
  const void* tips;
  const void* table;
  int ncodes;
  const double* weights;
  const void* mats;
  const void* freqs;
  int freqs_batched;
  const void* props;
  int props_batched;
  void* ws;
  size_t ws_bytes;
  int grad_mode;
  const void* mats_dt;
};

static int validate(const CallArgs& c) {
  HX_REQUIRE(c.plan != nullptr, HX_ERR_INVALID_ARGUMENT, "null plan");
  HX_REQUIRE(c.S >= 2 && c.S <= 64, HX_ERR_UNSUPPORTED, "state count must be 2..64");
  HX_REQUIRE(c.N >= 1 && c.C >= 1 && c.K >= 1, HX_ERR_INVALID_ARGUMENT, "empty dimension");
  if (c.S == 4) {
    HX_REQUIRE(c.C <= 16, HX_ERR_UNSUPPORTED, "more than 16 rate categories");
  } else {
    HX_REQUIRE(2 * c.C * (c.S <= 32 ? 32 : 64) <= 512, HX_ERR_UNSUPPORTED,
               "categories x padded states exceed the dense-state kernels (C*S_pad <= 256)");
  }
  HX_REQUIRE(dtype_ok(c.dtype), HX_ERR_INVALID_ARGUMENT, "bad dtype");
  HX_REQUIRE(c.tip_kind == HX_TIPS_CODES || c.tip_kind == HX_TIPS_DENSE, HX_ERR_INVALID_ARGUMENT,
             "bad tip kind");
  if (c.tip_kind == HX_TIPS_CODES)
    HX_REQUIRE(c.ncodes >= 1 && c.ncodes <= kMaxCodes && c.table != nullptr, HX_ERR_INVALID_ARGUMENT,
               "tip table must have 1..256 rows");
  HX_REQUIRE(c.tips && c.weights && c.mats && c.freqs && c.props && c.ws, HX_ERR_INVALID_ARGUMENT,
             "null pointer argument");
  const WsLayout w = ws_layout(c.plan, c.S, c.N, c.C, c.K, c.dtype, c.grad_mode);
  HX_REQUIRE(c.ws_bytes >= w.total, HX_ERR_WORKSPACE_TOO_SMALL, "workspace too small");
  return HX_OK;
}

template <typename R>
static PassArgs<R> make_args(const CallArgs& c) {
  const WsLayout w = ws_layout(c.plan, c.S, c.N, c.C, c.K, c.dtype, c.grad_mode);
  const Tiling t = make_tiling(c.N, c.C, c.K, c.S, c.dtype == HX_F64 ? 8 : 4);
  char* base = static_cast<char*>(c.ws);
  PassArgs<R> a;
  a.kids = c.plan->d_kids;
  a.T = c.plan->T;
  a.N = c.N;
  a.C = c.C;
  a.K = c.K;
  a.B = 2 * c.plan->T - 2;
  a.ntiles = t.ntiles;
  a.tiles_per_chunk = t.tiles_per_chunk;
  a.nchunks = t.nchunks;
  a.np = c.S == 4 ? t.ntiles * 64 : c.N;
  a.mats = static_cast<const R*>(c.mats);
  a.codes = c.tip_kind == HX_TIPS_CODES ? static_cast<const uint8_t*>(c.tips) : nullptr;
  a.table = static_cast<const R*>(c.table);
  a.ncodes = c.ncodes;
  a.dense = c.tip_kind == HX_TIPS_DENSE ? static_cast<const R*>(c.tips) : nullptr;
  a.partials = reinterpret_cast<R*>(base + w.partials);
  a.scales = reinterpret_cast<int32_t*>(base + w.scales);
  a.codes_pad = reinterpret_cast<uint8_t*>(base + w.scales);
  a.freqs = static_cast<const R*>(c.freqs);
  a.freqs_ks = c.freqs_batched ? c.S : 0;
  a.props = static_cast<const R*>(c.props);
  a.props_ks = c.props_batched ? c.C : 0;
  a.weights = c.weights;
  a.rootsums = reinterpret_cast<double*>(base + w.rootsums);
  a.gpart = reinterpret_cast<R*>(base + w.gpart);
  a.mats_dt = c.grad_mode == 2 ? static_cast<const R*>(c.mats_dt) : nullptr;
  a.gtau = reinterpret_cast<double*>(base + w.gpart);
  a.ppad = use_ppad(c.S, c.dtype, c.grad_mode) ? reinterpret_cast<R*>(base + w.ppad) : nullptr;
  a.split = c.dtype == HX_F32_BF16X3;
  a.drop_mode = walk_drop_mode(c.S, c.C, c.N, c.K, c.dtype, c.tip_kind);
  a.cher = c.plan->d_cher;
  a.ncher = c.plan->ncher;
  a.pcodes = c.S == 4 && sizeof(R) == 4 ? reinterpret_cast<uint16_t*>(base + w.pcodes) : nullptr;
  return a;
}

template <typename R>
static int run_post(const CallArgs& c, double* out_l, double* out_f, double* out_p, hipStream_t st) {
  const PassArgs<R> a = make_args<R>(c);
  const hx_plan_s* plan = c.plan;
  int nparts = a.nchunks;
  if (c.S == 4) {
    const int rc = launch_post_dna<R>(a, plan->d_walk, plan->d_segtab, plan->nseg, plan->I,
                                      plan->walk_need, c.tip_kind, st);
    if (rc != HX_OK) return rc;
    nparts = a.ntiles;
  } else {
    const int H = (int)plan->post_off.size() - 1;
    for (int l = 0; l < H; ++l) {
      const int off = plan->post_off[l], width = plan->post_off[l + 1] - off;
      const int rc =
          launch_dense<R>(true, a, plan->d_post + off, width, c.tip_kind, l == H - 1, c.S, st);
      if (rc != HX_OK) return rc;
    }
  }
  // dense states: the pre-order reads up to 16 C S + 64 elements past a slab
  const WsLayout w = ws_layout(c.plan, c.S, c.N, c.C, c.K, c.dtype, c.grad_mode);
  const size_t reach = c.S == 4 ? 0 : 16 * (size_t)c.C * c.S + 64;
  const size_t root_slabs = (size_t)c.K * c.N * c.C * c.S;
  uint32_t* tail = reinterpret_cast<uint32_t*>(static_cast<char*>(c.ws) + w.tail);
  uint32_t* head = reinterpret_cast<uint32_t*>(a.partials + (size_t)(plan->root - plan->T) * root_slabs);
  const int nwords = (int)(sizeof(R) / 4);
  k_reduce_root<<<dim3(c.K, 1 + c.S + c.C), 256, 0, st>>>(
      a.rootsums, c.K, nparts, c.S, c.C, out_l, out_f, out_p, tail, (int)reach * nwords, head,
      (int)(reach < root_slabs ? reach : root_slabs) * nwords);
  HX_HIP_TRY(hipGetLastError());
  return HX_OK;
}

template <typename R>
static int run_pre(const CallArgs& c, double* out, hipStream_t st) {
  const PassArgs<R> a = make_args<R>(c);
  const hx_plan_s* plan = c.plan;
  const int D = (int)plan->pre_off.size() - 1;
  if (a.ppad != nullptr) {
    k_pad_mats<R><<<(unsigned)((size_t)c.K * a.B * c.C), 256, 0, st>>>(a.mats, a.ppad, c.S,
                                                                         a.split);
    HX_HIP_TRY(hipGetLastError());
  }
  if constexpr (std::is_same<R, float>::value) {
    if (use_pre_walk(c.S, c.C, 4, c.N, c.K)) {
      const int rc = launch_pre_walk(a, plan->d_prew, plan->I, c.tip_kind, st);
      if (rc != HX_OK) return rc;
    }
  }
  for (int d = 0; d < D && !(sizeof(R) == 4 && use_pre_walk(c.S, c.C, 4, c.N, c.K)); ++d) {
    const int off = plan->pre_off[d], width = plan->pre_off[d + 1] - off;
    const int rc = c.S == 4
                       ? launch_pre_dna<R>(a, plan->d_pre + off, width, c.tip_kind, d == 0, st)
                       : launch_dense<R>(false, a, plan->d_pre + off, width, c.tip_kind, d == 0,
                                         c.S, st);
    if (rc != HX_OK) return rc;
  }
  if (c.grad_mode == 2) {
    const size_t total = (size_t)c.K * a.B * c.C;
    k_reduce_tau<<<(unsigned)((total + 255) / 256), 256, 0, st>>>(a.gtau, out, c.K, a.B, a.nchunks,
                                                                  c.C);
  } else {
    const int CSS = c.C * c.S * c.S;
    const size_t total = (size_t)c.K * a.B * CSS;
    k_reduce_g<R><<<(unsigned)((total + 255) / 256), 256, 0, st>>>(a.gpart, out, c.K, a.B,
                                                                   a.nchunks, CSS);
  }
  HX_HIP_TRY(hipGetLastError());
  return HX_OK;
}

}  // namespace hx

extern "C" {

size_t hx_workspace_bytes(hx_plan_t plan, int S, int N, int C, int K, int dtype, int grad_mode) {
  if (!plan || grad_mode < 0 || grad_mode > 2) return 0;
  return hx::ws_layout(plan, S, N, C, K, dtype, grad_mode).total;
}

int hx_loglik(hx_plan_t plan, int S, int N, int C, int K, int dtype, int tip_kind, const void* tips,
              const void* tip_table, int ncodes, const double* weights, const void* mats,
              const void* freqs, int freqs_batched, const void* props, int props_batched,
              void* workspace, size_t workspace_bytes, double* out_loglik, double* out_dfreqs,
              double* out_dprops, hx_stream_t stream) {
  hx::CallArgs c{plan, S, N, C, K, dtype, tip_kind, tips, tip_table, ncodes, weights, mats,
                 freqs, freqs_batched, props, props_batched, workspace, workspace_bytes, 0, nullptr};
  int rc = hx::validate(c);
  if (rc != HX_OK) return rc;
  HX_REQUIRE(out_loglik != nullptr, HX_ERR_INVALID_ARGUMENT, "null output");
  hipStream_t st = reinterpret_cast<hipStream_t>(stream);
  if (dtype == HX_F64) return hx::run_post<double>(c, out_loglik, out_dfreqs, out_dprops, st);
  return hx::run_post<float>(c, out_loglik, out_dfreqs, out_dprops, st);
}

int hx_loglik_grad_mats(hx_plan_t plan, int S, int N, int C, int K, int dtype, int tip_kind,
                        const void* tips, const void* tip_table, int ncodes, const double* weights,
                        const void* mats, const void* freqs, int freqs_batched, const void* props,
                        int props_batched, void* workspace, size_t workspace_bytes,
                        double* out_dmats, hx_stream_t stream) {
  hx::CallArgs c{plan, S, N, C, K, dtype, tip_kind, tips, tip_table, ncodes, weights, mats,
                 freqs, freqs_batched, props, props_batched, workspace, workspace_bytes, 1, nullptr};
  int rc = hx::validate(c);
  if (rc != HX_OK) return rc;
  HX_REQUIRE(out_dmats != nullptr, HX_ERR_INVALID_ARGUMENT, "null output");
  hipStream_t st = reinterpret_cast<hipStream_t>(stream);
  if (dtype == HX_F64) return hx::run_pre<double>(c, out_dmats, st);
  return hx::run_pre<float>(c, out_dmats, st);
}

int hx_loglik_grad_tau(hx_plan_t plan, int S, int N, int C, int K, int dtype, int tip_kind,
                       const void* tips, const void* tip_table, int ncodes, const double* weights,
                       const void* mats, const void* mats_dtau, const void* freqs,
                       int freqs_batched, const void* props, int props_batched, void* workspace,
                       size_t workspace_bytes, double* out_dtau, hx_stream_t stream) {
  hx::CallArgs c{plan, S, N, C, K, dtype, tip_kind, tips, tip_table, ncodes, weights, mats,
                 freqs, freqs_batched, props, props_batched, workspace, workspace_bytes, 2,
                 mats_dtau};
  int rc = hx::validate(c);
  if (rc != HX_OK) return rc;
  HX_REQUIRE(out_dtau != nullptr && mats_dtau != nullptr, HX_ERR_INVALID_ARGUMENT,
             "null output or dP/dtau");
  hipStream_t st = reinterpret_cast<hipStream_t>(stream);
  if (dtype == HX_F64) return hx::run_pre<double>(c, out_dtau, st);
  return hx::run_pre<float>(c, out_dtau, st);
}

}  // extern "C"

// Felsenstein pruning on gfx950: post-order partials, fused root reduction and
// the pre-order adjoint that yields d loglik / d mats.
//
// Reference: torchtree/evolution/tree_likelihood.py (calculate_treelikelihood*,
// lines 14-278). The reference walks post_indexing in Python and issues two
// batched matmuls and one multiply per node; here one launch covers every
// node of a tree level for every sample, and each thread owns one
// (sample, pattern, category) lane for the whole level.
//
// Layout in HBM (R = float or double):
//   partials [slot][K][N][C][S]   slot = node - T, pattern-major so the C*S
//                                 values of one pattern are contiguous and a
//                                 wave reads 64 consecutive S-vectors
//   scales   [slot][K][N] int32   cumulative power-of-two exponent of the
//                                 subtree below the node
//   rootsums [K][chunk][1+S+C]    per-chunk FP64 partial sums at the root
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

template <typename R, int S>
__device__ __forceinline__ void vload(const R* __restrict__ p, R (&x)[S]) {
  if constexpr ((S * sizeof(R)) % 16 == 0) {
    const uint4* q = reinterpret_cast<const uint4*>(p);
#pragma unroll
    for (int i = 0; i < (int)(S * sizeof(R) / 16); ++i) {
      uint4 v = q[i];
      __builtin_memcpy(reinterpret_cast<char*>(x) + 16 * i, &v, 16);
    }
  } else {
#pragma unroll
    for (int s = 0; s < S; ++s) x[s] = p[s];
  }
}

template <typename R, int S>
__device__ __forceinline__ void vstore(R* __restrict__ p, const R (&x)[S]) {
  if constexpr ((S * sizeof(R)) % 16 == 0) {
    uint4* q = reinterpret_cast<uint4*>(p);
#pragma unroll
    for (int i = 0; i < (int)(S * sizeof(R) / 16); ++i) {
      uint4 v;
      __builtin_memcpy(&v, reinterpret_cast<const char*>(x) + 16 * i, 16);
      q[i] = v;
    }
  } else {
#pragma unroll
    for (int s = 0; s < S; ++s) p[s] = x[s];
  }
}

__device__ __forceinline__ float fmax_(float a, float b) { return fmaxf(a, b); }
__device__ __forceinline__ double fmax_(double a, double b) { return fmax(a, b); }
__device__ __forceinline__ float frexp_(float m, int* e) { return frexpf(m, e); }
__device__ __forceinline__ double frexp_(double m, int* e) { return frexp(m, e); }
__device__ __forceinline__ float ldexp_(float x, int e) { return ldexpf(x, e); }
__device__ __forceinline__ double ldexp_(double x, int e) { return ldexp(x, e); }

// Partial of child `ch` for lane (k, n, c); zero for padding lanes.
template <typename R, int S, int TIPS>
__device__ __forceinline__ void load_child(const PassArgs<R>& a, const R* tbl, int ch, int k, int n,
                                           int c, bool val, R (&x)[S], int& E) {
  E = 0;
  if (!val) {
#pragma unroll
    for (int s = 0; s < S; ++s) x[s] = R(0);
    return;
  }
  if (ch < a.T) {
    if constexpr (TIPS == HX_TIPS_CODES) {
      const int code = a.codes[(size_t)ch * a.N + n];
#pragma unroll
      for (int s = 0; s < S; ++s) x[s] = tbl[code * S + s];
    } else {
      vload<R, S>(a.dense + ((size_t)ch * a.N + n) * S, x);
    }
  } else {
    const size_t row = ((size_t)(ch - a.T) * a.K + k) * a.N + n;
    vload<R, S>(a.partials + row * a.C * S + (size_t)c * S, x);
    E = a.scales[row];
  }
}

template <typename R, int S>
__device__ __forceinline__ void matvec(const R (&P)[S * S], const R (&x)[S], R (&y)[S]) {
#pragma unroll
  for (int i = 0; i < S; ++i) {
    R acc = R(0);
#pragma unroll
    for (int j = 0; j < S; ++j) acc = fma(P[i * S + j], x[j], acc);
    y[i] = acc;
  }
}

template <typename R, int S>
__device__ __forceinline__ void matvec_t(const R (&P)[S * S], const R (&x)[S], R (&y)[S]) {
#pragma unroll
  for (int j = 0; j < S; ++j) {
    R acc = R(0);
#pragma unroll
    for (int i = 0; i < S; ++i) acc = fma(P[i * S + j], x[i], acc);
    y[j] = acc;
  }
}

// Bring the (C,S)-max of one pattern into [0.5, 1) by a power of two; the CP
// lanes of a pattern are adjacent. Returns the exponent removed.
template <typename R, int S, int CP>
__device__ __forceinline__ int rescale(R (&x)[S]) {
  R m = R(0);
#pragma unroll
  for (int s = 0; s < S; ++s) m = fmax_(m, x[s]);
#pragma unroll
  for (int o = 1; o < CP; o <<= 1) m = fmax_(m, __shfl_xor(m, o));
  int e = 0;
  if (m > R(0) && isfinite(m)) {
    frexp_(m, &e);
#pragma unroll
    for (int s = 0; s < S; ++s) x[s] = ldexp_(x[s], -e);
  }
  return e;
}

template <typename R, int S>
__device__ __forceinline__ void load_mat(const R* __restrict__ p, R (&P)[S * S]) {
#pragma unroll
  for (int i = 0; i < S * S; ++i) P[i] = p[i];
}

template <typename R, int S, int TIPS>
__device__ __forceinline__ void stage_table(const PassArgs<R>& a, R* tbl) {
  if constexpr (TIPS == HX_TIPS_CODES) {
    for (int i = threadIdx.x; i < a.ncodes * S; i += kBlock) tbl[i] = a.table[i];
    __syncthreads();
  }
}

// Sum over the pattern lanes of a wave (lanes that share c), then over waves
// through LDS. On return lanes 0..CP-1 of the block hold the totals in `v`
// (indexed by their c).
template <int CP, int NV>
__device__ __forceinline__ void block_sum_by_c(double (&v)[NV], double* red) {
  const int lane = threadIdx.x % kWave, wave = threadIdx.x / kWave;
#pragma unroll
  for (int q = 0; q < NV; ++q)
#pragma unroll
    for (int o = CP; o < kWave; o <<= 1) v[q] += __shfl_xor(v[q], o);
  if (lane < CP)
#pragma unroll
    for (int q = 0; q < NV; ++q) red[(wave * CP + lane) * NV + q] = v[q];
  __syncthreads();
  if (threadIdx.x < CP) {
#pragma unroll
    for (int q = 0; q < NV; ++q) {
      double t = 0.0;
      for (int w = 0; w < kBlock / kWave; ++w) t += red[(w * CP + threadIdx.x) * NV + q];
      v[q] = t;
    }
  }
}

// ---------------------------------------------------------------- post-order
template <typename R, int S, int CP, int TIPS, bool ROOT>
__global__ __launch_bounds__(kBlock) void k_post(PassArgs<R> a, const int32_t* __restrict__ nodes) {
  constexpr int PPB = kBlock / CP;
  __shared__ R tbl[TIPS == HX_TIPS_CODES ? kMaxCodes * S : 1];
  __shared__ double red[ROOT ? (kBlock / kWave) * CP * (2 + S) : 1];
  const int c = threadIdx.x % CP, pl = threadIdx.x / CP;
  const int chunk = blockIdx.x, k = blockIdx.y;
  const int v = nodes[blockIdx.z];
  const int left = a.kids[2 * v], right = a.kids[2 * v + 1];
  const bool cval = c < a.C;
  const int cc = cval ? c : 0;
  stage_table<R, S, TIPS>(a, tbl);

  R Pl[S * S], Pr[S * S];
  load_mat<R, S>(a.mats + (((size_t)k * a.B + left) * a.C + cc) * S * S, Pl);
  load_mat<R, S>(a.mats + (((size_t)k * a.B + right) * a.C + cc) * S * S, Pr);

  R pi[S];
  R pc = R(0);
  double acc[2 + S];  // log-lik, dfreqs[S], dprops (this lane's c)
  if constexpr (ROOT) {
#pragma unroll
    for (int s = 0; s < S; ++s) pi[s] = a.freqs[(size_t)k * a.freqs_ks + s];
    pc = cval ? a.props[(size_t)k * a.props_ks + c] : R(0);
#pragma unroll
    for (int q = 0; q < 2 + S; ++q) acc[q] = 0.0;
  }

  const int t0 = chunk * a.tiles_per_chunk;
  const int t1 = min(t0 + a.tiles_per_chunk, a.ntiles);
  for (int tile = t0; tile < t1; ++tile) {
    const int n = tile * PPB + pl;
    const bool nval = n < a.N;
    const bool val = nval && cval;
    R xl[S], xr[S], yl[S], yr[S], x[S];
    int El, Er;
    load_child<R, S, TIPS>(a, tbl, left, k, n, cc, val, xl, El);
    load_child<R, S, TIPS>(a, tbl, right, k, n, cc, val, xr, Er);
    matvec<R, S>(Pl, xl, yl);
    matvec<R, S>(Pr, xr, yr);
#pragma unroll
    for (int s = 0; s < S; ++s) x[s] = yl[s] * yr[s];
    const int e = rescale<R, S, CP>(x);
    const int E = El + Er + e;
    if constexpr (!ROOT) {
      const size_t row = ((size_t)(v - a.T) * a.K + k) * a.N + n;
      if (val) vstore<R, S>(a.partials + row * a.C * S + (size_t)c * S, x);
      if (nval && c == 0) a.scales[row] = E;
    } else {
      // site likelihood sum_c p_c pi.x_c (tree_likelihood.py:71-74, 215-220)
      R px = R(0);
#pragma unroll
      for (int s = 0; s < S; ++s) px = fma(pi[s], x[s], px);
      R lik = pc * px;
#pragma unroll
      for (int o = 1; o < CP; o <<= 1) lik += __shfl_xor(lik, o);
      if (val) {
        const double w = a.weights[n];
        const double dl = (double)lik;
        const double wi = w / dl;
        if (c == 0) acc[0] += w * (log(dl) + (double)E * kLn2);
#pragma unroll
        for (int s = 0; s < S; ++s) acc[1 + s] += wi * (double)(pc * x[s]);
        acc[1 + S] += wi * (double)px;
      }
    }
  }
  if constexpr (ROOT) {
    block_sum_by_c<CP, 2 + S>(acc, red);
    __shared__ double tot[2 + S];
    if (threadIdx.x == 0)
      for (int q = 0; q < 1 + S; ++q) tot[q] = 0.0;
    __syncthreads();
    if (threadIdx.x < CP) {
      for (int q = 0; q < 1 + S; ++q) atomicAdd(&tot[q], acc[q]);
    }
    __syncthreads();
    double* out = a.rootsums + ((size_t)k * a.nchunks + chunk) * (1 + S + a.C);
    if (threadIdx.x == 0)
      for (int q = 0; q < 1 + S; ++q) out[q] = tot[q];
    if (threadIdx.x < a.C) out[1 + S + threadIdx.x] = acc[1 + S];
  }
}

// ----------------------------------------------------------------- pre-order
// For parent p with children a, b and pre-order partial U_p (the root uses
// freqs): Ut_a = U_p * (P_b L_b), Ut_b = U_p * (P_a L_a); the pattern's
// scaled likelihood is sum_c p_c Ut_a.(P_a L_a); the branch statistic
// d loglik / d P_a[c][i][j] += w p_c Ut_a[i] L_a[j] / lik; the children's
// pre-order partials P_a^T Ut_a overwrite their post-order partials.
template <typename R, int S, int CP, int TIPS, bool ROOT>
__global__ __launch_bounds__(kBlock) void k_pre(PassArgs<R> a, const int32_t* __restrict__ nodes) {
  constexpr int PPB = kBlock / CP;
  constexpr int NG = 2 * S * S;
  __shared__ R tbl[TIPS == HX_TIPS_CODES ? kMaxCodes * S : 1];
  __shared__ double red[(kBlock / kWave) * CP * NG];
  const int c = threadIdx.x % CP, pl = threadIdx.x / CP;
  const int chunk = blockIdx.x, k = blockIdx.y;
  const int p = nodes[blockIdx.z];
  const int ca = a.kids[2 * p], cb = a.kids[2 * p + 1];
  const bool cval = c < a.C;
  const int cc = cval ? c : 0;
  stage_table<R, S, TIPS>(a, tbl);

  R Pa[S * S], Pb[S * S];
  load_mat<R, S>(a.mats + (((size_t)k * a.B + ca) * a.C + cc) * S * S, Pa);
  load_mat<R, S>(a.mats + (((size_t)k * a.B + cb) * a.C + cc) * S * S, Pb);
  const R pc = cval ? a.props[(size_t)k * a.props_ks + c] : R(0);
  R pi[S];
  if constexpr (ROOT) {
#pragma unroll
    for (int s = 0; s < S; ++s) pi[s] = a.freqs[(size_t)k * a.freqs_ks + s];
  }
  R ga[S * S], gb[S * S];
#pragma unroll
  for (int i = 0; i < S * S; ++i) ga[i] = gb[i] = R(0);

  const int t0 = chunk * a.tiles_per_chunk;
  const int t1 = min(t0 + a.tiles_per_chunk, a.ntiles);
  for (int tile = t0; tile < t1; ++tile) {
    const int n = tile * PPB + pl;
    const bool nval = n < a.N;
    const bool val = nval && cval;
    R u[S], la[S], lb[S], pla[S], plb[S], ua[S], ub[S];
    int dummy;
    if constexpr (ROOT) {
#pragma unroll
      for (int s = 0; s < S; ++s) u[s] = val ? pi[s] : R(0);
    } else {
      const size_t row = ((size_t)(p - a.T) * a.K + k) * a.N + n;
      if (val) {
        vload<R, S>(a.partials + row * a.C * S + (size_t)c * S, u);
      } else {
#pragma unroll
        for (int s = 0; s < S; ++s) u[s] = R(0);
      }
    }
    load_child<R, S, TIPS>(a, tbl, ca, k, n, cc, val, la, dummy);
    load_child<R, S, TIPS>(a, tbl, cb, k, n, cc, val, lb, dummy);
    matvec<R, S>(Pa, la, pla);
    matvec<R, S>(Pb, lb, plb);
    R d = R(0);
#pragma unroll
    for (int s = 0; s < S; ++s) {
      ua[s] = u[s] * plb[s];
      ub[s] = u[s] * pla[s];
      d = fma(ua[s], pla[s], d);
    }
    d *= pc;
#pragma unroll
    for (int o = 1; o < CP; o <<= 1) d += __shfl_xor(d, o);
    const R coef = val ? (R)a.weights[n] / d * pc : R(0);
#pragma unroll
    for (int i = 0; i < S; ++i) {
      const R wa = coef * ua[i], wb = coef * ub[i];
#pragma unroll
      for (int j = 0; j < S; ++j) {
        ga[i * S + j] = fma(wa, la[j], ga[i * S + j]);
        gb[i * S + j] = fma(wb, lb[j], gb[i * S + j]);
      }
    }
    if (ca >= a.T) {
      R na[S];
      matvec_t<R, S>(Pa, ua, na);
      rescale<R, S, CP>(na);
      const size_t row = ((size_t)(ca - a.T) * a.K + k) * a.N + n;
      if (val) vstore<R, S>(a.partials + row * a.C * S + (size_t)c * S, na);
    }
    if (cb >= a.T) {
      R nb[S];
      matvec_t<R, S>(Pb, ub, nb);
      rescale<R, S, CP>(nb);
      const size_t row = ((size_t)(cb - a.T) * a.K + k) * a.N + n;
      if (val) vstore<R, S>(a.partials + row * a.C * S + (size_t)c * S, nb);
    }
  }
  double g[NG];
#pragma unroll
  for (int i = 0; i < S * S; ++i) {
    g[i] = (double)ga[i];
    g[S * S + i] = (double)gb[i];
  }
  block_sum_by_c<CP, NG>(g, red);
  if (a.mats_dt != nullptr) {
    // d loglik / d tau_{b,c} = sum_ij G_bc[i][j] dP_bc/dtau[i][j]
    if (threadIdx.x < a.C) {
      const R* qa = a.mats_dt + (((size_t)k * a.B + ca) * a.C + threadIdx.x) * S * S;
      const R* qb = a.mats_dt + (((size_t)k * a.B + cb) * a.C + threadIdx.x) * S * S;
      double sa = 0.0, sb = 0.0;
#pragma unroll
      for (int i = 0; i < S * S; ++i) {
        sa = fma(g[i], (double)qa[i], sa);
        sb = fma(g[S * S + i], (double)qb[i], sb);
      }
      a.gtau[(((size_t)ca * a.K + k) * a.nchunks + chunk) * a.C + threadIdx.x] = sa;
      a.gtau[(((size_t)cb * a.K + k) * a.nchunks + chunk) * a.C + threadIdx.x] = sb;
    }
  } else if (threadIdx.x < a.C) {
    const size_t css = (size_t)a.C * S * S;
    R* oa = a.gpart + (((size_t)ca * a.K + k) * a.nchunks + chunk) * css + (size_t)threadIdx.x * S * S;
    R* ob = a.gpart + (((size_t)cb * a.K + k) * a.nchunks + chunk) * css + (size_t)threadIdx.x * S * S;
#pragma unroll
    for (int i = 0; i < S * S; ++i) {
      oa[i] = (R)g[i];
      ob[i] = (R)g[S * S + i];
    }
  }
}

// ------------------------------------------------------------ chunk reductions
__global__ void k_reduce_root(const double* __restrict__ rootsums, int K, int nchunks, int S, int C,
                              double* out_l, double* out_f, double* out_p) {
  const int Q = 1 + S + C;
  const int idx = blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= K * Q) return;
  const int k = idx / Q, q = idx % Q;
  double t = 0.0;
  for (int ch = 0; ch < nchunks; ++ch) t += rootsums[((size_t)k * nchunks + ch) * Q + q];
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
  double t = 0.0;
  for (int ch = 0; ch < nchunks; ++ch) t += (double)src[(size_t)ch * CSS];
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
  for (int ch = 0; ch < nchunks; ++ch) t += src[(size_t)ch * C];
  dtau[idx] = t;
}

// ------------------------------------------------------------------- host side
struct WsLayout {
  size_t partials, scales, rootsums, gpart, total;
};

// grad_mode: 0 loglik only, 1 d loglik / d mats, 2 d loglik / d tau.
static WsLayout ws_layout(const hx_plan_s* plan, int S, int N, int C, int K, int dtype,
                          int grad_mode) {
  const size_t rb = dtype == HX_F64 ? 8 : 4;
  const Tiling t = make_tiling(N, C, K, S, (int)rb);
  const size_t I = plan->I, B = 2 * (size_t)plan->T - 2;
  WsLayout w;
  size_t off = 0;
  w.partials = off;
  off = align_up(off + I * K * (size_t)N * C * S * rb, 256);
  w.scales = off;
  off = align_up(off + I * K * (size_t)N * 4, 256);
  w.rootsums = off;
  off = align_up(off + (size_t)K * t.nchunks * (1 + S + C) * 8, 256);
  w.gpart = off;
  if (grad_mode == 1) off = align_up(off + B * K * (size_t)t.nchunks * C * S * S * rb, 256);
  if (grad_mode == 2) off = align_up(off + B * K * (size_t)t.nchunks * C * 8, 256);
  w.total = off;
  return w;
}

template <typename R, int S, int CP>
static void launch_post(const PassArgs<R>& a, const int32_t* nodes, int width, int tips, bool root,
                        hipStream_t st) {
  dim3 grid(a.nchunks, a.K, width);
  if (tips == HX_TIPS_CODES) {
    if (root) k_post<R, S, CP, HX_TIPS_CODES, true><<<grid, kBlock, 0, st>>>(a, nodes);
    else k_post<R, S, CP, HX_TIPS_CODES, false><<<grid, kBlock, 0, st>>>(a, nodes);
  } else {
    if (root) k_post<R, S, CP, HX_TIPS_DENSE, true><<<grid, kBlock, 0, st>>>(a, nodes);
    else k_post<R, S, CP, HX_TIPS_DENSE, false><<<grid, kBlock, 0, st>>>(a, nodes);
  }
}

template <typename R, int S, int CP>
static void launch_pre(const PassArgs<R>& a, const int32_t* nodes, int width, int tips, bool root,
                       hipStream_t st) {
  dim3 grid(a.nchunks, a.K, width);
  if (tips == HX_TIPS_CODES) {
    if (root) k_pre<R, S, CP, HX_TIPS_CODES, true><<<grid, kBlock, 0, st>>>(a, nodes);
    else k_pre<R, S, CP, HX_TIPS_CODES, false><<<grid, kBlock, 0, st>>>(a, nodes);
  } else {
    if (root) k_pre<R, S, CP, HX_TIPS_DENSE, true><<<grid, kBlock, 0, st>>>(a, nodes);
    else k_pre<R, S, CP, HX_TIPS_DENSE, false><<<grid, kBlock, 0, st>>>(a, nodes);
  }
}

template <typename R, int S>
static int dispatch_cp(bool post, const PassArgs<R>& a, const int32_t* nodes, int width, int tips,
                       bool root, hipStream_t st) {
  const int cp = next_pow2(a.C);
  switch (cp) {
#define HX_CASE(CPV)                                                      \
  case CPV:                                                               \
    if (post) launch_post<R, S, CPV>(a, nodes, width, tips, root, st);    \
    else launch_pre<R, S, CPV>(a, nodes, width, tips, root, st);          \
    break;
    HX_CASE(1)
    HX_CASE(2)
    HX_CASE(4)
    HX_CASE(8)
    HX_CASE(16)
#undef HX_CASE
    default:
      set_error("more than 16 rate categories");
      return HX_ERR_UNSUPPORTED;
  }
  return HX_OK;
}

template <typename R>
static int launch_dense(bool post, const PassArgs<R>& a, const int32_t* nodes, int width, int tips,
                        bool root, int S, hipStream_t st) {
  if constexpr (std::is_same<R, float>::value) {
    if (use_mfma(a.C, S, HX_F32)) return launch_mfma(post, a, nodes, width, tips, root, S, st);
  }
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
  HX_REQUIRE(c.dtype == HX_F32 || c.dtype == HX_F64, HX_ERR_INVALID_ARGUMENT, "bad dtype");
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
  a.mats = static_cast<const R*>(c.mats);
  a.codes = c.tip_kind == HX_TIPS_CODES ? static_cast<const uint8_t*>(c.tips) : nullptr;
  a.table = static_cast<const R*>(c.table);
  a.ncodes = c.ncodes;
  a.dense = c.tip_kind == HX_TIPS_DENSE ? static_cast<const R*>(c.tips) : nullptr;
  a.partials = reinterpret_cast<R*>(base + w.partials);
  a.scales = reinterpret_cast<int32_t*>(base + w.scales);
  a.freqs = static_cast<const R*>(c.freqs);
  a.freqs_ks = c.freqs_batched ? c.S : 0;
  a.props = static_cast<const R*>(c.props);
  a.props_ks = c.props_batched ? c.C : 0;
  a.weights = c.weights;
  a.rootsums = reinterpret_cast<double*>(base + w.rootsums);
  a.gpart = reinterpret_cast<R*>(base + w.gpart);
  a.mats_dt = c.grad_mode == 2 ? static_cast<const R*>(c.mats_dt) : nullptr;
  a.gtau = reinterpret_cast<double*>(base + w.gpart);
  return a;
}

template <typename R>
static int run_post(const CallArgs& c, double* out_l, double* out_f, double* out_p, hipStream_t st) {
  const PassArgs<R> a = make_args<R>(c);
  const hx_plan_s* plan = c.plan;
  const int H = (int)plan->post_off.size() - 1;
  for (int l = 0; l < H; ++l) {
    const int off = plan->post_off[l], width = plan->post_off[l + 1] - off;
    int rc = c.S == 4
                 ? dispatch_cp<R, 4>(true, a, plan->d_post + off, width, c.tip_kind, l == H - 1, st)
                 : launch_dense<R>(true, a, plan->d_post + off, width, c.tip_kind, l == H - 1, c.S, st);
    if (rc != HX_OK) return rc;
  }
  const int Q = 1 + c.S + c.C;
  const int total = c.K * Q;
  k_reduce_root<<<(total + 255) / 256, 256, 0, st>>>(a.rootsums, c.K, a.nchunks, c.S, c.C, out_l,
                                                     out_f, out_p);
  HX_HIP_TRY(hipGetLastError());
  return HX_OK;
}

template <typename R>
static int run_pre(const CallArgs& c, double* out, hipStream_t st) {
  const PassArgs<R> a = make_args<R>(c);
  const hx_plan_s* plan = c.plan;
  const int D = (int)plan->pre_off.size() - 1;
  for (int d = 0; d < D; ++d) {
    const int off = plan->pre_off[d], width = plan->pre_off[d + 1] - off;
    int rc = c.S == 4
                 ? dispatch_cp<R, 4>(false, a, plan->d_pre + off, width, c.tip_kind, d == 0, st)
                 : launch_dense<R>(false, a, plan->d_pre + off, width, c.tip_kind, d == 0, c.S, st);
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

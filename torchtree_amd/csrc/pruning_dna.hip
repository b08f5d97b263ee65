// Four-state (nucleotide) pruning kernels: one fused post-order walk per
// pattern tile and a pre-order adjoint that covers two tree levels per launch.
//
// Reference: calculate_treelikelihood_discrete* (torchtree/evolution/
// tree_likelihood.py:40-278) at S = 4, where each node update is two 4x4
// matrix-vector products per (sample, pattern, category): HBM-bound work.
//
// Block = C waves; wave c handles rate category c and lane l pattern
// 64 * tile + l, so the P matrices of a wave are uniform (scalar loads, SGPR
// operands) and every partial access of a wave is one contiguous 1 KiB (FP32)
// run of the category-major layout
//   partials [slot][K][C][N][4]   slot = node - T
// The (C, S)-maximum that rescales a pattern and the per-pattern likelihood
// are exchanged between the C waves through LDS.
//
// Post-order: one block walks the whole tree for its 64-pattern tile. Each
// internal partial is written to HBM once (the adjoint needs it) and kept on
// a per-thread LDS stack until its parent consumes it, so no partial is read
// back from HBM. The walk order (larger-need child first) bounds the stack at
// about log2(T) entries; entries beyond the LDS depth `dl` are read back from
// HBM with their exponent from `scales`.
//
// Pre-order: a block takes one node p at an even depth and handles the
// branches to its children and grandchildren. U at the children stays in
// registers, so only every second level of pre-order partials touches HBM,
// and each branch statistic d loglik / d P accumulates in registers over
// the block's chunk of tiles.
#include "pruning.hpp"

#include <math.h>

#include <type_traits>

namespace hx {

namespace {

template <typename R>
struct alignas(4 * sizeof(R)) V4 {
  R x[4];
};

template <typename R>
__device__ __forceinline__ V4<R> ld4(const R* p) {
  V4<R> v;
  if constexpr (sizeof(R) == 4) {
    const float4 q = *reinterpret_cast<const float4*>(p);
    v.x[0] = q.x, v.x[1] = q.y, v.x[2] = q.z, v.x[3] = q.w;
  } else {
    const double2 q0 = reinterpret_cast<const double2*>(p)[0];
    const double2 q1 = reinterpret_cast<const double2*>(p)[1];
    v.x[0] = q0.x, v.x[1] = q0.y, v.x[2] = q1.x, v.x[3] = q1.y;
  }
  return v;
}

template <typename R>
__device__ __forceinline__ void st4(R* p, const V4<R>& v) {
  if constexpr (sizeof(R) == 4) {
    *reinterpret_cast<float4*>(p) = make_float4(v.x[0], v.x[1], v.x[2], v.x[3]);
  } else {
    reinterpret_cast<double2*>(p)[0] = make_double2(v.x[0], v.x[1]);
    reinterpret_cast<double2*>(p)[1] = make_double2(v.x[2], v.x[3]);
  }
}

template <typename R>
__device__ __forceinline__ V4<R> zero4() {
  V4<R> v;
#pragma unroll
  for (int s = 0; s < 4; ++s) v.x[s] = R(0);
  return v;
}

__device__ __forceinline__ int uni(int x) { return __builtin_amdgcn_readfirstlane(x); }

__device__ __forceinline__ float fmax_(float a, float b) { return fmaxf(a, b); }
__device__ __forceinline__ double fmax_(double a, double b) { return fmax(a, b); }
__device__ __forceinline__ float frexp_(float m, int* e) { return frexpf(m, e); }
__device__ __forceinline__ double frexp_(double m, int* e) { return frexp(m, e); }
__device__ __forceinline__ float ldexp_(float x, int e) { return ldexpf(x, e); }
__device__ __forceinline__ double ldexp_(double x, int e) { return ldexp(x, e); }

// P of (sample k, branch b, category c), a wave-uniform 4x4 block.
template <typename R>
__device__ __forceinline__ void load_p(const R* __restrict__ mats, int B, int C, int k, int b, int c,
                                       R (&P)[16]) {
  const R* p = mats + (((size_t)k * B + b) * C + c) * 16;
#pragma unroll
  for (int i = 0; i < 16; ++i) P[i] = p[i];
}

template <typename R>
__device__ __forceinline__ V4<R> mv(const R (&P)[16], const V4<R>& x) {
  V4<R> y;
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    R acc = R(0);
#pragma unroll
    for (int j = 0; j < 4; ++j) acc = fma(P[4 * i + j], x.x[j], acc);
    y.x[i] = acc;
  }
  return y;
}

template <typename R>
__device__ __forceinline__ V4<R> mv_t(const R (&P)[16], const V4<R>& x) {
  V4<R> y;
#pragma unroll
  for (int j = 0; j < 4; ++j) {
    R acc = R(0);
#pragma unroll
    for (int i = 0; i < 4; ++i) acc = fma(P[4 * i + j], x.x[i], acc);
    y.x[j] = acc;
  }
  return y;
}

template <typename R>
__device__ __forceinline__ R max4(const V4<R>& v) {
  return fmax_(fmax_(v.x[0], v.x[1]), fmax_(v.x[2], v.x[3]));
}

// Exponent that brings the pattern maximum M into [0.5, 1); 0 for M = 0 or
// non-finite M (the partial then carries its value unscaled).
template <typename R>
__device__ __forceinline__ int scale_exp(R M) {
  int e = 0;
  if (M > R(0) && isfinite(M)) frexp_(M, &e);
  return e;
}

template <typename R>
__device__ __forceinline__ void scale4(V4<R>& v, int e) {
  if (e != 0) {
#pragma unroll
    for (int s = 0; s < 4; ++s) v.x[s] = ldexp_(v.x[s], -e);
  }
}

__device__ __forceinline__ size_t prow(int T, int K, int C, int N, int v, int k, int c) {
  return (((size_t)(v - T) * K + k) * C + c) * (size_t)N;
}

// Tip partial of tip `ch` at pattern n (zero for padding lanes).
template <typename R, int TIPS>
__device__ __forceinline__ V4<R> tip4(const PassArgs<R>& a, const V4<R>* tbl, int ch, int n,
                                      bool val) {
  if (!val) return zero4<R>();
  if constexpr (TIPS == HX_TIPS_CODES) {
    return tbl[a.codes[(size_t)ch * a.N + n]];
  } else {
    return ld4(a.dense + ((size_t)ch * a.N + n) * 4);
  }
}

// Post-order partial of node ch (tip or internal) at pattern n.
template <typename R, int TIPS>
__device__ __forceinline__ V4<R> node4(const PassArgs<R>& a, const V4<R>* tbl, int ch, int k, int c,
                                       int n, bool val) {
  if (ch < a.T) return tip4<R, TIPS>(a, tbl, ch, n, val);
  if (!val) return zero4<R>();
  return ld4(a.partials + (prow(a.T, a.K, a.C, a.N, ch, k, c) + n) * 4);
}

template <typename R, int TIPS>
__device__ __forceinline__ void stage_tbl(const PassArgs<R>& a, V4<R>* tbl, int nt) {
  if constexpr (TIPS == HX_TIPS_CODES) {
    for (int i = threadIdx.x; i < a.ncodes; i += nt) tbl[i] = ld4(a.table + 4 * i);
  }
}

// Sums 16 per-lane values over the 64 lanes of the calling wave, in double.
// `buf` is this wave's LDS scratch of 16 x 64 R. Lane 4 v (v = 0..15) returns
// the total of value v; the other lanes return partial sums.
template <typename R>
__device__ __forceinline__ double wave_sum16(const R (&g)[16], R* buf) {
  const int lane = threadIdx.x & 63;
#pragma unroll
  for (int j = 0; j < 16; ++j) buf[j * 64 + lane] = g[j];
  __builtin_amdgcn_wave_barrier();
  const int v = lane >> 2, q = lane & 3;
  const R* row = buf + v * 64 + q * 16;
  double s = 0.0;
#pragma unroll
  for (int i = 0; i < 16; i += 4) {
    const V4<R> t = ld4(row + i);
    s += (double)t.x[0];
    s += (double)t.x[1];
    s += (double)t.x[2];
    s += (double)t.x[3];
  }
  __builtin_amdgcn_wave_barrier();
  s += __shfl_xor(s, 1);
  s += __shfl_xor(s, 2);
  return s;
}

// Writes the reduced statistic of branch b (category c) for one chunk: the 16
// entries of d loglik / d P, or their contraction with dP/dtau.
template <typename R>
__device__ __forceinline__ void flush_branch(const PassArgs<R>& a, const R* __restrict__ mats_dt,
                                             const R (&g)[16], R* buf, int b, int k, int c,
                                             int chunk) {
  const int lane = threadIdx.x & 63;
  const double s = wave_sum16(g, buf);
  const bool owner = (lane & 3) == 0;
  const int v = lane >> 2;
  if (mats_dt != nullptr) {
    const R* q = mats_dt + (((size_t)k * a.B + b) * a.C + c) * 16;
    double t = owner ? s * (double)q[v] : 0.0;
#pragma unroll
    for (int o = 4; o < 64; o <<= 1) t += __shfl_xor(t, o);
    if (lane == 0) a.gtau[(((size_t)b * a.K + k) * a.nchunks + chunk) * a.C + c] = t;
  } else if (owner) {
    a.gpart[((((size_t)b * a.K + k) * a.nchunks + chunk) * a.C + c) * 16 + v] = (R)s;
  }
}

// Outer-product accumulation g += w u (x) l.
template <typename R>
__device__ __forceinline__ void acc_outer(R (&g)[16], R w, const V4<R>& u, const V4<R>& l) {
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    const R wu = w * u.x[i];
#pragma unroll
    for (int j = 0; j < 4; ++j) g[4 * i + j] = fma(wu, l.x[j], g[4 * i + j]);
  }
}

}  // namespace

// ------------------------------------------------------------------ post-order
// prog: int4 per internal node in walk order, {v, left, right, slots} with
// slots = (left slot + 1) | (right slot + 1) << 8 | (output slot) << 16 and
// slot 0 meaning "tip". Root last. grid = (ntiles, K), block = 64 C.
template <typename R, int TIPS>
__global__ __launch_bounds__(1024) void k_post_walk(PassArgs<R> a, const R* __restrict__ mats,
                                                    const int4* __restrict__ prog, int nprog,
                                                    int dl) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  const int C = a.C, nt = C * 64;
  const int c = uni(threadIdx.x >> 6), lane = threadIdx.x & 63;
  const int tile = blockIdx.x, k = blockIdx.y;
  const int n = tile * 64 + lane;
  const bool val = n < a.N;
  V4<R>* stk = reinterpret_cast<V4<R>*>(smem_raw);              // [dl][nt]
  int* est = reinterpret_cast<int*>(stk + (size_t)dl * nt);     // [dl][nt]
  R* mx = reinterpret_cast<R*>(est + (size_t)dl * nt);          // [2][nt]
  double* red = reinterpret_cast<double*>(mx + 2 * nt);         // [C][6]
  V4<R>* tbl = reinterpret_cast<V4<R>*>(red + 6 * C);           // [ncodes]
  stage_tbl<R, TIPS>(a, tbl, nt);
  __syncthreads();

  int par = 0;
  V4<R> y;
  int E = 0;
  for (int i = 0; i < nprog; ++i) {
    const int4 ins = prog[i];
    const int v = ins.x, l = ins.y, r = ins.z;
    const int sl = (ins.w & 0xff) - 1, sr = ((ins.w >> 8) & 0xff) - 1, so = ins.w >> 16;
    V4<R> xl, xr;
    int El = 0, Er = 0;
    if (sl < 0) {
      xl = tip4<R, TIPS>(a, tbl, l, n, val);
    } else if (sl < dl) {
      xl = stk[sl * nt + threadIdx.x];
      El = est[sl * nt + threadIdx.x];
    } else {
      xl = node4<R, TIPS>(a, tbl, l, k, c, n, val);
      El = val ? a.scales[((size_t)(l - a.T) * a.K + k) * a.N + n] : 0;
    }
    if (sr < 0) {
      xr = tip4<R, TIPS>(a, tbl, r, n, val);
    } else if (sr < dl) {
      xr = stk[sr * nt + threadIdx.x];
      Er = est[sr * nt + threadIdx.x];
    } else {
      xr = node4<R, TIPS>(a, tbl, r, k, c, n, val);
      Er = val ? a.scales[((size_t)(r - a.T) * a.K + k) * a.N + n] : 0;
    }
    R P[16];
    load_p(mats, a.B, C, k, l, c, P);
    const V4<R> yl = mv(P, xl);
    load_p(mats, a.B, C, k, r, c, P);
    const V4<R> yr = mv(P, xr);
#pragma unroll
    for (int s = 0; s < 4; ++s) y.x[s] = yl.x[s] * yr.x[s];
    mx[par * nt + threadIdx.x] = max4(y);
    __syncthreads();
    R M = R(0);
    for (int cc = 0; cc < C; ++cc) M = fmax_(M, mx[par * nt + cc * 64 + lane]);
    par ^= 1;
    const int e = scale_exp(M);
    scale4(y, e);
    E = El + Er + e;
    if (i + 1 < nprog) {
      if (val) st4(a.partials + (prow(a.T, a.K, C, a.N, v, k, c) + n) * 4, y);
      if (so < dl) {
        stk[so * nt + threadIdx.x] = y;
        est[so * nt + threadIdx.x] = E;
      } else if (c == 0 && val) {
        a.scales[((size_t)(v - a.T) * a.K + k) * a.N + n] = E;
      }
    }
  }

  // root: site likelihood sum_c p_c pi.x_c (tree_likelihood.py:71-74, 215-220)
  const R pc = a.props[(size_t)k * a.props_ks + c];
  R pi[4];
#pragma unroll
  for (int s = 0; s < 4; ++s) pi[s] = a.freqs[(size_t)k * a.freqs_ks + s];
  R px = R(0);
#pragma unroll
  for (int s = 0; s < 4; ++s) px = fma(pi[s], y.x[s], px);
  mx[par * nt + threadIdx.x] = pc * px;
  __syncthreads();
  R lik = R(0);
  for (int cc = 0; cc < C; ++cc) lik += mx[par * nt + cc * 64 + lane];
  double acc[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};  // loglik, dfreqs[4], dprops[c]
  if (val) {
    const double w = a.weights[n];
    const double dl_ = (double)lik;
    const double wi = w / dl_;
    if (c == 0) acc[0] = w * (log(dl_) + (double)E * kLn2);
#pragma unroll
    for (int s = 0; s < 4; ++s) acc[1 + s] = wi * (double)(pc * y.x[s]);
    acc[5] = wi * (double)px;
  }
#pragma unroll
  for (int q = 0; q < 6; ++q)
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) acc[q] += __shfl_xor(acc[q], o);
  if (lane == 0)
#pragma unroll
    for (int q = 0; q < 6; ++q) red[c * 6 + q] = acc[q];
  __syncthreads();
  double* out = a.rootsums + ((size_t)k * a.ntiles + tile) * (5 + C);
  if (threadIdx.x < 5) {
    double t = 0.0;
    for (int cc = 0; cc < C; ++cc) t += red[cc * 6 + threadIdx.x];
    out[threadIdx.x] = t;
  }
  if (threadIdx.x < C) out[5 + threadIdx.x] = red[threadIdx.x * 6 + 5];
}

// ------------------------------------------------------------------- pre-order
// Group rooted at p (root: U = freqs). For parent q with children a, b:
//   Ut_a = U_q * (P_b L_b), Ut_b = U_q * (P_a L_a)
//   lik  = sum_c p_c Ut_a.(P_a L_a)                    (scaled site likelihood)
//   d loglik / d P_a[c][i][j] += w p_c Ut_a[i] L_a[j] / lik
//   U_a  = P_a^T Ut_a, rescaled (scale-free: every term is a ratio)
// LV = 2 also applies this to the children's children, with the children's U
// kept in registers; only the U of the last level is written (over L, in place).
// NTMAX: block-size bound (64 C rounded up), which sets the register budget.
template <typename R, int TIPS, bool ROOT, int LV, int NTMAX>
__global__ __launch_bounds__(NTMAX) void k_pre_grp(
    PassArgs<R> a, const R* __restrict__ mats, const R* __restrict__ mats_dt,
    const int32_t* __restrict__ nodes) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  const int C = a.C, nt = C * 64;
  const int c = uni(threadIdx.x >> 6), lane = threadIdx.x & 63;
  const int chunk = blockIdx.x, k = blockIdx.y;
  const int p = nodes[blockIdx.z];
  R* xs = reinterpret_cast<R*>(smem_raw);                 // [2][4][nt] exchange
  R* wbuf = xs + 8 * nt;                                  // [C][16][64] reductions
  V4<R>* tbl = reinterpret_cast<V4<R>*>(wbuf + (size_t)C * 1024);
  stage_tbl<R, TIPS>(a, tbl, nt);
  __syncthreads();

  int ch[2], gk[2][2];
  bool in_c[2], in_g[2][2];
#pragma unroll
  for (int h = 0; h < 2; ++h) {
    ch[h] = a.kids[2 * p + h];
    in_c[h] = ch[h] >= a.T;
#pragma unroll
    for (int j = 0; j < 2; ++j) {
      gk[h][j] = (LV == 2 && in_c[h]) ? a.kids[2 * ch[h] + j] : -1;
      in_g[h][j] = gk[h][j] >= a.T;
    }
  }
  const R pc = a.props[(size_t)k * a.props_ks + c];
  V4<R> pi;
  if constexpr (ROOT) {
#pragma unroll
    for (int s = 0; s < 4; ++s) pi.x[s] = a.freqs[(size_t)k * a.freqs_ks + s];
  }
  R g[2][16];
  R gg[LV == 2 ? 2 : 1][2][16];
#pragma unroll
  for (int i = 0; i < 16; ++i) {
    g[0][i] = g[1][i] = R(0);
#pragma unroll
    for (int h = 0; h < (LV == 2 ? 2 : 1); ++h) gg[h][0][i] = gg[h][1][i] = R(0);
  }

  int par = 0;
  // Exchange NQ values per lane with the other category waves of the block.
  auto xsum = [&](R* vals, int nq, bool take_max) {
#pragma unroll
    for (int q = 0; q < 4; ++q)
      if (q < nq) xs[(par * 4 + q) * nt + threadIdx.x] = vals[q];
    __syncthreads();
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      if (q < nq) {
        R t = R(0);
        for (int cc = 0; cc < C; ++cc) {
          const R u = xs[(par * 4 + q) * nt + cc * 64 + lane];
          t = take_max ? fmax_(t, u) : t + u;
        }
        vals[q] = t;
      }
    }
    par ^= 1;
  };

  const int t0 = chunk * a.tiles_per_chunk;
  const int t1 = min(t0 + a.tiles_per_chunk, a.ntiles);
  for (int tile = t0; tile < t1; ++tile) {
    const int n = tile * 64 + lane;
    const bool val = n < a.N;
    const R w = val ? (R)a.weights[n] : R(0);
    V4<R> U;
    if constexpr (ROOT) {
      U = val ? pi : zero4<R>();
    } else {
      U = val ? ld4(a.partials + (prow(a.T, a.K, C, a.N, p, k, c) + n) * 4) : zero4<R>();
    }
    V4<R> L[2], PL[2], Ut[2];
    R P[16];
#pragma unroll
    for (int h = 0; h < 2; ++h) {
      L[h] = node4<R, TIPS>(a, tbl, ch[h], k, c, n, val);
      load_p(mats, a.B, C, k, ch[h], c, P);
      PL[h] = mv(P, L[h]);
    }
    R d[4];
    d[0] = R(0);
#pragma unroll
    for (int s = 0; s < 4; ++s) {
      Ut[0].x[s] = U.x[s] * PL[1].x[s];
      Ut[1].x[s] = U.x[s] * PL[0].x[s];
      d[0] = fma(Ut[0].x[s], PL[0].x[s], d[0]);
    }
    d[0] *= pc;
    xsum(d, 1, false);
    const R coef = val ? w / d[0] * pc : R(0);
    acc_outer(g[0], coef, Ut[0], L[0]);
    acc_outer(g[1], coef, Ut[1], L[1]);
    // children's pre-order partials
    V4<R> Uc[2];
    R m[4];
#pragma unroll
    for (int h = 0; h < 2; ++h) {
      m[h] = R(0);
      if (in_c[h]) {
        load_p(mats, a.B, C, k, ch[h], c, P);
        Uc[h] = mv_t(P, Ut[h]);
        m[h] = max4(Uc[h]);
      }
    }
    if (in_c[0] || in_c[1]) {
      xsum(m, 2, true);
#pragma unroll
      for (int h = 0; h < 2; ++h) {
        if (in_c[h]) {
          scale4(Uc[h], scale_exp(m[h]));
          if constexpr (LV == 1) {
            if (val) st4(a.partials + (prow(a.T, a.K, C, a.N, ch[h], k, c) + n) * 4, Uc[h]);
          }
        }
      }
    }
    if constexpr (LV == 2) {
      if (in_c[0] || in_c[1]) {
        V4<R> Lg[2][2], PLg[2][2], Utg[2][2];
#pragma unroll
        for (int h = 0; h < 2; ++h) {
          d[h] = R(0);
          if (in_c[h]) {
#pragma unroll
            for (int j = 0; j < 2; ++j) {
              Lg[h][j] = node4<R, TIPS>(a, tbl, gk[h][j], k, c, n, val);
              load_p(mats, a.B, C, k, gk[h][j], c, P);
              PLg[h][j] = mv(P, Lg[h][j]);
            }
#pragma unroll
            for (int s = 0; s < 4; ++s) {
              Utg[h][0].x[s] = Uc[h].x[s] * PLg[h][1].x[s];
              Utg[h][1].x[s] = Uc[h].x[s] * PLg[h][0].x[s];
              d[h] = fma(Utg[h][0].x[s], PLg[h][0].x[s], d[h]);
            }
            d[h] *= pc;
          }
        }
        xsum(d, 2, false);
        bool any_g = false;
#pragma unroll
        for (int h = 0; h < 2; ++h) {
          if (in_c[h]) {
            const R ch_coef = val ? w / d[h] * pc : R(0);
            acc_outer(gg[h][0], ch_coef, Utg[h][0], Lg[h][0]);
            acc_outer(gg[h][1], ch_coef, Utg[h][1], Lg[h][1]);
          }
#pragma unroll
          for (int j = 0; j < 2; ++j) {
            m[2 * h + j] = R(0);
            if (in_g[h][j]) {
              any_g = true;
              load_p(mats, a.B, C, k, gk[h][j], c, P);
              Utg[h][j] = mv_t(P, Utg[h][j]);
              m[2 * h + j] = max4(Utg[h][j]);
            }
          }
        }
        if (any_g) {
          xsum(m, 4, true);
#pragma unroll
          for (int h = 0; h < 2; ++h)
#pragma unroll
            for (int j = 0; j < 2; ++j)
              if (in_g[h][j]) {
                scale4(Utg[h][j], scale_exp(m[2 * h + j]));
                if (val)
                  st4(a.partials + (prow(a.T, a.K, C, a.N, gk[h][j], k, c) + n) * 4, Utg[h][j]);
              }
        }
      }
    }
  }

  R* buf = wbuf + (size_t)c * 1024;
#pragma unroll
  for (int h = 0; h < 2; ++h) flush_branch(a, mats_dt, g[h], buf, ch[h], k, c, chunk);
  if constexpr (LV == 2) {
#pragma unroll
    for (int h = 0; h < 2; ++h)
      if (in_c[h])
#pragma unroll
        for (int j = 0; j < 2; ++j) flush_branch(a, mats_dt, gg[h][j], buf, gk[h][j], k, c, chunk);
  }
}

// ------------------------------------------------------------------ launchers
static size_t tbl_bytes(int tips, int ncodes, size_t rb) {
  return tips == HX_TIPS_CODES ? (size_t)ncodes * 4 * rb : 0;
}

template <typename R>
size_t dna_walk_lds(int C, int dl, int tips, int ncodes) {
  const size_t nt = (size_t)C * 64, rb = sizeof(R);
  return (size_t)dl * nt * (4 * rb + 4) + 2 * nt * rb + 6 * (size_t)C * 8 +
         tbl_bytes(tips, ncodes, rb);
}

template <typename R>
int dna_walk_depth(int C, int need, int tips, int ncodes) {
  int dl = need;
  while (dl > 0 && dna_walk_lds<R>(C, dl, tips, ncodes) > kDnaLdsBudget) --dl;
  return dl;
}

template <typename R>
int launch_post_dna(const PassArgs<R>& a, const int32_t* prog, int nprog, int need, int tips,
                    hipStream_t st) {
  const int dl = dna_walk_depth<R>(a.C, need, tips, a.ncodes);
  const size_t lds = dna_walk_lds<R>(a.C, dl, tips, a.ncodes);
  dim3 grid(a.ntiles, a.K);
  const int threads = a.C * 64;
  const int4* pg = reinterpret_cast<const int4*>(prog);
  if (tips == HX_TIPS_CODES)
    k_post_walk<R, HX_TIPS_CODES><<<grid, threads, lds, st>>>(a, a.mats, pg, nprog, dl);
  else
    k_post_walk<R, HX_TIPS_DENSE><<<grid, threads, lds, st>>>(a, a.mats, pg, nprog, dl);
  HX_HIP_TRY(hipGetLastError());
  return HX_OK;
}

template <typename R>
int dna_pre_levels(int C) {
  return (std::is_same<R, float>::value && C <= 4) ? 2 : 1;
}

template <typename R>
int launch_pre_dna(const PassArgs<R>& a, const int32_t* nodes, int width, int tips, bool root,
                   hipStream_t st) {
  const size_t nt = (size_t)a.C * 64, rb = sizeof(R);
  const size_t lds = 8 * nt * rb + (size_t)a.C * 1024 * rb + tbl_bytes(tips, a.ncodes, rb);
  dim3 grid(a.nchunks, a.K, width);
  const int threads = a.C * 64;
  const int lv = dna_pre_levels<R>(a.C);
#define HX_PRE(TV, RV, LVV, NTV) \
  k_pre_grp<R, TV, RV, LVV, NTV><<<grid, threads, lds, st>>>(a, a.mats, a.mats_dt, nodes)
#define HX_PRE_T(LVV, NTV)                                                               \
  if (tips == HX_TIPS_CODES) {                                                           \
    if (root) HX_PRE(HX_TIPS_CODES, true, LVV, NTV); else HX_PRE(HX_TIPS_CODES, false, LVV, NTV); \
  } else {                                                                               \
    if (root) HX_PRE(HX_TIPS_DENSE, true, LVV, NTV); else HX_PRE(HX_TIPS_DENSE, false, LVV, NTV); \
  }
  if (lv == 2) {
    if constexpr (std::is_same<R, float>::value) HX_PRE_T(2, 256)
  } else if (threads <= 512) {
    HX_PRE_T(1, 512)
  } else {
    HX_PRE_T(1, 1024)
  }
#undef HX_PRE_T
#undef HX_PRE
  HX_HIP_TRY(hipGetLastError());
  return HX_OK;
}

template int launch_post_dna<float>(const PassArgs<float>&, const int32_t*, int, int, int,
                                    hipStream_t);
template int launch_post_dna<double>(const PassArgs<double>&, const int32_t*, int, int, int,
                                     hipStream_t);
template int launch_pre_dna<float>(const PassArgs<float>&, const int32_t*, int, int, bool,
                                   hipStream_t);
template int launch_pre_dna<double>(const PassArgs<double>&, const int32_t*, int, int, bool,
                                    hipStream_t);
template int dna_pre_levels<float>(int);
template int dna_pre_levels<double>(int);

}  // namespace hx

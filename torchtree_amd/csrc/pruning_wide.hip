// Pruning kernels for dense state spaces (S != 4: 20-state amino acids,
// 61-state codons, any S <= 64).
//
// Reference: the same calculate_treelikelihood* functions
// (torchtree/evolution/tree_likelihood.py:40-278) at S = 20 / 61, where each
// node update is a dense S x S by S x N contraction per (sample, category).
//
// Block = one (sample, node) and a tile of NT patterns for every category;
// thread (h, c, i) owns child h, category c and state i (S padded to SP = 32
// or 64 lanes per (h, c) group). Each thread keeps row i of P_h (and, for the
// pre-order, column i and row i of d loglik / d P_h accumulated over the
// whole pattern chunk) in registers; the tile's child partials sit in LDS and
// are read as broadcasts. Global layouts are those of pruning.hip
// ([slot][K][N][C][S] partials with S unpadded).
#include "pruning.hpp"

#include <math.h>

namespace hx {

namespace {

constexpr int kWideMaxThreads = 512;

__device__ __forceinline__ float wfmax(float a, float b) { return fmaxf(a, b); }
__device__ __forceinline__ double wfmax(double a, double b) { return fmax(a, b); }
__device__ __forceinline__ float wfrexp(float m, int* e) { return frexpf(m, e); }
__device__ __forceinline__ double wfrexp(double m, int* e) { return frexp(m, e); }
__device__ __forceinline__ float wldexp(float x, int e) { return ldexpf(x, e); }
__device__ __forceinline__ double wldexp(double x, int e) { return ldexp(x, e); }

template <int SP, typename T>
__device__ __forceinline__ T group_max(T v) {
#pragma unroll
  for (int o = 1; o < SP; o <<= 1) v = wfmax(v, __shfl_xor(v, o));
  return v;
}

template <int SP, typename T>
__device__ __forceinline__ T group_sum(T v) {
#pragma unroll
  for (int o = 1; o < SP; o <<= 1) v += __shfl_xor(v, o);
  return v;
}

// Child partial element (pattern nn, category cc, state s) of child `ch`.
template <typename R, int TIPS>
__device__ __forceinline__ R child_elem(const PassArgs<R>& a, int ch, int k, int nn, int cc, int s,
                                        int S) {
  if (ch < a.T) {
    if constexpr (TIPS == HX_TIPS_CODES) {
      const int code = a.codes[(size_t)ch * a.N + nn];
      return a.table[(size_t)code * S + s];
    } else {
      return a.dense[((size_t)ch * a.N + nn) * S + s];
    }
  }
  return a.partials[((((size_t)(ch - a.T) * a.K + k) * a.N + nn) * a.C + cc) * S + s];
}

// Fill Lt[h][n][c][s] (h = 0, 1) for the tile starting at nbase.
template <typename R, int SP, int TIPS>
__device__ __forceinline__ void fill_children(const PassArgs<R>& a, R* Lt, int ch0, int ch1, int k,
                                              int nbase, int NT, int S) {
  const int G = a.C * SP;
  const int total = 2 * NT * G;
  for (int q = threadIdx.x; q < total; q += blockDim.x) {
    const int hh = q / (NT * G);
    const int rem = q - hh * NT * G;
    const int n = rem / G;
    const int r2 = rem - n * G;
    const int cc = r2 / SP, s = r2 - cc * SP;
    const int nn = nbase + n;
    R val = R(0);
    if (nn < a.N && s < S) val = child_elem<R, TIPS>(a, hh ? ch1 : ch0, k, nn, cc, s, S);
    Lt[q] = val;
  }
}

}  // namespace


// ---------------------------------------------------------------- post-order
template <typename R, int SP, int TIPS, bool ROOT>
__global__ __launch_bounds__(kWideMaxThreads) void k_post_wide(PassArgs<R> a,
                                                               const int32_t* __restrict__ nodes,
                                                               int S, int NT) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  const int C = a.C, G = C * SP;
  R* Lt = reinterpret_cast<R*>(smem_raw);  // [2][NT][G]
  R* Yt = Lt + 2 * NT * G;                 // [2][NT][G]
  R* redm = Yt + 2 * NT * G;               // [NT][C]
  double* rd = reinterpret_cast<double*>(redm + ((NT * C + 1) & ~1));  // [NT][C] (root)
  int* Et = reinterpret_cast<int*>(rd + NT * C);                       // [2][NT]
  int* Er = Et + 2 * NT;                                               // [NT] root exponents

  const int t = threadIdx.x;
  const int h = t / G, r = t - h * G, c = r / SP, i = r - c * SP;
  const int chunk = blockIdx.x, k = blockIdx.y, v = nodes[blockIdx.z];
  const int ch0 = a.kids[2 * v], ch1 = a.kids[2 * v + 1];
  const int chh = h ? ch1 : ch0;

  R prow[SP];
  {
    const R* pm = a.mats + (((size_t)k * a.B + chh) * C + c) * S * S;
#pragma unroll
    for (int j = 0; j < SP; ++j) prow[j] = (i < S && j < S) ? pm[i * S + j] : R(0);
  }
  R pi_i = R(0), pc = R(0);
  double acc_l = 0.0, acc_f = 0.0, acc_p = 0.0;
  if constexpr (ROOT) {
    pi_i = i < S ? a.freqs[(size_t)k * a.freqs_ks + i] : R(0);
    pc = a.props[(size_t)k * a.props_ks + c];
  }

  const int t0 = chunk * a.tiles_per_chunk;
  const int t1 = min(t0 + a.tiles_per_chunk, a.ntiles);
  for (int tile = t0; tile < t1; ++tile) {
    const int nbase = tile * NT;
    fill_children<R, SP, TIPS>(a, Lt, ch0, ch1, k, nbase, NT, S);
    if (t < 2 * NT) {
      const int hh = t / NT, n = t - hh * NT, nn = nbase + n;
      const int chq = hh ? ch1 : ch0;
      Et[t] = (nn < a.N && chq >= a.T) ? a.scales[((size_t)(chq - a.T) * a.K + k) * a.N + nn] : 0;
    }
    __syncthreads();
    for (int n = 0; n < NT; ++n) {
      const R* lv = Lt + (h * NT + n) * G + c * SP;
      R y = R(0);
#pragma unroll
      for (int j = 0; j < SP; ++j) y = fma(prow[j], lv[j], y);
      Yt[(h * NT + n) * G + c * SP + i] = y;
    }
    __syncthreads();
    if (h == 0) {
      for (int n = 0; n < NT; ++n) {
        const R x = Yt[n * G + c * SP + i] * Yt[(NT + n) * G + c * SP + i];
        const R m = group_max<SP>(x);
        if (i == 0) redm[n * C + c] = m;
      }
    }
    __syncthreads();
    if (h == 0) {
      for (int n = 0; n < NT; ++n) {
        const int nn = nbase + n;
        R x = Yt[n * G + c * SP + i] * Yt[(NT + n) * G + c * SP + i];
        R m = R(0);
        for (int cc = 0; cc < C; ++cc) m = wfmax(m, redm[n * C + cc]);
        int e = 0;
        if (m > R(0) && isfinite(m)) {
          wfrexp(m, &e);
          x = wldexp(x, -e);
        }
        const int E = Et[n] + Et[NT + n] + e;
        if constexpr (!ROOT) {
          if (nn < a.N) {
            const size_t row = ((size_t)(v - a.T) * a.K + k) * a.N + nn;
            if (i < S) a.partials[(row * C + c) * S + i] = x;
            if (c == 0 && i == 0) a.scales[row] = E;
          }
        } else {
          // px_c = pi . x_c; lik = sum_c p_c px_c (tree_likelihood.py:71-74, 215-220)
          const R px = group_sum<SP>(pi_i * x);
          Yt[n * G + c * SP + i] = x;  // keep the scaled root partial for the gradients
          if (i == 0) rd[n * C + c] = (double)px;
          if (c == 0 && i == 0) Er[n] = E;
        }
      }
    }
    if constexpr (ROOT) {
      __syncthreads();
      if (h == 0) {
        for (int n = 0; n < NT; ++n) {
          const int nn = nbase + n;
          if (nn >= a.N) continue;
          double lik = 0.0;
          for (int cc = 0; cc < C; ++cc) {
            const R pcc = a.props[(size_t)k * a.props_ks + cc];
            lik += (double)(pcc * (R)rd[n * C + cc]);
          }
          const double w = a.weights[nn];
          const double wi = w / lik;
          if (c == 0 && i == 0) acc_l += w * (log(lik) + (double)Er[n] * 0.69314718055994530942);
          if (i < S) acc_f += wi * (double)(pc * Yt[n * G + c * SP + i]);
          if (i == 0) acc_p += wi * rd[n * C + c];
        }
      }
    }
    __syncthreads();
  }
  if constexpr (ROOT) {
    // acc_f: sum over c per state; acc_p per c; acc_l single. Reuse Lt as scratch.
    double* red = reinterpret_cast<double*>(Lt);
    if (h == 0) red[c * SP + i] = acc_f;
    __syncthreads();
    double* out = a.rootsums + ((size_t)k * a.nchunks + chunk) * (1 + S + C);
    if (h == 0 && c == 0 && i < S) {
      double tsum = 0.0;
      for (int cc = 0; cc < C; ++cc) tsum += red[cc * SP + i];
      out[1 + i] = tsum;
    }
    if (h == 0 && i == 0) out[1 + S + c] = acc_p;
    if (h == 0 && c == 0 && i == 0) out[0] = acc_l;
  }
}

// ----------------------------------------------------------------- pre-order
template <typename R, int SP, int TIPS, bool ROOT>
__global__ __launch_bounds__(kWideMaxThreads) void k_pre_wide(PassArgs<R> a,
                                                              const int32_t* __restrict__ nodes,
                                                              int S, int NT) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  // FP64 rows of 64 padded states do not fit next to the gradient row: then
  // row i of P is read from memory (L1/L2) instead of registers.
  constexpr bool kColRegs = sizeof(R) == 4 || SP == 32;
  constexpr int kNtMax = 8;
  const int C = a.C, G = C * SP;
  R* Lt = reinterpret_cast<R*>(smem_raw);  // [2][NT][G] children post-order partials
  R* Yt = Lt + 2 * NT * G;                 // [2][NT][G] P_h L_h, then new pre-order partials
  R* Ut = Yt + 2 * NT * G;                 // [NT][G]    parent pre-order partial
  R* Wt = Ut + NT * G;                     // [2][NT][G] Ut_h = U_p * (P_other L_other)
  R* dens = Wt + 2 * NT * G;               // [NT][C]
  R* mx = dens + NT * C;                   // [2][NT][C]

  const int t = threadIdx.x;
  const int h = t / G, r = t - h * G, c = r / SP, i = r - c * SP;
  const int chunk = blockIdx.x, k = blockIdx.y, p = nodes[blockIdx.z];
  const int ch0 = a.kids[2 * p], ch1 = a.kids[2 * p + 1];
  const int chh = h ? ch1 : ch0;
  const R* pm = a.mats + (((size_t)k * a.B + chh) * C + c) * S * S;

  R prow[kColRegs ? SP : 1];
  if constexpr (kColRegs) {
#pragma unroll
    for (int j = 0; j < SP; ++j) prow[j] = (i < S && j < S) ? pm[i * S + j] : R(0);
  }
  const R pc = a.props[(size_t)k * a.props_ks + c];
  R g[SP];
#pragma unroll
  for (int j = 0; j < SP; ++j) g[j] = R(0);

  const int t0 = chunk * a.tiles_per_chunk;
  const int t1 = min(t0 + a.tiles_per_chunk, a.ntiles);
  for (int tile = t0; tile < t1; ++tile) {
    const int nbase = tile * NT;
    fill_children<R, SP, TIPS>(a, Lt, ch0, ch1, k, nbase, NT, S);
    for (int q = t; q < NT * G; q += blockDim.x) {
      const int n = q / G, r2 = q - n * G, cc = r2 / SP, s = r2 - cc * SP;
      const int nn = nbase + n;
      R val = R(0);
      if (nn < a.N && s < S) {
        if constexpr (ROOT) {
          val = a.freqs[(size_t)k * a.freqs_ks + s];
        } else {
          val = a.partials[((((size_t)(p - a.T) * a.K + k) * a.N + nn) * C + cc) * S + s];
        }
      }
      Ut[q] = val;
    }
    __syncthreads();
    for (int n = 0; n < NT; ++n) {
      const R* lv = Lt + (h * NT + n) * G + c * SP;
      R y = R(0);
      if constexpr (kColRegs) {
#pragma unroll
        for (int j = 0; j < SP; ++j) y = fma(prow[j], lv[j], y);
      } else {
        if (i < S)
          for (int j = 0; j < S; ++j) y = fma(pm[i * S + j], lv[j], y);
      }
      Yt[(h * NT + n) * G + c * SP + i] = y;
    }
    __syncthreads();
    for (int n = 0; n < NT; ++n) {
      const R u = Ut[n * G + c * SP + i];
      const R wt = u * Yt[((1 - h) * NT + n) * G + c * SP + i];
      Wt[(h * NT + n) * G + c * SP + i] = wt;
      if (h == 0) {
        const R d = group_sum<SP>(wt * Yt[n * G + c * SP + i]);
        if (i == 0) dens[n * C + c] = pc * d;
      }
    }
    __syncthreads();
    for (int n = 0; n < NT; ++n) {
      const int nn = nbase + n;
      R den = R(0);
      for (int cc = 0; cc < C; ++cc) den += dens[n * C + cc];
      const R coef = nn < a.N ? (R)a.weights[nn] / den * pc : R(0);
      const R cu = coef * Wt[(h * NT + n) * G + c * SP + i];
      const R* lv = Lt + (h * NT + n) * G + c * SP;
#pragma unroll
      for (int j = 0; j < SP; ++j) g[j] = fma(cu, lv[j], g[j]);
    }
    if (chh >= a.T) {
      // new pre-order partial: nu[n][i] = sum_j P_h[j][i] Ut_h[n][j]; column i of
      // P_h is read once per tile with coalesced loads (lanes i are contiguous)
      R nu[kNtMax];
#pragma unroll
      for (int n = 0; n < kNtMax; ++n) nu[n] = R(0);
      for (int j = 0; j < S; ++j) {
        const R pji = i < S ? pm[j * S + i] : R(0);
#pragma unroll
        for (int n = 0; n < kNtMax; ++n)
          if (n < NT) nu[n] = fma(pji, Wt[(h * NT + n) * G + c * SP + j], nu[n]);
      }
#pragma unroll
      for (int n = 0; n < kNtMax; ++n) {
        if (n < NT) {
          Yt[(h * NT + n) * G + c * SP + i] = nu[n];
          const R m = group_max<SP>(nu[n]);
          if (i == 0) mx[(h * NT + n) * C + c] = m;
        }
      }
    }
    __syncthreads();
    if (chh >= a.T) {
      for (int n = 0; n < NT; ++n) {
        const int nn = nbase + n;
        R m = R(0);
        for (int cc = 0; cc < C; ++cc) m = wfmax(m, mx[(h * NT + n) * C + cc]);
        R nu = Yt[(h * NT + n) * G + c * SP + i];
        if (m > R(0) && isfinite(m)) {
          int e;
          wfrexp(m, &e);
          nu = wldexp(nu, -e);
        }
        if (nn < a.N && i < S)
          a.partials[((((size_t)(chh - a.T) * a.K + k) * a.N + nn) * C + c) * S + i] = nu;
      }
    }
    __syncthreads();
  }
  if (a.mats_dt != nullptr) {
    // d loglik / d tau_{b,c}: row i of G_h^c against row i of dP/dtau, summed over i
    double s = 0.0;
    if (i < S) {
      const R* qrow = a.mats_dt + (((size_t)k * a.B + chh) * C + c) * S * S + (size_t)i * S;
#pragma unroll
      for (int j = 0; j < SP; ++j)
        if (j < S) s = fma((double)g[j], (double)qrow[j], s);
    }
    s = group_sum<SP>(s);
    if (i == 0) a.gtau[(((size_t)chh * a.K + k) * a.nchunks + chunk) * C + c] = s;
  } else if (i < S) {
    const size_t css = (size_t)C * S * S;
    R* o = a.gpart + (((size_t)chh * a.K + k) * a.nchunks + chunk) * css + ((size_t)c * S + i) * S;
#pragma unroll
    for (int j = 0; j < SP; ++j)
      if (j < S) o[j] = g[j];
  }
}

static size_t wide_lds(int C, int S, int NT, int rb, bool pre) {
  const int SP = S <= 32 ? 32 : 64;
  const size_t G = (size_t)C * SP;
  if (pre) return (7 * NT * G + 3 * NT * C) * rb + 64;
  return (4 * NT * G + NT * C + 2) * rb + NT * C * 8 + 3 * NT * 4 + 64;
}

template <typename R>
int launch_post_wide(const PassArgs<R>& a, const int32_t* nodes, int width, int tips, bool root,
                     int S, hipStream_t st) {
  const int SP = S <= 32 ? 32 : 64;
  const int threads = 2 * a.C * SP;
  if (threads > kWideMaxThreads) {
    set_error("categories * padded states too large for the dense-state kernels");
    return HX_ERR_UNSUPPORTED;
  }
  const int NT = wide_nt(a.C, S, (int)sizeof(R));
  const size_t lds = wide_lds(a.C, S, NT, (int)sizeof(R), false);
  dim3 grid(a.nchunks, a.K, width);
#define HX_LAUNCH_POST(SPV, TV, RV) \
  k_post_wide<R, SPV, TV, RV><<<grid, threads, lds, st>>>(a, nodes, S, NT)
  if (SP == 32) {
    if (tips == HX_TIPS_CODES) { if (root) HX_LAUNCH_POST(32, HX_TIPS_CODES, true); else HX_LAUNCH_POST(32, HX_TIPS_CODES, false); }
    else { if (root) HX_LAUNCH_POST(32, HX_TIPS_DENSE, true); else HX_LAUNCH_POST(32, HX_TIPS_DENSE, false); }
  } else {
    if (tips == HX_TIPS_CODES) { if (root) HX_LAUNCH_POST(64, HX_TIPS_CODES, true); else HX_LAUNCH_POST(64, HX_TIPS_CODES, false); }
    else { if (root) HX_LAUNCH_POST(64, HX_TIPS_DENSE, true); else HX_LAUNCH_POST(64, HX_TIPS_DENSE, false); }
  }
#undef HX_LAUNCH_POST
  return HX_OK;
}

template <typename R>
int launch_pre_wide(const PassArgs<R>& a, const int32_t* nodes, int width, int tips, bool root,
                    int S, hipStream_t st) {
  const int SP = S <= 32 ? 32 : 64;
  const int threads = 2 * a.C * SP;
  if (threads > kWideMaxThreads) {
    set_error("categories * padded states too large for the dense-state kernels");
    return HX_ERR_UNSUPPORTED;
  }
  const int NT = wide_nt(a.C, S, (int)sizeof(R));
  const size_t lds = wide_lds(a.C, S, NT, (int)sizeof(R), true);
  dim3 grid(a.nchunks, a.K, width);
#define HX_LAUNCH_PRE(SPV, TV, RV) \
  k_pre_wide<R, SPV, TV, RV><<<grid, threads, lds, st>>>(a, nodes, S, NT)
  if (SP == 32) {
    if (tips == HX_TIPS_CODES) { if (root) HX_LAUNCH_PRE(32, HX_TIPS_CODES, true); else HX_LAUNCH_PRE(32, HX_TIPS_CODES, false); }
    else { if (root) HX_LAUNCH_PRE(32, HX_TIPS_DENSE, true); else HX_LAUNCH_PRE(32, HX_TIPS_DENSE, false); }
  } else {
    if (tips == HX_TIPS_CODES) { if (root) HX_LAUNCH_PRE(64, HX_TIPS_CODES, true); else HX_LAUNCH_PRE(64, HX_TIPS_CODES, false); }
    else { if (root) HX_LAUNCH_PRE(64, HX_TIPS_DENSE, true); else HX_LAUNCH_PRE(64, HX_TIPS_DENSE, false); }
  }
#undef HX_LAUNCH_PRE
  return HX_OK;
}

template int launch_post_wide<float>(const PassArgs<float>&, const int32_t*, int, int, bool, int,
                                     hipStream_t);
template int launch_post_wide<double>(const PassArgs<double>&, const int32_t*, int, int, bool, int,
                                      hipStream_t);
template int launch_pre_wide<float>(const PassArgs<float>&, const int32_t*, int, int, bool, int,
                                    hipStream_t);
template int launch_pre_wide<double>(const PassArgs<double>&, const int32_t*, int, int, bool, int,
                                     hipStream_t);

}  // namespace hx

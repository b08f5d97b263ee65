// FP32 pruning kernels for dense state spaces on the matrix cores (MFMA).
//
// Reference: calculate_treelikelihood_discrete* (torchtree/evolution/
// tree_likelihood.py:40-278) at S = 20 (amino acids) and 61 (codons), where
// every node update is a dense S x S by S x N contraction.
//
// Block = one (sample, node) and tiles of 16 patterns; wave (h, c) owns child
// h and category c (2*C waves). With lane = nn + 16 g, the three products of a
// tile map onto v_mfma_f32_16x16x4f32 as
//   forward    Y[m][i]  = sum_j L[m][j] P[i][j]        A = L (patterns x states),
//                                                       B = P^T held in registers
//   gradient   G[i][j] += sum_m U'[m][i] L[m][j]        contraction over patterns;
//                                                       A is the wave's own U' registers
//   transpose  U[m][j]  = sum_i Ut[m][i] P[i][j]        A from an LDS transpose of Ut
// S is padded to SP = 16 * MT states; K index k = 4 ks + g of a step ks maps to
// state g * (4 MT) + ks (forward / transpose) or to pattern 4 g + ks (gradient),
// so each lane's A operand is a contiguous run and the accumulator layout of
// one product is the operand layout of the next. The 16x16x4 f32 layout used:
// A[i][k] from lane i + 16 k, B[k][j] from lane j + 16 k, D[i][j] in lane
// j + 16 (i / 4), register i % 4 (checked on the device by
// hx_selftest_mfma_layout).
#include "pruning.hpp"

#include <math.h>

namespace hx {

namespace {

typedef float f4 __attribute__((ext_vector_type(4)));
constexpr int kTile = 16;

__device__ __forceinline__ f4 mma(float a, float b, f4 c) {
  return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
}

__device__ __forceinline__ float row_max16(float v) {
#pragma unroll
  for (int o = 1; o < 16; o <<= 1) v = fmaxf(v, __shfl_xor(v, o));
  return v;
}

template <typename T>
__device__ __forceinline__ T row_sum16(T v) {
#pragma unroll
  for (int o = 1; o < 16; o <<= 1) v += __shfl_xor(v, o);
  return v;
}

// Child partial of child ch at absolute pattern n, category c, state s (0 outside).
template <int TIPS>
__device__ __forceinline__ float child_val(const PassArgs<float>& a, int ch, int k, int n, int c,
                                           int s, int S) {
  if (n >= a.N || s >= S) return 0.f;
  if (ch < a.T) {
    if constexpr (TIPS == HX_TIPS_CODES) {
      return a.table[(size_t)a.codes[(size_t)ch * a.N + n] * S + s];
    } else {
      return a.dense[((size_t)ch * a.N + n) * S + s];
    }
  }
  return a.partials[((((size_t)(ch - a.T) * a.K + k) * a.N + n) * a.C + c) * S + s];
}

}  // namespace

// ---------------------------------------------------------------- post-order
template <int MT, int TIPS, bool ROOT>
__global__ __launch_bounds__(512) void k_post_mfma(PassArgs<float> a,
                                                   const int32_t* __restrict__ nodes, int S) {
  constexpr int SP = 16 * MT, KS = 4 * MT;
  extern __shared__ __align__(16) unsigned char smem_raw[];
  const int C = a.C;
  float* Yx = reinterpret_cast<float*>(smem_raw);       // [C][16][SP] child-1 products
  float* mx = Yx + C * kTile * SP;                      // [C][16]
  double* pxs = reinterpret_cast<double*>(mx + C * kTile);  // [C][16]
  double* red = pxs + C * kTile;                        // [C][SP] root reductions
  int* Et = reinterpret_cast<int*>(red + C * SP);       // [2][16]
  int* Er = Et + 2 * kTile;                             // [16]

  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int h = wave / C, c = wave - h * C;
  const int nn = lane & 15, g = lane >> 4;
  const int chunk = blockIdx.x, k = blockIdx.y, v = nodes[blockIdx.z];
  const int ch0 = a.kids[2 * v], ch1 = a.kids[2 * v + 1];
  const int chh = h ? ch1 : ch0;

  float Pf[MT][KS];  // Pf[t][ks] = P[16 t + nn][g KS + ks]
  {
    const float* pm = a.mats + (((size_t)k * a.B + chh) * C + c) * S * S;
#pragma unroll
    for (int t = 0; t < MT; ++t)
#pragma unroll
      for (int ks = 0; ks < KS; ++ks) {
        const int i = 16 * t + nn, j = g * KS + ks;
        Pf[t][ks] = (i < S && j < S) ? pm[i * S + j] : 0.f;
      }
  }
  float pi_t[MT];
  float pc = 0.f;
  double acc_l = 0.0, acc_p = 0.0, acc_f[MT];
  if constexpr (ROOT) {
#pragma unroll
    for (int t = 0; t < MT; ++t) {
      const int i = 16 * t + nn;
      pi_t[t] = i < S ? a.freqs[(size_t)k * a.freqs_ks + i] : 0.f;
      acc_f[t] = 0.0;
    }
    pc = a.props[(size_t)k * a.props_ks + c];
  }

  const int t0 = chunk * a.tiles_per_chunk;
  const int t1 = min(t0 + a.tiles_per_chunk, a.ntiles);
  for (int tile = t0; tile < t1; ++tile) {
    const int nbase = tile * kTile;
    if (threadIdx.x < 2 * kTile) {
      const int hh = threadIdx.x / kTile, m = threadIdx.x - hh * kTile, n = nbase + m;
      const int chq = hh ? ch1 : ch0;
      Et[threadIdx.x] =
          (n < a.N && chq >= a.T) ? a.scales[((size_t)(chq - a.T) * a.K + k) * a.N + n] : 0;
    }
    f4 acc[MT];
#pragma unroll
    for (int t = 0; t < MT; ++t) acc[t] = f4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int ks = 0; ks < KS; ++ks) {
      const float av = child_val<TIPS>(a, chh, k, nbase + nn, c, g * KS + ks, S);
#pragma unroll
      for (int t = 0; t < MT; ++t) acc[t] = mma(av, Pf[t][ks], acc[t]);
    }
    // lane (nn, g): acc[t][r] = Y[m = 4 g + r][i = 16 t + nn]
    if (h == 1) {
#pragma unroll
      for (int t = 0; t < MT; ++t)
#pragma unroll
        for (int r = 0; r < 4; ++r) Yx[(c * kTile + 4 * g + r) * SP + 16 * t + nn] = acc[t][r];
    }
    __syncthreads();
    float x[MT][4];
    if (h == 0) {
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        float m_ = 0.f;
#pragma unroll
        for (int t = 0; t < MT; ++t) {
          x[t][r] = acc[t][r] * Yx[(c * kTile + 4 * g + r) * SP + 16 * t + nn];
          m_ = fmaxf(m_, x[t][r]);
        }
        m_ = row_max16(m_);
        if (nn == 0) mx[c * kTile + 4 * g + r] = m_;
      }
    }
    __syncthreads();
    if (h == 0) {
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int m = 4 * g + r, n = nbase + m;
        float M = 0.f;
        for (int cc = 0; cc < C; ++cc) M = fmaxf(M, mx[cc * kTile + m]);
        int e = 0;
        if (M > 0.f && isfinite(M)) {
          frexpf(M, &e);
#pragma unroll
          for (int t = 0; t < MT; ++t) x[t][r] = ldexpf(x[t][r], -e);
        }
        const int E = Et[m] + Et[kTile + m] + e;
        if constexpr (!ROOT) {
          if (n < a.N) {
            const size_t row = ((size_t)(v - a.T) * a.K + k) * a.N + n;
#pragma unroll
            for (int t = 0; t < MT; ++t)
              if (16 * t + nn < S) a.partials[(row * C + c) * S + 16 * t + nn] = x[t][r];
            if (c == 0 && nn == 0) a.scales[row] = E;
          }
        } else {
          float p_ = 0.f;
#pragma unroll
          for (int t = 0; t < MT; ++t) p_ = fmaf(pi_t[t], x[t][r], p_);
          p_ = row_sum16(p_);
          if (nn == 0) pxs[c * kTile + m] = (double)p_;
          if (c == 0 && nn == 0) Er[m] = E;
        }
      }
    }
    if constexpr (ROOT) {
      __syncthreads();
      if (h == 0) {
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          const int m = 4 * g + r, n = nbase + m;
          if (n >= a.N) continue;
          double lik = 0.0;
          for (int cc = 0; cc < C; ++cc)
            lik += (double)(a.props[(size_t)k * a.props_ks + cc] * (float)pxs[cc * kTile + m]);
          const double w = a.weights[n];
          const double wi = w / lik;
          if (c == 0 && nn == 0) acc_l += w * (log(lik) + (double)Er[m] * kLn2);
#pragma unroll
          for (int t = 0; t < MT; ++t) acc_f[t] += wi * (double)(pc * x[t][r]);
          if (nn == 0) acc_p += wi * pxs[c * kTile + m];
        }
      }
    }
    __syncthreads();
  }
  if constexpr (ROOT) {
    if (h == 0) {
#pragma unroll
      for (int t = 0; t < MT; ++t) {
        double f = acc_f[t];
        f += __shfl_xor(f, 16);
        f += __shfl_xor(f, 32);
        if (g == 0) red[c * SP + 16 * t + nn] = f;
      }
      acc_p += __shfl_xor(acc_p, 16);
      acc_p += __shfl_xor(acc_p, 32);
      acc_l += __shfl_xor(acc_l, 16);
      acc_l += __shfl_xor(acc_l, 32);
    }
    __syncthreads();
    double* out = a.rootsums + ((size_t)k * a.nchunks + chunk) * (1 + S + C);
    if (h == 0 && c == 0) {
      for (int s = lane; s < S; s += 64) {
        double tsum = 0.0;
        for (int cc = 0; cc < C; ++cc) tsum += red[cc * SP + s];
        out[1 + s] = tsum;
      }
      if (lane == 0) out[0] = acc_l;
    }
    if (h == 0 && lane == 0) out[1 + S + c] = acc_p;
  }
}

// ----------------------------------------------------------------- pre-order
template <int MT, int TIPS, bool ROOT>
__global__ __launch_bounds__(512) void k_pre_mfma(PassArgs<float> a,
                                                  const int32_t* __restrict__ nodes, int S) {
  constexpr int SP = 16 * MT, KS = 4 * MT;
  extern __shared__ __align__(16) unsigned char smem_raw[];
  const int C = a.C;
  float* Yx = reinterpret_cast<float*>(smem_raw);  // [2][C][16][SP]: P L, then Ut transposes
  float* dens = Yx + 2 * C * kTile * SP;           // [C][16]
  float* mxx = dens + C * kTile;                   // [2][C][16]

  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int h = wave / C, c = wave - h * C;
  const int nn = lane & 15, g = lane >> 4;
  const int chunk = blockIdx.x, k = blockIdx.y, p = nodes[blockIdx.z];
  const int ch0 = a.kids[2 * p], ch1 = a.kids[2 * p + 1];
  const int chh = h ? ch1 : ch0;
  const bool internal = chh >= a.T;
  const float* pm = a.mats + (((size_t)k * a.B + chh) * C + c) * S * S;

  // Pf[t][ks] = P[16 t + nn][g KS + ks]; at 64 padded states the gradient
  // accumulators take those registers and P is re-read from L1/L2 per tile
  constexpr bool kPfRegs = MT < 4;
  float Pf[kPfRegs ? MT : 1][kPfRegs ? KS : 1];
  if constexpr (kPfRegs) {
#pragma unroll
    for (int t = 0; t < MT; ++t)
#pragma unroll
      for (int ks = 0; ks < KS; ++ks) {
        const int i = 16 * t + nn, j = g * KS + ks;
        Pf[t][ks] = (i < S && j < S) ? pm[i * S + j] : 0.f;
      }
  }
  const float pc = a.props[(size_t)k * a.props_ks + c];
  f4 G[MT][MT];  // G[Mt][Nt] lane (nn, g) reg r = dlogL/dP[16 Mt + 4 g + r][16 Nt + nn]
#pragma unroll
  for (int i = 0; i < MT; ++i)
#pragma unroll
    for (int j = 0; j < MT; ++j) G[i][j] = f4{0.f, 0.f, 0.f, 0.f};
  float* mine = Yx + (size_t)(h * C + c) * kTile * SP;
  const float* other = Yx + (size_t)((1 - h) * C + c) * kTile * SP;

  const int t0 = chunk * a.tiles_per_chunk;
  const int t1 = min(t0 + a.tiles_per_chunk, a.ntiles);
  for (int tile = t0; tile < t1; ++tile) {
    const int nbase = tile * kTile;
    // forward P_h L_h
    f4 acc[MT];
#pragma unroll
    for (int t = 0; t < MT; ++t) acc[t] = f4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int ks = 0; ks < KS; ++ks) {
      const float av = child_val<TIPS>(a, chh, k, nbase + nn, c, g * KS + ks, S);
#pragma unroll
      for (int t = 0; t < MT; ++t) {
        float bv;
        if constexpr (kPfRegs) {
          bv = Pf[t][ks];
        } else {
          const int i = 16 * t + nn, j = g * KS + ks;
          bv = (i < S && j < S) ? pm[i * S + j] : 0.f;
        }
        acc[t] = mma(av, bv, acc[t]);
      }
    }
#pragma unroll
    for (int t = 0; t < MT; ++t)
#pragma unroll
      for (int r = 0; r < 4; ++r) mine[(4 * g + r) * SP + 16 * t + nn] = acc[t][r];
    __syncthreads();
    // Ut_h = U_p * (P_other L_other), in the accumulator layout
    float u[MT][4];
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int m = 4 * g + r, n = nbase + m;
      float d = 0.f;
#pragma unroll
      for (int t = 0; t < MT; ++t) {
        const int i = 16 * t + nn;
        float up = 0.f;
        if (n < a.N && i < S) {
          if constexpr (ROOT) {
            up = a.freqs[(size_t)k * a.freqs_ks + i];
          } else {
            up = a.partials[((((size_t)(p - a.T) * a.K + k) * a.N + n) * C + c) * S + i];
          }
        }
        u[t][r] = up * other[m * SP + i];
        d = fmaf(u[t][r], acc[t][r], d);
      }
      if (h == 0) {
        d = row_sum16(d);
        if (nn == 0) dens[c * kTile + m] = pc * d;
      }
    }
    __syncthreads();
    // gradient rows: G += (coef Ut)^T L over the tile's 16 patterns
    float coef[4];
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int m = 4 * g + r, n = nbase + m;
      float den = 0.f;
      for (int cc = 0; cc < C; ++cc) den += dens[cc * kTile + m];
      coef[r] = n < a.N ? (float)a.weights[n] / den * pc : 0.f;
    }
#pragma unroll
    for (int ks = 0; ks < 4; ++ks) {
      // K index (ks, g) = pattern 4 g + ks: the A value of lane (nn, g) is its own
      // register u[Mt][ks] scaled by coef[ks]; B = L[4 g + ks][16 Nt + nn]
      float bv[MT];
#pragma unroll
      for (int nt = 0; nt < MT; ++nt)
        bv[nt] = child_val<TIPS>(a, chh, k, nbase + 4 * g + ks, c, 16 * nt + nn, S);
#pragma unroll
      for (int mt = 0; mt < MT; ++mt) {
        const float av = coef[ks] * u[mt][ks];
#pragma unroll
        for (int nt = 0; nt < MT; ++nt) G[mt][nt] = mma(av, bv[nt], G[mt][nt]);
      }
    }
    // transpose Ut through this wave's own LDS rows (P L rows are no longer read)
#pragma unroll
    for (int t = 0; t < MT; ++t)
#pragma unroll
      for (int r = 0; r < 4; ++r) mine[(4 * g + r) * SP + 16 * t + nn] = u[t][r];
    __syncthreads();
    f4 nu[MT];
    if (internal) {
#pragma unroll
      for (int t = 0; t < MT; ++t) nu[t] = f4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
      for (int ks = 0; ks < KS; ++ks) {
        const int i = g * KS + ks;  // K index (ks, g) = state g KS + ks
        const float av = mine[nn * SP + i];
#pragma unroll
        for (int t = 0; t < MT; ++t) {
          const int j = 16 * t + nn;
          const float bv = (i < S && j < S) ? pm[i * S + j] : 0.f;
          nu[t] = mma(av, bv, nu[t]);
        }
      }
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        float m_ = 0.f;
#pragma unroll
        for (int t = 0; t < MT; ++t) m_ = fmaxf(m_, nu[t][r]);
        m_ = row_max16(m_);
        if (nn == 0) mxx[(h * C + c) * kTile + 4 * g + r] = m_;
      }
    }
    __syncthreads();
    if (internal) {
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int m = 4 * g + r, n = nbase + m;
        float M = 0.f;
        for (int cc = 0; cc < C; ++cc) M = fmaxf(M, mxx[(h * C + cc) * kTile + m]);
        int e = 0;
        if (M > 0.f && isfinite(M)) frexpf(M, &e);
        if (n < a.N) {
          const size_t row = ((size_t)(chh - a.T) * a.K + k) * a.N + n;
#pragma unroll
          for (int t = 0; t < MT; ++t)
            if (16 * t + nn < S) a.partials[(row * C + c) * S + 16 * t + nn] = ldexpf(nu[t][r], -e);
        }
      }
    }
    __syncthreads();
  }
  if (a.mats_dt != nullptr) {
    const float* q = a.mats_dt + (((size_t)k * a.B + chh) * C + c) * S * S;
    double s = 0.0;
#pragma unroll
    for (int mt = 0; mt < MT; ++mt)
#pragma unroll
      for (int nt = 0; nt < MT; ++nt)
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          const int i = 16 * mt + 4 * g + r, j = 16 * nt + nn;
          if (i < S && j < S) s = fma((double)G[mt][nt][r], (double)q[i * S + j], s);
        }
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) s += __shfl_xor(s, o);
    if (lane == 0) a.gtau[(((size_t)chh * a.K + k) * a.nchunks + chunk) * C + c] = s;
  } else {
    const size_t css = (size_t)C * S * S;
    float* o = a.gpart + (((size_t)chh * a.K + k) * a.nchunks + chunk) * css + (size_t)c * S * S;
#pragma unroll
    for (int mt = 0; mt < MT; ++mt)
#pragma unroll
      for (int nt = 0; nt < MT; ++nt)
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          const int i = 16 * mt + 4 * g + r, j = 16 * nt + nn;
          if (i < S && j < S) o[i * S + j] = G[mt][nt][r];
        }
  }
}

bool use_mfma(int C, int S, int dtype) { return dtype == HX_F32 && S != 4 && C >= 1 && C <= 4; }

static size_t mfma_lds(int C, int S, bool pre) {
  const int MT = S <= 16 ? 1 : (S <= 32 ? 2 : 4);
  const int SP = 16 * MT;
  if (pre) return (size_t)(2 * C * kTile * SP + 3 * C * kTile) * 4 + 64;
  return (size_t)(C * kTile * SP + C * kTile) * 4 + (size_t)(C * kTile + C * SP) * 8 + 3 * kTile * 4 +
         64;
}

int launch_mfma(bool post, const PassArgs<float>& a, const int32_t* nodes, int width, int tips,
                bool root, int S, hipStream_t st) {
  const int MT = S <= 16 ? 1 : (S <= 32 ? 2 : 4);
  const int threads = 2 * a.C * 64;
  const size_t lds = mfma_lds(a.C, S, !post);
  dim3 grid(a.nchunks, a.K, width);
#define HX_MF(KER, MTV, TV, RV) KER<MTV, TV, RV><<<grid, threads, lds, st>>>(a, nodes, S)
#define HX_MF_ALL(KER, MTV)                                                     \
  if (tips == HX_TIPS_CODES) {                                                  \
    if (root) HX_MF(KER, MTV, HX_TIPS_CODES, true); else HX_MF(KER, MTV, HX_TIPS_CODES, false); \
  } else {                                                                      \
    if (root) HX_MF(KER, MTV, HX_TIPS_DENSE, true); else HX_MF(KER, MTV, HX_TIPS_DENSE, false); \
  }
  if (post) {
    if (MT == 1) { HX_MF_ALL(k_post_mfma, 1) } else if (MT == 2) { HX_MF_ALL(k_post_mfma, 2) } else { HX_MF_ALL(k_post_mfma, 4) }
  } else {
    if (MT == 1) { HX_MF_ALL(k_pre_mfma, 1) } else if (MT == 2) { HX_MF_ALL(k_pre_mfma, 2) } else { HX_MF_ALL(k_pre_mfma, 4) }
  }
#undef HX_MF_ALL
#undef HX_MF
  return HX_OK;
}

// ------------------------------------------------------- layout self-test
// D = A B for one 16x16x4 f32 MFMA with A[i][k] = 1 + i + 16 k and
// B[k][j] = 1 + 0.5 j - k, read back through the assumed lane layouts.
__global__ void k_mfma_layout(float* out) {
  const int lane = threadIdx.x;
  const int i = lane & 15, kk = lane >> 4;
  const float av = 1.f + i + 16.f * kk;         // A[i][k], lane = i + 16 k
  const float bv = 1.f + 0.5f * i - (float)kk;  // B[k][j], lane = j + 16 k (j = lane & 15)
  f4 d = mma(av, bv, f4{0.f, 0.f, 0.f, 0.f});
#pragma unroll
  for (int r = 0; r < 4; ++r) out[(4 * (lane >> 4) + r) * 16 + (lane & 15)] = d[r];
}

}  // namespace hx

extern "C" int hx_selftest_mfma_layout(int* mismatches, hx_stream_t stream) {
  HX_REQUIRE(mismatches != nullptr, HX_ERR_INVALID_ARGUMENT, "null argument");
  hipStream_t st = reinterpret_cast<hipStream_t>(stream);
  float* d = nullptr;
  HX_HIP_TRY(hipMalloc(&d, 256 * sizeof(float)));
  hx::k_mfma_layout<<<1, 64, 0, st>>>(d);
  float host[256];
  hipError_t e = hipMemcpyAsync(host, d, sizeof(host), hipMemcpyDeviceToHost, st);
  if (e == hipSuccess) e = hipStreamSynchronize(st);
  hipFree(d);
  if (e != hipSuccess) {
    hx::set_error(std::string("mfma self-test: ") + hipGetErrorString(e));
    return HX_ERR_HIP;
  }
  int bad = 0;
  for (int i = 0; i < 16; ++i)
    for (int j = 0; j < 16; ++j) {
      double ref = 0.0;
      for (int kk = 0; kk < 4; ++kk) ref += (1.0 + i + 16.0 * kk) * (1.0 + 0.5 * j - kk);
      if (fabs(host[i * 16 + j] - ref) > 1e-3 * (1.0 + fabs(ref))) ++bad;
    }
  *mismatches = bad;
  return HX_OK;
}

// Pruning kernels for dense state spaces on the matrix cores (MFMA), FP32
// (v_mfma_f32_16x16x4f32) and FP64 (v_mfma_f64_16x16x4f64).
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
//                                                       A is the wave's own U' registers,
//                                                       B the child's tile staged in LDS
//   transpose  U[m][j]  = sum_i Ut[m][i] P[i][j]        A from an LDS transpose of Ut
// S is padded to SP = 16 * MT states; K index k = 4 ks + g of a step ks maps to
// state g * KV + ks (forward / transpose) or to pattern 4 g + ks (gradient),
// so each lane's A operand is a contiguous run and the accumulator layout of
// one product is the operand layout of the next. The contractions over states
// take KV steps with 4 KV >= S: KV = 4 MT in general and 5 for 17..20 states
// (amino acids), which skips the padded states' steps. The 16x16x4 f32 layout used:
// A[i][k] from lane i + 16 k, B[k][j] from lane j + 16 k, D[i][j] in lane
// j + 16 (i / 4), register i % 4 for FP32 and lane j + 16 (i % 4), register
// i / 4 for FP64 (acc_row; checked on the device by hx_selftest_mfma_layout).
#include "pruning.hpp"

#include <math.h>

namespace hx {

namespace {

template <typename R>
using vec4_t = R __attribute__((ext_vector_type(4)));
// four floats at a 4-byte aligned address (one 16-byte load)
typedef float f32x4_t __attribute__((ext_vector_type(4)));
typedef f32x4_t f32x4u_t __attribute__((aligned(4)));
constexpr int kTile = 16;

// One 16x16x4 product D += A B on the matrix cores; FP32 and FP64 use the
// same operand and accumulator lane layout (checked on the device by
// hx_selftest_mfma_layout).
__device__ __forceinline__ vec4_t<float> mma(float a, float b, vec4_t<float> c) {
  return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
}
__device__ __forceinline__ vec4_t<double> mma(double a, double b, vec4_t<double> c) {
  return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
}

// ---- split-precision products on the BF16 matrix cores
//
// A float x is carried as hi = bf16(x) and lo = bf16(x - hi) (both rounded to nearest even;
// x - hi is exact in FP32), which keeps 16 significant bits: |x - hi - lo| <= 2^-17 |x|.
// A product of two such operands is formed as lo_a hi_b + hi_a lo_b + hi_a hi_b on
// v_mfma_f32_16x16x32_bf16 (8 operand values per lane, K = 32) or v_mfma_f32_16x16x16_bf16
// (4 per lane, K = 16), accumulating in FP32; the dropped lo_a lo_b term is below 2^-16 of
// the product. Operand and accumulator lanes follow the FP32 instruction's: A[i][k] in lane
// i + 16 (k / n), element k % n, B[k][j] in lane j + 16 (k / n), and the FP32 16x16
// accumulator (checked on the device by hx_selftest_mfma_layout). Every caller builds A and
// B with the same map from (lane group, element) to the contraction index, so the sum runs
// over the same index pairs as the FP32 products.
typedef __bf16 bf16x8_t __attribute__((ext_vector_type(8)));
typedef __bf16 bf16x4_t __attribute__((ext_vector_type(4)));
template <int N>
struct BfPack;
template <>
struct BfPack<8> {
  bf16x8_t hi, lo;
};
template <>
struct BfPack<4> {
  bf16x4_t hi, lo;
};

// hi and lo halves of x, as one 32-bit word (hi in the low 16 bits): the form kept in LDS and
// in the padded copies of P, so that an operand is assembled without converting again
__device__ __forceinline__ uint32_t bf_pair(float x) {
  const __bf16 hi = (__bf16)x;
  const __bf16 lo = (__bf16)(x - (float)hi);
  return (uint32_t)__builtin_bit_cast(uint16_t, hi) |
         ((uint32_t)__builtin_bit_cast(uint16_t, lo) << 16);
}
__device__ __forceinline__ __bf16 pair_hi(uint32_t p) {
  return __builtin_bit_cast(__bf16, (uint16_t)(p & 0xffffu));
}
__device__ __forceinline__ __bf16 pair_lo(uint32_t p) {
  return __builtin_bit_cast(__bf16, (uint16_t)(p >> 16));
}

// Four hi/lo words from 16 aligned bytes. The whole vector is converted, then indexed:
// a bit cast of one element of a vector lvalue reads element 0 whatever the index
// (hipcc turns __builtin_bit_cast(uint32_t, v[e]) into a read of v's first word).
typedef uint32_t u32x4_t __attribute__((ext_vector_type(4)));
__device__ __forceinline__ u32x4_t load_pairs4(const float* p) {
  return __builtin_bit_cast(u32x4_t, *reinterpret_cast<const vec4_t<float>*>(p));
}

template <int N>
__device__ __forceinline__ void pack_set(BfPack<N>& b, int e, float x) {
  const __bf16 hi = (__bf16)x;
  b.hi[e] = hi;
  b.lo[e] = (__bf16)(x - (float)hi);
}
template <int N>
__device__ __forceinline__ void pack_set_pair(BfPack<N>& b, int e, uint32_t p) {
  b.hi[e] = pair_hi(p);
  b.lo[e] = pair_lo(p);
}

__device__ __forceinline__ vec4_t<float> mma3(const BfPack<8>& a, const BfPack<8>& b,
                                              vec4_t<float> c) {
  c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a.lo, b.hi, c, 0, 0, 0);
  c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a.hi, b.lo, c, 0, 0, 0);
  return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a.hi, b.hi, c, 0, 0, 0);
}
__device__ __forceinline__ vec4_t<float> mma3(const BfPack<4>& a, const BfPack<4>& b,
                                              vec4_t<float> c) {
  c = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a.lo, b.hi, c, 0, 0, 0);
  c = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a.hi, b.lo, c, 0, 0, 0);
  return __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a.hi, b.hi, c, 0, 0, 0);
}

// The state contractions (forward, transpose) take KV values per lane; in the split form
// they go as packs of KB = 8 (one K = 32 instruction covers four lane groups of 8 states)
// or KB = 4 for KV = 4, NB packs per lane, values past KV zero.
template <int KV>
struct SplitShape {
  static constexpr int KB = KV <= 4 ? 4 : 8;
  static constexpr int NB = (KV + KB - 1) / KB;
};

__device__ __forceinline__ float fmax_(float a, float b) { return fmaxf(a, b); }
__device__ __forceinline__ double fmax_(double a, double b) { return fmax(a, b); }
__device__ __forceinline__ float frexp_(float m, int* e) { return frexpf(m, e); }
__device__ __forceinline__ double frexp_(double m, int* e) { return frexp(m, e); }
__device__ __forceinline__ float ldexp_(float x, int e) { return ldexpf(x, e); }
__device__ __forceinline__ double ldexp_(double x, int e) { return ldexp(x, e); }

// Row of the 16x16 accumulator held by lane group g (lane = column + 16 g)
// in register r: FP32 keeps rows 4g..4g+3 in a lane, FP64 rows g, g+4, g+8,
// g+12 (checked on the device by hx_selftest_mfma_layout).
template <typename R>
__device__ __forceinline__ constexpr int acc_row(int g, int r) {
  return sizeof(R) == 4 ? 4 * g + r : g + 4 * r;
}

template <typename R>
__device__ __forceinline__ R row_max16(R v) {
#pragma unroll
  for (int o = 1; o < 16; o <<= 1) v = fmax_(v, __shfl_xor(v, o));
  return v;
}

template <typename T>
__device__ __forceinline__ T row_sum16(T v) {
#pragma unroll
  for (int o = 1; o < 16; o <<= 1) v += __shfl_xor(v, o);
  return v;
}

// The same 16-lane reductions for FP32 through DPP row permutes (no LDS
// round trip): quad_perm [1,0,3,2] and [2,3,0,1] reduce each quad, the
// half-row mirror pairs quads, the row mirror pairs half-rows, so every lane
// ends with its row's result. Every lane of the wave must be active.
template <int CTRL>
__device__ __forceinline__ float dpp_f32(float v) {
  // bound_ctrl set: the form the compiler folds into the consuming add / max
  // (every source lane of these permutes is valid, so it never applies)
  return __builtin_bit_cast(float,
                            __builtin_amdgcn_mov_dpp(__builtin_bit_cast(int, v), CTRL, 0xF, 0xF, true));
}
// The maximum is taken on the bit patterns as signed integers, which order
// non-negative floats as floats do and put every negative value and -0 below
// +0; the callers' inputs are maxima started at +0, so the result is the
// float maximum (and the integer max folds the permute into one instruction).
template <int CTRL>
__device__ __forceinline__ int dpp_i32(int x) {
  return __builtin_amdgcn_mov_dpp(x, CTRL, 0xF, 0xF, true);
}
__device__ __forceinline__ float row_max16_dpp(float v) {
  int x = __builtin_bit_cast(int, v);
  x = max(x, dpp_i32<0xB1>(x));
  x = max(x, dpp_i32<0x4E>(x));
  x = max(x, dpp_i32<0x141>(x));
  x = max(x, dpp_i32<0x140>(x));
  return __builtin_bit_cast(float, x);
}
__device__ __forceinline__ float row_sum16_dpp(float v) {
  v += dpp_f32<0xB1>(v);
  v += dpp_f32<0x4E>(v);
  v += dpp_f32<0x141>(v);
  return v + dpp_f32<0x140>(v);
}
template <typename R>
__device__ __forceinline__ R rmax16(R v) {
  if constexpr (sizeof(R) == 4) return row_max16_dpp(v);
  else return row_max16(v);
}
template <typename R>
__device__ __forceinline__ R rsum16(R v) {
  if constexpr (sizeof(R) == 4) return row_sum16_dpp(v);
  else return row_sum16(v);
}
// w / den: the hardware reciprocal (1 ulp) in FP32, a division in FP64
__device__ __forceinline__ float div_(float w, float den) { return w * __builtin_amdgcn_rcpf(den); }
__device__ __forceinline__ double div_(double w, double den) { return w / den; }

// States s0 .. s0 + KS - 1 of child ch at absolute pattern n, category c (0
// outside): one code load for a tip, 16-byte loads for an internal child's
// row when S, s0 and KS are multiples of the vector width (rows then start
// 16-byte aligned and each load fills whole elements of out).
template <typename R, int TIPS, int KS>
__device__ __forceinline__ void child_row(const PassArgs<R>& a, int ch, int k, int n, int c, int s0,
                                          int S, R (&out)[KS]) {
#pragma unroll
  for (int q = 0; q < KS; ++q) out[q] = R(0);
  if (n >= a.N || s0 >= S) return;
  const R* src;
  if (ch < a.T) {
    if constexpr (TIPS == HX_TIPS_CODES) src = a.table + (size_t)a.codes[(size_t)ch * a.N + n] * S + s0;
    else src = a.dense + ((size_t)ch * a.N + n) * S + s0;
  } else {
    src = a.partials + ((((size_t)(ch - a.T) * a.K + k) * a.N + n) * a.C + c) * S + s0;
  }
  constexpr int VW = 16 / sizeof(R);  // elements per 16-byte load
  if (KS % VW == 0 && S % VW == 0 && s0 % VW == 0) {
#pragma unroll
    for (int q = 0; q < KS; q += VW) {
      if (s0 + q + VW <= S) {
        if constexpr (sizeof(R) == 4) {
          const float4 v = *reinterpret_cast<const float4*>(src + q);
          out[q] = v.x, out[q + 1] = v.y, out[q + 2] = v.z, out[q + 3] = v.w;
        } else {
          const double2 v = *reinterpret_cast<const double2*>(src + q);
          out[q] = v.x, out[q + 1] = v.y;
        }
      }
    }
  } else {
#pragma unroll
    for (int q = 0; q < KS; ++q)
      if (s0 + q < S) out[q] = src[q];
  }
}

// The forward product's P operand, held in registers for a chunk, and the product
// acc[t] += L P^T over the lane's KV states g KV .. g KV + KV - 1:
//   plain: f[t][ks] = P[16 t + nn][g KV + ks], KV steps of the 16x16x4 instruction;
//   split: the same values as BF16 hi/lo packs b[t][ks / KB] (element ks % KB, zero past
//          KV), NB split products per output tile.
template <typename R, int MT, int KV, bool SPLIT>
struct FwdP {
  static_assert(!SPLIT || sizeof(R) == 4, "split products take FP32 operands");
  static constexpr int KB = SplitShape<KV>::KB, NB = SplitShape<KV>::NB;
  R f[SPLIT ? 1 : MT][SPLIT ? 1 : KV];
  BfPack<KB> b[SPLIT ? MT : 1][SPLIT ? NB : 1];

  __device__ __forceinline__ void load(const R* pm, int S, int nn, int g) {
#pragma unroll
    for (int t = 0; t < MT; ++t)
#pragma unroll
      for (int q = 0; q < (SPLIT ? NB * KB : KV); ++q) {
        const int i = 16 * t + nn, j = g * KV + q;
        const R v = (q < KV && i < S && j < S) ? pm[i * S + j] : R(0);
        if constexpr (SPLIT) pack_set(b[t][q / KB], q % KB, v);
        else f[t][q] = v;
      }
  }

  __device__ __forceinline__ void product(const R (&arow)[KV], vec4_t<R> (&acc)[MT]) const {
    if constexpr (SPLIT) {
      BfPack<KB> ap[NB];
#pragma unroll
      for (int q = 0; q < NB * KB; ++q) pack_set(ap[q / KB], q % KB, q < KV ? arow[q] : 0.f);
#pragma unroll
      for (int bb = 0; bb < NB; ++bb)
#pragma unroll
        for (int t = 0; t < MT; ++t) acc[t] = mma3(ap[bb], b[t][bb], acc[t]);
    } else {
#pragma unroll
      for (int ks = 0; ks < KV; ++ks)
#pragma unroll
        for (int t = 0; t < MT; ++t) acc[t] = mma(arow[ks], f[t][ks], acc[t]);
    }
  }
};

}  // namespace

// ---------------------------------------------------------------- post-order
// Launched for the root (ROOT = true: the per-chunk root sums of loglik, d/dfreqs
// and d/dprops); the levels below use k_post_mfma_r.
template <typename R, int MT, int KV, int TIPS, bool ROOT, bool SPLIT = false>
__global__ __launch_bounds__(512) void k_post_mfma(PassArgs<R> a,
                                                   const int32_t* __restrict__ nodes, int S) {
  constexpr int SP = 16 * MT;
  extern __shared__ __align__(16) unsigned char smem_raw[];
  const int C = a.C;
  R* Yx = reinterpret_cast<R*>(smem_raw);       // [C][16][SP] child-1 products
  R* mx = Yx + C * kTile * SP;                      // [C][16]
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

  FwdP<R, MT, KV, SPLIT> fp;  // P[16 t + nn][g KV + ks]
  fp.load(a.mats + (((size_t)k * a.B + chh) * C + c) * S * S, S, nn, g);
  R pi_t[MT];
  R pc = R(0);
  double acc_l = 0.0, acc_p = 0.0, acc_f[MT];
  if constexpr (ROOT) {
#pragma unroll
    for (int t = 0; t < MT; ++t) {
      const int i = 16 * t + nn;
      pi_t[t] = i < S ? a.freqs[(size_t)k * a.freqs_ks + i] : R(0);
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
    vec4_t<R> acc[MT];
#pragma unroll
    for (int t = 0; t < MT; ++t) acc[t] = vec4_t<R>{R(0), R(0), R(0), R(0)};
    R arow[KV];
    child_row<R, TIPS, KV>(a, chh, k, nbase + nn, c, g * KV, S, arow);
    fp.product(arow, acc);
    // lane (nn, g): acc[t][r] = Y[m = 4 g + r][i = 16 t + nn]
    if (h == 1) {
#pragma unroll
      for (int t = 0; t < MT; ++t)
#pragma unroll
        for (int r = 0; r < 4; ++r) Yx[(c * kTile + acc_row<R>(g, r)) * SP + 16 * t + nn] = acc[t][r];
    }
    lds_barrier();
    R x[MT][4];
    if (h == 0) {
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        R m_ = R(0);
#pragma unroll
        for (int t = 0; t < MT; ++t) {
          x[t][r] = acc[t][r] * Yx[(c * kTile + acc_row<R>(g, r)) * SP + 16 * t + nn];
          m_ = fmax_(m_, x[t][r]);
        }
        m_ = row_max16(m_);
        if (nn == 0) mx[c * kTile + acc_row<R>(g, r)] = m_;
      }
    }
    lds_barrier();
    if (h == 0) {
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int m = acc_row<R>(g, r), n = nbase + m;
        R M = R(0);
        for (int cc = 0; cc < C; ++cc) M = fmax_(M, mx[cc * kTile + m]);
        int e = 0;
        if (M > R(0) && isfinite(M)) {
          frexp_(M, &e);
#pragma unroll
          for (int t = 0; t < MT; ++t) x[t][r] = ldexp_(x[t][r], -e);
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
          R p_ = R(0);
#pragma unroll
          for (int t = 0; t < MT; ++t) p_ = fma(pi_t[t], x[t][r], p_);
          p_ = row_sum16(p_);
          if (nn == 0) pxs[c * kTile + m] = (double)p_;
          if (c == 0 && nn == 0) Er[m] = E;
        }
      }
    }
    if constexpr (ROOT) {
      lds_barrier();
      if (h == 0) {
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          const int m = acc_row<R>(g, r), n = nbase + m;
          if (n >= a.N) continue;
          double lik = 0.0;
          for (int cc = 0; cc < C; ++cc)
            lik += (double)(a.props[(size_t)k * a.props_ks + cc] * (R)pxs[cc * kTile + m]);
          const double w = a.weights[n];
          const double wi = w / lik;
          if (c == 0 && nn == 0) acc_l += w * (log(lik) + (double)Er[m] * kLn2);
#pragma unroll
          for (int t = 0; t < MT; ++t) acc_f[t] += wi * (double)(pc * x[t][r]);
          if (nn == 0) acc_p += wi * pxs[c * kTile + m];
        }
      }
    }
    lds_barrier();
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
    lds_barrier();
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

// Post-order below the root (P in registers at up to 64 padded states). Every global load
// of a tile is issued on every path so the compiler's waits are exact: child
// rows and states are clamped to the child's own last row and state and then
// zeroed by a multiply (every value read is a written partial), each wave
// loads the children's scales of the tile's rows (wave 0 stages them in LDS,
// double-buffered by tile parity so two barriers per tile suffice), and a tip
// child's codes for the chunk are staged in LDS once. Same products as
// k_post_mfma.
template <typename R, int MT, int KV, int TIPS, bool SPLIT = false>
__global__ __launch_bounds__(512) void k_post_mfma_r(PassArgs<R> a,
                                                     const int32_t* __restrict__ nodes, int S) {
  constexpr int SP = 16 * MT;
  // LDS row stride of the product tiles: rows 4g + r of the accumulator layout start 16g banks
  // apart (SP would put the four lane groups g on the same banks)
  constexpr int LS = SP + 4;
  extern __shared__ __align__(16) unsigned char smem_raw[];
  const int C = a.C;
  R* Yx = reinterpret_cast<R*>(smem_raw);              // [C][16][LS] child-1 products
  R* mx = Yx + C * kTile * LS;                         // [16][4] per-row category maxima
  int* Et = reinterpret_cast<int*>(mx + 4 * kTile);    // [2 parity][2 children][16]
  uint8_t* ctile = reinterpret_cast<uint8_t*>(Et + 4 * kTile);  // [2][tpc * 16] codes

  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int lane = threadIdx.x & 63;
  const int h = wave / C, c = wave - h * C;
  const int nn = lane & 15, g = lane >> 4;
  // node indices in scalar registers (the parent slab's descriptor must be wave-uniform)
  const int chunk = blockIdx.x, k = blockIdx.y;
  const int v = __builtin_amdgcn_readfirstlane(nodes[blockIdx.z]);
  const int ch0 = __builtin_amdgcn_readfirstlane(a.kids[2 * v]);
  const int ch1 = __builtin_amdgcn_readfirstlane(a.kids[2 * v + 1]);
  const int chh = h ? ch1 : ch0;
  const bool internal = chh >= a.T;
  const int CS = C * S;
  const size_t slab = (size_t)a.N * CS;
  const R* cslab = a.partials + ((size_t)(internal ? chh - a.T : 0) * a.K + k) * slab;
  R* vslab = a.partials + ((size_t)(v - a.T) * a.K + k) * slab;
  const int t0 = chunk * a.tiles_per_chunk;
  const int t1 = min(t0 + a.tiles_per_chunk, a.ntiles);
  const int npat = a.tiles_per_chunk * kTile;

  if constexpr (TIPS == HX_TIPS_CODES) {
    for (int q = threadIdx.x; q < 2 * npat; q += blockDim.x) {
      const int hh = q >= npat, m = q - hh * npat, n = t0 * kTile + m;
      const int chq = hh ? ch1 : ch0;
      ctile[q] = (chq < a.T && n < a.N) ? a.codes[(size_t)chq * a.N + n] : 0;
    }
  }
  // categories C .. 3 of mx stay zero, so a row's maximum reads all four entries
  for (int q = threadIdx.x; q < 4 * kTile; q += blockDim.x) mx[q] = R(0);

  FwdP<R, MT, KV, SPLIT> fp;  // P[16 t + nn][g KV + ks]
  fp.load(a.mats + (((size_t)k * a.B + chh) * C + c) * S * S, S, nn, g);
  bool stok[MT];
#pragma unroll
  for (int t = 0; t < MT; ++t) stok[t] = 16 * t + nn < S;
  const uint32_t rstep = sizeof(R) == 4 ? CS : 4 * CS;
  const uint32_t ub = (uint32_t)(acc_row<R>(g, 0) * CS + c * S + nn);
  const uint8_t* mycodes = ctile + h * npat;
  // children's scales: lane (q = lane & 15, child (lane >> 4) & 1); a tip's reads
  // land on slab 0 of the scales and are multiplied by 0
  const int chq = ((lane >> 4) & 1) ? ch1 : ch0;
  const int32_t* sq = a.scales + ((size_t)(chq >= a.T ? chq - a.T : 0) * a.K + k) * a.N;
  const float qmask = chq >= a.T ? 1.f : 0.f;
  lds_barrier();

  // a tile's child rows, loaded as they are (rows clamped to N - 1; states past S read the
  // row's last state, or with 16-byte loads the words after the row, which the padding after
  // the partials keeps inside the workspace); the values past S or N are zeroed with a select
  // when the tile starts. With split products the rows are issued one tile ahead (128 taxa,
  // 10k patterns, 61 states: 1.79 -> 1.64 ms); the longer FP32 products hide their latency
  // (2.01 -> 2.07 ms when issued ahead), so FP32 issues them at the tile start
  auto post_rows = [&](int tl, R (&row)[KV]) {
    const int ncl = min(tl * kTile + nn, a.N - 1);
    if constexpr (sizeof(R) == 4 && KV % 4 == 0) {
      if (internal) {
        const R* rp = cslab + (uint32_t)(ncl * CS + c * S + g * KV);
#pragma unroll
        for (int q = 0; q < KV; q += 4) {
          const f32x4u_t w = *reinterpret_cast<const f32x4u_t*>(rp + q);
#pragma unroll
          for (int e = 0; e < 4; ++e) row[q + e] = w[e];
        }
        return;
      }
    }
    const R* bp;
    uint32_t base, last;  // first element and last state of the row read
    if (internal) {
      bp = cslab, base = (uint32_t)(ncl * CS + c * S);
    } else if constexpr (TIPS == HX_TIPS_CODES) {
      bp = a.table, base = (uint32_t)(mycodes[(tl - t0) * kTile + nn] * S);
    } else {
      bp = a.dense + (size_t)chh * a.N * S, base = (uint32_t)(ncl * S);
    }
    last = base + S - 1;
    base += g * KV;
#pragma unroll
    for (int q = 0; q < KV; ++q) row[q] = bp[min(base + q, last)];
  };
  R arow[KV];
  if constexpr (SPLIT) post_rows(t0, arow);

  for (int tile = t0; tile < t1; ++tile) {
    const int nbase = tile * kTile;
    const uint32_t tb = (uint32_t)nbase * (uint32_t)CS;
    if constexpr (!SPLIT) post_rows(tile, arow);
    {
      const bool rok = nbase + nn < a.N;
#pragma unroll
      for (int q = 0; q < KV; ++q) arow[q] = (rok && g * KV + q < S) ? arow[q] : R(0);
    }
    const float sv = (float)sq[min(nbase + (lane & 15), a.N - 1)] * qmask;
    vec4_t<R> acc[MT];
#pragma unroll
    for (int t = 0; t < MT; ++t) acc[t] = vec4_t<R>{R(0), R(0), R(0), R(0)};
    fp.product(arow, acc);
    // the next tile's rows (the last tile re-reads its own, unused)
    if constexpr (SPLIT) post_rows(min(tile + 1, t1 - 1), arow);
    // lane (nn, g): acc[t][r] = Y[m = acc_row(g, r)][i = 16 t + nn]
    if (h == 1) {
#pragma unroll
      for (int t = 0; t < MT; ++t)
#pragma unroll
        for (int r = 0; r < 4; ++r) Yx[(c * kTile + acc_row<R>(g, r)) * LS + 16 * t + nn] = acc[t][r];
    }
    int* et = Et + (tile & 1) * 2 * kTile;
    if (wave == 0 && lane < 2 * kTile) et[lane] = (int)sv;
    lds_barrier();
    R x[MT][4];
    if (h == 0) {
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        R m_ = R(0);
#pragma unroll
        for (int t = 0; t < MT; ++t) {
          x[t][r] = acc[t][r] * Yx[(c * kTile + acc_row<R>(g, r)) * LS + 16 * t + nn];
          m_ = fmax_(m_, x[t][r]);
        }
        m_ = rmax16(m_);
        if (nn == 0) mx[acc_row<R>(g, r) * 4 + c] = m_;
      }
    }
    lds_barrier();
    if (h == 0) {
      const __amdgpu_buffer_rsrc_t vrs = __builtin_amdgcn_make_buffer_rsrc(
          vslab + (size_t)__builtin_amdgcn_readfirstlane(nbase) * CS, (short)0,
          kTile * CS * (int)sizeof(R), 0x00020000);
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int m = acc_row<R>(g, r), n = nbase + m;
        const vec4_t<R> mv = *reinterpret_cast<const vec4_t<R>*>(mx + m * 4);
        const R M = fmax_(fmax_(mv[0], mv[1]), fmax_(mv[2], mv[3]));
        int e = 0;
        if (M > R(0) && isfinite(M)) frexp_(M, &e);
        const int E = et[m] + et[kTile + m] + e;
        if (n < a.N) {
#pragma unroll
          for (int t = 0; t < MT; ++t)
            if (stok[t]) {
              if constexpr (sizeof(R) == 4) {
                // a tile's rows through a descriptor at the tile: lane offset of row
                // acc_row(g, r) in bytes, 64 t bytes as the scalar offset
                __builtin_amdgcn_raw_buffer_store_b32(
                    __builtin_bit_cast(uint32_t, ldexp_(x[t][r], -e)), vrs,
                    (int)((ub + r * rstep) * (uint32_t)sizeof(R)), 64 * t, 0);
              } else {
                vslab[tb + ub + r * rstep + 16 * t] = ldexp_(x[t][r], -e);
              }
            }
          if (c == 0 && nn == 0) a.scales[((size_t)(v - a.T) * a.K + k) * a.N + n] = E;
        }
      }
    }
    // no third barrier: h = 1 waves rewrite Yx only after the second barrier,
    // by which the h = 0 waves have read it; mx is rewritten after the next
    // first barrier; Et alternates between two buffers
  }
}

// ----------------------------------------------------------------- pre-order
template <typename R, int MT, int KV, int TIPS, bool ROOT, bool SPLIT = false>
__global__ __launch_bounds__(512) __attribute__((amdgpu_waves_per_eu(
    sizeof(R) == 4 && MT <= 2 ? 4 : 1))) void k_pre_mfma(PassArgs<R> a,
                                                  const int32_t* __restrict__ nodes, int S) {
  static_assert(!SPLIT, "k_pre_mfma serves FP64 at 64 padded states only");
  constexpr int SP = 16 * MT;
  extern __shared__ __align__(16) unsigned char smem_raw[];
  const int C = a.C;
  R* Yx = reinterpret_cast<R*>(smem_raw);  // [2][C][16][SP]: P L, then Ut transposes
  R* dens = Yx + 2 * C * kTile * SP;           // [C][16]
  R* mxx = dens + C * kTile;                   // [2][C][16]
  R* Lx = mxx + 2 * C * kTile;                 // [2][C][16][SP]: the child's partial tile

  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int h = wave / C, c = wave - h * C;
  const int nn = lane & 15, g = lane >> 4;
  const int chunk = blockIdx.x, k = blockIdx.y, p = nodes[blockIdx.z];
  const int ch0 = a.kids[2 * p], ch1 = a.kids[2 * p + 1];
  const int chh = h ? ch1 : ch0;
  const bool internal = chh >= a.T;
  const R* pm = a.mats + (((size_t)k * a.B + chh) * C + c) * S * S;

  // Pf[t][ks] = P[16 t + nn][g KV + ks]; at 64 padded states the gradient
  // accumulators take those registers and P is re-read from L1/L2 per tile
  constexpr bool kPfRegs = MT < 4;
  R Pf[kPfRegs ? MT : 1][kPfRegs ? KV : 1];
  if constexpr (kPfRegs) {
#pragma unroll
    for (int t = 0; t < MT; ++t)
#pragma unroll
      for (int ks = 0; ks < KV; ++ks) {
        const int i = 16 * t + nn, j = g * KV + ks;
        Pf[t][ks] = (i < S && j < S) ? pm[i * S + j] : R(0);
      }
  }
  const R pc = a.props[(size_t)k * a.props_ks + c];
  vec4_t<R> G[MT][MT];  // G[Mt][Nt] lane (nn, g) reg r = dlogL/dP[16 Mt + 4 g + r][16 Nt + nn]
#pragma unroll
  for (int i = 0; i < MT; ++i)
#pragma unroll
    for (int j = 0; j < MT; ++j) G[i][j] = vec4_t<R>{R(0), R(0), R(0), R(0)};
  R* mine = Yx + (size_t)(h * C + c) * kTile * SP;
  const R* other = Yx + (size_t)((1 - h) * C + c) * kTile * SP;
  R* lt = Lx + (size_t)(h * C + c) * kTile * SP;  // this wave's L tile [16 patterns][SP]
  if (4 * KV < SP) {
    // states 4 KV .. SP - 1 of the tile are never written: zero them once
#pragma unroll
    for (int q = lane; q < kTile * SP; q += 64) lt[q] = R(0);
  }

  const int t0 = chunk * a.tiles_per_chunk;
  const int t1 = min(t0 + a.tiles_per_chunk, a.ntiles);
  for (int tile = t0; tile < t1; ++tile) {
    const int nbase = tile * kTile;
    // forward P_h L_h
    // the parent's adjoint for this tile, issued with the child rows so the
    // two loads overlap (rows m = acc_row(g, r), states 16 t + nn)
    R up[MT][4];
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int n = nbase + acc_row<R>(g, r);
#pragma unroll
      for (int t = 0; t < MT; ++t) {
        const int i = 16 * t + nn;
        up[t][r] = R(0);
        if (n < a.N && i < S) {
          if constexpr (ROOT) {
            up[t][r] = a.freqs[(size_t)k * a.freqs_ks + i];
          } else {
            up[t][r] = a.partials[((((size_t)(p - a.T) * a.K + k) * a.N + n) * C + c) * S + i];
          }
        }
      }
    }
    vec4_t<R> acc[MT];
#pragma unroll
    for (int t = 0; t < MT; ++t) acc[t] = vec4_t<R>{R(0), R(0), R(0), R(0)};
    R arow[KV];
    child_row<R, TIPS, KV>(a, chh, k, nbase + nn, c, g * KV, S, arow);
    // the tile of L, read back transposed as the statistic's B operand (this
    // wave's own rows: no barrier)
#pragma unroll
    for (int q = 0; q < KV; ++q) lt[nn * SP + g * KV + q] = arow[q];
#pragma unroll
    for (int ks = 0; ks < KV; ++ks) {
      const R av = arow[ks];
#pragma unroll
      for (int t = 0; t < MT; ++t) {
        R bv;
        if constexpr (kPfRegs) {
          bv = Pf[t][ks];
        } else {
          const int i = 16 * t + nn, j = g * KV + ks;
          bv = (i < S && j < S) ? pm[i * S + j] : R(0);
        }
        acc[t] = mma(av, bv, acc[t]);
      }
    }
#pragma unroll
    for (int t = 0; t < MT; ++t)
#pragma unroll
      for (int r = 0; r < 4; ++r) mine[(acc_row<R>(g, r)) * SP + 16 * t + nn] = acc[t][r];
    __syncthreads();
    // Ut_h = U_p * (P_other L_other), in the accumulator layout
    R u[MT][4];
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int m = acc_row<R>(g, r);
      R d = R(0);
#pragma unroll
      for (int t = 0; t < MT; ++t) {
        const int i = 16 * t + nn;
        u[t][r] = up[t][r] * other[m * SP + i];
        d = fma(u[t][r], acc[t][r], d);
      }
      if (h == 0) {
        d = row_sum16(d);
        if (nn == 0) dens[c * kTile + m] = pc * d;
      }
    }
    __syncthreads();
    // gradient rows: G += (coef Ut)^T L over the tile's 16 patterns
    R coef[4];
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int m = acc_row<R>(g, r), n = nbase + m;
      R den = R(0);
      for (int cc = 0; cc < C; ++cc) den += dens[cc * kTile + m];
      coef[r] = n < a.N ? (R)a.weights[n] / den * pc : R(0);
    }
#pragma unroll
    for (int ks = 0; ks < 4; ++ks) {
      // K index (ks, g) = pattern 4 g + ks: the A value of lane (nn, g) is its own
      // register u[Mt][ks] scaled by coef[ks]; B = L[4 g + ks][16 Nt + nn]
      R bv[MT];
#pragma unroll
      for (int nt = 0; nt < MT; ++nt) bv[nt] = lt[acc_row<R>(g, ks) * SP + 16 * nt + nn];
#pragma unroll
      for (int mt = 0; mt < MT; ++mt) {
        const R av = coef[ks] * u[mt][ks];
#pragma unroll
        for (int nt = 0; nt < MT; ++nt) G[mt][nt] = mma(av, bv[nt], G[mt][nt]);
      }
    }
    // transpose Ut through this wave's own LDS rows: the other child's waves
    // read these P L rows before the previous barrier, and only this wave
    // reads the transpose, so no barrier is needed here
#pragma unroll
    for (int t = 0; t < MT; ++t)
#pragma unroll
      for (int r = 0; r < 4; ++r) mine[(acc_row<R>(g, r)) * SP + 16 * t + nn] = u[t][r];
    vec4_t<R> nu[MT];
    if (internal) {
#pragma unroll
      for (int t = 0; t < MT; ++t) nu[t] = vec4_t<R>{R(0), R(0), R(0), R(0)};
#pragma unroll
      for (int ks = 0; ks < KV; ++ks) {
        const int i = g * KV + ks;  // K index (ks, g) = state g KV + ks
        const R av = mine[nn * SP + i];
#pragma unroll
        for (int t = 0; t < MT; ++t) {
          const int j = 16 * t + nn;
          const R bv = (i < S && j < S) ? pm[i * S + j] : R(0);
          nu[t] = mma(av, bv, nu[t]);
        }
      }
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        R m_ = R(0);
#pragma unroll
        for (int t = 0; t < MT; ++t) m_ = fmax_(m_, nu[t][r]);
        m_ = row_max16(m_);
        if (nn == 0) mxx[(h * C + c) * kTile + acc_row<R>(g, r)] = m_;
      }
    }
    __syncthreads();
    if (internal) {
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int m = acc_row<R>(g, r), n = nbase + m;
        R M = R(0);
        for (int cc = 0; cc < C; ++cc) M = fmax_(M, mxx[(h * C + cc) * kTile + m]);
        int e = 0;
        if (M > R(0) && isfinite(M)) frexp_(M, &e);
        if (n < a.N) {
          const size_t row = ((size_t)(chh - a.T) * a.K + k) * a.N + n;
#pragma unroll
          for (int t = 0; t < MT; ++t)
            if (16 * t + nn < S) a.partials[(row * C + c) * S + 16 * t + nn] = ldexp_(nu[t][r], -e);
        }
      }
    }
    // no barrier: the next tile's first barrier orders this tile's reads of
    // dens and mxx before their next writes, and each wave rewrites only its
    // own rows of Yx and Lx before it
  }
  if (a.mats_dt != nullptr) {
    const R* q = a.mats_dt + (((size_t)k * a.B + chh) * C + c) * S * S;
    double s = 0.0;
#pragma unroll
    for (int mt = 0; mt < MT; ++mt)
#pragma unroll
      for (int nt = 0; nt < MT; ++nt)
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          const int i = 16 * mt + acc_row<R>(g, r), j = 16 * nt + nn;
          if (i < S && j < S) s = fma((double)G[mt][nt][r], (double)q[i * S + j], s);
        }
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) s += __shfl_xor(s, o);
    if (lane == 0) a.gtau[(((size_t)chh * a.K + k) * a.nchunks + chunk) * C + c] = s;
  } else {
    const size_t css = (size_t)C * S * S;
    R* o = a.gpart + (((size_t)chh * a.K + k) * a.nchunks + chunk) * css + (size_t)c * S * S;
#pragma unroll
    for (int mt = 0; mt < MT; ++mt)
#pragma unroll
      for (int nt = 0; nt < MT; ++nt)
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          const int i = 16 * mt + acc_row<R>(g, r), j = 16 * nt + nn;
          if (i < S && j < S) o[i * S + j] = G[mt][nt][r];
        }
  }
}

// Child rows of one tile for k_pre_mfma_r: states g KV .. g KV + KV - 1 of
// pattern tile * 16 + nn, every load issued (clamped caller buffers, padded
// workspace). The values are left as loaded, so that nothing waits on them
// until pre_rows_mask, at the start of the tile that uses them, zeroes the
// out-of-range ones by a multiply.
template <typename R, int KV, int TIPS>
__device__ __forceinline__ void pre_rows(const PassArgs<R>& a, bool internal, const R* cslab,
                                         int chh, const uint8_t* mycodes, int tile, int t0,
                                         int nn, int g, int S, int CS, uint32_t cb,
                                         R (&arow)[KV]) {
  const int nbase = tile * kTile;
  if constexpr (sizeof(R) == 4 && KV % 4 == 0) {
    // workspace rows: KV / 4 loads of 16 bytes (4-byte aligned: a category's row starts at
    // c S floats), no clamp needed
    if (internal) {
      const R* rp = cslab + (uint32_t)nbase * (uint32_t)CS + cb;
#pragma unroll
      for (int q = 0; q < KV; q += 4) {
        const f32x4u_t v = *reinterpret_cast<const f32x4u_t*>(rp + q);
#pragma unroll
        for (int e = 0; e < 4; ++e) arow[q + e] = v[e];
      }
      return;
    }
  }
  const R* bp;
  uint32_t base, lim;
  if (internal) {
    bp = cslab, base = (uint32_t)nbase * (uint32_t)CS + cb, lim = 0xffffffffu;
  } else if constexpr (TIPS == HX_TIPS_CODES) {
    const int code = mycodes[(tile - t0) * kTile + nn];
    bp = a.table, base = (uint32_t)(code * S + g * KV), lim = (uint32_t)(a.ncodes * S - 1);
  } else {
    bp = a.dense + (size_t)chh * a.N * S;
    base = (uint32_t)(min(nbase + nn, a.N - 1) * S + g * KV), lim = (uint32_t)(a.N * S - 1);
  }
#pragma unroll
  for (int q = 0; q < KV; ++q) arow[q] = bp[min(base + q, lim)];
}
template <typename R, int KV>
__device__ __forceinline__ void pre_rows_mask(int N, int tile, int nn, int g, int S,
                                              R (&arow)[KV]) {
  const bool nok = tile * kTile + nn < N;
#pragma unroll
  for (int q = 0; q < KV; ++q) arow[q] *= (nok && g * KV + q < S) ? R(1) : R(0);
}

// The 16 blocks of 16 bytes per lane of one padded P copy, consumed in order
// with D of them in flight: block b + D is issued when block b is taken.
template <int D>
struct PRing {
  u32x4_t r[D];
  template <typename F>
  __device__ __forceinline__ void start(F& ld, int part) {
#pragma unroll
    for (int b = 0; b < D; ++b) r[b] = ld(part, b);
  }
  template <typename F>
  __device__ __forceinline__ u32x4_t take(F& ld, int part, int b) {
    const u32x4_t v = r[b % D];
    if (b + D < 16) r[b % D] = ld(part, b + D);
    return v;
  }
};
constexpr int kPD = 4;

// Pre-order for at most 32 padded states with both P operands in registers
// (the forward's P^T and the transpose's P) for the whole chunk; at 64 padded
// states (FP32 only) the 64 gradient accumulators leave no room for them, and
// each tile reads them as aligned 16-byte loads from zero-padded 64 x 64 copies
// of P and P^T (ppad). Each tile
// issues the same global loads on every path, so the compiler's waits are
// exact: rows past N and states past S of the workspace partials read the
// padding that follows them, reads of caller buffers (codes, tables, dense
// tips, weights) use clamped indices, and every such value is then zeroed.
// Addresses are a (node, sample) slab base plus 32-bit element offsets,
// (n C + c) S + s. A tip child's codes for the chunk are staged in LDS once.
// Same products and barriers as k_pre_mfma.
template <typename R, int MT, int KV, int TIPS, bool ROOT, bool SPLIT = false>
__global__ __launch_bounds__(512) __attribute__((amdgpu_waves_per_eu(
    sizeof(R) == 4 && MT <= 2 ? (SPLIT && KV == 8 ? 3 : 4) : 1))) void k_pre_mfma_r(
    PassArgs<R> a, const int32_t* __restrict__ nodes, int S) {
  static_assert(MT <= 2 || (sizeof(R) == 4 && KV == 16), "64 padded states: FP32, 16 steps");
  // split products: the statistic's L tile, the transpose's P tile and the padded P copies
  // hold each value's BF16 hi/lo pair in its 32-bit slot
  using SSh = SplitShape<KV>;
  constexpr int KB = SSh::KB, NB = SSh::NB;
  constexpr bool kPReg = MT <= 2;
  constexpr int SP = 16 * MT;
  // LDS row stride of the [16][SP] tiles: at 32 padded states, SP + 4 starts rows 4g + r of the
  // accumulator layout 16g banks apart (SP puts the four lane groups g on the same banks); at 64
  // the extra address registers would cost the kernel spills, so the tiles stay dense
  constexpr int LS = MT <= 2 ? SP + 4 : SP;
  extern __shared__ __align__(16) unsigned char smem_raw[];
  const int C = a.C;
  R* Yx = reinterpret_cast<R*>(smem_raw);  // [2][C][16][LS]: P L, then Ut transposes
  R* dens = Yx + 2 * C * kTile * LS;       // [16][4]: per row, one entry per category
  R* mxx = dens + 4 * kTile;               // [2][16][4]
  R* Lx = mxx + 8 * kTile;                 // [2][C][16][LS]: the child's partial tile
  constexpr int kPtS = SP + 1;             // row stride of the P tiles (no bank conflicts)
  R* PtL = Lx + 2 * C * kTile * LS;        // [2][C][4 KV][SP + 1]: the transpose's P (MT <= 2)
  uint8_t* ctile = reinterpret_cast<uint8_t*>(PtL + (MT <= 2 ? 2 * C * 4 * KV * kPtS : 0));
  // [2][tpc * 16] codes

  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int lane = threadIdx.x & 63;
  const int h = wave / C, c = wave - h * C;
  const int nn = lane & 15, g = lane >> 4;
  // node indices read once per wave into scalar registers (the slab and P descriptors
  // below must be wave-uniform)
  const int chunk = blockIdx.x, k = blockIdx.y;
  const int p = __builtin_amdgcn_readfirstlane(nodes[blockIdx.z]);
  const int ch0 = __builtin_amdgcn_readfirstlane(a.kids[2 * p]);
  const int ch1 = __builtin_amdgcn_readfirstlane(a.kids[2 * p + 1]);
  const int chh = h ? ch1 : ch0;
  const bool internal = chh >= a.T;
  const int CS = C * S;
  const size_t slab = (size_t)a.N * CS;
  const R* pslab = a.partials + ((size_t)(p - a.T) * a.K + k) * slab;
  R* cslab = a.partials + ((size_t)(internal ? chh - a.T : 0) * a.K + k) * slab;
  const R* pm = a.mats + (((size_t)k * a.B + chh) * C + c) * S * S;
  const int t0 = chunk * a.tiles_per_chunk;
  const int t1 = min(t0 + a.tiles_per_chunk, a.ntiles);
  const int npat = a.tiles_per_chunk * kTile;

  if constexpr (TIPS == HX_TIPS_CODES) {
    // the chunk's codes of each tip child (0 past N)
    for (int q = threadIdx.x; q < 2 * npat; q += blockDim.x) {
      const int hh = q >= npat, m = q - hh * npat, n = t0 * kTile + m;
      const int chq = hh ? ch1 : ch0;
      ctile[q] = (chq < a.T && n < a.N) ? a.codes[(size_t)chq * a.N + n] : 0;
    }
  }

  // Pf[t][ks] = P[16 t + nn][g KV + ks] (forward), Pt[t][ks] = P[g KV + ks][16 t + nn]
  // (transpose)
  // the transpose's operand P[g KV + ks][16 t + nn] lives in this wave's LDS tile
  FwdP<R, kPReg ? MT : 1, kPReg ? KV : 1, SPLIT && kPReg> fp;
  R* ptw = PtL + (size_t)(h * C + c) * 4 * KV * kPtS;
  if constexpr (kPReg) {
    fp.load(pm, S, nn, g);
#pragma unroll
    for (int t = 0; t < MT; ++t)
#pragma unroll
      for (int ks = 0; ks < KV; ++ks) {
        const int i = 16 * t + nn, j = g * KV + ks;
        const R v = (i < S && j < S) ? pm[j * S + i] : R(0);
        if constexpr (SPLIT) ptw[j * kPtS + i] = __builtin_bit_cast(float, bf_pair(v));
        else ptw[j * kPtS + i] = v;
      }
  }
  // padded copies of P ([0]) and P^T ([1]) in lane order: lane (nn, g) finds M[16 t + nn][g KV
  // + s .. s + 3] of either at its own 16 bytes of block 4 t + s / 4 (256 floats a block), or
  // in the split form the hi and lo operands of states g KV + 8 b .. + 7 at blocks 4 t + 2 b
  // and 4 t + 2 b + 1; read as one raw buffer (a lane offset, and the block offset as a
  // scalar; the matrix index made provably wave-uniform, so the descriptor stays in scalar
  // registers)
  const R* ppb = kPReg ? nullptr
                       : a.ppad + (size_t)__builtin_amdgcn_readfirstlane((k * a.B + chh) * C + c) *
                                      (2 * 64 * 64);
  const __amdgpu_buffer_rsrc_t prs = __builtin_amdgcn_make_buffer_rsrc(
      const_cast<R*>(ppb), (short)0, kPReg ? 0 : 2 * 64 * 64 * (int)sizeof(R), 0x00020000);
  // 16 bytes of block `blk` of copy `part` (0: P, 1: P^T)
  auto pp4 = [&](int part, int blk) -> u32x4_t {
    return __builtin_amdgcn_raw_buffer_load_b128(prs, 16 * lane, (16 * part + blk) * 1024, 0);
  };
  const R pc = a.props[(size_t)k * a.props_ks + c];
  vec4_t<R> G[MT][MT];  // G[Mt][Nt] lane (nn, g) reg r = dlogL/dP[16 Mt + acc_row(g, r)][16 Nt + nn]
#pragma unroll
  for (int i = 0; i < MT; ++i)
#pragma unroll
    for (int j = 0; j < MT; ++j) G[i][j] = vec4_t<R>{R(0), R(0), R(0), R(0)};
  R* mine = Yx + (size_t)(h * C + c) * kTile * LS;
  const R* other = Yx + (size_t)((1 - h) * C + c) * kTile * LS;
  R* lt = Lx + (size_t)(h * C + c) * kTile * LS;  // this wave's L tile [16 patterns][LS]
  if (4 * KV < SP) {
    // states 4 KV .. SP - 1 of the tile are never written
    for (int q = lane; q < kTile * LS; q += 64) lt[q] = R(0);
  }
  // categories C .. 3 of dens and mxx stay zero, so sums and maxima over a
  // row read all four entries
  for (int q = threadIdx.x; q < 12 * kTile; q += blockDim.x) dens[q] = R(0);
  bool stok[MT];
#pragma unroll
  for (int t = 0; t < MT; ++t) stok[t] = 16 * t + nn < S;
  R upr[MT];  // root: the frequencies
  if constexpr (ROOT) {
#pragma unroll
    for (int t = 0; t < MT; ++t) {
      const R v = a.freqs[(size_t)k * a.freqs_ks + (stok[t] ? 16 * t + nn : 0)];
      upr[t] = stok[t] ? v : R(0);
    }
  }
  // lane offsets within a slab: rows acc_row(g, r) = acc_row(g, 0) + r (FP32) or + 4 r (FP64)
  const uint32_t rstep = sizeof(R) == 4 ? CS : 4 * CS;
  const uint32_t ub = (uint32_t)(acc_row<R>(g, 0) * CS + c * S + nn);
  const uint32_t cb = (uint32_t)(nn * CS + c * S + g * KV);
  const uint8_t* mycodes = ctile + h * npat;
  lds_barrier();
  // the first tile's child rows; each tile then issues the next tile's rows
  // after its forward product, so their latency overlaps the rest of the tile
  R arow[KV];
  pre_rows<R, KV, TIPS>(a, internal, cslab, chh, mycodes, t0, t0, nn, g, S, CS, cb, arow);
  // the parent's adjoint (zero past N and S) and the weights of a tile
  // in FP32 a tile's rows of a slab are read and written through a buffer descriptor at the
  // tile (scalar registers), with the lane's byte offset of row acc_row(g, r) in tvo[r] and
  // 64 t bytes as the scalar offset: no 64-bit address arithmetic per load or store
  uint32_t tvo[4];
#pragma unroll
  for (int r = 0; r < 4; ++r) tvo[r] = (ub + r * rstep) * (uint32_t)sizeof(R);
  auto tile_rsrc = [&](const R* sl, int tl) {
    return __builtin_amdgcn_make_buffer_rsrc(
        const_cast<R*>(sl + (size_t)__builtin_amdgcn_readfirstlane(tl * kTile) * CS), (short)0,
        kTile * CS * (int)sizeof(R), 0x00020000);
  };
  auto load_up = [&](int tl, R (&up)[MT][4], R (&wt)[4]) {
    const int nb = tl * kTile;
    const uint32_t tbl = (uint32_t)nb * (uint32_t)CS;
    if constexpr (!ROOT && sizeof(R) == 4) {
      const __amdgpu_buffer_rsrc_t urs = tile_rsrc(pslab, tl);
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const bool ok = nb + acc_row<R>(g, r) < a.N;
#pragma unroll
        for (int t = 0; t < MT; ++t)
          up[t][r] = __builtin_bit_cast(
                         R, __builtin_amdgcn_raw_buffer_load_b32(urs, (int)tvo[r], 64 * t, 0)) *
                     ((ok && stok[t]) ? R(1) : R(0));
      }
    } else {
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const bool ok = nb + acc_row<R>(g, r) < a.N;
#pragma unroll
        for (int t = 0; t < MT; ++t) {
          if constexpr (ROOT) {
            up[t][r] = ok ? upr[t] : R(0);
          } else {
            up[t][r] = pslab[tbl + ub + r * rstep + 16 * t] * ((ok && stok[t]) ? R(1) : R(0));
          }
        }
      }
    }
#pragma unroll
    for (int r = 0; r < 4; ++r) wt[r] = (R)a.weights[min(nb + acc_row<R>(g, r), a.N - 1)];
  };
  // at 64 padded states with split products the next tile's adjoint and weights are issued
  // with its child rows, after the forward product, so that the forward's P loads are not
  // queued behind loads from HBM issued just before them (128 taxa, 10k patterns, 61 states:
  // 3.30 against 3.83 ms; with FP32 products, whose longer forward hides those loads, it
  // cost time: 4.42 against 3.86 ms)
  constexpr bool kUpAhead = !kPReg && SPLIT;
  R upn[MT][4], wtn[4];
  if constexpr (kUpAhead) load_up(t0, upn, wtn);

  for (int tile = t0; tile < t1; ++tile) {
    const int nbase = tile * kTile;
    const uint32_t tb = (uint32_t)nbase * (uint32_t)CS;
    pre_rows_mask<R, KV>(a.N, tile, nn, g, S, arow);
    bool rok[4];
#pragma unroll
    for (int r = 0; r < 4; ++r) rok[r] = nbase + acc_row<R>(g, r) < a.N;
    R up[MT][4], wt[4];
    if constexpr (kUpAhead) {
#pragma unroll
      for (int t = 0; t < MT; ++t)
#pragma unroll
        for (int r = 0; r < 4; ++r) up[t][r] = upn[t][r];
#pragma unroll
      for (int r = 0; r < 4; ++r) wt[r] = wtn[r];
    } else {
      // the parent's adjoint, then the weights (the child rows were issued by the
      // previous tile)
      load_up(tile, up, wt);
    }

    // forward P_h L_h; the tile of L is kept as the statistic's B operand
    vec4_t<R> acc[MT];
#pragma unroll
    for (int t = 0; t < MT; ++t) acc[t] = vec4_t<R>{R(0), R(0), R(0), R(0)};
#pragma unroll
    for (int q = 0; q < KV; ++q) {
      if constexpr (SPLIT) lt[nn * LS + g * KV + q] = __builtin_bit_cast(float, bf_pair(arow[q]));
      else lt[nn * LS + g * KV + q] = arow[q];
    }
    if constexpr (kPReg) {
      fp.product(arow, acc);
    } else if constexpr (SPLIT) {
      // 64 padded states: KV = 16 = two packs of 8; blocks 4 t + 2 bb and 4 t + 2 bb + 1 of
      // the padded copy are the hi and lo operands of pack bb
      BfPack<8> ap[2];
#pragma unroll
      for (int q = 0; q < 16; ++q) pack_set(ap[q / 8], q % 8, arow[q]);
      PRing<kPD> ring;
      ring.start(pp4, 0);
#pragma unroll
      for (int t = 0; t < MT; ++t)
#pragma unroll
        for (int bb = 0; bb < 2; ++bb) {
          BfPack<8> bp;
          bp.hi = __builtin_bit_cast(bf16x8_t, ring.take(pp4, 0, 4 * t + 2 * bb));
          bp.lo = __builtin_bit_cast(bf16x8_t, ring.take(pp4, 0, 4 * t + 2 * bb + 1));
          acc[t] = mma3(ap[bb], bp, acc[t]);
        }
    } else {
      PRing<kPD> ring;
      ring.start(pp4, 0);
#pragma unroll
      for (int t = 0; t < MT; ++t)
#pragma unroll
        for (int q = 0; q < KV; q += 4) {
          const vec4_t<R> pv = __builtin_bit_cast(vec4_t<R>, ring.take(pp4, 0, 4 * t + q / 4));
#pragma unroll
          for (int e = 0; e < 4; ++e) acc[t] = mma(arow[q + e], pv[e], acc[t]);
        }
    }
#pragma unroll
    for (int t = 0; t < MT; ++t)
#pragma unroll
      for (int r = 0; r < 4; ++r) mine[acc_row<R>(g, r) * LS + 16 * t + nn] = acc[t][r];
    // the next tile's child rows (the last tile re-reads its own, unused)
    pre_rows<R, KV, TIPS>(a, internal, cslab, chh, mycodes, min(tile + 1, t1 - 1), t0, nn, g, S, CS,
                          cb, arow);
    if constexpr (kUpAhead) load_up(min(tile + 1, t1 - 1), upn, wtn);
    lds_barrier();
    // Ut_h = U_p * (P_other L_other), in the accumulator layout; the row maxima
    // of Ut over (category, state) set the adjoint's rescale exponent
    R u[MT][4];
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int m = acc_row<R>(g, r);
      R d = R(0), um = R(0);
#pragma unroll
      for (int t = 0; t < MT; ++t) {
        u[t][r] = up[t][r] * other[m * LS + 16 * t + nn];
        d = fma(u[t][r], acc[t][r], d);
        um = fmax_(um, u[t][r]);
      }
      if (h == 0) {
        d = rsum16(d);
        if (nn == 0) dens[m * 4 + c] = pc * d;
      }
      if (internal) {
        um = rmax16(um);
        if (nn == 0) mxx[(h * kTile + m) * 4 + c] = um;
      }
    }
    lds_barrier();
    // gradient rows: G += (coef Ut)^T L over the tile's 16 patterns
    R coef[4];
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int m = acc_row<R>(g, r);
      const vec4_t<R> dv = *reinterpret_cast<const vec4_t<R>*>(dens + m * 4);
      const R den = (dv[0] + dv[1]) + (dv[2] + dv[3]);
      coef[r] = rok[r] ? div_(wt[r], den) * pc : R(0);
    }
    if constexpr (SPLIT) {
      // one K = 16 split product per output tile: lane group g holds the patterns
      // acc_row(g, 0..3) = 4 g .. 4 g + 3 as elements 0..3 of A and B
      BfPack<4> bp[MT];
#pragma unroll
      for (int nt = 0; nt < MT; ++nt)
#pragma unroll
        for (int ks = 0; ks < 4; ++ks)
          pack_set_pair(bp[nt], ks, __builtin_bit_cast(
                                        uint32_t, lt[acc_row<R>(g, ks) * LS + 16 * nt + nn]));
#pragma unroll
      for (int mt = 0; mt < MT; ++mt) {
        BfPack<4> ap;
#pragma unroll
        for (int ks = 0; ks < 4; ++ks) pack_set(ap, ks, coef[ks] * u[mt][ks]);
#pragma unroll
        for (int nt = 0; nt < MT; ++nt) G[mt][nt] = mma3(ap, bp[nt], G[mt][nt]);
      }
    } else {
#pragma unroll
      for (int ks = 0; ks < 4; ++ks) {
        R bv[MT];
#pragma unroll
        for (int nt = 0; nt < MT; ++nt) bv[nt] = lt[acc_row<R>(g, ks) * LS + 16 * nt + nn];
#pragma unroll
        for (int mt = 0; mt < MT; ++mt) {
          const R av = coef[ks] * u[mt][ks];
#pragma unroll
          for (int nt = 0; nt < MT; ++nt) G[mt][nt] = mma(av, bv[nt], G[mt][nt]);
        }
      }
    }
    // transpose Ut through this wave's own LDS rows (only this wave reads them)
#pragma unroll
    for (int t = 0; t < MT; ++t)
#pragma unroll
      for (int r = 0; r < 4; ++r) mine[acc_row<R>(g, r) * LS + 16 * t + nn] = u[t][r];
    vec4_t<R> nu[MT];
    if (internal) {
#pragma unroll
      for (int t = 0; t < MT; ++t) nu[t] = vec4_t<R>{R(0), R(0), R(0), R(0)};
      if constexpr (SPLIT) {
        // K index (pack q / KB, element q % KB, lane group g) = state g KV + q, as in the
        // forward product; the P operand comes as hi/lo pairs from the LDS tile (KV <= 8)
        // or the padded copies (KV = 16)
        BfPack<KB> ap[NB];
#pragma unroll
        for (int q = 0; q < NB * KB; ++q)
          pack_set(ap[q / KB], q % KB, q < KV ? mine[nn * LS + g * KV + q] : 0.f);
        PRing<kPReg ? 1 : kPD> ring;
        if constexpr (!kPReg) ring.start(pp4, 1);
#pragma unroll
        for (int t = 0; t < MT; ++t)
#pragma unroll
          for (int bb = 0; bb < NB; ++bb) {  // one pack of P live at a time
            BfPack<KB> bp;
            if constexpr (kPReg) {
#pragma unroll
              for (int e = 0; e < KB; ++e) {
                const int q = bb * KB + e;
                pack_set_pair(bp, e, q < KV ? __builtin_bit_cast(
                                                  uint32_t, ptw[(g * KV + q) * kPtS + 16 * t + nn])
                                            : 0u);
              }
            } else {
              // KB = 8: blocks 4 t + 2 bb and 4 t + 2 bb + 1 are the hi and lo operands
              static_assert(kPReg || KB == 8, "64 padded states: packs of 8");
              if constexpr (!kPReg) {
                bp.hi = __builtin_bit_cast(bf16x8_t, ring.take(pp4, 1, 4 * t + 2 * bb));
                bp.lo = __builtin_bit_cast(bf16x8_t, ring.take(pp4, 1, 4 * t + 2 * bb + 1));
              }
            }
            nu[t] = mma3(ap[bb], bp, nu[t]);
          }
      } else if constexpr (kPReg) {
#pragma unroll
        for (int ks = 0; ks < KV; ++ks) {
          const R av = mine[nn * LS + g * KV + ks];  // K index (ks, g) = state g KV + ks
#pragma unroll
          for (int t = 0; t < MT; ++t) nu[t] = mma(av, ptw[(g * KV + ks) * kPtS + 16 * t + nn], nu[t]);
        }
      } else {
        PRing<kPD> ring;
        ring.start(pp4, 1);
#pragma unroll
        for (int t = 0; t < MT; ++t)
#pragma unroll
          for (int q = 0; q < KV; q += 4) {
            const vec4_t<R> pv = __builtin_bit_cast(vec4_t<R>, ring.take(pp4, 1, 4 * t + q / 4));
#pragma unroll
            for (int e = 0; e < 4; ++e) nu[t] = mma(mine[nn * LS + g * KV + q + e], pv[e], nu[t]);
          }
      }
    }
    // no barrier before the stores: the exponent comes from the Ut maxima shared
    // at the previous barrier. U = P^T Ut is at most S times the largest Ut, so
    // U 2^-e stays below 2 S; a power-of-two scale is exact and cancels in every
    // term of G and of the likelihood ratios below this node
    if (internal) {
      const __amdgpu_buffer_rsrc_t crs = tile_rsrc(cslab, tile);
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int m = acc_row<R>(g, r);
        const vec4_t<R> mv = *reinterpret_cast<const vec4_t<R>*>(mxx + (h * kTile + m) * 4);
        const R M = fmax_(fmax_(mv[0], mv[1]), fmax_(mv[2], mv[3]));
        int e = 0;
        if (M > R(0) && isfinite(M)) frexp_(M, &e);
#pragma unroll
        for (int t = 0; t < MT; ++t)
          if (rok[r] && stok[t]) {
            if constexpr (sizeof(R) != 4) {
              cslab[tb + ub + r * rstep + 16 * t] = ldexp_(nu[t][r], -e);
            } else {
              __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(uint32_t, ldexp_(nu[t][r], -e)),
                                                    crs, (int)tvo[r], 64 * t, 0);
            }
          }
      }
    }
    // no barrier: the next tile's first barrier orders this tile's reads of
    // dens and mxx before their next writes; each wave rewrites only its own
    // rows of Yx (read by the sibling wave before the second barrier) and Lx
  }
  if (a.mats_dt != nullptr) {
    const R* q = a.mats_dt + (((size_t)k * a.B + chh) * C + c) * S * S;
    double sacc = 0.0;
#pragma unroll
    for (int mt = 0; mt < MT; ++mt)
#pragma unroll
      for (int nt = 0; nt < MT; ++nt)
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          const int i = 16 * mt + acc_row<R>(g, r), j = 16 * nt + nn;
          if (i < S && j < S) sacc = fma((double)G[mt][nt][r], (double)q[i * S + j], sacc);
        }
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) sacc += __shfl_xor(sacc, o);
    if (lane == 0) a.gtau[(((size_t)chh * a.K + k) * a.nchunks + chunk) * C + c] = sacc;
  } else {
    const size_t css = (size_t)C * S * S;
    R* o = a.gpart + (((size_t)chh * a.K + k) * a.nchunks + chunk) * css + (size_t)c * S * S;
#pragma unroll
    for (int mt = 0; mt < MT; ++mt)
#pragma unroll
      for (int nt = 0; nt < MT; ++nt)
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          const int i = 16 * mt + acc_row<R>(g, r), j = 16 * nt + nn;
          if (i < S && j < S) o[i * S + j] = G[mt][nt][r];
        }
  }
}

template <typename R>
static size_t mfma_lds(int C, int S, bool pre, int tiles_per_chunk) {
  const int MT = S <= 16 ? 1 : (S <= 32 ? 2 : 4);
  const size_t SP = 16 * MT, rb = sizeof(R);
  // k_pre_mfma_r (MT <= 2) pads the per-row category entries to 4 and stages
  // two children's codes for the chunk
  if (pre && (MT <= 2 || rb == 4)) {
    const int KV = MT == 2 && S <= 20 ? 5 : 4 * MT;
    const size_t pt = MT <= 2 ? 2 * (size_t)C * 4 * KV * (SP + 1) : 0;  // transpose P tiles
    // the [16][SP] tiles are stored with row stride SP + 4 at MT <= 2 (k_pre_mfma_r)
    return (4 * C * kTile * (MT <= 2 ? SP + 4 : SP) + 12 * kTile + pt) * rb + 64 +
           2 * (size_t)tiles_per_chunk * kTile;
  }
  if (pre) return (4 * C * kTile * SP + 3 * C * kTile) * rb + 64;
  return (C * kTile * SP + C * kTile) * rb + (size_t)(C * kTile + C * SP) * 8 + 3 * kTile * 4 + 64;
}

template <typename R, bool SPLIT>
static int launch_mfma_t(bool post, const PassArgs<R>& a, const int32_t* nodes, int width,
                         int tips, bool root, int S, hipStream_t st) {
  const int MT = S <= 16 ? 1 : (S <= 32 ? 2 : 4);
  const int threads = 2 * a.C * 64;
  const size_t lds = mfma_lds<R>(a.C, S, !post, a.tiles_per_chunk);
  HX_REQUIRE(lds <= 160 * 1024, HX_ERR_UNSUPPORTED, "dense-state kernel LDS exceeds 160 KiB");
  // k_pre_mfma_r addresses a (node, sample) slab of partials with 32-bit
  // element offsets, reading up to 16 C S + 64 elements past its end
  HX_REQUIRE((size_t)(a.N + 16) * a.C * S + 64 < ((size_t)1 << 31), HX_ERR_UNSUPPORTED,
             "dense-state slab exceeds 31-bit offsets");
  dim3 grid(a.nchunks, a.K, width);
#define HX_MF(KER, MTV, KVV, TV, RV)                                                         \
  do {                                                                                       \
    auto kern = KER<R, MTV, KVV, TV, RV, SPLIT>;                                             \
    if (lds > 64 * 1024)                                                                     \
      HX_HIP_TRY(hipFuncSetAttribute(reinterpret_cast<const void*>(kern),                    \
                                     hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds)); \
    kern<<<grid, threads, lds, st>>>(a, nodes, S);                                           \
  } while (0)
#define HX_MF_ALL(KER, MTV, KVV)                                                             \
  if (tips == HX_TIPS_CODES) {                                                               \
    if (root) HX_MF(KER, MTV, KVV, HX_TIPS_CODES, true);                                     \
    else HX_MF(KER, MTV, KVV, HX_TIPS_CODES, false);                                         \
  } else {                                                                                   \
    if (root) HX_MF(KER, MTV, KVV, HX_TIPS_DENSE, true);                                     \
    else HX_MF(KER, MTV, KVV, HX_TIPS_DENSE, false);                                         \
  }
// the root launch of the post-order (k_post_mfma_r serves every other level)
#define HX_MF_ROOT(MTV, KVV)                                                                 \
  if (tips == HX_TIPS_CODES) HX_MF(k_post_mfma, MTV, KVV, HX_TIPS_CODES, true);              \
  else HX_MF(k_post_mfma, MTV, KVV, HX_TIPS_DENSE, true);
  if (post && !root) {
    // scale tiles, per-row maxima, the chunk's codes
    const size_t lds_r = (size_t)a.C * kTile * (16 * MT + 4) * sizeof(R) + 4 * kTile * sizeof(R) +
                         4 * kTile * 4 + 2 * (size_t)a.tiles_per_chunk * kTile + 64;
    HX_REQUIRE(lds_r <= 64 * 1024, HX_ERR_UNSUPPORTED, "dense-state post-order LDS exceeds 64 KiB");
    // MT 4: C <= 4, so C x 16 x 64 products take at most 16 KiB (32 KiB in FP64)
#define HX_MFR(MTV, KVV)                                                                     \
  if (tips == HX_TIPS_CODES)                                                                 \
    k_post_mfma_r<R, MTV, KVV, HX_TIPS_CODES, SPLIT><<<grid, threads, lds_r, st>>>(a, nodes, \
                                                                                   S);       \
  else                                                                                       \
    k_post_mfma_r<R, MTV, KVV, HX_TIPS_DENSE, SPLIT><<<grid, threads, lds_r, st>>>(a, nodes, \
                                                                                   S);
    if (MT == 1) {
      HX_MFR(1, 4)
    } else if (MT == 4) {
      HX_MFR(4, 16)
    } else if (S <= 20) {
      HX_MFR(2, 5)
    } else {
      HX_MFR(2, 8)
    }
#undef HX_MFR
  } else if (post) {
    if (MT == 1) {
      HX_MF_ROOT(1, 4)
    } else if (MT == 2) {
      if (S <= 20) { HX_MF_ROOT(2, 5) } else { HX_MF_ROOT(2, 8) }
    } else {
      HX_MF_ROOT(4, 16)
    }
  } else if (MT == 1) {
    HX_MF_ALL(k_pre_mfma_r, 1, 4)
  } else if (MT == 2) {
    if (S <= 20) { HX_MF_ALL(k_pre_mfma_r, 2, 5) } else { HX_MF_ALL(k_pre_mfma_r, 2, 8) }
  } else if constexpr (sizeof(R) == 4) {
    HX_REQUIRE(a.ppad != nullptr, HX_ERR_INVALID_ARGUMENT, "padded P copies missing");
    HX_MF_ALL(k_pre_mfma_r, 4, 16)
  } else {
    HX_MF_ALL(k_pre_mfma, 4, 16)
  }
#undef HX_MF_ROOT
#undef HX_MF_ALL
#undef HX_MF
  HX_HIP_TRY(hipGetLastError());
  return HX_OK;
}

template <typename R>
int launch_mfma(bool post, const PassArgs<R>& a, const int32_t* nodes, int width, int tips,
                bool root, int S, hipStream_t st) {
  if constexpr (std::is_same<R, float>::value) {
    if (a.split) return launch_mfma_t<float, true>(post, a, nodes, width, tips, root, S, st);
  }
  return launch_mfma_t<R, false>(post, a, nodes, width, tips, root, S, st);
}

template int launch_mfma<float>(bool, const PassArgs<float>&, const int32_t*, int, int, bool, int,
                                hipStream_t);
template int launch_mfma<double>(bool, const PassArgs<double>&, const int32_t*, int, int, bool, int,
                                 hipStream_t);

// ------------------------------------------------------- layout self-test
// D = A B for one 16x16x4 MFMA with A[i][k] = 1 + i + 16 k and
// B[k][j] = 1 + 0.5 j - k, read back through the assumed lane layouts.
template <typename R>
__global__ void k_mfma_layout(R* out) {
  const int lane = threadIdx.x;
  const int i = lane & 15, kk = lane >> 4;
  const R av = R(1) + R(i) + R(16) * R(kk);        // A[i][k], lane = i + 16 k
  const R bv = R(1) + R(0.5) * R(i) - R(kk);       // B[k][j], lane = j + 16 k (j = lane & 15)
  vec4_t<R> d = mma(av, bv, vec4_t<R>{R(0), R(0), R(0), R(0)});
#pragma unroll
  for (int r = 0; r < 4; ++r) out[acc_row<R>(lane >> 4, r) * 16 + (lane & 15)] = d[r];
}

// The same check for the two BF16 instructions of the split products: A[i][k] = (i + k) % 7 - 3
// and B[k][j] = (2 k + j) % 5 - 2 (exact in BF16, exact sums in FP32) with k = N g + e in
// lane i + 16 g (A) or j + 16 g (B), element e; and one full split product (mma3) of
// non-representable values against its FP64 value.
template <int N>
__global__ void k_mfma_layout_bf16(float* out) {
  const int lane = threadIdx.x;
  const int i = lane & 15, g = lane >> 4;
  BfPack<N> a, b, a3, b3;
#pragma unroll
  for (int e = 0; e < N; ++e) {
    const int k = N * g + e;
    a.hi[e] = (__bf16)(float)((i + k) % 7 - 3);
    b.hi[e] = (__bf16)(float)((2 * k + i) % 5 - 2);
    pack_set(a3, e, 1.0f + 0.01f * (float)i + 0.003f * (float)k);
    pack_set(b3, e, 0.7f - 0.002f * (float)k + 0.013f * (float)i);
  }
  vec4_t<float> d = {0.f, 0.f, 0.f, 0.f}, d3 = {0.f, 0.f, 0.f, 0.f};
  if constexpr (N == 8) d = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a.hi, b.hi, d, 0, 0, 0);
  else d = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a.hi, b.hi, d, 0, 0, 0);
  d3 = mma3(a3, b3, d3);
#pragma unroll
  for (int r = 0; r < 4; ++r) {
    out[acc_row<float>(g, r) * 16 + i] = d[r];
    out[256 + acc_row<float>(g, r) * 16 + i] = d3[r];
  }
}

// The 16-block 4x4x1 FP32 product of the four-state pre-order walk's branch statistic:
// A[b][i] in lane 4b + i, B[b][j] in lane 4b + j, D[b][i][j] in register i of lane 4b + j.
// A[b][i] = 1 + i + 4b and B[b][j] = 1 + 0.5 j - 0.25 b (products exact in FP32).
__global__ void k_mfma_layout_4x4(float* out) {
  const int lane = threadIdx.x;
  const int b = lane >> 2, x = lane & 3;
  const float av = 1.f + (float)x + 4.f * (float)b;
  const float bv = 1.f + 0.5f * (float)x - 0.25f * (float)b;
  const vec4_t<float> d = __builtin_amdgcn_mfma_f32_4x4x1f32(av, bv, vec4_t<float>{0.f, 0.f, 0.f, 0.f}, 0, 0, 0);
#pragma unroll
  for (int r = 0; r < 4; ++r) out[(b * 4 + r) * 4 + x] = d[r];
}

static int mfma_layout_check_4x4(hipStream_t st, int* bad) {
  float* d = nullptr;
  HX_HIP_TRY(hipMalloc(&d, 256 * sizeof(float)));
  k_mfma_layout_4x4<<<1, 64, 0, st>>>(d);
  float host[256];
  hipError_t e = hipMemcpyAsync(host, d, sizeof(host), hipMemcpyDeviceToHost, st);
  if (e == hipSuccess) e = hipStreamSynchronize(st);
  hipFree(d);
  if (e != hipSuccess) {
    set_error(std::string("mfma 4x4 self-test: ") + hipGetErrorString(e));
    return HX_ERR_HIP;
  }
  for (int b = 0; b < 16; ++b)
    for (int i = 0; i < 4; ++i)
      for (int j = 0; j < 4; ++j) {
        const double ref = (1.0 + i + 4.0 * b) * (1.0 + 0.5 * j - 0.25 * b);
        if ((double)host[(b * 4 + i) * 4 + j] != ref) ++*bad;
      }
  return HX_OK;
}

template <int N>
static int mfma_layout_check_bf16(hipStream_t st, int* bad) {
  float* d = nullptr;
  HX_HIP_TRY(hipMalloc(&d, 512 * sizeof(float)));
  k_mfma_layout_bf16<N><<<1, 64, 0, st>>>(d);
  float host[512];
  hipError_t e = hipMemcpyAsync(host, d, sizeof(host), hipMemcpyDeviceToHost, st);
  if (e == hipSuccess) e = hipStreamSynchronize(st);
  hipFree(d);
  if (e != hipSuccess) {
    set_error(std::string("mfma bf16 self-test: ") + hipGetErrorString(e));
    return HX_ERR_HIP;
  }
  for (int i = 0; i < 16; ++i)
    for (int j = 0; j < 16; ++j) {
      double ref = 0.0, ref3 = 0.0;
      for (int k = 0; k < 4 * N; ++k) {
        ref += (double)((i + k) % 7 - 3) * (double)((2 * k + j) % 5 - 2);
        ref3 += (1.0 + 0.01 * i + 0.003 * k) * (0.7 - 0.002 * k + 0.013 * j);
      }
      if ((double)host[i * 16 + j] != ref) ++*bad;
      // a single BF16 product would be off by ~4e-3; the split form by ~1e-5
      if (fabs((double)host[256 + i * 16 + j] - ref3) > 1e-4 * fabs(ref3)) ++*bad;
    }
  return HX_OK;
}

template <typename R>
static int mfma_layout_check(hipStream_t st, int* bad) {
  R* d = nullptr;
  HX_HIP_TRY(hipMalloc(&d, 256 * sizeof(R)));
  k_mfma_layout<R><<<1, 64, 0, st>>>(d);
  R host[256];
  hipError_t e = hipMemcpyAsync(host, d, sizeof(host), hipMemcpyDeviceToHost, st);
  if (e == hipSuccess) e = hipStreamSynchronize(st);
  hipFree(d);
  if (e != hipSuccess) {
    set_error(std::string("mfma self-test: ") + hipGetErrorString(e));
    return HX_ERR_HIP;
  }
  for (int i = 0; i < 16; ++i)
    for (int j = 0; j < 16; ++j) {
      double ref = 0.0;
      for (int kk = 0; kk < 4; ++kk) ref += (1.0 + i + 16.0 * kk) * (1.0 + 0.5 * j - kk);
      if (fabs((double)host[i * 16 + j] - ref) > 1e-3 * (1.0 + fabs(ref))) ++*bad;
    }
  return HX_OK;
}

}  // namespace hx

extern "C" int hx_selftest_mfma_layout(int* mismatches, hx_stream_t stream) {
  HX_REQUIRE(mismatches != nullptr, HX_ERR_INVALID_ARGUMENT, "null argument");
  hipStream_t st = reinterpret_cast<hipStream_t>(stream);
  int bad = 0;
  int rc = hx::mfma_layout_check<float>(st, &bad);
  if (rc == HX_OK) rc = hx::mfma_layout_check<double>(st, &bad);
  if (rc == HX_OK) rc = hx::mfma_layout_check_bf16<8>(st, &bad);
  if (rc == HX_OK) rc = hx::mfma_layout_check_bf16<4>(st, &bad);
  if (rc == HX_OK) rc = hx::mfma_layout_check_4x4(st, &bad);
  *mismatches = bad;
  return rc;
}

// Four-state (nucleotide) pruning kernels: one fused post-order walk of the
// whole tree per pattern tile, and the pre-order adjoint either as one walk
// of the whole tree or one tree level per launch.
//
// Reference: calculate_treelikelihood_discrete* (torchtree/evolution/
// tree_likelihood.py:40-278) at S = 4, where each node update is two 4x4
// matrix-vector products per (sample, pattern, category): HBM-bound work.
//
// Partials use the category-major layout
//   partials [slot][K][C][np][4]  slot = node - T, np = 64 x tiles
// so a wave's accesses to one category are contiguous 16-byte vectors, and
// the padding lanes of the last tile have slots of their own.
//
// Post-order (k_post_walk): lane = pattern with all C categories in the lane,
// one wave per (tile, sample), walking the tree program; each internal
// partial is written once and read back only by the lane that wrote it, when
// a child waited on the stack at a depth not kept in registers. Pre-order:
// k_pre_walk (FP32, C <= 4, enough work to fill the chip) walks the tree
// depth first with the adjoints in registers; otherwise k_pre runs one tree
// level per launch with thread = (pattern, category).
#include "pruning.hpp"
#include "walk.hpp"

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

// Non-temporal load (streaming hint: the line is a first candidate for
// eviction), for data read once so that lines re-read soon stay in L2.
template <typename R>
__device__ __forceinline__ V4<R> ld4_nt(const R* p) {
  V4<R> v;
  if constexpr (sizeof(R) == 4) {
    using f4 = float __attribute__((ext_vector_type(4)));
    const f4 q = __builtin_nontemporal_load(reinterpret_cast<const f4*>(p));
    v.x[0] = q[0], v.x[1] = q[1], v.x[2] = q[2], v.x[3] = q[3];
  } else {
    using d2 = double __attribute__((ext_vector_type(2)));
    const d2 q0 = __builtin_nontemporal_load(reinterpret_cast<const d2*>(p));
    const d2 q1 = __builtin_nontemporal_load(reinterpret_cast<const d2*>(p) + 1);
    v.x[0] = q0[0], v.x[1] = q0[1], v.x[2] = q1[0], v.x[3] = q1[1];
  }
  return v;
}

template <typename R>
__device__ __forceinline__ V4<R> zero4() {
  V4<R> v;
#pragma unroll
  for (int s = 0; s < 4; ++s) v.x[s] = R(0);
  return v;
}

__device__ __forceinline__ int uni(int x) { return __builtin_amdgcn_readfirstlane(x); }

// Read-only data read with wave-uniform addresses (the walk program, P
// matrices) goes through the constant address space so the compiler always
// uses scalar loads; those are counted apart from vector memory, so waiting
// for them never waits for the partials' stores.
template <typename T>
using cptr = const __attribute__((address_space(4))) T*;
template <typename T>
__device__ __forceinline__ cptr<T> to_const(const T* p) {
  return (cptr<T>)p;
}

__device__ __forceinline__ float fmax_(float a, float b) { return fmaxf(a, b); }
__device__ __forceinline__ double fmax_(double a, double b) { return fmax(a, b); }
__device__ __forceinline__ float frexp_(float m, int* e) { return frexpf(m, e); }
__device__ __forceinline__ double frexp_(double m, int* e) { return frexp(m, e); }
__device__ __forceinline__ float ldexp_(float x, int e) { return ldexpf(x, e); }
__device__ __forceinline__ double ldexp_(double x, int e) { return ldexp(x, e); }

// The 4x4 block at p (16 consecutive values, wave-uniform address).
template <typename R>
__device__ __forceinline__ void load_p16(cptr<R> p, R (&P)[16]) {
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

template <typename R, int TIPS>
__device__ __forceinline__ void stage_tbl(const PassArgs<R>& a, V4<R>* tbl, int nt) {
  if constexpr (TIPS == HX_TIPS_CODES) {
    for (int i = threadIdx.x; i < a.ncodes; i += nt) tbl[i] = ld4(a.table + 4 * i);
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
// A block of W waves walks the whole tree for one 64-pattern tile; wave w
// handles sample blockIdx.y * W + w and lane l pattern 64 tile + l with all C
// categories of it, so the rescaling maximum and the root likelihood never
// leave the lane and the waves never synchronise after staging. The tile's
// tip codes (T x 64 bytes) are staged in LDS once per block; P of each
// (branch, category) is wave-uniform (scalar loads, SGPR operands).
//
// prog: kWalkInts ints per internal node in walk order (walk.hpp). A child is
// a tip, the previous entry's result (registers), or an earlier result that
// waited on the stack: its partial is read back from the copy this lane
// wrote to HBM (in L2 in practice), issued one node ahead and before that
// node's stores so the wait for it does not wait for them, and its exponent
// comes from the lane's LDS stack. grid = (ntiles, ceil(K / W)), block = 64 W.
// CM: categories rounded up to a power of two (EXACT: C == CM).
// No stack depth stays in registers: keeping depths 2..4 there (the program's
// register mode 1, walk.hpp) takes the four-category FP32 walk from 69 to 114
// VGPRs, 4 waves per SIMD instead of 6 (the LDS limit), and the walk is
// latency-bound, so the extra waves gain more than the read-backs cost
// (measured at 512 taxa, 100,000 patterns, K = 32: 21.2 -> 20.0 ms).
// a.drop_mode >= 0: the cherries whose drop bit (walk.hpp) is set are not
// stored; their store instruction is still issued, as a raw buffer store
// with an empty range (the hardware discards it), so every entry issues the
// same vector memory operations and the waits stay exact.
__device__ __forceinline__ void store_row_f32(const __amdgpu_buffer_rsrc_t& r, unsigned off,
                                              const V4<float>& v) {
  typedef uint32_t u4 __attribute__((ext_vector_type(4)));
  const u4 w = {__float_as_uint(v.x[0]), __float_as_uint(v.x[1]), __float_as_uint(v.x[2]),
                __float_as_uint(v.x[3])};
  __builtin_amdgcn_raw_buffer_store_b128(w, r, (int)off, 0, 0);
}

template <typename R, int TIPS, int CM, bool EXACT>
__global__ __launch_bounds__(CM >= 8 ? 256 : 1024) void k_post_walk(PassArgs<R> a, const R* __restrict__ mats,
                                                    const int4* __restrict__ prog, int nprog,
                                                    int ne, int W) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  const int C = EXACT ? CM : a.C;
  const int wave = uni(threadIdx.x >> 6), lane = threadIdx.x & 63;
  const int tile = blockIdx.x, k = blockIdx.y * W + wave;
  const int n = tile * 64 + lane;
  const bool val = n < a.N;
  uint8_t* lc = smem_raw;                                              // [T][64]
  const size_t lc_bytes = TIPS == HX_TIPS_CODES ? (size_t)a.T * 64 : 0;
  V4<R>* tbl = reinterpret_cast<V4<R>*>(smem_raw + lc_bytes);          // [ncodes]
  int* est = reinterpret_cast<int*>(tbl + (TIPS == HX_TIPS_CODES ? a.ncodes : 0)) +
             (size_t)wave * ne * 64;                                   // [W][ne][64]
  const int nt = blockDim.x;
  if constexpr (TIPS == HX_TIPS_CODES) {
    const size_t stride = (size_t)a.ntiles * 64;
    for (int i = threadIdx.x; i < a.T * 4; i += nt) {
      const int row = i >> 2, q = i & 3;
      *reinterpret_cast<uint4*>(lc + row * 64 + q * 16) =
          *reinterpret_cast<const uint4*>(a.codes_pad + row * stride + tile * 64 + q * 16);
    }
  }
  stage_tbl<R, TIPS>(a, tbl, nt);
  __syncthreads();
  if (k >= a.K) return;

  // element offsets: node slot s, category c of this lane's pattern at
  // base + s * sstride + c * cstride (layout [slot][K][C][np][4], np = 64 ntiles,
  // so the padding lanes of the last tile have slots of their own)
  const size_t cstride = (size_t)a.np * 4;
  const size_t sstride = (size_t)a.K * C * cstride;
  R* base = a.partials + (size_t)k * C * cstride + (size_t)n * 4;
  const cptr<R> mk = to_const(mats) + (size_t)k * a.B * C * 16;
  const cptr<int> cprog = to_const(reinterpret_cast<const int*>(prog));
  auto entry = [&](int i, int4& f0, int4& f1) {
    const cptr<int> q = cprog + (size_t)kWalkInts * i;
    f0.x = q[0], f0.y = q[1], f0.z = q[2], f0.w = q[3];
    f1.x = q[4], f1.y = q[5], f1.z = q[6], f1.w = q[7];
  };

  auto tip = [&](int ch) -> V4<R> {
    if constexpr (TIPS == HX_TIPS_CODES) return tbl[lc[ch * 64 + lane]];
    else return val ? ld4(a.dense + ((size_t)ch * a.N + n) * 4) : zero4<R>();
  };
  // Issued every node, for the stacked child of the next entry or, when it
  // has none, from this sample's P blocks (a few hot cache lines, value
  // unused): every iteration then issues the same vector memory operations,
  // so the wait for these loads counts only the older ones and not this
  // node's stores, and the placeholder loads never reach HBM.
  auto stacked_of = [&](const int4& f0, const int4& f1) {
    if ((f1.z & 0xff) == 2) return f0.y;
    if (((f1.z >> 8) & 0xff) == 2) return f0.z;
    return -1;
  };
  const R* hot = mats + (size_t)k * a.B * C * 16 + (lane & 3) * 4;
  auto readback = [&](int ch, V4<R>(&x)[CM]) {
    const bool real = ch >= 0;
    const R* p = base + (size_t)(real ? ch - a.T : 0) * sstride;
#pragma unroll
    for (int c = 0; c < CM; ++c)
      if (EXACT || c < C) x[c] = ld4(real ? p + (size_t)c * cstride : hot);
  };

  // The first entry of a walk has two tip children, so xs is first read
  // after a loop iteration has issued its read-back and stores.
  V4<R> y[CM], xs[CM];
  int E = 0;
  int4 e0, e1;
  entry(0, e0, e1);
  // entries 0 .. nprog-2 store their partial; the last entry is the root
  for (int i = 0; i < nprog; ++i) {
    const bool last = i + 1 == nprog;
    const int v = e0.x, l = e0.y, r = e0.z;
    const int sl = (e0.w & 0xff) - 1, sr = ((e0.w >> 8) & 0xff) - 1, so = e0.w >> 16;
    const int kl = e1.z & 0xff, kr = (e1.z >> 8) & 0xff;
    int4 f0 = e0, f1 = e1;
    if (i + 1 < nprog) entry(i + 1, f0, f1);
    V4<R> xl[CM], xr[CM];
    int El = 0, Er = 0;
    if (kl == 0) {
      const V4<R> t = tip(l);
#pragma unroll
      for (int c = 0; c < CM; ++c) xl[c] = t;
    } else if (kl == 1) {
#pragma unroll
      for (int c = 0; c < CM; ++c) xl[c] = y[c];
      El = E;
    } else {
#pragma unroll
      for (int c = 0; c < CM; ++c) xl[c] = xs[c];
      El = est[sl * 64 + lane];
    }
    if (kr == 0) {
      const V4<R> t = tip(r);
#pragma unroll
      for (int c = 0; c < CM; ++c) xr[c] = t;
    } else if (kr == 1) {
#pragma unroll
      for (int c = 0; c < CM; ++c) xr[c] = y[c];
      Er = E;
    } else {
#pragma unroll
      for (int c = 0; c < CM; ++c) xr[c] = xs[c];
      Er = est[sr * 64 + lane];
    }
    // P of (l, c) and (r, c); the loads of category c + 1 are issued before
    // category c is computed
    R M = R(0);
    R Pl[2][16], Pr[2][16];
    load_p16(mk + (size_t)l * C * 16, Pl[0]);
    load_p16(mk + (size_t)r * C * 16, Pr[0]);
#pragma unroll
    for (int c = 0; c < CM; ++c) {
      if (EXACT || c < C) {
        if (c + 1 < CM && (EXACT || c + 1 < C)) {
          load_p16(mk + ((size_t)l * C + c + 1) * 16, Pl[(c + 1) & 1]);
          load_p16(mk + ((size_t)r * C + c + 1) * 16, Pr[(c + 1) & 1]);
        }
        const V4<R> yl = mv(Pl[c & 1], xl[c]);
        const V4<R> yr = mv(Pr[c & 1], xr[c]);
#pragma unroll
        for (int s = 0; s < 4; ++s) y[c].x[s] = yl.x[s] * yr.x[s];
        M = fmax_(M, max4(y[c]));
      } else {
        y[c] = zero4<R>();
      }
    }
    const int e = scale_exp(M);
#pragma unroll
    for (int c = 0; c < CM; ++c) scale4(y[c], e);
    E = El + Er + e;
    if (last) break;
    // the next entry's stacked child, before this node's stores
    readback(stacked_of(f0, f1), xs);
    if constexpr (sizeof(R) == 4) {
      // this node's C rows of this sample as one raw buffer (empty for a dropped cherry)
      const bool drop = a.drop_mode >= 0 && ((e1.w >> a.drop_mode) & 1);
      const R* node = a.partials + (size_t)(v - a.T) * sstride + (size_t)k * C * cstride;
      const __amdgpu_buffer_rsrc_t rs = __builtin_amdgcn_make_buffer_rsrc(
          const_cast<R*>(node), (short)0, drop ? 0 : (int)((size_t)C * cstride * sizeof(R)),
          0x00020000);
#pragma unroll
      for (int c = 0; c < CM; ++c)
        if (EXACT || c < C) store_row_f32(rs, (unsigned)((c * cstride + (size_t)n * 4) * sizeof(R)), y[c]);
    } else {
      R* p = base + (size_t)(v - a.T) * sstride;
#pragma unroll
      for (int c = 0; c < CM; ++c)
        if (EXACT || c < C) st4(p + (size_t)c * cstride, y[c]);
    }
    est[so * 64 + lane] = E;
    e0 = f0;
    e1 = f1;
  }

  // root: site likelihood sum_c p_c pi.x_c (tree_likelihood.py:71-74, 215-220)
  R pi[4];
#pragma unroll
  for (int s = 0; s < 4; ++s) pi[s] = a.freqs[(size_t)k * a.freqs_ks + s];
  R px[CM];
  R lik = R(0);
#pragma unroll
  for (int c = 0; c < CM; ++c) {
    px[c] = R(0);
    if (c < C) {
#pragma unroll
      for (int s = 0; s < 4; ++s) px[c] = fma(pi[s], y[c].x[s], px[c]);
      lik += a.props[(size_t)k * a.props_ks + c] * px[c];
    }
  }
  double acc[5 + CM];  // loglik, dfreqs[4], dprops[c]
#pragma unroll
  for (int q = 0; q < 5 + CM; ++q) acc[q] = 0.0;
  if (val) {
    const double w = a.weights[n];
    const double dl_ = (double)lik;
    const double wi = w / dl_;
    acc[0] = w * (log(dl_) + (double)E * kLn2);
#pragma unroll
    for (int c = 0; c < CM; ++c) {
      if (c < C) {
        const R pc = a.props[(size_t)k * a.props_ks + c];
#pragma unroll
        for (int s = 0; s < 4; ++s) acc[1 + s] += wi * (double)(pc * y[c].x[s]);
        acc[5 + c] = wi * (double)px[c];
      }
    }
  }
#pragma unroll
  for (int q = 0; q < 5 + CM; ++q)
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) acc[q] += __shfl_xor(acc[q], o);
  double* out = a.rootsums + ((size_t)k * a.ntiles + tile) * (5 + C);
  if (lane == 0) {
#pragma unroll
    for (int q = 0; q < 5 + CM; ++q)
      if (q < 5 + C) out[q] = acc[q];
  }
}

// ------------------------------------------------------------------- pre-order
// For parent p with children a, b and pre-order partial U_p (root: freqs):
//   Ut_a = U_p * (P_b L_b), Ut_b = U_p * (P_a L_a)
//   lik  = sum_c p_c Ut_a.(P_a L_a)                    (scaled site likelihood)
//   d loglik / d P_a[c][i][j] += w p_c Ut_a[i] L_a[j] / lik
//   U_a  = P_a^T Ut_a, rescaled (scale-free: every term is a ratio), written
//          over L_a in place
// Thread (pattern, category) with the CP = next_pow2(C) categories of a
// pattern in adjacent lanes: the likelihood sum and the rescaling maximum are
// shuffles within those lanes, so the tile loop has no barriers. The two P
// blocks are staged in LDS once per block; each branch statistic accumulates
// in registers over the block's chunk of 64-pattern tiles.
template <typename R>
struct PreTile {
  V4<R> U;
  V4<R> L[2];
  int code[2];
};

template <int CP, typename T>
__device__ __forceinline__ T qsum(T v) {
#pragma unroll
  for (int o = 1; o < CP; o <<= 1) v += __shfl_xor(v, o);
  return v;
}

template <int CP, typename T>
__device__ __forceinline__ T qmax(T v) {
#pragma unroll
  for (int o = 1; o < CP; o <<= 1) v = fmax_(v, __shfl_xor(v, o));
  return v;
}

// Entries of a 16-entry statistic transposed per pass so that the buffer
// [64][CP][NE + 1] stays within 36 KiB.
constexpr int pre_pass_entries(int cp, int rb) {
  int ne = 16;
  while (ne > 1 && (size_t)64 * cp * (ne + 1) * rb > 36 * 1024) ne >>= 1;
  return ne;
}

template <typename R, int TIPS, bool ROOT, int CP>
__global__ __launch_bounds__(64 * CP) void k_pre(PassArgs<R> a, const R* __restrict__ mats,
                                                 const R* __restrict__ mats_dt,
                                                 const int32_t* __restrict__ nodes) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  const int C = a.C, nt = 64 * CP;
  const int t = threadIdx.x, c = t % CP, pl = t / CP;
  const bool cval = c < C;
  const int cc = cval ? c : 0;
  const int chunk = blockIdx.x, k = blockIdx.y;
  const int p = nodes[blockIdx.z];
  int ch[2];
  bool in_c[2];
#pragma unroll
  for (int h = 0; h < 2; ++h) {
    ch[h] = a.kids[2 * p + h];
    in_c[h] = ch[h] >= a.T;
  }
  R* Ps = reinterpret_cast<R*>(smem_raw);                         // [2][C][16]
  V4<R>* tbl = reinterpret_cast<V4<R>*>(Ps + 2 * C * 16);         // [ncodes]
  // gb: [64 patterns][CP][NE + 1] statistic transpose for the chunk
  // reduction, NE of the 16 entries per pass
  constexpr int NE = pre_pass_entries(CP, sizeof(R));
  R* gb = reinterpret_cast<R*>(tbl + (TIPS == HX_TIPS_CODES ? a.ncodes : 0));
  for (int i = t; i < 2 * C * 16; i += nt) {
    const int h = i / (C * 16), rem = i - h * C * 16;
    Ps[i] = mats[((size_t)k * a.B + ch[h]) * C * 16 + rem];
  }
  stage_tbl<R, TIPS>(a, tbl, nt);

  const R pc = cval ? a.props[(size_t)k * a.props_ks + c] : R(0);
  V4<R> pi;
  if constexpr (ROOT) {
#pragma unroll
    for (int s = 0; s < 4; ++s) pi.x[s] = a.freqs[(size_t)k * a.freqs_ks + s];
  }
  R g[2][16];
#pragma unroll
  for (int i = 0; i < 16; ++i) g[0][i] = g[1][i] = R(0);

  auto fetch = [&](int tile, PreTile<R>& in) {
    const int n = tile * 64 + pl;
    in.U = zero4<R>();
    if constexpr (!ROOT) {
      if (cval) in.U = ld4(a.partials + (prow(a.T, a.K, C, a.np, p, k, cc) + n) * 4);
    }
#pragma unroll
    for (int h = 0; h < 2; ++h) {
      in.L[h] = zero4<R>();
      in.code[h] = 0;
      if (in_c[h]) {
        if (cval) in.L[h] = ld4(a.partials + (prow(a.T, a.K, C, a.np, ch[h], k, cc) + n) * 4);
      } else if constexpr (TIPS == HX_TIPS_CODES) {
        if (n < a.N) in.code[h] = a.codes[(size_t)ch[h] * a.N + n];
      } else {
        if (n < a.N) in.L[h] = ld4(a.dense + ((size_t)ch[h] * a.N + n) * 4);
      }
    }
  };
  auto getP = [&](int h, R(&P)[16]) {
    const R* src = Ps + (h * C + cc) * 16;
#pragma unroll
    for (int i = 0; i < 16; i += 4) {
      const V4<R> v = ld4(src + i);
      P[i] = v.x[0], P[i + 1] = v.x[1], P[i + 2] = v.x[2], P[i + 3] = v.x[3];
    }
  };

  // No software prefetch of the next tile: its registers cost a wave per
  // SIMD of occupancy, which hides the load latency better (measured).
  const int t0 = chunk * a.tiles_per_chunk;
  const int t1 = min(t0 + a.tiles_per_chunk, a.ntiles);
  __syncthreads();  // P blocks and code table staged
  for (int tile = t0; tile < t1; ++tile) {
    PreTile<R> cur;
    fetch(tile, cur);
    const int n = tile * 64 + pl;
    const bool val = n < a.N && cval;
    const R w = val ? (R)a.weights[n] : R(0);
    V4<R> U = cur.U;
    if constexpr (ROOT) U = cval ? pi : zero4<R>();
    V4<R> PL[2], Ut[2];
    R P[16];
#pragma unroll
    for (int h = 0; h < 2; ++h) {
      if constexpr (TIPS == HX_TIPS_CODES) {
        if (!in_c[h]) cur.L[h] = tbl[cur.code[h]];
      }
      getP(h, P);
      PL[h] = mv(P, cur.L[h]);
    }
    R d = R(0);
#pragma unroll
    for (int s = 0; s < 4; ++s) {
      Ut[0].x[s] = U.x[s] * PL[1].x[s];
      Ut[1].x[s] = U.x[s] * PL[0].x[s];
      d = fma(Ut[0].x[s], PL[0].x[s], d);
    }
    d = qsum<CP>(d * pc);
    const R coef = val ? w / d * pc : R(0);
    acc_outer(g[0], coef, Ut[0], cur.L[0]);
    acc_outer(g[1], coef, Ut[1], cur.L[1]);
#pragma unroll
    for (int h = 0; h < 2; ++h) {
      if (in_c[h]) {
        getP(h, P);
        V4<R> Uc = mv_t(P, Ut[h]);
        scale4(Uc, scale_exp(qmax<CP>(max4(Uc))));
        if (cval) st4(a.partials + (prow(a.T, a.K, C, a.np, ch[h], k, cc) + n) * 4, Uc);
      }
    }
  }

  // Chunk reduction of each branch statistic over the tile's 64 patterns:
  // transpose through LDS, thread (ci = (c, entry), part) sums 16 patterns,
  // then the 4 parts combine by shuffles (in double).
  auto flush = [&](const R(&gv)[16], int b) {
    const int part = t & 3, ci = t >> 2, c2 = ci >> 4, e = ci & 15;
    double sum = 0.0;
#pragma unroll
    for (int e0 = 0; e0 < 16; e0 += NE) {
      __syncthreads();
#pragma unroll
      for (int i = 0; i < NE; ++i) gb[(pl * CP + c) * (NE + 1) + i] = gv[e0 + i];
      __syncthreads();
      if (e >= e0 && e < e0 + NE) {
#pragma unroll
        for (int j = 0; j < 16; ++j)
          sum += (double)gb[((part * 16 + j) * CP + c2) * (NE + 1) + (e - e0)];
      }
    }
    sum += __shfl_xor(sum, 1);
    sum += __shfl_xor(sum, 2);
    if (mats_dt != nullptr) {
      // lanes (c2, e, part) of one category c2 form one wave: contract over e
      double tv = (part == 0 && c2 < C)
                      ? sum * (double)mats_dt[(((size_t)k * a.B + b) * C + c2) * 16 + e] : 0.0;
#pragma unroll
      for (int o = 4; o < 64; o <<= 1) tv += __shfl_xor(tv, o);
      if ((t & 63) == 0 && c2 < C) a.gtau[(((size_t)b * a.K + k) * a.nchunks + chunk) * C + c2] = tv;
    } else if (part == 0 && c2 < C) {
      a.gpart[((((size_t)b * a.K + k) * a.nchunks + chunk) * C + c2) * 16 + e] = (R)sum;
    }
  };
#pragma unroll
  for (int h = 0; h < 2; ++h) flush(g[h], ch[h]);
}

// -------------------------------------------------------------- pre-order walk
// Lane = pattern with all C categories in the lane, as in the post-order walk.
// A block's W waves walk the same program (walk.hpp, build_pre_walk) for W
// consecutive tiles of one sample, grid = (ceil(ntiles / W), K), so every
// wave is at the same entry and the branch statistics of the W tiles are
// summed in LDS once per entry. Entry (p; a, b) with p's adjoint U (the
// frequencies at the root):
//   m_h  = P_h L_h,  Ut_a = U * m_b,  Ut_b = U * m_a        (per category)
//   lik  = sum_c p_c sum_s U m_a m_b                         (scaled site likelihood)
//   G_h[c][i][j] = sum over patterns of (w p_c / lik) Ut_h[c][i] L_h[c][j]
//   U_h  = P_h^T Ut_h with one power-of-two rescale over all categories
// (tree_likelihood.py:40-278 through autograd). Every term of G is a ratio,
// so the scales of U and L cancel. The pattern sum of G runs on the matrix
// cores: each wave writes Ut (rows 4c + i) and (w / lik) L (rows 4c + j) to
// LDS, pattern minor, and reads them back as the operands of 16
// v_mfma_f32_4x4x1_16b_f32, whose block 4c + q sums category c's 4 x 4 outer
// products over patterns 16q .. 16q + 15; a row rotation then adds the four
// blocks q of each category. Adjoints waiting on the walk's stack live in registers for
// depths below RS and in the pushed child's own partials slot (in place, its
// post-order partial has been read) deeper down.
constexpr int kXS = 68;  // row stride (floats) of the LDS pattern transposes

// v + (v from the lane `ROT` below in the same row of 16 lanes, cyclically): DPP row_ror
template <int ROT>
__device__ __forceinline__ float add_row_ror(float v) {
  static_assert(ROT > 0 && ROT < 16, "row rotation");
  return v + __builtin_bit_cast(float, __builtin_amdgcn_mov_dpp(__builtin_bit_cast(int, v),
                                                                 0x120 + ROT, 0xF, 0xF, true));
}

template <int CM>
struct PreIn {
  V4<float> L[2][CM];
  V4<float> Up[CM];
  V4<float> tipv[2];  // dense tips: the tip's partial
  int code[2];
  float dt;  // dP/dtau entry this thread contracts with (d loglik / d tau mode)
};

template <int TIPS, int CM, bool EXACT, int RS, bool TAU>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(
    !EXACT ? 1 : (RS == 0 ? 4 : 3)))) void k_pre_walk(PassArgs<float> a, const float* __restrict__ mats,
                                                  const float* __restrict__ mats_dt,
                                                  const int* __restrict__ prog, int nprog) {
  using R = float;
  using f4 = float __attribute__((ext_vector_type(4)));
  constexpr int W = kPreWalkW;  // waves (tiles) per workgroup
  using H0 = std::integral_constant<int, 0>;
  using H1 = std::integral_constant<int, 1>;
  extern __shared__ __align__(16) unsigned char smem_raw[];
  const int C = EXACT ? CM : a.C;
  const int wave = uni(threadIdx.x >> 6), lane = threadIdx.x & 63;
  const int chunk = blockIdx.x, k = blockIdx.y;
  const int tile = chunk * W + wave;
  // waves past the last tile repeat it with zero weight and store nothing
  const bool tval = tile < a.ntiles;
  const int tl = tval ? tile : a.ntiles - 1;
  const int n = tl * 64 + lane;
  const bool val = tval && n < a.N;
  float* xt = reinterpret_cast<float*>(smem_raw) + (size_t)wave * 2 * 16 * kXS;  // [2][16][kXS]
  float* red = reinterpret_cast<float*>(smem_raw) + (size_t)W * 2 * 16 * kXS;    // [2][2][W][64]
  V4<R>* tbl = reinterpret_cast<V4<R>*>(red + 4 * W * 64);
  for (int i = lane; i < 2 * 16 * kXS; i += 64) xt[i] = 0.f;  // rows of absent categories
  stage_tbl<R, TIPS>(a, tbl, blockDim.x);
  __syncthreads();

  const R w = val ? (R)a.weights[n] : R(0);
  R pc[CM];
#pragma unroll
  for (int c = 0; c < CM; ++c) pc[c] = (EXACT || c < C) ? a.props[(size_t)k * a.props_ks + c] : R(0);
  V4<R> pi;
#pragma unroll
  for (int s = 0; s < 4; ++s) pi.x[s] = a.freqs[(size_t)k * a.freqs_ks + s];

  // Vector loads and stores address a wave-uniform base plus 16 bytes per
  // lane: the (node, category) row of this tile, or for placeholders one
  // cache line of this sample's P blocks that every wave keeps reading.
  const size_t cstride = (size_t)a.np * 4;
  const R* hot = mats + (size_t)k * a.B * C * 16;
  // lane byte offsets, unsigned so the loads use a scalar base plus a 32-bit
  // vector offset
  const unsigned off_real = lane * 16u, off_hot = (lane & 3) * 16u;
  auto at = [](const R* base, unsigned off) {
    return reinterpret_cast<const R*>(reinterpret_cast<const char*>(base) + off);
  };
  const cptr<R> mk = to_const(mats) + (size_t)k * a.B * C * 16;
  const cptr<int> cprog = to_const(prog);
  auto entry = [&](int i, int4& f0, int4& f1, int4& f2, int4& f3) {
    const cptr<int> q = cprog + (size_t)kPreInts * i;
    f0.x = q[0], f0.y = q[1], f0.z = q[2], f0.w = q[3];
    f1.x = q[4], f1.y = q[5], f1.z = q[6], f1.w = q[7];
    f2.x = q[8], f2.y = q[9], f2.z = q[10], f2.w = q[11];
    f3.x = q[12], f3.y = q[13], f3.z = q[14], f3.w = q[15];
  };
  // child h of an entry is a cherry whose partial the post-order walk did not store
  const int dmode = TIPS == HX_TIPS_CODES ? a.drop_mode : -1;
  auto dropped = [&](const int4& f1, int h) {
    return dmode >= 0 && ((f1.w >> (2 * h + dmode)) & 1);
  };
  // Every entry issues the same vector loads (placeholders where the entry
  // has nothing to read), so the wait for one entry's data counts the next
  // entry's loads exactly and never drains them.
  // Uniform base of a node's rows for this tile and sample; a child or pop
  // with nothing to read uses the hot line with a category stride of zero.
  const R* wbase = a.partials + (size_t)k * C * cstride + (size_t)tl * 256;
  const size_t sstride = (size_t)a.K * C * cstride;
  // 16-bit codes: a tip's code, or a dropped cherry's two codes (rows named by the entry)
  const uint16_t* cbase = a.pcodes + (size_t)tl * 64 + lane;
  auto fetch = [&](const int4& f0, const int4& f1, const int4& f2, PreIn<CM>& in) {
    if constexpr (TAU) {
      // issued first: the copy of this value into the loop-carried register
      // then waits for this load only
      const int CS = C * 16, t2 = threadIdx.x % (2 * CS);
      const int h = t2 >= CS, e = t2 - h * CS;
      const int b = h ? f0.z : f0.y;
      in.dt = mats_dt[(((size_t)k * a.B + b) * C + (e >> 4)) * 16 + (e & 15)];
    }
    {
      const bool gpop = (f0.w & 3) == 2 && (f0.w >> 8) >= RS;
      const R* pb = gpop ? wbase + (size_t)(f0.x - a.T) * sstride : hot;
      const size_t cs = gpop ? cstride : 0;
      const unsigned off = gpop ? off_real : off_hot;
#pragma unroll
      for (int c = 0; c < CM; ++c)
        if (EXACT || c < C) in.Up[c] = ld4(at(pb + c * cs, off));
    }
#pragma unroll
    for (int h = 0; h < 2; ++h) {
      const int ch = h ? f0.z : f0.y;
      const bool tip = (f1.z >> h) & 1;
      const bool coded = tip || dropped(f1, h);  // from codes: nothing to read here
      const R* pb = coded ? hot : wbase + (size_t)(ch - a.T) * sstride;
      const size_t cs = coded ? 0 : cstride;
      const unsigned off = coded ? off_hot : off_real;
      // the code first: a tip or a rebuilt cherry waits for it alone
      if constexpr (TIPS == HX_TIPS_CODES) {
        in.code[h] = cbase[(size_t)(coded ? (h ? f2.y : f2.x) : 0) * a.np];
      } else {
        const int nn = n < a.N ? n : a.N - 1;
        in.tipv[h] = ld4(tip ? a.dense + ((size_t)ch * a.N + nn) * 4 : at(hot, off_hot));
      }
      // a post-order partial is read once: streaming, so that pushed
      // adjoints stay in L2 until their pop
#pragma unroll
      for (int c = 0; c < CM; ++c)
        if (EXACT || c < C) in.L[h][c] = ld4_nt(at(pb + c * cs, off));
    }
  };


  // register stack: depths 0 and 1 (RS of them) in s0 and s1, moved by
  // selects on the wave-uniform depth
  static_assert(RS >= 0 && RS <= 2, "at most two register slots");
  V4<R> U[CM], s0[CM], s1[CM];
#pragma unroll
  for (int c = 0; c < CM; ++c) U[c] = s0[c] = s1[c] = zero4<R>();
  float* xs = xt;              // rows 4c + i: (w p_c / lik) Ut[c][i]
  float* ls = xt + 16 * kXS;    // rows 4c + j: L[c][j] (a tip: rows j)
  const int g4 = lane >> 4;

  // One entry per trip. Part 1 consumes the prefetched data (the adjoint, m,
  // the statistic's transposes, a pushed adjoint's store); then the next
  // entry's loads are issued into the same registers; part 2 forms the
  // statistic on the matrix cores, sums it over the block's tiles and forms
  // the continuing adjoint. Vector memory counters complete in order, so a
  // store only delays a later wait when it is issued shortly before the
  // awaited loads: the push store is issued early in part 1, and the
  // statistic store (part 2, after the loads) is issued by every thread on
  // every path so the wait for the next entry never has to cover it.
  // The statistic of thread t: entry t mod 2CS of child h's G.
  const int CS = C * 16, t2 = threadIdx.x % (2 * CS);
  const int sh = t2 >= CS, se = t2 - sh * CS;
  const R spc = a.props[(size_t)k * a.props_ks + (se >> 4)];  // p_c of this thread's entry
  auto stat_store = [&](const int4& q0, float s, float dt) {
    const int b = sh ? q0.z : q0.y;
    if constexpr (TAU) {
      double tv = (double)s * (double)dt;
#pragma unroll
      for (int o = 1; o < 16; o <<= 1) tv += __shfl_xor(tv, o);
      a.gtau[(((size_t)b * a.K + k) * a.nchunks + chunk) * C + (se >> 4)] = tv;
    } else {
      a.gpart[(((size_t)b * a.K + k) * a.nchunks + chunk) * CS + se] = s;
    }
  };
  PreIn<CM> cur;
  int4 e0, e1, e2, e3;
  entry(0, e0, e1, e2, e3);
  fetch(e0, e1, e2, cur);
  // the store every trip ends with, here a zero that entry 0 overwrites, so
  // the loop is entered with the same outstanding memory operations (in the
  // same order: no load may be scheduled after it)
  __builtin_amdgcn_sched_barrier(0);
  stat_store(e0, 0.f, 0.f);
  for (int i = 0; i < nprog; ++i) {
    int4 f0 = e0, f1 = e1, f2 = e2, f3 = e3;
    if (i + 1 < nprog) entry(i + 1, f0, f1, f2, f3);
    const int ch[2] = {e0.y, e0.z};
    const int usrc = e0.w, cont = e1.x, push = e1.y, tips = e1.z;

    // ---------------------------------------------------------------- part 1
    // the parent's adjoint: the frequencies, the previous entry's
    // continuation (already in U) or a pop
    if (usrc == 0) {
#pragma unroll
      for (int c = 0; c < CM; ++c) U[c] = pi;
    } else if (usrc != 1) {
      const int d = usrc >> 8;
      if (d >= RS) {
#pragma unroll
        for (int c = 0; c < CM; ++c) U[c] = cur.Up[c];
      } else {
#pragma unroll
        for (int c = 0; c < CM; ++c)
#pragma unroll
          for (int q = 0; q < 4; ++q) U[c].x[q] = d == 0 ? s0[c].x[q] : s1[c].x[q];
      }
    }
    // m_h = P_h L_h
    V4<R> m[2][CM], tv[2];
    auto child = [&](auto hc) {
      constexpr int h = decltype(hc)::value;
      if ((tips >> h) & 1) {
        if constexpr (TIPS == HX_TIPS_CODES) tv[h] = tbl[cur.code[h]];
        else tv[h] = cur.tipv[h];
#pragma unroll
        for (int c = 0; c < CM; ++c) {
          if (EXACT || c < C) {
            R P[16];
            load_p16(mk + ((size_t)ch[h] * C + c) * 16, P);
            m[h][c] = mv(P, tv[h]);
          }
        }
      } else {
        if constexpr (TIPS == HX_TIPS_CODES) {
          if (dropped(e1, h)) {
            // the cherry's partial before the post-order walk's power-of-two rescale: every
            // term below is a ratio in it, so the result is the same as from the stored rows
            const int code = cur.code[h];
            const V4<R> t0 = tbl[code & 0xff], t1 = tbl[code >> 8];
            const int g0 = h ? e3.x : e2.z, g1 = h ? e3.y : e2.w;
#pragma unroll
            for (int c = 0; c < CM; ++c) {
              if (EXACT || c < C) {
                R P[16];
                load_p16(mk + ((size_t)g0 * C + c) * 16, P);
                const V4<R> y0 = mv(P, t0);
                load_p16(mk + ((size_t)g1 * C + c) * 16, P);
                const V4<R> y1 = mv(P, t1);
                // through the placeholder rows just loaded (finite P values, times 0): the
                // rows are then awaited one category at a time, as on the stored path, and
                // the two paths join with the same outstanding loads
#pragma unroll
                for (int s = 0; s < 4; ++s)
                  cur.L[h][c].x[s] = fmaf(0.f, cur.L[h][c].x[s], y0.x[s] * y1.x[s]);
              }
            }
          }
        }
#pragma unroll
        for (int c = 0; c < CM; ++c) {
          if (EXACT || c < C) {
            R P[16];
            load_p16(mk + ((size_t)ch[h] * C + c) * 16, P);
            m[h][c] = mv(P, cur.L[h][c]);
          }
        }
      }
    };
    child(H0{});
    child(H1{});
    float dt = 0.f;
    if constexpr (TAU) dt = cur.dt;
    V4<R> Ut[2][CM];
    R lik = R(0);
#pragma unroll
    for (int c = 0; c < CM; ++c) {
      if (EXACT || c < C) {
        R t = R(0);
#pragma unroll
        for (int s = 0; s < 4; ++s) {
          Ut[0][c].x[s] = U[c].x[s] * m[1][c].x[s];
          Ut[1][c].x[s] = U[c].x[s] * m[0][c].x[s];
          t = fma(Ut[0][c].x[s], m[0][c].x[s], t);
        }
        lik = fma(pc[c], t, lik);
      }
    }
    // w / lik goes into the B operand (L: 4 values for a tip) and p_c is
    // applied to the summed statistic
    const R coef = (val && lik > R(0)) ? w * __builtin_amdgcn_rcpf(lik) : R(0);
    // The statistic of child h: its rows to the wave's LDS transposes, then
    // 16 products on the matrix cores; child 1's rows overwrite child 0's
    // after child 0's operand reads (LDS operations of a wave stay in order).
    auto stat_rows = [&](auto hc) {
      constexpr int h = decltype(hc)::value;
      const bool tip = (tips >> h) & 1;
#pragma unroll
      for (int c = 0; c < CM; ++c) {
        if (EXACT || c < C) {
#pragma unroll
          for (int s = 0; s < 4; ++s) xs[(4 * c + s) * kXS + lane] = Ut[h][c].x[s];
          if (!tip) {
#pragma unroll
            for (int s = 0; s < 4; ++s) ls[(4 * c + s) * kXS + lane] = coef * cur.L[h][c].x[s];
          }
        }
      }
      if (tip) {
#pragma unroll
        for (int s = 0; s < 4; ++s) ls[s * kXS + lane] = coef * tv[h].x[s];
      }
    };
    // Block b = 4c + q of the 4x4x1 product: lane 4b + i gives A[b][i] = Ut[c][i] and lane
    // 4b + j gives B[b][j] = (w / lik) L[c][j] of pattern 16q + m in product m; register i of
    // lane 4b + j accumulates G[c][i][j] over those patterns.
    const int srow = 4 * g4 + (lane & 3), scol = 16 * ((lane >> 2) & 3);
    auto stat_mfma = [&](int h) {
      __builtin_amdgcn_wave_barrier();
      const int rl = ((tips >> h) & 1) ? (lane & 3) : srow;
      f4 g = f4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        const f4 A = *reinterpret_cast<const f4*>(xs + srow * kXS + scol + 4 * u);
        const f4 Bv = *reinterpret_cast<const f4*>(ls + rl * kXS + scol + 4 * u);
#pragma unroll
        for (int e = 0; e < 4; ++e) g = __builtin_amdgcn_mfma_f32_4x4x1f32(A[e], Bv[e], g, 0, 0, 0);
      }
      __builtin_amdgcn_wave_barrier();
      return g;
    };
    // The four pattern groups of each category (lanes 16c + 4q + j) are added by row
    // rotations; lane 16c + j then holds G_h[c][i][j] in register i and writes it to the
    // wave's slot of the block's sum, double-buffered by the entry's parity (the other
    // buffer's readers have all passed the previous entry's barrier).
    const int par = i & 1;
    auto stat_put = [&](int h, f4 g) {
#pragma unroll
      for (int r = 0; r < 4; ++r) g[r] = add_row_ror<8>(add_row_ror<4>(g[r]));
      if (((lane >> 2) & 3) == 0 && g4 < C) {
        float* dst = red + ((par * 2 + h) * W + wave) * 64 + g4 * 16 + (lane & 3);
#pragma unroll
        for (int r = 0; r < 4; ++r) dst[4 * r] = g[r];
      }
    };
    stat_rows(H0{});
    stat_put(0, stat_mfma(0));
    stat_rows(H1{});
    // adjoints of internal children, computed straight into their registers
    // (the child index is a compile-time constant in each branch)
    auto adjoint = [&](auto hc, V4<R>(&Un)[CM]) {
      constexpr int h = decltype(hc)::value;
      R M = R(0);
#pragma unroll
      for (int c = 0; c < CM; ++c) {
        if (EXACT || c < C) {
          R P[16];
          load_p16(mk + ((size_t)ch[h] * C + c) * 16, P);
          Un[c] = mv_t(P, Ut[h][c]);
          M = fmax_(M, max4(Un[c]));
        } else {
          Un[c] = zero4<R>();
        }
      }
      const int ex = scale_exp(M);
#pragma unroll
      for (int c = 0; c < CM; ++c) scale4(Un[c], ex);
    };
    auto push_adjoint = [&](auto hc, int d) {
      constexpr int h = decltype(hc)::value;
      V4<R> Un[CM];
      adjoint(hc, Un);
      if (d >= RS) {
        if (tval) {
          R* dst = const_cast<R*>(wbase) + (size_t)(ch[h] - a.T) * sstride;
#pragma unroll
          for (int c = 0; c < CM; ++c)
            if (EXACT || c < C) st4(const_cast<R*>(at(dst + c * cstride, off_real)), Un[c]);
        }
      } else {
#pragma unroll
        for (int c = 0; c < CM; ++c)
#pragma unroll
          for (int q = 0; q < 4; ++q) {
            s0[c].x[q] = d == 0 ? Un[c].x[q] : s0[c].x[q];
            s1[c].x[q] = d == 1 ? Un[c].x[q] : s1[c].x[q];
          }
      }
    };
    if (push >= 0) {
      if ((push & 0xff) == 0) push_adjoint(H0{}, push >> 8);
      else push_adjoint(H1{}, push >> 8);
    }
    if (cont == 0) adjoint(H0{}, U);
    else if (cont == 1) adjoint(H1{}, U);

    // ------------------------------------------------ the next entry's loads
    fetch(f0, f1, f2, cur);

    // ---------------------------------------------------------------- part 2
    stat_put(1, stat_mfma(1));
    lds_barrier();
    {
      // thread t sums its entry over the W tiles; threads with the same entry
      // store the same value to the same address
      float s = 0.f;
      for (int q = 0; q < W; ++q) s += red[((par * 2 + sh) * W + q) * 64 + se];
      stat_store(e0, s * spc, dt);
    }
    e0 = f0;
    e1 = f1;
    e2 = f2;
    e3 = f3;
  }
}

// Occupancy and stack: the d loglik / d mats walk with tip codes fits 128
// VGPRs with every push in memory (4 waves per SIMD); the d loglik / d tau
// walk with tip codes keeps one register stack slot at four categories and
// the others two, at 3 waves per SIMD (168 VGPRs) without spills (a spill's scratch
// traffic would break the in-order waits).
template <int CM, bool EXACT>
static void pre_walk_kernel(int tips, bool tau,
                            void (*&kern)(PassArgs<float>, const float*, const float*, const int*,
                                          int)) {
  if (tau)
    kern = tips == HX_TIPS_CODES ? k_pre_walk<HX_TIPS_CODES, CM, EXACT, (CM == 4 ? 1 : 2), true>
                                 : k_pre_walk<HX_TIPS_DENSE, CM, EXACT, 2, true>;
  else
    kern = tips == HX_TIPS_CODES ? k_pre_walk<HX_TIPS_CODES, CM, EXACT, 0, false>
                                 : k_pre_walk<HX_TIPS_DENSE, CM, EXACT, 2, false>;
}

// ------------------------------------------------------------------ launchers
static size_t tbl_bytes(int tips, int ncodes, size_t rb) {
  return tips == HX_TIPS_CODES ? (size_t)ncodes * 4 * rb : 0;
}

// Copies tip codes [T][N] into rows of ntiles * 64 bytes (zero padded) so the
// walk stages a tile's codes with aligned 16-byte loads.
// grid (ceil(npad / 1024), T): thread j4 writes bytes 4 j4 .. 4 j4 + 3 of row blockIdx.y as
// one 32-bit store (npad is a multiple of 64, so the stores are aligned)
__global__ void k_pad_codes(const uint8_t* __restrict__ codes, uint8_t* __restrict__ out, int T,
                            int N, int npad) {
  const int t = blockIdx.y;
  const int j = 4 * (blockIdx.x * blockDim.x + threadIdx.x);
  if (j >= npad) return;
  const uint8_t* src = codes + (size_t)t * N;
  uint32_t v = 0;
#pragma unroll
  for (int q = 0; q < 4; ++q)
    if (j + q < N) v |= (uint32_t)src[j + q] << (8 * q);
  *reinterpret_cast<uint32_t*>(out + (size_t)t * npad + j) = v;
}

template <typename R>
size_t dna_walk_lds(int T, int W, int need, int tips, int ncodes) {
  return (tips == HX_TIPS_CODES ? (size_t)T * 64 : 0) + tbl_bytes(tips, ncodes, sizeof(R)) +
         (size_t)W * need * 64 * 4;
}

template <typename R>
int launch_post_dna(const PassArgs<R>& a, const int32_t* prog, const int32_t* segtab, int nseg,
                    int nprog, int need, int tips, hipStream_t st) {
  (void)segtab;
  (void)nseg;
  // W samples share the staged codes; 8 keeps their P matrices (K B C 64 B
  // per sample) within L2; many categories need a smaller block (registers)
  const int wmax = a.C > 4 ? 4 : 8;
  int W = a.K < wmax ? a.K : wmax;
  while (W > 1 && dna_walk_lds<R>(a.T, W, need, tips, a.ncodes) > kDnaWalkLdsMax) W >>= 1;
  const size_t lds = dna_walk_lds<R>(a.T, W, need, tips, a.ncodes);
  HX_REQUIRE(lds <= kDnaWalkLdsMax, HX_ERR_UNSUPPORTED,
             "post-order walk: tip codes of the tile exceed LDS (too many taxa)");
  HX_REQUIRE(sizeof(R) == 8 || (size_t)a.C * a.np * 4 * sizeof(R) < ((size_t)1 << 31),
             HX_ERR_UNSUPPORTED, "post-order walk: a node's rows of one sample exceed 2 GiB");
  // mode 1 would drop cherries that a walk with register depths takes from registers; this
  // walk keeps none, so only mode 0 matches it
  HX_REQUIRE(a.drop_mode < 0 || (sizeof(R) == 4 && tips == HX_TIPS_CODES && a.drop_mode == 0),
             HX_ERR_UNSUPPORTED, "post-order walk: drop mode without the register depths it assumes");
  if (tips == HX_TIPS_CODES) {
    const size_t npad = (size_t)a.ntiles * 64;
    HX_REQUIRE(a.T <= 65535, HX_ERR_UNSUPPORTED, "post-order walk: too many taxa");
    k_pad_codes<<<dim3((unsigned)((npad / 4 + 255) / 256), (unsigned)a.T), 256, 0, st>>>(
        a.codes, a.codes_pad, a.T, a.N, (int)npad);
  }
  dim3 grid(a.ntiles, (a.K + W - 1) / W);
  const int4* pg = reinterpret_cast<const int4*>(prog);
#define HX_WALK(CMV, EX)                                                                     \
  do {                                                                                       \
    auto kern = tips == HX_TIPS_CODES ? k_post_walk<R, HX_TIPS_CODES, CMV, EX>               \
                                      : k_post_walk<R, HX_TIPS_DENSE, CMV, EX>;              \
    if (lds > 64 * 1024)                                                                     \
      HX_HIP_TRY(hipFuncSetAttribute(reinterpret_cast<const void*>(kern),                    \
                                     hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds)); \
    kern<<<grid, 64 * W, lds, st>>>(a, a.mats, pg, nprog, need, W);                          \
  } while (0)
  switch (a.C) {
    case 1: HX_WALK(1, true); break;
    case 2: HX_WALK(2, true); break;
    case 3: HX_WALK(4, false); break;
    case 4: HX_WALK(4, true); break;
    default:
      if (a.C <= 8) HX_WALK(8, false);
      else HX_WALK(16, false);
  }
#undef HX_WALK
  HX_HIP_TRY(hipGetLastError());
  return HX_OK;
}

template <typename R>
int launch_pre_dna(const PassArgs<R>& a, const int32_t* nodes, int width, int tips, bool root,
                   hipStream_t st) {
  const int cp = next_pow2(a.C);
  const size_t rb = sizeof(R);
  const int ne = pre_pass_entries(cp, (int)rb);
  const size_t lds = (size_t)2 * a.C * 16 * rb + tbl_bytes(tips, a.ncodes, rb) +
                     (size_t)64 * cp * (ne + 1) * rb;
  HX_REQUIRE(lds <= kDnaLdsBudget, HX_ERR_UNSUPPORTED, "pre-order LDS exceeds the 64 KiB budget");
  dim3 grid(a.nchunks, a.K, width);
  const int threads = 64 * cp;
#define HX_PRE(TV, RV, CPV) \
  k_pre<R, TV, RV, CPV><<<grid, threads, lds, st>>>(a, a.mats, a.mats_dt, nodes)
#define HX_PRE_T(CPV)                                                                        \
  if (tips == HX_TIPS_CODES) {                                                               \
    if (root) HX_PRE(HX_TIPS_CODES, true, CPV); else HX_PRE(HX_TIPS_CODES, false, CPV);      \
  } else {                                                                                   \
    if (root) HX_PRE(HX_TIPS_DENSE, true, CPV); else HX_PRE(HX_TIPS_DENSE, false, CPV);      \
  }
  switch (cp) {
    case 1: HX_PRE_T(1) break;
    case 2: HX_PRE_T(2) break;
    case 4: HX_PRE_T(4) break;
    case 8: HX_PRE_T(8) break;
    default: HX_PRE_T(16) break;
  }
#undef HX_PRE_T
#undef HX_PRE
  HX_HIP_TRY(hipGetLastError());
  return HX_OK;
}

// The pre-order walk's 16-bit codes [T + ncher][np]: row t < T is tip t's code, row T + j
// cherry j's two codes (first tip's | second's << 8); padding patterns are 0.
// grid (ceil(np / 1024), T + ncher): thread j4 writes patterns 4 j4 .. 4 j4 + 3 of row
// blockIdx.y as one 64-bit store (np is a multiple of 64, so the stores are aligned)
__global__ void k_pair_codes(const uint8_t* __restrict__ codes, const int32_t* __restrict__ cher,
                             uint16_t* __restrict__ out, int T, int ncher, int N, int np) {
  const int r = blockIdx.y;
  const int j = 4 * (blockIdx.x * blockDim.x + threadIdx.x);
  if (j >= np) return;
  const bool tip = r < T;
  const uint8_t* s0 = codes + (size_t)(tip ? r : cher[2 * (r - T)]) * N;
  const uint8_t* s1 = codes + (size_t)(tip ? r : cher[2 * (r - T) + 1]) * N;
  uint64_t v = 0;
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    if (j + q < N) {
      const uint32_t c = tip ? (uint32_t)s0[j + q] : ((uint32_t)s0[j + q] | ((uint32_t)s1[j + q] << 8));
      v |= (uint64_t)c << (16 * q);
    }
  }
  *reinterpret_cast<uint64_t*>(out + (size_t)r * np + j) = v;
}

int launch_pre_walk(const PassArgs<float>& a, const int32_t* prog, int nprog, int tips,
                    hipStream_t st) {
  const int W = kPreWalkW;
  HX_REQUIRE(a.C <= kPreWalkMaxC && a.nchunks == (a.ntiles + W - 1) / W, HX_ERR_UNSUPPORTED,
             "pre-order walk: chunking does not match the workgroup's tiles");
  if (tips == HX_TIPS_CODES) {
    HX_REQUIRE(a.pcodes != nullptr && 2 * a.ncher <= a.T, HX_ERR_UNSUPPORTED,
               "pre-order walk: no room for the 16-bit codes");
    HX_REQUIRE(a.T + a.ncher <= 65535, HX_ERR_UNSUPPORTED, "pre-order walk: too many code rows");
    k_pair_codes<<<dim3((unsigned)((a.np / 4 + 255) / 256), (unsigned)(a.T + a.ncher)), 256, 0,
                   st>>>(a.codes, a.cher, a.pcodes, a.T, a.ncher, a.N, a.np);
    HX_HIP_TRY(hipGetLastError());
  }
  const size_t lds = (size_t)W * 2 * 16 * kXS * sizeof(float) + (size_t)4 * W * 64 * sizeof(float) +
                     tbl_bytes(tips, a.ncodes, sizeof(float));
  HX_REQUIRE(lds <= kDnaWalkLdsMax, HX_ERR_UNSUPPORTED, "pre-order walk: LDS");
  void (*kern)(PassArgs<float>, const float*, const float*, const int*, int) = nullptr;
  const bool tau = a.mats_dt != nullptr;
  switch (a.C) {
    case 1: pre_walk_kernel<1, true>(tips, tau, kern); break;
    case 2: pre_walk_kernel<2, true>(tips, tau, kern); break;
    case 3: pre_walk_kernel<4, false>(tips, tau, kern); break;
    default: pre_walk_kernel<4, true>(tips, tau, kern); break;
  }
  if (lds > 64 * 1024)
    HX_HIP_TRY(hipFuncSetAttribute(reinterpret_cast<const void*>(kern),
                                   hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
  kern<<<dim3(a.nchunks, a.K), 64 * W, lds, st>>>(a, a.mats, a.mats_dt, prog, nprog);
  HX_HIP_TRY(hipGetLastError());
  return HX_OK;
}

template int launch_post_dna<float>(const PassArgs<float>&, const int32_t*, const int32_t*, int,
                                    int, int, int, hipStream_t);
template int launch_post_dna<double>(const PassArgs<double>&, const int32_t*, const int32_t*, int,
                                     int, int, int, hipStream_t);
template int launch_pre_dna<float>(const PassArgs<float>&, const int32_t*, int, int, bool,
                                   hipStream_t);
template int launch_pre_dna<double>(const PassArgs<double>&, const int32_t*, int, int, bool,
                                    hipStream_t);

}  // namespace hx

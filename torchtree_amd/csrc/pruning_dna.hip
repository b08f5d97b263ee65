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
//   partials [slot][K][C][np][4]  slot = node - T, np = 64 x tiles
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

// Workgroup barrier that orders LDS accesses only. __syncthreads() also
// releases global memory at workgroup scope, which waits for every vector
// memory operation in flight (the stores of partials, prefetched loads) at
// each barrier; the exchanges here go through LDS, and a thread only reads
// back global data it wrote itself, so that wait is not needed.
__device__ __forceinline__ void lds_barrier() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
  __builtin_amdgcn_s_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
}

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
  const cptr<R> p = to_const(mats) + (((size_t)k * B + b) * C + c) * 16;
#pragma unroll
  for (int i = 0; i < 16; ++i) P[i] = p[i];
}

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
  return ld4(a.partials + (prow(a.T, a.K, a.C, a.np, ch, k, c) + n) * 4);
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
  // has none, for slot 0 (a valid address, value unused): every iteration
  // then issues the same vector memory operations, so the wait for these
  // loads counts only the older ones and not this node's stores.
  auto stacked_of = [&](const int4& f0, const int4& f1) {
    if ((f1.z & 0xff) == 2) return f0.y;
    if (((f1.z >> 8) & 0xff) == 2) return f0.z;
    return a.T;
  };
  auto readback = [&](int ch, V4<R>(&x)[CM]) {
    const R* p = base + (size_t)(ch - a.T) * sstride;
#pragma unroll
    for (int c = 0; c < CM; ++c)
      if (EXACT || c < C) x[c] = ld4(p + (size_t)c * cstride);
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
    R* p = base + (size_t)(v - a.T) * sstride;
#pragma unroll
    for (int c = 0; c < CM; ++c)
      if (EXACT || c < C) st4(p + (size_t)c * cstride, y[c]);
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
// Group rooted at p (root: U = freqs). For parent q with children a, b:
//   Ut_a = U_q * (P_b L_b), Ut_b = U_q * (P_a L_a)
//   lik  = sum_c p_c Ut_a.(P_a L_a)                    (scaled site likelihood)
//   d loglik / d P_a[c][i][j] += w p_c Ut_a[i] L_a[j] / lik
//   U_a  = P_a^T Ut_a, rescaled (scale-free: every term is a ratio)
// LV = 2 also applies this to the children's children, with the children's U
// kept in registers; only the U of the last level is written (over L, in place).
// Every global load of a tile (U, the children's and grandchildren's L or tip
// codes) is issued one tile ahead, before the barriers of the current tile.
// CM: C rounded up to a power of two (block-size bound and unrolled category
// loops).
template <typename R, int LV>
struct PreTile {
  V4<R> U;
  V4<R> Lc[2];
  V4<R> Lg[LV == 2 ? 2 : 1][2];
  int cc[2];
  int cg[LV == 2 ? 2 : 1][2];
};

template <typename R, int TIPS, bool ROOT, int LV, int CM>
__global__ __launch_bounds__(64 * CM) void k_pre_grp(
    PassArgs<R> a, const R* __restrict__ mats, const R* __restrict__ mats_dt,
    const int32_t* __restrict__ nodes, int wslots) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  constexpr int NQ = LV == 2 ? 4 : 2;  // values exchanged at once
  const int C = a.C, nt = C * 64;
  const int c = uni(threadIdx.x >> 6), lane = threadIdx.x & 63;
  const int chunk = blockIdx.x, k = blockIdx.y;
  const int p = nodes[blockIdx.z];
  constexpr int NX = 64 * CM;  // exchange rows: unused categories hold 0
  R* xs = reinterpret_cast<R*>(smem_raw);                 // [2][NQ][NX] exchange
  R* wbuf = xs + 2 * NQ * NX;                             // [wslots][16][64] reductions
  V4<R>* tbl = reinterpret_cast<V4<R>*>(wbuf + (size_t)wslots * 1024);
  stage_tbl<R, TIPS>(a, tbl, nt);
  if constexpr (CM > 1) {
    for (int i = threadIdx.x; i < 2 * NQ * (NX - nt); i += nt) {
      const int row = i / (NX - nt), col = i % (NX - nt);
      xs[row * NX + nt + col] = R(0);
    }
  }

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

  // node `nd` (internal: its partial, tip: its code or dense row) at pattern n
  auto fetch_node = [&](int nd, int n, bool val, V4<R>& L, int& code) {
    L = zero4<R>();
    code = 0;
    if (!val) return;
    if (nd >= a.T) {
      L = ld4(a.partials + (prow(a.T, a.K, C, a.np, nd, k, c) + n) * 4);
    } else if constexpr (TIPS == HX_TIPS_CODES) {
      code = a.codes[(size_t)nd * a.N + n];
    } else {
      L = ld4(a.dense + ((size_t)nd * a.N + n) * 4);
    }
  };
  auto fetch = [&](int tile, PreTile<R, LV>& in) {
    const int n = tile * 64 + lane;
    const bool val = n < a.N;
    in.U = zero4<R>();
    if constexpr (!ROOT) {
      if (val) in.U = ld4(a.partials + (prow(a.T, a.K, C, a.np, p, k, c) + n) * 4);
    }
#pragma unroll
    for (int h = 0; h < 2; ++h) fetch_node(ch[h], n, val, in.Lc[h], in.cc[h]);
    if constexpr (LV == 2) {
#pragma unroll
      for (int h = 0; h < 2; ++h)
        if (in_c[h])
#pragma unroll
          for (int j = 0; j < 2; ++j) fetch_node(gk[h][j], n, val, in.Lg[h][j], in.cg[h][j]);
    }
  };
  // tip rows come from the LDS table once the codes have arrived
  auto resolve = [&](int nd, V4<R>& L, int code) {
    if constexpr (TIPS == HX_TIPS_CODES) {
      if (nd < a.T) L = tbl[code];
    }
  };

  int par = 0;
  // Exchange up to four values per lane with the other category waves.
  auto xsum = [&](R* vals, int nq, bool take_max) {
#pragma unroll
    for (int q = 0; q < NQ; ++q)
      if (q < nq) xs[(par * NQ + q) * NX + threadIdx.x] = vals[q];
    lds_barrier();
#pragma unroll
    for (int q = 0; q < NQ; ++q) {
      if (q < nq) {
        R u[CM];
#pragma unroll
        for (int cc = 0; cc < CM; ++cc) u[cc] = xs[(par * NQ + q) * NX + cc * 64 + lane];
        R t = u[0];
#pragma unroll
        for (int cc = 1; cc < CM; ++cc) t = take_max ? fmax_(t, u[cc]) : t + u[cc];
        vals[q] = t;
      }
    }
    par ^= 1;
  };

  const int t0 = chunk * a.tiles_per_chunk;
  const int t1 = min(t0 + a.tiles_per_chunk, a.ntiles);
  PreTile<R, LV> nxt;
  if (t0 < t1) fetch(t0, nxt);
  lds_barrier();  // code table staged
  for (int tile = t0; tile < t1; ++tile) {
    PreTile<R, LV> cur = nxt;
    if (tile + 1 < t1) fetch(tile + 1, nxt);
    const int n = tile * 64 + lane;
    const bool val = n < a.N;
    const R w = val ? (R)a.weights[n] : R(0);
    V4<R> U = cur.U;
    if constexpr (ROOT) U = val ? pi : zero4<R>();
    V4<R> PL[2], Ut[2];
    R P[16];
#pragma unroll
    for (int h = 0; h < 2; ++h) {
      resolve(ch[h], cur.Lc[h], cur.cc[h]);
      load_p(mats, a.B, C, k, ch[h], c, P);
      PL[h] = mv(P, cur.Lc[h]);
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
    acc_outer(g[0], coef, Ut[0], cur.Lc[0]);
    acc_outer(g[1], coef, Ut[1], cur.Lc[1]);
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
            if (val) st4(a.partials + (prow(a.T, a.K, C, a.np, ch[h], k, c) + n) * 4, Uc[h]);
          }
        }
      }
    }
    if constexpr (LV == 2) {
      if (in_c[0] || in_c[1]) {
        V4<R> PLg[2][2], Utg[2][2];
#pragma unroll
        for (int h = 0; h < 2; ++h) {
          d[h] = R(0);
          if (in_c[h]) {
#pragma unroll
            for (int j = 0; j < 2; ++j) {
              resolve(gk[h][j], cur.Lg[h][j], cur.cg[h][j]);
              load_p(mats, a.B, C, k, gk[h][j], c, P);
              PLg[h][j] = mv(P, cur.Lg[h][j]);
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
            acc_outer(gg[h][0], ch_coef, Utg[h][0], cur.Lg[h][0]);
            acc_outer(gg[h][1], ch_coef, Utg[h][1], cur.Lg[h][1]);
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
                  st4(a.partials + (prow(a.T, a.K, C, a.np, gk[h][j], k, c) + n) * 4, Utg[h][j]);
              }
        }
      }
    }
  }

  // Per-wave reduction buffers; when fewer than C fit, waves take turns.
  R* buf = wbuf + (size_t)(c % wslots) * 1024;
  const int rounds = (C + wslots - 1) / wslots;
  for (int rd = 0; rd < rounds; ++rd) {
    if (c / wslots == rd) {
#pragma unroll
      for (int h = 0; h < 2; ++h) flush_branch(a, mats_dt, g[h], buf, ch[h], k, c, chunk);
      if constexpr (LV == 2) {
#pragma unroll
        for (int h = 0; h < 2; ++h)
          if (in_c[h])
#pragma unroll
            for (int j = 0; j < 2; ++j)
              flush_branch(a, mats_dt, gg[h][j], buf, gk[h][j], k, c, chunk);
      }
    }
    if (rounds > 1) lds_barrier();
  }
}

// ------------------------------------------------------------------ launchers
static size_t tbl_bytes(int tips, int ncodes, size_t rb) {
  return tips == HX_TIPS_CODES ? (size_t)ncodes * 4 * rb : 0;
}

// Copies tip codes [T][N] into rows of ntiles * 64 bytes (zero padded) so the
// walk stages a tile's codes with aligned 16-byte loads.
__global__ void k_pad_codes(const uint8_t* __restrict__ codes, uint8_t* __restrict__ out, int T,
                            int N, int npad) {
  const size_t idx = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= (size_t)T * npad) return;
  const int t = (int)(idx / npad), j = (int)(idx % npad);
  out[idx] = j < N ? codes[(size_t)t * N + j] : 0;
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
  if (tips == HX_TIPS_CODES) {
    const size_t npad = (size_t)a.ntiles * 64, total = (size_t)a.T * npad;
    k_pad_codes<<<(unsigned)((total + 255) / 256), 256, 0, st>>>(a.codes, a.codes_pad, a.T, a.N,
                                                                 (int)npad);
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
int dna_pre_levels(int C) {
  return (std::is_same<R, float>::value && C <= 4) ? 2 : 1;
}

template <typename R>
int launch_pre_dna(const PassArgs<R>& a, const int32_t* nodes, int width, int tips, bool root,
                   hipStream_t st) {
  const size_t nt = (size_t)a.C * 64, rb = sizeof(R);
  const int lv = dna_pre_levels<R>(a.C);
  const size_t nx = (size_t)next_pow2(a.C) * 64;
  const size_t fixed = 2 * (lv == 2 ? 4 : 2) * nx * rb + tbl_bytes(tips, a.ncodes, rb);
  int wslots = a.C;
  while (wslots > 1 && fixed + (size_t)wslots * 1024 * rb > kDnaLdsBudget) --wslots;
  const size_t lds = fixed + (size_t)wslots * 1024 * rb;
  HX_REQUIRE(lds <= kDnaLdsBudget, HX_ERR_UNSUPPORTED, "pre-order LDS exceeds the 64 KiB budget");
  dim3 grid(a.nchunks, a.K, width);
  const int threads = a.C * 64;
#define HX_PRE(TV, RV, LVV, CMV) \
  k_pre_grp<R, TV, RV, LVV, CMV><<<grid, threads, lds, st>>>(a, a.mats, a.mats_dt, nodes, wslots)
#define HX_PRE_T(LVV, CMV)                                                               \
  if (tips == HX_TIPS_CODES) {                                                           \
    if (root) HX_PRE(HX_TIPS_CODES, true, LVV, CMV); else HX_PRE(HX_TIPS_CODES, false, LVV, CMV); \
  } else {                                                                               \
    if (root) HX_PRE(HX_TIPS_DENSE, true, LVV, CMV); else HX_PRE(HX_TIPS_DENSE, false, LVV, CMV); \
  }
  const int cm = next_pow2(a.C);
  if (lv == 2) {
    if constexpr (std::is_same<R, float>::value) {
      if (cm == 1) { HX_PRE_T(2, 1) } else if (cm == 2) { HX_PRE_T(2, 2) } else { HX_PRE_T(2, 4) }
    }
  } else {
    switch (cm) {
      case 1: HX_PRE_T(1, 1) break;
      case 2: HX_PRE_T(1, 2) break;
      case 4: HX_PRE_T(1, 4) break;
      case 8: HX_PRE_T(1, 8) break;
      default: HX_PRE_T(1, 16) break;
    }
  }
#undef HX_PRE_T
#undef HX_PRE
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
template int dna_pre_levels<float>(int);
template int dna_pre_levels<double>(int);

}  // namespace hx

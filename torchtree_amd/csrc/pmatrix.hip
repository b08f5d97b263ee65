// Transition matrices P = L diag(exp(lambda * t * r_c)) R from an FP64
// eigensystem, and the chain rule back from d loglik / d P.
//
// Reference: SymmetricSubstitutionModel.p_t (substitution_model/abstract.py:57-79)
// computes the same product with diag_embed and two batched matmuls, and gets
// its gradient from autograd through torch.linalg.eigh (abstract.py:79).
// Here the exponentials and products stay in FP64; the output is cast to the
// pruning precision. The backward uses the Frechet derivative of exp in the
// eigenbasis with divided differences of exp, which is exact at repeated
// eigenvalues (where the eigh backward is undefined).
#include "common.hpp"

#include <math.h>

namespace hx {

// One thread per matrix element; S is small (<= 8), so each element
// recomputes its S exponentials.
template <typename O>
__global__ void k_pmatrix_elem(int S, int K, int B, int C, int eb, const double* __restrict__ lam,
                               const double* __restrict__ L, const double* __restrict__ R,
                               const double* __restrict__ bls, const double* __restrict__ rates,
                               int rb, O* __restrict__ out, int order) {
  const size_t idx = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  const size_t SS = (size_t)S * S;
  const size_t total = (size_t)K * B * C * SS;
  if (idx >= total) return;
  const int ij = (int)(idx % SS);
  const size_t mat = idx / SS;
  const int c = (int)(mat % C);
  const size_t kb = mat / C;
  const int b = (int)(kb % B), k = (int)(kb / B);
  const int e = eb ? k : 0;
  const double tau = bls[(size_t)k * B + b] * rates[(size_t)(rb ? k : 0) * C + c];
  const int i = ij / S, j = ij % S;
  const double* le = L + (size_t)e * SS;
  const double* re = R + (size_t)e * SS;
  const double* la = lam + (size_t)e * S;
  double acc = 0.0;
  for (int m = 0; m < S; ++m) {
    const double d = order ? la[m] * exp(la[m] * tau) : exp(la[m] * tau);
    acc = fma(le[i * S + m] * d, re[m * S + j], acc);
  }
  out[idx] = (O)acc;
}

// Dense spectra (17..64 states): P = L diag(e) R on the FP64 matrix cores, one block of 256
// threads per matrix. S is padded to SP = 16 MT with zeros; L diag(e) and R are staged in LDS
// with row stride SP + 1, and 16x16 tiles are formed with v_mfma_f64_16x16x4f64 (A[i][k] from
// lane i + 16 k, B[k][j] from lane j + 16 k, D[i][j] in lane j + 16 (i % 4), register i / 4).
// The same values as k_pmatrix_block up to the order of the sums.
template <typename O, int MT>
__global__ __launch_bounds__(256) void k_pmatrix_mfma(int S, int K, int B, int C, int eb,
                                                      const double* __restrict__ lam,
                                                      const double* __restrict__ L,
                                                      const double* __restrict__ R,
                                                      const double* __restrict__ bls,
                                                      const double* __restrict__ rates, int rb,
                                                      O* __restrict__ out, int order) {
  typedef double d4 __attribute__((ext_vector_type(4)));
  constexpr int SP = 16 * MT, LD = SP + 1;
  extern __shared__ double sm[];
  double* Ab = sm;            // [SP][LD]: L[i][m] e[m]
  double* Bb = sm + SP * LD;  // [SP][LD]: R[m][j]
  double* ex = sm + 2 * SP * LD;
  const size_t mat = blockIdx.x;
  const int c = (int)(mat % C);
  const size_t kb = mat / C;
  const int b = (int)(kb % B), k = (int)(kb / B);
  const int e = eb ? k : 0;
  const size_t SS = (size_t)S * S;
  const double tau = bls[(size_t)k * B + b] * rates[(size_t)(rb ? k : 0) * C + c];
  const int tid = threadIdx.x;
  if (tid < S) {
    const double lm = lam[(size_t)e * S + tid];
    ex[tid] = order ? lm * exp(lm * tau) : exp(lm * tau);
  }
  __syncthreads();
  const double* le = L + (size_t)e * SS;
  const double* re = R + (size_t)e * SS;
  for (int q = tid; q < SP * SP; q += 256) {
    const int i = q / SP, j = q % SP;
    const bool in = i < S && j < S;
    Ab[i * LD + j] = in ? le[i * S + j] * ex[j] : 0.0;
    Bb[i * LD + j] = in ? re[i * S + j] : 0.0;
  }
  __syncthreads();
  const int wave = tid >> 6, lane = tid & 63, nn = lane & 15, g4 = lane >> 4;
  O* o = out + mat * SS;
  for (int t = wave; t < MT * MT; t += 4) {
    const int ti = t / MT, tj = t % MT;
    d4 d = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
    for (int ks = 0; ks < SP / 4; ++ks) {
      const int m = 4 * ks + g4;
      d = __builtin_amdgcn_mfma_f64_16x16x4f64(Ab[(16 * ti + nn) * LD + m],
                                               Bb[m * LD + 16 * tj + nn], d, 0, 0, 0);
    }
    const int j = 16 * tj + nn;
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int i = 16 * ti + g4 + 4 * r;
      if (i < S && j < S) o[(size_t)i * S + j] = (O)d[r];
    }
  }
}

// One block per matrix for larger S: exponentials staged in LDS.
template <typename O>
__global__ void k_pmatrix_block(int S, int K, int B, int C, int eb, const double* __restrict__ lam,
                                const double* __restrict__ L, const double* __restrict__ R,
                                const double* __restrict__ bls, const double* __restrict__ rates,
                                int rb, O* __restrict__ out, int order) {
  __shared__ double ex[64];
  const size_t mat = blockIdx.x;
  const int c = (int)(mat % C);
  const size_t kb = mat / C;
  const int b = (int)(kb % B), k = (int)(kb / B);
  const int e = eb ? k : 0;
  const size_t SS = (size_t)S * S;
  const double tau = bls[(size_t)k * B + b] * rates[(size_t)(rb ? k : 0) * C + c];
  if (threadIdx.x < S) {
    const double lm = lam[(size_t)e * S + threadIdx.x];
    ex[threadIdx.x] = order ? lm * exp(lm * tau) : exp(lm * tau);
  }
  __syncthreads();
  const double* le = L + (size_t)e * SS;
  const double* re = R + (size_t)e * SS;
  for (int ij = threadIdx.x; ij < S * S; ij += blockDim.x) {
    const int i = ij / S, j = ij % S;
    double acc = 0.0;
    for (int m = 0; m < S; ++m) acc = fma(le[i * S + m] * ex[m], re[m * S + j], acc);
    out[mat * SS + ij] = (O)acc;
  }
}

// (exp(a) - exp(b)) / (a - b) = exp(m) * (1 - exp(-d)) / d with m = max(a, b), d = |a - b|
// (ea = exp(a), eb = exp(b)). Factoring out the larger exponent keeps every factor finite:
// exp(m) is the larger of the two, and (1 - exp(-d)) / d lies in (0, 1]. The form
// exp(b) * expm1(a - b) / (a - b) gives 0 * inf for b < -709 with a - b > 709.
__device__ __forceinline__ double exp_divdiff(double a, double b, double ea, double eb) {
  const double d = fabs(a - b);
  const double em = a > b ? ea : eb;
  if (d == 0.0) return em;
  return em * (-expm1(-d) / d);
}

// Small S: one thread per (k, b, c). Y = L^T G R^T, dtau = sum_m lam_m e^{lam_m tau} Y_mm,
// xpart[k][b][c] = tau * (Y o F) with F the divided differences of exp(lam * tau).
template <int S>
__global__ void k_pgrad_small(int K, int B, int C, int eb, const double* __restrict__ lam,
                              const double* __restrict__ L, const double* __restrict__ R,
                              const double* __restrict__ bls, const double* __restrict__ rates,
                              int rb, const double* __restrict__ dmats, double* __restrict__ dtau,
                              double* __restrict__ xpart) {
  constexpr int SS = S * S;
  const size_t idx = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= (size_t)K * B * C) return;
  const int c = (int)(idx % C);
  const size_t kb = idx / C;
  const int b = (int)(kb % B), k = (int)(kb / B);
  const int e = eb ? k : 0;
  const double tau = bls[(size_t)k * B + b] * rates[(size_t)(rb ? k : 0) * C + c];
  const double* le = L + (size_t)e * SS;
  const double* re = R + (size_t)e * SS;
  const double* la = lam + (size_t)e * S;
  double G[SS], tmp[SS];
#pragma unroll
  for (int q = 0; q < SS; ++q) G[q] = dmats[idx * SS + q];
#pragma unroll
  for (int i = 0; i < S; ++i)
#pragma unroll
    for (int n = 0; n < S; ++n) {
      double acc = 0.0;
#pragma unroll
      for (int j = 0; j < S; ++j) acc = fma(G[i * S + j], re[n * S + j], acc);
      tmp[i * S + n] = acc;
    }
  double a[S], ex[S];
#pragma unroll
  for (int m = 0; m < S; ++m) {
    a[m] = la[m] * tau;
    ex[m] = exp(a[m]);
  }
  double dt = 0.0;
#pragma unroll
  for (int m = 0; m < S; ++m)
#pragma unroll
    for (int n = 0; n < S; ++n) {
      double y = 0.0;
#pragma unroll
      for (int i = 0; i < S; ++i) y = fma(le[i * S + m], tmp[i * S + n], y);
      if (m == n) dt = fma(la[m] * ex[m], y, dt);
      xpart[idx * SS + m * S + n] = tau * y * exp_divdiff(a[m], a[n], ex[m], ex[n]);
    }
  dtau[idx] = dt;
}

// Any S <= 64: one block per (k, b, c), staged through LDS.
__global__ void k_pgrad_block(int S, int K, int B, int C, int eb, const double* __restrict__ lam,
                              const double* __restrict__ L, const double* __restrict__ R,
                              const double* __restrict__ bls, const double* __restrict__ rates,
                              int rb, const double* __restrict__ dmats, double* __restrict__ dtau,
                              double* __restrict__ xpart) {
  extern __shared__ double sm[];
  const int SS = S * S;
  double* G = sm;            // [S][S], reused for Y
  double* tmp = sm + SS;     // [S][S]
  double* ex = sm + 2 * SS;  // [S]
  const size_t idx = blockIdx.x;
  const int c = (int)(idx % C);
  const size_t kb = idx / C;
  const int b = (int)(kb % B), k = (int)(kb / B);
  const int e = eb ? k : 0;
  const double* le = L + (size_t)e * SS;
  const double* re = R + (size_t)e * SS;
  const double* la = lam + (size_t)e * S;
  const double tau = bls[(size_t)k * B + b] * rates[(size_t)(rb ? k : 0) * C + c];
  const double* g = dmats + idx * SS;
  for (int ij = threadIdx.x; ij < SS; ij += blockDim.x) G[ij] = g[ij];
  if (threadIdx.x < S) ex[threadIdx.x] = exp(la[threadIdx.x] * tau);
  __syncthreads();
  for (int in = threadIdx.x; in < SS; in += blockDim.x) {
    const int i = in / S, n = in % S;
    double acc = 0.0;
    for (int j = 0; j < S; ++j) acc = fma(G[i * S + j], re[n * S + j], acc);
    tmp[in] = acc;
  }
  __syncthreads();
  for (int mn = threadIdx.x; mn < SS; mn += blockDim.x) {
    const int m = mn / S, n = mn % S;
    double acc = 0.0;
    for (int i = 0; i < S; ++i) acc = fma(le[i * S + m], tmp[i * S + n], acc);
    G[mn] = acc;
  }
  __syncthreads();
  double* X = xpart + idx * SS;
  for (int mn = threadIdx.x; mn < SS; mn += blockDim.x) {
    const int m = mn / S, n = mn % S;
    X[mn] = tau * G[mn] * exp_divdiff(la[m] * tau, la[n] * tau, ex[m], ex[n]);
  }
  if (threadIdx.x == 0) {
    double t = 0.0;
    for (int m = 0; m < S; ++m) t = fma(la[m] * ex[m], G[m * S + m], t);
    dtau[idx] = t;
  }
}

// Dense spectra (17..64 states): the same chain rule with both products on the FP64 matrix
// cores. One block of 256 threads per (k, b, c); S is padded to SP = 16 MT with zeros and G, R,
// then L and the products are staged in LDS with row stride SP + 1. Products of 16x16 tiles
// (v_mfma_f64_16x16x4f64): A[i][k] from lane i + 16 k, B[k][j] from lane j + 16 k, D[i][j] in
// lane j + 16 (i % 4), register i / 4 (the layout hx_selftest_mfma_layout checks). The results
// equal k_pgrad_block's up to the order of the sums.
template <int MT>
__global__ __launch_bounds__(256) void k_pgrad_mfma(int S, int K, int B, int C, int eb,
                                                    const double* __restrict__ lam,
                                                    const double* __restrict__ L,
                                                    const double* __restrict__ R,
                                                    const double* __restrict__ bls,
                                                    const double* __restrict__ rates, int rb,
                                                    const double* __restrict__ dmats,
                                                    double* __restrict__ dtau,
                                                    double* __restrict__ xpart) {
  typedef double d4 __attribute__((ext_vector_type(4)));
  constexpr int SP = 16 * MT, LD = SP + 1;
  extern __shared__ double sm[];
  double* Ab = sm;                // [SP][LD]: G, then L
  double* Bb = sm + SP * LD;      // [SP][LD]: R, then Y
  double* Tb = sm + 2 * SP * LD;  // [SP][LD]: G R^T
  double* ex = sm + 3 * SP * LD;  // [SP]
  const int SS = S * S;
  const size_t idx = blockIdx.x;
  const int c = (int)(idx % C);
  const size_t kb = idx / C;
  const int b = (int)(kb % B), k = (int)(kb / B);
  const int e = eb ? k : 0;
  const double* le = L + (size_t)e * SS;
  const double* re = R + (size_t)e * SS;
  const double* la = lam + (size_t)e * S;
  const double tau = bls[(size_t)k * B + b] * rates[(size_t)(rb ? k : 0) * C + c];
  const double* g = dmats + idx * SS;
  const int tid = threadIdx.x;
  for (int q = tid; q < SP * SP; q += 256) {
    const int i = q / SP, j = q % SP;
    const bool in = i < S && j < S;
    Ab[i * LD + j] = in ? g[i * S + j] : 0.0;
    Bb[i * LD + j] = in ? re[i * S + j] : 0.0;
  }
  if (tid < S) ex[tid] = exp(la[tid] * tau);
  __syncthreads();
  const int wave = tid >> 6, lane = tid & 63, nn = lane & 15, g4 = lane >> 4;
  // T[i][n] = sum_j G[i][j] R[n][j]
  for (int t = wave; t < MT * MT; t += 4) {
    const int ti = t / MT, tn = t % MT;
    d4 d = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
    for (int ks = 0; ks < SP / 4; ++ks) {
      const int j = 4 * ks + g4;
      d = __builtin_amdgcn_mfma_f64_16x16x4f64(Ab[(16 * ti + nn) * LD + j],
                                               Bb[(16 * tn + nn) * LD + j], d, 0, 0, 0);
    }
#pragma unroll
    for (int r = 0; r < 4; ++r) Tb[(16 * ti + g4 + 4 * r) * LD + 16 * tn + nn] = d[r];
  }
  __syncthreads();
  for (int q = tid; q < SP * SP; q += 256) {
    const int i = q / SP, j = q % SP;
    Ab[i * LD + j] = (i < S && j < S) ? le[i * S + j] : 0.0;
  }
  __syncthreads();
  // Y[m][n] = sum_i L[i][m] T[i][n], into Bb (R is no longer read)
  for (int t = wave; t < MT * MT; t += 4) {
    const int tm = t / MT, tn = t % MT;
    d4 d = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
    for (int ks = 0; ks < SP / 4; ++ks) {
      const int i = 4 * ks + g4;
      d = __builtin_amdgcn_mfma_f64_16x16x4f64(Ab[i * LD + 16 * tm + nn],
                                               Tb[i * LD + 16 * tn + nn], d, 0, 0, 0);
    }
#pragma unroll
    for (int r = 0; r < 4; ++r) Bb[(16 * tm + g4 + 4 * r) * LD + 16 * tn + nn] = d[r];
  }
  __syncthreads();
  double* X = xpart + idx * SS;
  for (int mn = tid; mn < SS; mn += 256) {
    const int m = mn / S, n = mn % S;
    X[mn] = tau * Bb[m * LD + n] * exp_divdiff(la[m] * tau, la[n] * tau, ex[m], ex[n]);
  }
  if (tid == 0) {
    double t = 0.0;
    for (int m = 0; m < S; ++m) t = fma(la[m] * ex[m], Bb[m * LD + m], t);
    dtau[idx] = t;
  }
}

__device__ __forceinline__ double block_reduce_sum(double v, double* red) {
  red[threadIdx.x] = v;
  __syncthreads();
  for (int o = blockDim.x / 2; o > 0; o >>= 1) {
    if ((int)threadIdx.x < o) red[threadIdx.x] += red[threadIdx.x + o];
    __syncthreads();
  }
  return red[0];
}

__global__ void k_pgrad_bls(int K, int B, int C, const double* __restrict__ rates, int rb,
                            const double* __restrict__ dtau, double* __restrict__ dbls) {
  const size_t idx = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= (size_t)K * B) return;
  const int k = (int)(idx / B);
  double t = 0.0;
  for (int c = 0; c < C; ++c) t += rates[(size_t)(rb ? k : 0) * C + c] * dtau[idx * C + c];
  dbls[idx] = t;
}

// Block per (k, c): sum over branches of bls * dtau.
__global__ void k_pgrad_rates(int K, int B, int C, const double* __restrict__ bls,
                              const double* __restrict__ dtau, double* __restrict__ drates) {
  __shared__ double red[256];
  const int k = blockIdx.x / C, c = blockIdx.x % C;
  double t = 0.0;
  for (int b = threadIdx.x; b < B; b += blockDim.x)
    t += bls[(size_t)k * B + b] * dtau[((size_t)k * B + b) * C + c];
  t = block_reduce_sum(t, red);
  if (threadIdx.x == 0) drates[blockIdx.x] = t;
}

// X[e][mn] = sum over the samples of e, branches and categories of xpart[row][mn], with a
// fixed order. For S * S dividing 256 (four states) and room for at least 8 parts: two passes.
// Pass 1, grid (NB, E): block b sums its share of e's rows into part[e][b][mn]; its 256 threads
// read consecutive elements of the row-major [rows][SS] array, thread t keeping column t % SS
// of every (256 / SS)-th row from row t / SS, and the threads of one column are then added in
// LDS. Pass 2 adds the NB parts in order. Otherwise (dense states: S * S = 400 or 3721 entries,
// enough blocks by themselves) one block per (e, mn), k_pgrad_xsum.
__global__ void k_pgrad_xsum_part(int SS, int K, int B, int C, int eb, int NB,
                                  const double* __restrict__ xpart, double* __restrict__ part) {
  __shared__ double red[256];
  const int e = blockIdx.y, b = blockIdx.x;
  const size_t per = (size_t)B * C;
  const size_t r0 = eb ? (size_t)e * per : 0;
  const size_t rows = eb ? per : (size_t)K * per;
  const size_t rpb = (rows + NB - 1) / NB;
  const size_t lo = r0 + (size_t)b * rpb, hi = r0 + min(rows, (size_t)(b + 1) * rpb);
  double* out = part + ((size_t)e * NB + b) * SS;
  const int t = threadIdx.x;
  const int col = t % SS, step = 256 / SS;
  double acc = 0.0;
  for (size_t r = lo + t / SS; r < hi; r += step) acc += xpart[r * SS + col];
  red[t] = acc;
  __syncthreads();
  if (t < SS) {
    double v = 0.0;
    for (int q = 0; q < step; ++q) v += red[q * SS + t];
    out[t] = v;
  }
}

__global__ void k_pgrad_xsum(int SS, int K, int B, int C, int eb, const double* __restrict__ xpart,
                             double* __restrict__ X) {
  __shared__ double red[256];
  const int e = blockIdx.x / SS, mn = blockIdx.x % SS;
  const size_t per = (size_t)B * C;
  const size_t r0 = eb ? (size_t)e * per : 0;
  const size_t r1 = eb ? r0 + per : (size_t)K * per;
  double t = 0.0;
  for (size_t r = r0 + threadIdx.x; r < r1; r += blockDim.x) t += xpart[r * SS + mn];
  t = block_reduce_sum(t, red);
  if (threadIdx.x == 0) X[blockIdx.x] = t;
}

__global__ void k_pgrad_xsum_fin(int SS, int E, int NB, const double* __restrict__ part,
                                 double* __restrict__ X) {
  const size_t idx = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= (size_t)E * SS) return;
  const size_t e = idx / SS, mn = idx % SS;
  double v = 0.0;
  for (int b = 0; b < NB; ++b) v += part[(e * NB + b) * SS + mn];
  X[idx] = v;
}

// Per eigensystem e: dQ = R^T X L^T.
__global__ void k_pgrad_q(int S, const double* __restrict__ L, const double* __restrict__ R,
                          const double* __restrict__ Xall, double* __restrict__ dq) {
  extern __shared__ double sm[];
  const int SS = S * S;
  double* X = sm;
  double* tmp = sm + SS;
  const int e = blockIdx.x;
  for (int mn = threadIdx.x; mn < SS; mn += blockDim.x) X[mn] = Xall[(size_t)e * SS + mn];
  __syncthreads();
  const double* le = L + (size_t)e * SS;
  const double* re = R + (size_t)e * SS;
  for (int mj = threadIdx.x; mj < SS; mj += blockDim.x) {
    const int m = mj / S, j = mj % S;
    double acc = 0.0;
    for (int n = 0; n < S; ++n) acc = fma(X[m * S + n], le[j * S + n], acc);
    tmp[mj] = acc;
  }
  __syncthreads();
  for (int ij = threadIdx.x; ij < SS; ij += blockDim.x) {
    const int i = ij / S, j = ij % S;
    double acc = 0.0;
    for (int m = 0; m < S; ++m) acc = fma(re[m * S + i], tmp[m * S + j], acc);
    dq[(size_t)e * SS + ij] = acc;
  }
}

// dQ = R^T X L^T for 17..64 states on the FP64 matrix cores: one block of 256 threads per
// eigensystem, S padded to SP = 16 MT, X and L, then R and X L^T staged in LDS (row stride
// SP + 1); the same products as k_pgrad_q up to the order of the sums.
template <int MT>
__global__ __launch_bounds__(256) void k_pgrad_q_mfma(int S, const double* __restrict__ L,
                                                      const double* __restrict__ R,
                                                      const double* __restrict__ Xall,
                                                      double* __restrict__ dq) {
  typedef double d4 __attribute__((ext_vector_type(4)));
  constexpr int SP = 16 * MT, LD = SP + 1;
  extern __shared__ double sm[];
  double* Ab = sm;                // [SP][LD]: X, then R
  double* Bb = sm + SP * LD;      // [SP][LD]: L
  double* Tb = sm + 2 * SP * LD;  // [SP][LD]: X L^T
  const int SS = S * S;
  const int e = blockIdx.x, tid = threadIdx.x;
  const double* xe = Xall + (size_t)e * SS;
  const double* le = L + (size_t)e * SS;
  const double* re = R + (size_t)e * SS;
  for (int q = tid; q < SP * SP; q += 256) {
    const int i = q / SP, j = q % SP;
    const bool in = i < S && j < S;
    Ab[i * LD + j] = in ? xe[i * S + j] : 0.0;
    Bb[i * LD + j] = in ? le[i * S + j] : 0.0;
  }
  __syncthreads();
  const int wave = tid >> 6, lane = tid & 63, nn = lane & 15, g4 = lane >> 4;
  // T[m][j] = sum_n X[m][n] L[j][n]
  for (int t = wave; t < MT * MT; t += 4) {
    const int tm = t / MT, tj = t % MT;
    d4 d = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
    for (int ks = 0; ks < SP / 4; ++ks) {
      const int n = 4 * ks + g4;
      d = __builtin_amdgcn_mfma_f64_16x16x4f64(Ab[(16 * tm + nn) * LD + n],
                                               Bb[(16 * tj + nn) * LD + n], d, 0, 0, 0);
    }
#pragma unroll
    for (int r = 0; r < 4; ++r) Tb[(16 * tm + g4 + 4 * r) * LD + 16 * tj + nn] = d[r];
  }
  __syncthreads();
  for (int q = tid; q < SP * SP; q += 256) {
    const int i = q / SP, j = q % SP;
    Ab[i * LD + j] = (i < S && j < S) ? re[i * S + j] : 0.0;
  }
  __syncthreads();
  // dQ[i][j] = sum_m R[m][i] T[m][j]
  double* o = dq + (size_t)e * SS;
  for (int t = wave; t < MT * MT; t += 4) {
    const int ti = t / MT, tj = t % MT;
    d4 d = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
    for (int ks = 0; ks < SP / 4; ++ks) {
      const int m = 4 * ks + g4;
      d = __builtin_amdgcn_mfma_f64_16x16x4f64(Ab[m * LD + 16 * ti + nn],
                                               Tb[m * LD + 16 * tj + nn], d, 0, 0, 0);
    }
    const int j = 16 * tj + nn;
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int i = 16 * ti + g4 + 4 * r;
      if (i < S && j < S) o[i * S + j] = d[r];
    }
  }
}

}  // namespace hx

extern "C" {

static int pmatrix_impl(int S, int K, int B, int C, int eigen_batched, const double* lambda,
                        const double* left, const double* right, const double* bls,
                        const double* rates, int rates_batched, void* mats, int dtype,
                        hx_stream_t stream, int order) {
  HX_REQUIRE(S >= 2 && S <= 64 && K >= 1 && B >= 1 && C >= 1, HX_ERR_INVALID_ARGUMENT,
             "bad dimensions (S must be 2..64)");
  HX_REQUIRE(lambda && left && right && bls && rates && mats, HX_ERR_INVALID_ARGUMENT,
             "null pointer argument");
  HX_REQUIRE(hx::dtype_ok(dtype), HX_ERR_INVALID_ARGUMENT, "bad dtype");
  hipStream_t st = reinterpret_cast<hipStream_t>(stream);
  const size_t nmat = (size_t)K * B * C;
  if (S <= 8) {
    const size_t total = nmat * S * S;
    const unsigned grid = (unsigned)((total + 255) / 256);
    if (dtype == HX_F64)
      hx::k_pmatrix_elem<double><<<grid, 256, 0, st>>>(S, K, B, C, eigen_batched, lambda, left,
                                                        right, bls, rates, rates_batched,
                                                        static_cast<double*>(mats), order);
    else
      hx::k_pmatrix_elem<float><<<grid, 256, 0, st>>>(S, K, B, C, eigen_batched, lambda, left,
                                                       right, bls, rates, rates_batched,
                                                       static_cast<float*>(mats), order);
  } else if (S > 16) {
    HX_REQUIRE(nmat < (1u << 31), HX_ERR_UNSUPPORTED, "too many matrices");
#define HX_PM(OT, MTV)                                                                       \
  do {                                                                                       \
    const size_t smm = (2 * (size_t)(16 * MTV) * (16 * MTV + 1) + 16 * MTV) * sizeof(double); \
    if (smm > 64 * 1024)                                                                     \
      HX_HIP_TRY(hipFuncSetAttribute(reinterpret_cast<const void*>(hx::k_pmatrix_mfma<OT, MTV>), \
                                     hipFuncAttributeMaxDynamicSharedMemorySize, (int)smm)); \
    hx::k_pmatrix_mfma<OT, MTV><<<(unsigned)nmat, 256, smm, st>>>(                           \
        S, K, B, C, eigen_batched, lambda, left, right, bls, rates, rates_batched,           \
        static_cast<OT*>(mats), order);                                                      \
  } while (0)
    if (dtype == HX_F64) {
      if (S <= 32) HX_PM(double, 2);
      else HX_PM(double, 4);
    } else {
      if (S <= 32) HX_PM(float, 2);
      else HX_PM(float, 4);
    }
#undef HX_PM
  } else {
    HX_REQUIRE(nmat < (1u << 31), HX_ERR_UNSUPPORTED, "too many matrices");
    if (dtype == HX_F64)
      hx::k_pmatrix_block<double><<<(unsigned)nmat, 256, 0, st>>>(
          S, K, B, C, eigen_batched, lambda, left, right, bls, rates, rates_batched,
          static_cast<double*>(mats), order);
    else
      hx::k_pmatrix_block<float><<<(unsigned)nmat, 256, 0, st>>>(
          S, K, B, C, eigen_batched, lambda, left, right, bls, rates, rates_batched,
          static_cast<float*>(mats), order);
  }
  HX_HIP_TRY(hipGetLastError());
  return HX_OK;
}

int hx_pmatrix(int S, int K, int B, int C, int eigen_batched, const double* lambda,
               const double* left, const double* right, const double* bls, const double* rates,
               int rates_batched, void* mats, int dtype, hx_stream_t stream) {
  return pmatrix_impl(S, K, B, C, eigen_batched, lambda, left, right, bls, rates, rates_batched,
                      mats, dtype, stream, 0);
}

int hx_pmatrix_dtau(int S, int K, int B, int C, int eigen_batched, const double* lambda,
                    const double* left, const double* right, const double* bls,
                    const double* rates, int rates_batched, void* mats_dtau, int dtype,
                    hx_stream_t stream) {
  return pmatrix_impl(S, K, B, C, eigen_batched, lambda, left, right, bls, rates, rates_batched,
                      mats_dtau, dtype, stream, 1);
}

static size_t pgrad_layout(int S, int K, int B, int C, size_t* o_xpart, size_t* o_x) {
  const size_t n = (size_t)K * B * C;
  size_t off = hx::align_up(n * 8, 256);
  *o_xpart = off;
  off += hx::align_up(n * (size_t)S * S * 8, 256);
  *o_x = off;
  off += hx::align_up((size_t)K * S * S * 8, 256);
  return off;
}

size_t hx_pmatrix_grad_scratch_bytes(int S, int K, int B, int C) {
  size_t a, b;
  return pgrad_layout(S, K, B, C, &a, &b);
}

int hx_pmatrix_grad(int S, int K, int B, int C, int eigen_batched, const double* lambda,
                    const double* left, const double* right, const double* bls, const double* rates,
                    int rates_batched, const double* dmats, double* out_dbls, double* out_drates,
                    double* out_dq, void* scratch, hx_stream_t stream) {
  HX_REQUIRE(S >= 2 && S <= 64 && K >= 1 && B >= 1 && C >= 1, HX_ERR_INVALID_ARGUMENT,
             "bad dimensions");
  HX_REQUIRE(lambda && left && right && bls && rates && dmats && scratch, HX_ERR_INVALID_ARGUMENT,
             "null pointer argument");
  hipStream_t st = reinterpret_cast<hipStream_t>(stream);
  size_t o_xpart, o_x;
  pgrad_layout(S, K, B, C, &o_xpart, &o_x);
  char* base = static_cast<char*>(scratch);
  double* dtau = reinterpret_cast<double*>(base);
  double* xpart = reinterpret_cast<double*>(base + o_xpart);
  double* X = reinterpret_cast<double*>(base + o_x);
  const size_t n = (size_t)K * B * C;
  switch (S) {
    case 4:
      hx::k_pgrad_small<4><<<(unsigned)((n + 127) / 128), 128, 0, st>>>(
          K, B, C, eigen_batched, lambda, left, right, bls, rates, rates_batched, dmats, dtau, xpart);
      break;
    default: {
      HX_REQUIRE(n < (1u << 31), HX_ERR_UNSUPPORTED, "too many matrices");
      if (S > 16) {
        // dense spectra: both products on the FP64 matrix cores (128 taxa, 10k patterns, 61
        // states: the chain-rule phase 0.38 -> 0.22 ms per step)
#define HX_PG(MTV)                                                                           \
  do {                                                                                       \
    const size_t smm = (3 * (size_t)(16 * MTV) * (16 * MTV + 1) + 16 * MTV) * sizeof(double); \
    if (smm > 64 * 1024)                                                                     \
      HX_HIP_TRY(hipFuncSetAttribute(reinterpret_cast<const void*>(hx::k_pgrad_mfma<MTV>),   \
                                     hipFuncAttributeMaxDynamicSharedMemorySize, (int)smm)); \
    hx::k_pgrad_mfma<MTV><<<(unsigned)n, 256, smm, st>>>(S, K, B, C, eigen_batched, lambda,  \
                                                         left, right, bls, rates,            \
                                                         rates_batched, dmats, dtau, xpart); \
  } while (0)
        if (S <= 32) HX_PG(2);
        else HX_PG(4);
#undef HX_PG
      } else {
        const size_t sm1 = (2 * (size_t)S * S + S + 1) * sizeof(double);
        hx::k_pgrad_block<<<(unsigned)n, 256, sm1, st>>>(S, K, B, C, eigen_batched, lambda, left,
                                                         right, bls, rates, rates_batched, dmats,
                                                         dtau, xpart);
      }
    }
  }
  HX_HIP_TRY(hipGetLastError());
  if (out_dbls)
    hx::k_pgrad_bls<<<(unsigned)((K * (size_t)B + 255) / 256), 256, 0, st>>>(K, B, C, rates,
                                                                              rates_batched, dtau,
                                                                              out_dbls);
  if (out_drates)
    hx::k_pgrad_rates<<<(unsigned)(K * C), 256, 0, st>>>(K, B, C, bls, dtau, out_drates);
  if (out_dq) {
    // the two-pass parts go into dtau's room (n doubles), which the launches above have consumed
    const int E = eigen_batched ? K : 1;
    const int SS = S * S;
    const size_t cap = n / ((size_t)E * SS);
    if (256 % SS == 0 && cap >= 8) {
      const int NB = (int)(cap < 64 ? cap : 64);
      hx::k_pgrad_xsum_part<<<dim3((unsigned)NB, (unsigned)E), 256, 0, st>>>(
          SS, K, B, C, eigen_batched, NB, xpart, dtau);
      hx::k_pgrad_xsum_fin<<<(unsigned)((E * (size_t)SS + 255) / 256), 256, 0, st>>>(SS, E, NB,
                                                                                     dtau, X);
    } else {
      hx::k_pgrad_xsum<<<(unsigned)(E * SS), 256, 0, st>>>(SS, K, B, C, eigen_batched, xpart, X);
    }
    if (S > 16) {
#define HX_PQ(MTV)                                                                           \
  do {                                                                                       \
    const size_t smm = 3 * (size_t)(16 * MTV) * (16 * MTV + 1) * sizeof(double);            \
    if (smm > 64 * 1024)                                                                     \
      HX_HIP_TRY(hipFuncSetAttribute(reinterpret_cast<const void*>(hx::k_pgrad_q_mfma<MTV>), \
                                     hipFuncAttributeMaxDynamicSharedMemorySize, (int)smm)); \
    hx::k_pgrad_q_mfma<MTV><<<E, 256, smm, st>>>(S, left, right, X, out_dq);                 \
  } while (0)
      if (S <= 32) HX_PQ(2);
      else HX_PQ(4);
#undef HX_PQ
    } else {
      const size_t sm2 = 2 * (size_t)S * S * sizeof(double);
      hx::k_pgrad_q<<<E, 256, sm2, st>>>(S, left, right, X, out_dq);
    }
  }
  HX_HIP_TRY(hipGetLastError());
  return HX_OK;
}

}  // extern "C"

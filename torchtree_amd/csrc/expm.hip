// Transition matrices of non-reversible models: P = exp(Q tau) for a general rate matrix.
//
// Reference: NonSymmetricSubstitutionModel.p_t (torchtree/evolution/substitution_model/
// abstract.py:88-94), which calls torch.matrix_exp on Q tau for every branch, category and
// sample, and its autograd backward.
//
// Method: scaling and squaring of the truncated Taylor series (Moler & Van Loan 1978; the
// Taylor form analysed by Al-Mohy & Higham 2009). X = A / 2^s with s the smallest power that
// brings the 1-norm of A to at most 1/2, T = sum_{j<=18} X^j / j! by Horner's rule
// (T <- I + X T / j for j = 18..1), then s squarings. With ||X||_1 <= 1/2 the truncation error
// is below ||X||^19 / 19! < 1e-22, under FP64 rounding. Everything runs in FP64; P is written
// in the call's precision.
//
// Backward: for f = <G, exp(tau Q)>, d f / d tau = <G, Q P> and d f / d Q = tau L(tau Q^T, G),
// where L(A, E) is the Frechet derivative of exp at A in the direction E, read from the upper
// right block of exp([[A, E], [0, A]]) (Mathias 1996). One block per (sample, branch,
// category) matrix, n x n tiles in LDS (n = S, or 2S for the block matrix), S <= 32.
#include "common.hpp"

#include <math.h>

namespace hx {
namespace {

constexpr int kExpmMaxS = 32;
constexpr int kTaylor = 18;

// C = A B for n x n matrices in LDS (row-major), all threads of the block.
__device__ void mat_mul(const double* A, const double* B, double* C, int n) {
  for (int e = threadIdx.x; e < n * n; e += blockDim.x) {
    const int i = e / n, j = e - i * n;
    double acc = 0.0;
    for (int k = 0; k < n; ++k) acc = fma(A[i * n + k], B[k * n + j], acc);
    C[e] = acc;
  }
}

// exp(M) for the n x n matrix in M (LDS); the result is left in M. W1, W2: n x n scratch.
// `nrm` is the 1-norm of M (computed by the caller, block-uniform).
__device__ void expm_inplace(double* M, double* W1, double* W2, int n, double nrm) {
  int s = 0;
  if (nrm > 0.5 && isfinite(nrm)) s = (int)fmin(ceil(log2(nrm / 0.5)), 1100.0);
  const double scale = ldexp(1.0, -s);
  for (int e = threadIdx.x; e < n * n; e += blockDim.x) {
    M[e] *= scale;
    W1[e] = (e / n == e % n) ? 1.0 : 0.0;  // T = I
  }
  __syncthreads();
  for (int j = kTaylor; j >= 1; --j) {
    mat_mul(M, W1, W2, n);  // X T
    __syncthreads();
    const double inv = 1.0 / (double)j;
    for (int e = threadIdx.x; e < n * n; e += blockDim.x)
      W1[e] = W2[e] * inv + ((e / n == e % n) ? 1.0 : 0.0);
    __syncthreads();
  }
  for (int q = 0; q < s; ++q) {
    mat_mul(W1, W1, W2, n);
    __syncthreads();
    for (int e = threadIdx.x; e < n * n; e += blockDim.x) W1[e] = W2[e];
    __syncthreads();
  }
  for (int e = threadIdx.x; e < n * n; e += blockDim.x) M[e] = W1[e];
  __syncthreads();
}

// 1-norm (maximum column sum of magnitudes) of an n x n matrix in LDS, into *out (block-wide).
__device__ void norm1(const double* M, int n, double* red, double* out) {
  for (int j = threadIdx.x; j < n; j += blockDim.x) {
    double t = 0.0;
    for (int i = 0; i < n; ++i) t += fabs(M[i * n + j]);
    red[j] = t;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    double m = 0.0;
    for (int j = 0; j < n; ++j) m = fmax(m, red[j]);
    *out = m;
  }
  __syncthreads();
}

// mats[k][b][c] = exp(Q_e tau) (order 0) or Q_e exp(Q_e tau) (order 1), tau = bls[k][b] rates[.][c].
template <typename O>
__global__ __launch_bounds__(256) void k_expm(int S, int K, int B, int C, int qb,
                                              const double* __restrict__ Q,
                                              const double* __restrict__ bls,
                                              const double* __restrict__ rates, int rb, O* out,
                                              int order) {
  extern __shared__ double sm[];
  const int SS = S * S;
  double* M = sm;
  double* W1 = M + SS;
  double* W2 = W1 + SS;
  double* red = W2 + SS;  // [S] + 1
  const size_t idx = blockIdx.x;
  const int c = (int)(idx % C);
  const size_t kb = idx / C;
  const int b = (int)(kb % B), k = (int)(kb / B);
  const double* q = Q + (size_t)(qb ? k : 0) * SS;
  const double tau = bls[(size_t)k * B + b] * rates[(size_t)(rb ? k : 0) * C + c];
  for (int e = threadIdx.x; e < SS; e += blockDim.x) M[e] = q[e] * tau;
  __syncthreads();
  norm1(M, S, red, red + S);
  expm_inplace(M, W1, W2, S, red[S]);
  O* o = out + idx * SS;
  if (order == 0) {
    for (int e = threadIdx.x; e < SS; e += blockDim.x) o[e] = (O)M[e];
  } else {
    for (int e = threadIdx.x; e < SS; e += blockDim.x) {
      const int i = e / S, j = e - i * S;
      double acc = 0.0;
      for (int m = 0; m < S; ++m) acc = fma(q[i * S + m], M[m * S + j], acc);
      o[e] = (O)acc;
    }
  }
}

// Per matrix (k, b, c) with G = dmats[k][b][c]: dtau = <G, Q P> and
// xpart = tau L(tau Q^T, G), from exp([[tau Q^T, G], [0, tau Q^T]]) of size 2S.
__global__ __launch_bounds__(256) void k_expm_grad(int S, int K, int B, int C, int qb,
                                                   const double* __restrict__ Q,
                                                   const double* __restrict__ bls,
                                                   const double* __restrict__ rates, int rb,
                                                   const double* __restrict__ dmats,
                                                   double* __restrict__ dtau,
                                                   double* __restrict__ xpart) {
  extern __shared__ double sm[];
  const int n = 2 * S, NN = n * n, SS = S * S;
  double* M = sm;
  double* W1 = M + NN;
  double* W2 = W1 + NN;
  double* red = W2 + NN;  // [n] + 1, then [blockDim] partial sums
  const size_t idx = blockIdx.x;
  const int c = (int)(idx % C);
  const size_t kb = idx / C;
  const int b = (int)(kb % B), k = (int)(kb / B);
  const double* q = Q + (size_t)(qb ? k : 0) * SS;
  const double tau = bls[(size_t)k * B + b] * rates[(size_t)(rb ? k : 0) * C + c];
  const double* g = dmats + idx * SS;
  for (int e = threadIdx.x; e < NN; e += blockDim.x) {
    const int i = e / n, j = e - i * n;
    double v = 0.0;
    if (i < S && j < S) v = q[j * S + i] * tau;                       // tau Q^T
    else if (i >= S && j >= S) v = q[(j - S) * S + (i - S)] * tau;    // tau Q^T
    else if (i < S && j >= S) v = g[i * S + (j - S)];                 // G
    M[e] = v;
  }
  __syncthreads();
  norm1(M, n, red, red + n);
  expm_inplace(M, W1, W2, n, red[n]);
  // upper-left block: exp(tau Q^T) = P^T; upper-right: L(tau Q^T, G)
  double part = 0.0;
  for (int e = threadIdx.x; e < SS; e += blockDim.x) {
    const int i = e / S, j = e - i * S;
    double qp = 0.0;  // (Q P)[i][j] = sum_m Q[i][m] P[m][j] = sum_m Q[i][m] P^T[j][m]
    for (int m = 0; m < S; ++m) qp = fma(q[i * S + m], M[j * n + m], qp);
    part = fma(g[e], qp, part);
    xpart[idx * SS + e] = tau * M[i * n + S + j];
  }
  double* ps = red + n + 1;
  ps[threadIdx.x] = part;
  __syncthreads();
  if (threadIdx.x == 0) {
    double t = 0.0;
    for (int w = 0; w < (int)blockDim.x; ++w) t += ps[w];
    dtau[idx] = t;
  }
}

// dQ[e][i][j] = sum over the samples of eigensystem e, branches and categories of xpart, in
// a fixed order.
__global__ void k_expm_qsum(int SS, int K, int B, int C, int qb, const double* __restrict__ xpart,
                            double* __restrict__ dq) {
  const int e = blockIdx.x / SS, ij = blockIdx.x - e * SS;
  const int k0 = qb ? e : 0, k1 = qb ? e + 1 : K;
  double acc = 0.0;
  for (int k = k0; k < k1; ++k)
    for (int bc = threadIdx.x; bc < B * C; bc += blockDim.x)
      acc += xpart[((size_t)k * B * C + bc) * SS + ij];
  __shared__ double red[256];
  red[threadIdx.x] = acc;
  __syncthreads();
  for (int o = blockDim.x / 2; o > 0; o >>= 1) {
    if ((int)threadIdx.x < o) red[threadIdx.x] += red[threadIdx.x + o];
    __syncthreads();
  }
  if (threadIdx.x == 0) dq[blockIdx.x] = red[0];
}

__global__ void k_expm_bls(int K, int B, int C, const double* __restrict__ rates, int rb,
                           const double* __restrict__ dtau, double* __restrict__ dbls) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= (size_t)K * B) return;
  const int k = (int)(i / B);
  double t = 0.0;
  for (int c = 0; c < C; ++c) t = fma(rates[(size_t)(rb ? k : 0) * C + c], dtau[i * C + c], t);
  dbls[i] = t;
}

__global__ void k_expm_rates(int K, int B, int C, const double* __restrict__ bls,
                             const double* __restrict__ dtau, double* __restrict__ drates) {
  const int k = blockIdx.x / C, c = blockIdx.x - k * C;
  double acc = 0.0;
  for (int b = threadIdx.x; b < B; b += blockDim.x)
    acc = fma(bls[(size_t)k * B + b], dtau[((size_t)k * B + b) * C + c], acc);
  __shared__ double red[256];
  red[threadIdx.x] = acc;
  __syncthreads();
  for (int o = blockDim.x / 2; o > 0; o >>= 1) {
    if ((int)threadIdx.x < o) red[threadIdx.x] += red[threadIdx.x + o];
    __syncthreads();
  }
  if (threadIdx.x == 0) drates[blockIdx.x] = red[0];
}

int expm_impl(int S, int K, int B, int C, int qb, const double* Q, const double* bls,
              const double* rates, int rb, void* mats, int dtype, hx_stream_t stream, int order) {
  HX_REQUIRE(S >= 2 && S <= kExpmMaxS && K >= 1 && B >= 1 && C >= 1, HX_ERR_INVALID_ARGUMENT,
             "bad dimensions (S must be 2..32 for the matrix exponential)");
  HX_REQUIRE(Q && bls && rates && mats, HX_ERR_INVALID_ARGUMENT, "null pointer argument");
  HX_REQUIRE(dtype_ok(dtype), HX_ERR_INVALID_ARGUMENT, "bad dtype");
  const size_t nmat = (size_t)K * B * C;
  HX_REQUIRE(nmat < (1u << 31), HX_ERR_UNSUPPORTED, "too many matrices");
  hipStream_t st = reinterpret_cast<hipStream_t>(stream);
  const size_t lds = (3 * (size_t)S * S + S + 1) * sizeof(double);
  const unsigned threads = S <= 8 ? 64 : 256;
  if (dtype == HX_F64)
    k_expm<double><<<(unsigned)nmat, threads, lds, st>>>(S, K, B, C, qb, Q, bls, rates, rb,
                                                          static_cast<double*>(mats), order);
  else
    k_expm<float><<<(unsigned)nmat, threads, lds, st>>>(S, K, B, C, qb, Q, bls, rates, rb,
                                                         static_cast<float*>(mats), order);
  HX_HIP_TRY(hipGetLastError());
  return HX_OK;
}

size_t expm_grad_layout(int S, int K, int B, int C, size_t* o_xpart) {
  const size_t n = (size_t)K * B * C;
  size_t off = align_up(n * 8, 256);  // dtau
  *o_xpart = off;
  off += align_up(n * (size_t)S * S * 8, 256);
  return off;
}

}  // namespace
}  // namespace hx

extern "C" {

int hx_pmatrix_expm(int S, int K, int B, int C, int q_batched, const double* Q, const double* bls,
                    const double* rates, int rates_batched, void* mats, int dtype,
                    hx_stream_t stream) {
  return hx::expm_impl(S, K, B, C, q_batched, Q, bls, rates, rates_batched, mats, dtype, stream,
                       0);
}

int hx_pmatrix_expm_dtau(int S, int K, int B, int C, int q_batched, const double* Q,
                         const double* bls, const double* rates, int rates_batched,
                         void* mats_dtau, int dtype, hx_stream_t stream) {
  return hx::expm_impl(S, K, B, C, q_batched, Q, bls, rates, rates_batched, mats_dtau, dtype,
                       stream, 1);
}

size_t hx_pmatrix_expm_grad_scratch_bytes(int S, int K, int B, int C) {
  size_t a;
  return hx::expm_grad_layout(S, K, B, C, &a);
}

int hx_pmatrix_expm_grad(int S, int K, int B, int C, int q_batched, const double* Q,
                         const double* bls, const double* rates, int rates_batched,
                         const double* dmats, double* out_dbls, double* out_drates, double* out_dq,
                         void* scratch, hx_stream_t stream) {
  HX_REQUIRE(S >= 2 && S <= hx::kExpmMaxS && K >= 1 && B >= 1 && C >= 1, HX_ERR_INVALID_ARGUMENT,
             "bad dimensions (S must be 2..32 for the matrix exponential)");
  HX_REQUIRE(Q && bls && rates && dmats && scratch, HX_ERR_INVALID_ARGUMENT,
             "null pointer argument");
  const size_t nmat = (size_t)K * B * C;
  HX_REQUIRE(nmat < (1u << 31), HX_ERR_UNSUPPORTED, "too many matrices");
  hipStream_t st = reinterpret_cast<hipStream_t>(stream);
  size_t o_xpart;
  hx::expm_grad_layout(S, K, B, C, &o_xpart);
  char* base = static_cast<char*>(scratch);
  double* dtau = reinterpret_cast<double*>(base);
  double* xpart = reinterpret_cast<double*>(base + o_xpart);
  const int n = 2 * S;
  const unsigned threads = 256;
  const size_t lds = (3 * (size_t)n * n + n + 1 + threads) * sizeof(double);
  if (lds > 64 * 1024)
    HX_HIP_TRY(hipFuncSetAttribute(reinterpret_cast<const void*>(hx::k_expm_grad),
                                   hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
  hx::k_expm_grad<<<(unsigned)nmat, threads, lds, st>>>(S, K, B, C, q_batched, Q, bls, rates,
                                                        rates_batched, dmats, dtau, xpart);
  HX_HIP_TRY(hipGetLastError());
  if (out_dbls)
    hx::k_expm_bls<<<(unsigned)((K * (size_t)B + 255) / 256), 256, 0, st>>>(
        K, B, C, rates, rates_batched, dtau, out_dbls);
  if (out_drates) hx::k_expm_rates<<<(unsigned)(K * C), 256, 0, st>>>(K, B, C, bls, dtau, out_drates);
  if (out_dq) {
    const int E = q_batched ? K : 1;
    hx::k_expm_qsum<<<(unsigned)(E * S * S), 256, 0, st>>>(S * S, K, B, C, q_batched, xpart,
                                                           out_dq);
  }
  HX_HIP_TRY(hipGetLastError());
  return HX_OK;
}

}  // extern "C"

// Resident likelihood jobs for the serving loop.
//
// A job keeps one alignment's tip codes (cut into pattern slices), weights,
// topology plan, FP64 eigensystem, category rates/proportions and
// frequencies in device memory. hx_job_eval runs one loglik (+ branch-length
// gradient) request of K samples on a caller-given stream and workspace:
//   P matrices once, then for each pattern slice post-order (+ pre-order),
//   accumulating log-likelihood and d loglik / d mats over slices in a fixed
//   order, then the chain rule to d loglik / d branch lengths.
// Pattern slices bound the workspace whatever N is; the sums are exact
// because patterns are independent (TreeLikelihoodModel._call,
// tree_likelihood.py:313-356, evaluates all N at once).
#include "common.hpp"

#include <algorithm>
#include <mutex>
#include <string>
#include <vector>

struct hx_job_s {
  hx_plan_t plan = nullptr;
  int device = 0;
  int S = 0, N = 0, C = 0, dtype = HX_F32, T = 0, B = 0;
  int slice = 0, nslices = 0;
  int ncodes = 0;
  std::vector<int> slice_n;        // patterns per slice
  void* d_buf = nullptr;           // one device allocation for everything resident
  std::vector<uint8_t*> d_codes;   // per slice [T][n]
  std::vector<double*> d_weights;  // per slice [n]
  void* d_table = nullptr;         // [ncodes][S] in dtype
  double* d_lambda = nullptr;      // [S]
  double* d_left = nullptr;        // [S][S]
  double* d_right = nullptr;       // [S][S]
  bool general = false;            // a general rate matrix (d_left holds Q) instead of an eigensystem
  double* d_rates = nullptr;       // [C]
  void* d_props = nullptr;         // [C] in dtype
  void* d_freqs = nullptr;         // [S] in dtype
  size_t resident = 0;
};

namespace hx {

__global__ void k_fill0(double* p, size_t n) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0.0;
}

__global__ void k_accum(double* __restrict__ dst, const double* __restrict__ src, size_t n) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) dst[i] += src[i];
}

static void fill0(double* p, size_t n, hipStream_t st) {
  if (n) k_fill0<<<(unsigned)((n + 255) / 256), 256, 0, st>>>(p, n);
}

static void accum(double* dst, const double* src, size_t n, hipStream_t st) {
  if (n) k_accum<<<(unsigned)((n + 255) / 256), 256, 0, st>>>(dst, src, n);
}

__global__ void k_dbls(int K, int B, int C, const double* __restrict__ rates,
                       const double* __restrict__ dtau, double* __restrict__ dbls) {
  const size_t idx = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= (size_t)K * B) return;
  double t = 0.0;
  for (int c = 0; c < C; ++c) t = fma(rates[c], dtau[idx * C + c], t);
  dbls[idx] = t;
}

// (JobWs and job_ws_shape: layout.hpp)
static JobWs job_ws(const hx_job_s* j, int K, int grad) {
  return job_ws_shape(j->plan->T, j->S, j->slice, j->C, K, j->dtype, grad);
}

// Workspace of one hx_job_eval_model request: P in the job's dtype, the request's frequencies
// and proportions cast to it, one pattern slice's pruning workspace (d loglik / d mats mode),
// that slice's d loglik / d mats in FP64 with the chain-rule scratch, and the per-slice sums
// and their accumulators.
struct ModelWs {
  size_t mats, freqs, props, prune, prune_bytes, dmats, pgrad, ll, dfr, dpr, dbls, drates, dq;
  size_t acc_dfr, acc_dpr, acc_dbls, acc_drates, acc_dq, total;
};

static ModelWs model_ws(const hx_job_s* j, int K, int eigen_batched) {
  const size_t rb = j->dtype == HX_F64 ? 8 : 4;
  const size_t S = j->S, C = j->C, B = j->B, E = eigen_batched ? K : 1;
  const size_t nm = (size_t)K * B * C * S * S;
  ModelWs w{};
  size_t off = 0;
  auto take = [&](size_t bytes) {
    const size_t at = off;
    off = align_up(off + bytes, 256);
    return at;
  };
  w.mats = take(nm * rb);
  w.freqs = take((size_t)K * S * rb);
  w.props = take((size_t)K * C * rb);
  w.prune_bytes = workspace_bytes_for(j->plan->T, j->S, j->slice, j->C, K, j->dtype, HX_GRAD_MATS);
  w.prune = take(w.prune_bytes);
  w.dmats = take(nm * 8);
  w.pgrad = take(hx_pmatrix_grad_scratch_bytes(j->S, K, j->B, j->C));
  w.ll = take((size_t)K * 8);
  w.dfr = take((size_t)K * S * 8);
  w.dpr = take((size_t)K * C * 8);
  w.dbls = take((size_t)K * B * 8);
  w.drates = take((size_t)K * C * 8);
  w.dq = take(E * S * S * 8);
  w.acc_dfr = take((size_t)K * S * 8);
  w.acc_dpr = take((size_t)K * C * 8);
  w.acc_dbls = take((size_t)K * B * 8);
  w.acc_drates = take((size_t)K * C * 8);
  w.acc_dq = take(E * S * S * 8);
  w.total = off;
  return w;
}

template <typename R>
__global__ void k_cast_rows(const double* __restrict__ src, R* __restrict__ dst, size_t n) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) dst[i] = (R)src[i];
}

static void cast_rows(const double* src, void* dst, size_t n, int dtype, hipStream_t st) {
  const unsigned g = (unsigned)((n + 255) / 256);
  if (dtype == HX_F64) k_cast_rows<double><<<g, 256, 0, st>>>(src, static_cast<double*>(dst), n);
  else k_cast_rows<float><<<g, 256, 0, st>>>(src, static_cast<float*>(dst), n);
}

// A job with the eigensystem (lambda, left, right) of a reversible Q, or with a general Q
// (q non-null: stored where `left` goes, lambda and right zero).
static int job_create(hx_plan_t plan, int S, int N, int C, int dtype, int slice_patterns,
                      const uint8_t* codes, const double* table, int ncodes, const double* weights,
                      const double* lambda, const double* left, const double* right,
                      const double* q, const double* rates, const double* props,
                      const double* freqs, int device, hx_job_t* out) {
  HX_REQUIRE(plan && codes && table && weights && (q || (lambda && left && right)) && rates &&
                 props && freqs && out,
             HX_ERR_INVALID_ARGUMENT, "null argument");
  HX_REQUIRE(S >= 2 && S <= 64 && N >= 1 && C >= 1 && ncodes >= 1 && ncodes <= hx::kMaxCodes,
             HX_ERR_INVALID_ARGUMENT, "bad dimensions");
  HX_REQUIRE(hx::dtype_ok(dtype), HX_ERR_INVALID_ARGUMENT, "bad dtype");
  // The kernels index the code table with the raw code, so every code must name a row.
  const uint8_t max_code = *std::max_element(codes, codes + (size_t)plan->T * N);
  HX_REQUIRE(max_code < ncodes, HX_ERR_INVALID_ARGUMENT, "tip code outside the code table");
  auto* j = new hx_job_s();
  j->plan = plan;
  j->device = device;
  j->S = S;
  j->N = N;
  j->C = C;
  j->dtype = dtype;
  j->T = plan->T;
  j->B = 2 * plan->T - 2;
  j->ncodes = ncodes;
  j->slice = slice_patterns > 0 ? std::min(slice_patterns, N) : N;
  j->nslices = (N + j->slice - 1) / j->slice;
  const size_t rb = dtype == HX_F64 ? 8 : 4;
  // host staging of the resident image
  std::vector<size_t> off_codes(j->nslices), off_w(j->nslices);
  size_t off = 0;
  for (int s = 0; s < j->nslices; ++s) {
    const int n = std::min(j->slice, N - s * j->slice);
    j->slice_n.push_back(n);
    off_codes[s] = off;
    off = hx::align_up(off + (size_t)j->T * n, 256);
    off_w[s] = off;
    off = hx::align_up(off + (size_t)n * 8, 256);
  }
  const size_t o_table = off;
  off = hx::align_up(off + (size_t)ncodes * S * rb, 256);
  const size_t o_lam = off;
  off = hx::align_up(off + (size_t)S * 8, 256);
  const size_t o_left = off;
  off = hx::align_up(off + (size_t)S * S * 8, 256);
  const size_t o_right = off;
  off = hx::align_up(off + (size_t)S * S * 8, 256);
  const size_t o_rates = off;
  off = hx::align_up(off + (size_t)C * 8, 256);
  const size_t o_props = off;
  off = hx::align_up(off + (size_t)C * rb, 256);
  const size_t o_freqs = off;
  off = hx::align_up(off + (size_t)S * rb, 256);
  std::vector<uint8_t> img(off, 0);
  for (int s = 0; s < j->nslices; ++s) {
    const int n = j->slice_n[s];
    for (int t = 0; t < j->T; ++t)
      std::copy(codes + (size_t)t * N + (size_t)s * j->slice,
                codes + (size_t)t * N + (size_t)s * j->slice + n, img.data() + off_codes[s] + (size_t)t * n);
    std::copy(reinterpret_cast<const uint8_t*>(weights + (size_t)s * j->slice),
              reinterpret_cast<const uint8_t*>(weights + (size_t)s * j->slice + n),
              img.data() + off_w[s]);
  }
  for (int q = 0; q < ncodes * S; ++q) {
    if (dtype == HX_F64) reinterpret_cast<double*>(img.data() + o_table)[q] = table[q];
    else reinterpret_cast<float*>(img.data() + o_table)[q] = (float)table[q];
  }
  if (q) {
    std::copy(q, q + S * S, reinterpret_cast<double*>(img.data() + o_left));
  } else {
    std::copy(lambda, lambda + S, reinterpret_cast<double*>(img.data() + o_lam));
    std::copy(left, left + S * S, reinterpret_cast<double*>(img.data() + o_left));
    std::copy(right, right + S * S, reinterpret_cast<double*>(img.data() + o_right));
  }
  std::copy(rates, rates + C, reinterpret_cast<double*>(img.data() + o_rates));
  for (int q = 0; q < C; ++q) {
    if (dtype == HX_F64) reinterpret_cast<double*>(img.data() + o_props)[q] = props[q];
    else reinterpret_cast<float*>(img.data() + o_props)[q] = (float)props[q];
  }
  for (int q = 0; q < S; ++q) {
    if (dtype == HX_F64) reinterpret_cast<double*>(img.data() + o_freqs)[q] = freqs[q];
    else reinterpret_cast<float*>(img.data() + o_freqs)[q] = (float)freqs[q];
  }
  int prev = 0;
  hipError_t e = hipGetDevice(&prev);
  if (e == hipSuccess) e = hipSetDevice(device);
  if (e == hipSuccess) e = hipMalloc(&j->d_buf, off);
  if (e == hipSuccess) e = hipMemcpy(j->d_buf, img.data(), off, hipMemcpyHostToDevice);
  hipSetDevice(prev);
  if (e != hipSuccess) {
    hx::set_error(std::string("job upload: ") + hipGetErrorString(e));
    if (j->d_buf) hipFree(j->d_buf);
    delete j;
    return e == hipErrorOutOfMemory ? HX_ERR_OUT_OF_MEMORY : HX_ERR_HIP;
  }
  char* base = static_cast<char*>(j->d_buf);
  for (int s = 0; s < j->nslices; ++s) {
    j->d_codes.push_back(reinterpret_cast<uint8_t*>(base + off_codes[s]));
    j->d_weights.push_back(reinterpret_cast<double*>(base + off_w[s]));
  }
  j->d_table = base + o_table;
  j->d_lambda = reinterpret_cast<double*>(base + o_lam);
  j->d_left = reinterpret_cast<double*>(base + o_left);
  j->d_right = reinterpret_cast<double*>(base + o_right);
  j->general = q != nullptr;
  j->d_rates = reinterpret_cast<double*>(base + o_rates);
  j->d_props = base + o_props;
  j->d_freqs = base + o_freqs;
  j->resident = off;
  *out = j;
  return HX_OK;
}

}  // namespace hx

extern "C" {

int hx_job_create(hx_plan_t plan, int S, int N, int C, int dtype, int slice_patterns,
                  const uint8_t* codes, const double* table, int ncodes, const double* weights,
                  const double* lambda, const double* left, const double* right,
                  const double* rates, const double* props, const double* freqs, int device,
                  hx_job_t* out) {
  HX_REQUIRE(lambda && left && right, HX_ERR_INVALID_ARGUMENT, "null argument");
  return hx::job_create(plan, S, N, C, dtype, slice_patterns, codes, table, ncodes, weights,
                        lambda, left, right, nullptr, rates, props, freqs, device, out);
}

int hx_job_create_general(hx_plan_t plan, int S, int N, int C, int dtype, int slice_patterns,
                          const uint8_t* codes, const double* table, int ncodes,
                          const double* weights, const double* q, const double* rates,
                          const double* props, const double* freqs, int device, hx_job_t* out) {
  HX_REQUIRE(q, HX_ERR_INVALID_ARGUMENT, "null argument");
  HX_REQUIRE(S <= 32, HX_ERR_UNSUPPORTED, "general rate matrices: at most 32 states");
  return hx::job_create(plan, S, N, C, dtype, slice_patterns, codes, table, ncodes, weights,
                        nullptr, nullptr, nullptr, q, rates, props, freqs, device, out);
}

int hx_job_destroy(hx_job_t job) {
  if (!job) return HX_OK;
  if (job->d_buf) hipFree(job->d_buf);
  delete job;
  return HX_OK;
}

size_t hx_job_resident_bytes(hx_job_t job) { return job ? job->resident : 0; }

size_t hx_job_workspace_bytes(hx_job_t job, int K, int want_grad) {
  if (!job || K < 1) return 0;
  return hx::job_ws(job, K, want_grad).total;
}

size_t hx_job_workspace_bytes_for(int taxa, int S, int slice_patterns, int C, int K, int dtype,
                                  int want_grad) {
  if (taxa < 2 || S < 2 || slice_patterns < 1 || C < 1 || K < 1) return 0;
  if (!hx::dtype_ok(dtype)) return 0;
  return hx::job_ws_shape(taxa, S, slice_patterns, C, K, dtype, want_grad).total;
}

}  // extern "C"

namespace hx {

// One request of hx_job_eval / hx_job_eval_noalloc in a workspace laid out by job_ws.
static int job_eval_in(hx_job_t job, int K, const double* bls, double* out_loglik,
                       double* out_dbls, const JobWs& w, void* workspace, hx_stream_t stream) {
  const int grad = out_dbls != nullptr;
  hipStream_t st = reinterpret_cast<hipStream_t>(stream);
  char* base = static_cast<char*>(workspace);
  void* mats = base + w.mats;
  void* mats_dt = grad ? base + w.mats_dt : nullptr;
  // P and dP/dtau: from the eigensystem, or by exp(Q tau) for a general Q
  int rc = job->general
               ? hx_pmatrix_expm(job->S, K, job->B, job->C, 0, job->d_left, bls, job->d_rates, 0,
                                 mats, job->dtype, stream)
               : hx_pmatrix(job->S, K, job->B, job->C, 0, job->d_lambda, job->d_left, job->d_right,
                            bls, job->d_rates, 0, mats, job->dtype, stream);
  if (rc != HX_OK) return rc;
  if (grad) {
    rc = job->general
             ? hx_pmatrix_expm_dtau(job->S, K, job->B, job->C, 0, job->d_left, bls, job->d_rates, 0,
                                    mats_dt, job->dtype, stream)
             : hx_pmatrix_dtau(job->S, K, job->B, job->C, 0, job->d_lambda, job->d_left,
                               job->d_right, bls, job->d_rates, 0, mats_dt, job->dtype, stream);
    if (rc != HX_OK) return rc;
  }
  double* ll_tmp = reinterpret_cast<double*>(base + w.loglik_tmp);
  const size_t nt = (size_t)K * job->B * job->C;
  double* dt_tmp = grad ? reinterpret_cast<double*>(base + w.dtau_tmp) : nullptr;
  double* dt_acc = grad ? reinterpret_cast<double*>(base + w.dtau_acc) : nullptr;
  hx::fill0(out_loglik, K, st);
  if (grad) hx::fill0(dt_acc, nt, st);
  for (int s = 0; s < job->nslices; ++s) {
    const int n = job->slice_n[s];
    rc = hx_loglik(job->plan, job->S, n, job->C, K, job->dtype, HX_TIPS_CODES, job->d_codes[s],
                   job->d_table, job->ncodes, job->d_weights[s], mats, job->d_freqs, 0,
                   job->d_props, 0, base + w.prune, w.prune_bytes, ll_tmp, nullptr, nullptr,
                   stream);
    if (rc != HX_OK) return rc;
    hx::accum(out_loglik, ll_tmp, K, st);
    if (grad) {
      rc = hx_loglik_grad_tau(job->plan, job->S, n, job->C, K, job->dtype, HX_TIPS_CODES,
                              job->d_codes[s], job->d_table, job->ncodes, job->d_weights[s], mats,
                              mats_dt, job->d_freqs, 0, job->d_props, 0, base + w.prune,
                              w.prune_bytes, dt_tmp, stream);
      if (rc != HX_OK) return rc;
      hx::accum(dt_acc, dt_tmp, nt, st);
    }
  }
  if (grad) {
    const size_t nb = (size_t)K * job->B;
    hx::k_dbls<<<(unsigned)((nb + 255) / 256), 256, 0, st>>>(K, job->B, job->C, job->d_rates,
                                                             dt_acc, out_dbls);
  }
  HX_HIP_TRY(hipGetLastError());
  return HX_OK;
}

struct ModelArgs {
  int eigen_batched;
  const double *lambda, *left, *right, *rates, *props, *freqs;
  int model_batched;
  double *out_loglik, *out_dbls, *out_drates, *out_dq, *out_dfreqs, *out_dprops;
};

// One request with its own eigensystem, category rates, proportions and frequencies. Per
// pattern slice: post-order (logL, d/dfreqs and d/dprops at the root), pre-order d/dmats,
// and the chain rule to d/dbls, d/drates and d/dQ, all linear in the slice's terms, so the
// slices' FP64 results are summed in slice order.
static int job_eval_model_in(hx_job_t job, int K, const double* bls, const ModelArgs& a,
                             const ModelWs& w, void* workspace, hx_stream_t stream) {
  hipStream_t st = reinterpret_cast<hipStream_t>(stream);
  char* base = static_cast<char*>(workspace);
  auto at = [&](size_t off) { return reinterpret_cast<double*>(base + off); };
  const int S = job->S, C = job->C, B = job->B;
  const int E = a.eigen_batched ? K : 1;
  const int mb = a.model_batched && K > 1;
  const int rows = mb ? K : 1;
  void* mats = base + w.mats;
  int rc = hx_pmatrix(S, K, B, C, a.eigen_batched, a.lambda, a.left, a.right, bls, a.rates, mb,
                      mats, job->dtype, stream);
  if (rc != HX_OK) return rc;
  cast_rows(a.freqs, base + w.freqs, (size_t)rows * S, job->dtype, st);
  cast_rows(a.props, base + w.props, (size_t)rows * C, job->dtype, st);
  const size_t nkb = (size_t)K * B, nks = (size_t)K * S, nkc = (size_t)K * C;
  const size_t nq = (size_t)E * S * S;
  fill0(a.out_loglik, K, st);
  fill0(at(w.acc_dfr), nks, st);
  fill0(at(w.acc_dpr), nkc, st);
  fill0(at(w.acc_dbls), nkb, st);
  fill0(at(w.acc_drates), nkc, st);
  fill0(at(w.acc_dq), nq, st);
  for (int s = 0; s < job->nslices; ++s) {
    const int n = job->slice_n[s];
    rc = hx_loglik(job->plan, S, n, C, K, job->dtype, HX_TIPS_CODES, job->d_codes[s],
                   job->d_table, job->ncodes, job->d_weights[s], mats, base + w.freqs, mb,
                   base + w.props, mb, base + w.prune, w.prune_bytes, at(w.ll), at(w.dfr),
                   at(w.dpr), stream);
    if (rc != HX_OK) return rc;
    rc = hx_loglik_grad_mats(job->plan, S, n, C, K, job->dtype, HX_TIPS_CODES, job->d_codes[s],
                             job->d_table, job->ncodes, job->d_weights[s], mats, base + w.freqs,
                             mb, base + w.props, mb, base + w.prune, w.prune_bytes, at(w.dmats),
                             stream);
    if (rc != HX_OK) return rc;
    rc = hx_pmatrix_grad(S, K, B, C, a.eigen_batched, a.lambda, a.left, a.right, bls, a.rates, mb,
                         at(w.dmats), at(w.dbls), at(w.drates), at(w.dq), base + w.pgrad, stream);
    if (rc != HX_OK) return rc;
    accum(a.out_loglik, at(w.ll), K, st);
    accum(at(w.acc_dfr), at(w.dfr), nks, st);
    accum(at(w.acc_dpr), at(w.dpr), nkc, st);
    accum(at(w.acc_dbls), at(w.dbls), nkb, st);
    accum(at(w.acc_drates), at(w.drates), nkc, st);
    accum(at(w.acc_dq), at(w.dq), nq, st);
  }
  struct Out { double* dst; size_t src, n; };
  const Out outs[] = {{a.out_dbls, w.acc_dbls, nkb}, {a.out_drates, w.acc_drates, nkc},
                      {a.out_dq, w.acc_dq, nq}, {a.out_dfreqs, w.acc_dfr, nks},
                      {a.out_dprops, w.acc_dpr, nkc}};
  for (const Out& o : outs)
    if (o.dst) HX_HIP_TRY(hipMemcpyAsync(o.dst, at(o.src), o.n * 8, hipMemcpyDeviceToDevice, st));
  HX_HIP_TRY(hipGetLastError());
  return HX_OK;
}

// Workspace of hx_job_eval_noalloc: a file-scope device array, one copy per device, held from
// the library's load to the end of the process. In host code `scratch` names a host-side
// stand-in, not a device address: kernels and copies get the per-device address that
// hipGetSymbolAddress returns (looked up once per device, with that device current).
__device__ __align__(256) unsigned char scratch[HX_NOALLOC_WORKSPACE_BYTES];

constexpr int kScratchDevices = 64;
static std::mutex g_scratch_mu;
static void* g_scratch_addr[kScratchDevices] = {};

static int scratch_address(int device, void** out) {
  HX_REQUIRE(device >= 0 && device < kScratchDevices, HX_ERR_INVALID_ARGUMENT,
             "device index out of range for the static workspace");
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (g_scratch_addr[device] == nullptr) {
    int prev = 0;
    HX_HIP_TRY(hipGetDevice(&prev));
    if (prev != device) HX_HIP_TRY(hipSetDevice(device));
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (prev != device) hipSetDevice(prev);
    if (e != hipSuccess || p == nullptr) {
      set_error(std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
      return HX_ERR_HIP;
    }
    g_scratch_addr[device] = p;
  }
  *out = g_scratch_addr[device];
  return HX_OK;
}

}  // namespace hx

extern "C" {

int hx_job_eval(hx_job_t job, int K, const double* bls, double* out_loglik, double* out_dbls,
                void* workspace, size_t workspace_bytes, hx_stream_t stream) {
  HX_REQUIRE(job && bls && out_loglik && workspace && K >= 1, HX_ERR_INVALID_ARGUMENT,
             "null argument");
  const hx::JobWs w = hx::job_ws(job, K, out_dbls != nullptr);
  HX_REQUIRE(workspace_bytes >= w.total, HX_ERR_WORKSPACE_TOO_SMALL, "workspace too small");
  return hx::job_eval_in(job, K, bls, out_loglik, out_dbls, w, workspace, stream);
}

size_t hx_job_model_workspace_bytes(hx_job_t job, int K, int eigen_batched) {
  if (!job || K < 1) return 0;
  return hx::model_ws(job, K, eigen_batched).total;
}

int hx_job_eval_model(hx_job_t job, int K, const double* bls, int eigen_batched,
                      const double* lambda, const double* left, const double* right,
                      const double* rates, const double* props, const double* freqs,
                      int model_batched, double* out_loglik, double* out_dbls, double* out_drates,
                      double* out_dq, double* out_dfreqs, double* out_dprops, void* workspace,
                      size_t workspace_bytes, hx_stream_t stream) {
  HX_REQUIRE(job && bls && lambda && left && right && rates && props && freqs && out_loglik &&
                 workspace && K >= 1,
             HX_ERR_INVALID_ARGUMENT, "null argument");
  HX_REQUIRE(!job->general, HX_ERR_UNSUPPORTED,
             "a job with a general rate matrix takes no per-request eigensystem");
  const hx::ModelWs w = hx::model_ws(job, K, eigen_batched);
  HX_REQUIRE(workspace_bytes >= w.total, HX_ERR_WORKSPACE_TOO_SMALL, "workspace too small");
  const hx::ModelArgs a{eigen_batched, lambda, left, right, rates, props, freqs, model_batched,
                        out_loglik, out_dbls, out_drates, out_dq, out_dfreqs, out_dprops};
  return hx::job_eval_model_in(job, K, bls, a, w, workspace, stream);
}

int hx_job_eval_noalloc(hx_job_t job, int K, const double* bls, double* out_loglik,
                        double* out_dbls, hx_stream_t stream) {
  HX_REQUIRE(job && bls && out_loglik && K >= 1, HX_ERR_INVALID_ARGUMENT, "null argument");
  const hx::JobWs w = hx::job_ws(job, K, out_dbls != nullptr);
  if (w.total > (size_t)HX_NOALLOC_WORKSPACE_BYTES) {
    hx::set_error("request needs " + std::to_string(w.total) +
                  " bytes of workspace; the static workspace holds " +
                  std::to_string((size_t)HX_NOALLOC_WORKSPACE_BYTES) +
                  " (register the job with a smaller pattern slice, or use hx_job_eval)");
    return HX_ERR_WORKSPACE_TOO_SMALL;
  }
  void* ws = nullptr;
  const int rc = hx::scratch_address(job->device, &ws);
  if (rc != HX_OK) return rc;
  return hx::job_eval_in(job, K, bls, out_loglik, out_dbls, w, ws, stream);
}

}  // extern "C"

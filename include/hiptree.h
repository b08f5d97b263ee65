/*
 * hiptree: MI355X-native tree-likelihood engine, C ABI.
 *
 * Every entry point takes plain pointers and sizes; device pointers are HIP
 * device memory owned by the caller unless stated otherwise. Calls are
 * asynchronous on the given stream and return an hx_status; nothing throws
 * across this boundary. A handle (plan, server) is externally synchronised.
 *
 * Array conventions (row-major, innermost last):
 *   T taxa, I = T-1 internal nodes, B = 2T-2 branches (branch b = node b),
 *   S states, N site patterns, C rate categories, K samples.
 *   postorder   int32 [I][3]   (node, left, right) in post-order, root last
 *   mats        R [K][B][C][S][S]   transition matrices, R = float | double
 *   freqs       R [K or 1][S]
 *   props       R [K or 1][C]
 *   weights     double [N]
 *   tip codes   uint8 [T][N] with a table R [ncodes][S] of tip partial rows,
 *               or dense tip partials R [T][N][S]. Every code must be below
 *               ncodes: the kernels index the table with the raw code
 *               (hx_job_create checks this on the host; device-resident codes
 *               passed to hx_loglik* are the caller's to keep in range).
 *
 * Reference interfaces replaced (torchtree 1.0.2, torchtree/evolution/):
 *   hx_loglik      -> calculate_treelikelihood (tree_likelihood.py:14-37),
 *                     calculate_treelikelihood_discrete (40-75),
 *                     ..._tip_states_discrete (78-131), ..._discrete_safe (134-183),
 *                     ..._discrete_rescaled (186-221),
 *                     ..._tip_states_discrete_rescaled (224-278)
 *   hx_loglik_grad_mats -> the autograd backward of the functions above
 *                     with respect to mats, freqs and props (the reference
 *                     relies on torch autograd; there is no explicit code)
 *   hx_pmatrix     -> SymmetricSubstitutionModel.p_t (substitution_model/abstract.py:57-79)
 *                     and EmpiricalSubstitutionModel.p_t (general.py:318-331)
 *   hx_pmatrix_grad -> autograd backward of p_t through torch.linalg.eigh
 *                     (abstract.py:79), expressed with the exp divided differences
 */
#ifndef HIPTREE_H
#define HIPTREE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct ihipStream_t* hx_stream_t; /* same object as hipStream_t */

typedef enum {
  HX_OK = 0,
  HX_ERR_INVALID_ARGUMENT = 1,
  HX_ERR_UNSUPPORTED = 2,
  HX_ERR_HIP = 3,
  HX_ERR_WORKSPACE_TOO_SMALL = 4,
  HX_ERR_BAD_TOPOLOGY = 5,
  HX_ERR_OUT_OF_MEMORY = 6,
  HX_ERR_BUDGET = 7,
  HX_ERR_NOT_FOUND = 8
} hx_status;

/* Precision of a call (the `dtype` argument). HX_F32 and HX_F64: partials,
 * P and every product in that type. HX_F32_BF16X3: partials and P in FP32 (R =
 * float in every array convention), and the dense-state contractions (S != 4,
 * at most 4 categories) on the BF16 matrix cores in split precision: each
 * operand x carried as hi = bf16(x) plus lo = bf16(x - hi), each product formed
 * as lo*hi + hi*lo + hi*hi with FP32 accumulation. Four-state calls run as
 * HX_F32. */
enum { HX_F32 = 0, HX_F64 = 1, HX_F32_BF16X3 = 2 };
enum { HX_TIPS_CODES = 0, HX_TIPS_DENSE = 1 };

/* Human-readable text for a status, plus the last error detail. */
const char* hx_error_string(int status);
const char* hx_last_error(void);
int hx_version(void);

/* ---------------------------------------------------------------- plans */
/* A plan holds one topology: the post-order triples, the height-level
 * schedule for the post-order pass and the depth-level schedule for the
 * pre-order pass, copied once to device memory owned by the plan.
 * Replaces the per-call Python loop over post_indexing
 * (tree_likelihood.py:67-70) and AbstractTreeModel.update_traversals
 * (tree_model.py:187-195). `postorder` is a host pointer. */
typedef struct hx_plan_s* hx_plan_t;
int hx_plan_create(int taxa, const int32_t* postorder, int device, hx_plan_t* plan);
int hx_plan_destroy(hx_plan_t plan);
/* levels: post-order height levels; depth: pre-order depth levels. */
int hx_plan_info(hx_plan_t plan, int* post_levels, int* pre_levels, int* max_width);

/* The four-state walk programs of a plan (no reference counterpart: accounting of what
 * the walks move): post-order stack entries, pre-order stack depth, cherries (internal
 * nodes whose children are both tips), and the cherries whose partial the post-order walk
 * does not store under register mode 0 and 1 (rebuilt by the pre-order walk). Any output
 * pointer may be NULL. */
int hx_plan_walk_info(hx_plan_t plan, int* post_need, int* pre_depth, int* cherries,
                      int* dropped_mode0, int* dropped_mode1);

/* Device workspace for one call. grad_mode: HX_GRAD_NONE (hx_loglik only),
 * HX_GRAD_MATS (then hx_loglik_grad_mats), HX_GRAD_TAU (then hx_loglik_grad_tau). */
enum { HX_GRAD_NONE = 0, HX_GRAD_MATS = 1, HX_GRAD_TAU = 2 };
size_t hx_workspace_bytes(hx_plan_t plan, int S, int N, int C, int K, int dtype, int grad_mode);

/* --------------------------------------------------------- likelihood */
/* Post-order pruning with exact power-of-two rescaling of every node,
 * fused root reduction. Outputs (device, double):
 *   out_loglik [K]      sum_n w_n log L_n
 *   out_dfreqs [K][S]   d loglik / d freqs through the root (may be NULL)
 *   out_dprops [K][C]   d loglik / d props (may be NULL)
 * The workspace keeps the post-order partials; pass the same workspace to
 * hx_loglik_grad_mats afterwards to get the gradient with respect to mats. */
int hx_loglik(hx_plan_t plan, int S, int N, int C, int K, int dtype,
              int tip_kind, const void* tips, const void* tip_table, int ncodes,
              const double* weights, const void* mats,
              const void* freqs, int freqs_batched,
              const void* props, int props_batched,
              void* workspace, size_t workspace_bytes,
              double* out_loglik, double* out_dfreqs, double* out_dprops,
              hx_stream_t stream);

/* Pre-order adjoint over the partials left in `workspace` by hx_loglik
 * (consumed: the partials are overwritten in place). Output:
 *   out_dmats [K][B][C][S][S] double = d loglik / d mats. */
int hx_loglik_grad_mats(hx_plan_t plan, int S, int N, int C, int K, int dtype,
                        int tip_kind, const void* tips, const void* tip_table, int ncodes,
                        const double* weights, const void* mats,
                        const void* freqs, int freqs_batched,
                        const void* props, int props_batched,
                        void* workspace, size_t workspace_bytes,
                        double* out_dmats, hx_stream_t stream);

/* Pre-order adjoint reduced to branch-length scale: out_dtau [K][B][C] double
 * = d loglik / d tau with tau = bls[k][b] * rates[k][c], i.e. the contraction
 * of d loglik / d mats with dP/dtau (`mats_dtau`, from hx_pmatrix_dtau, same
 * layout and dtype as mats). Needs a workspace sized for HX_GRAD_TAU; the
 * per-chunk state is C doubles per branch instead of C*S*S values, which is
 * what keeps 61-state requests with many samples inside a bounded workspace.
 * Replaces the same autograd backward as hx_loglik_grad_mats, projected on
 * the branch-length and rate-category directions (tree_likelihood.py:346,
 * mats = p_t(bls * rates)). */
int hx_loglik_grad_tau(hx_plan_t plan, int S, int N, int C, int K, int dtype,
                       int tip_kind, const void* tips, const void* tip_table, int ncodes,
                       const double* weights, const void* mats, const void* mats_dtau,
                       const void* freqs, int freqs_batched,
                       const void* props, int props_batched,
                       void* workspace, size_t workspace_bytes,
                       double* out_dtau, hx_stream_t stream);

/* ------------------------------------------------------------ P matrices */
/* mats[k][b][c] = L_e diag(exp(lambda_e * bls[k][b] * rates[k][c])) R_e with
 * e = k if eigen_batched else 0. lambda [E][S], left [E][S][S] (Pi^-1/2 V),
 * right [E][S][S] (V^T Pi^1/2), all double; bls [K][B], rates [K or 1][C]
 * double; mats written in `dtype`. */
int hx_pmatrix(int S, int K, int B, int C, int eigen_batched,
               const double* lambda, const double* left, const double* right,
               const double* bls, const double* rates, int rates_batched,
               void* mats, int dtype, hx_stream_t stream);

/* dP/dtau = L_e diag(lambda_e exp(lambda_e tau)) R_e, same arguments and
 * layout as hx_pmatrix (the derivative of SymmetricSubstitutionModel.p_t,
 * abstract.py:64-76, with respect to its branch-length argument). */
int hx_pmatrix_dtau(int S, int K, int B, int C, int eigen_batched,
                    const double* lambda, const double* left, const double* right,
                    const double* bls, const double* rates, int rates_batched,
                    void* mats_dtau, int dtype, hx_stream_t stream);

/* Chain rule from dmats = d loglik / d mats [K][B][C][S][S] (double) to
 *   out_dbls  [K][B]      d loglik / d bls
 *   out_drates[K][C]      d loglik / d rates (sum over branches)
 *   out_dq    [E][S][S]   d loglik / d Q (normalised Q, summed over samples
 *                         that share an eigensystem)
 * using P = exp(Q tau) and the divided differences of exp on the spectrum
 * (exact at repeated eigenvalues). `scratch` needs
 * hx_pmatrix_grad_scratch_bytes(S, K, B, C) bytes of device memory. */
size_t hx_pmatrix_grad_scratch_bytes(int S, int K, int B, int C);
int hx_pmatrix_grad(int S, int K, int B, int C, int eigen_batched,
                    const double* lambda, const double* left, const double* right,
                    const double* bls, const double* rates, int rates_batched,
                    const double* dmats, double* out_dbls, double* out_drates,
                    double* out_dq, void* scratch, hx_stream_t stream);

/* ------------------------------------------ P matrices of non-reversible models */
/* mats[k][b][c] = exp(Q_e tau), tau = bls[k][b] * rates[k or 0][c], e = k if
 * q_batched else 0, for any rate matrix Q [E][S][S] (double, normalised; S <= 32),
 * in `dtype`: scaling and squaring of a degree-18 Taylor polynomial in FP64.
 * Replaces NonSymmetricSubstitutionModel.p_t (substitution_model/abstract.py:88-94),
 * which calls torch.matrix_exp. */
int hx_pmatrix_expm(int S, int K, int B, int C, int q_batched, const double* Q,
                    const double* bls, const double* rates, int rates_batched, void* mats,
                    int dtype, hx_stream_t stream);
/* dP/dtau = Q_e exp(Q_e tau), same arguments and layout as hx_pmatrix_expm. */
int hx_pmatrix_expm_dtau(int S, int K, int B, int C, int q_batched, const double* Q,
                         const double* bls, const double* rates, int rates_batched,
                         void* mats_dtau, int dtype, hx_stream_t stream);
/* Chain rule for P = exp(Q tau) from dmats = d loglik / d mats [K][B][C][S][S]
 * (double): out_dbls [K][B], out_drates [K][C], out_dq [E][S][S] = sum of
 * tau L(tau Q^T, G) with L the Frechet derivative of exp, from the upper right
 * block of exp([[tau Q^T, G], [0, tau Q^T]]). The autograd backward of
 * torch.matrix_exp in NonSymmetricSubstitutionModel.p_t (abstract.py:92). `scratch`
 * needs hx_pmatrix_expm_grad_scratch_bytes(S, K, B, C) bytes of device memory. */
size_t hx_pmatrix_expm_grad_scratch_bytes(int S, int K, int B, int C);
int hx_pmatrix_expm_grad(int S, int K, int B, int C, int q_batched, const double* Q,
                         const double* bls, const double* rates, int rates_batched,
                         const double* dmats, double* out_dbls, double* out_drates, double* out_dq,
                         void* scratch, hx_stream_t stream);

/* ------------------------------------------------------------ resident jobs */
/* A job keeps one alignment resident on a device for repeated requests:
 * uint8 tip codes [T][N] and their table [ncodes][S] (host pointers, copied
 * and cut into slices of `slice_patterns` patterns), weights [N], the FP64
 * eigensystem (lambda [S], left [S][S], right [S][S] as in hx_pmatrix),
 * category rates and proportions [C] and frequencies [S] (host, double).
 * No reference counterpart as such: TreeLikelihoodModel keeps its tips in
 * the model object (tree_likelihood.py:293-311) and recomputes every input
 * per call; a job is the device-resident form of that state for serving. */
typedef struct hx_job_s* hx_job_t;
int hx_job_create(hx_plan_t plan, int S, int N, int C, int dtype, int slice_patterns,
                  const uint8_t* codes, const double* table, int ncodes, const double* weights,
                  const double* lambda, const double* left, const double* right,
                  const double* rates, const double* props, const double* freqs, int device,
                  hx_job_t* job);
/* The same job for a non-reversible model (no eigensystem): Q [S][S] the normalised
 * general rate matrix (host, double; S <= 32). Requests then form P = exp(Q tau) and
 * dP/dtau = Q P with hx_pmatrix_expm / hx_pmatrix_expm_dtau, the device form of
 * NonSymmetricSubstitutionModel.p_t (substitution_model/abstract.py:88-94); everything
 * else is as for hx_job_create. hx_job_eval_model refuses such a job. */
int hx_job_create_general(hx_plan_t plan, int S, int N, int C, int dtype, int slice_patterns,
                          const uint8_t* codes, const double* table, int ncodes,
                          const double* weights, const double* q, const double* rates,
                          const double* props, const double* freqs, int device, hx_job_t* job);
int hx_job_destroy(hx_job_t job);
size_t hx_job_resident_bytes(hx_job_t job);
/* Workspace of one hx_job_eval with K samples (with or without the gradient). */
size_t hx_job_workspace_bytes(hx_job_t job, int K, int want_grad);
/* What hx_job_workspace_bytes reports for a job of this shape (taxa, states,
 * pattern slice, categories, dtype), computed on the host without creating
 * the job: for choosing a slice before registration. No counterpart in the
 * reference, which allocates per call through torch. Returns 0 for invalid
 * arguments. */
size_t hx_job_workspace_bytes_for(int taxa, int S, int slice_patterns, int C, int K, int dtype,
                                  int want_grad);
/* One request: bls [K][B] (device, double; unrooted trees carry 0 in slot
 * 2T-3) -> out_loglik [K] and, if out_dbls is not NULL, out_dbls [K][B]
 * (device, double). Pattern slices run in order on `stream`; results do not
 * depend on other requests running concurrently on other streams. This is
 * TreeLikelihoodModel._call (tree_likelihood.py:313-356) plus the backward
 * with respect to branch lengths. */
int hx_job_eval(hx_job_t job, int K, const double* bls, double* out_loglik, double* out_dbls,
                void* workspace, size_t workspace_bytes, hx_stream_t stream);

/* One request that brings its own substitution and site model: P from the
 * given FP64 eigensystem (lambda [E][S], left [E][S][S], right [E][S][S] as in
 * hx_pmatrix, E = K if eigen_batched else 1), category rates [R][C],
 * proportions [R][C] and root frequencies [R][S] (R = K if model_batched else
 * 1), all device doubles, in place of the ones the job was registered with.
 * This is TreeLikelihoodModel._call (tree_likelihood.py:313-356) re-evaluating
 * p_t from the current parameters on every call (line 346), as an ADVI or HMC
 * job over a GTR model does, with its backward. Outputs (device doubles; all
 * but out_loglik may be NULL):
 *   out_loglik [K]; out_dbls [K][B]; out_drates [K][C];
 *   out_dq [E][S][S]   d loglik / d normalised Q (through P only);
 *   out_dfreqs [K][S]  d loglik / d root frequencies; out_dprops [K][C].
 * The gradients with respect to a model's own parameters (exchangeabilities,
 * frequencies through Q, a site-model shape) follow from these by the chain
 * rule on the host. Workspace: hx_job_model_workspace_bytes(job, K, eigen_batched). */
size_t hx_job_model_workspace_bytes(hx_job_t job, int K, int eigen_batched);
int hx_job_eval_model(hx_job_t job, int K, const double* bls, int eigen_batched,
                      const double* lambda, const double* left, const double* right,
                      const double* rates, const double* props, const double* freqs,
                      int model_batched, double* out_loglik, double* out_dbls, double* out_drates,
                      double* out_dq, double* out_dfreqs, double* out_dprops, void* workspace,
                      size_t workspace_bytes, hx_stream_t stream);

/* hx_job_eval without a caller workspace and without any device allocation.
 * No counterpart in the reference: for callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc). The workspace is a library-owned static device array of
 * HX_NOALLOC_WORKSPACE_BYTES per device, held until the process ends; a
 * request whose hx_job_workspace_bytes(job, K, out_dbls != NULL) exceeds it
 * returns HX_ERR_WORKSPACE_TOO_SMALL before anything is launched (register the
 * job with a smaller pattern slice to fit). Every call on one device shares
 * that device's array, which keeps the last call's contents: calls that
 * overlap on one GPU (other streams, other host threads) race. Results are
 * those of hx_job_eval on the same inputs. */
#define HX_NOALLOC_WORKSPACE_BYTES (64u << 20)
int hx_job_eval_noalloc(hx_job_t job, int K, const double* bls, double* out_loglik,
                        double* out_dbls, hx_stream_t stream);

/* Checks on the device the operand/accumulator lane layout the MFMA
 * pruning kernels assume for v_mfma_f32_16x16x4f32; *mismatches = number of
 * wrong output elements (0 expected). No reference counterpart. */
int hx_selftest_mfma_layout(int* mismatches, hx_stream_t stream);

#ifdef __cplusplus
}
#endif
#endif /* HIPTREE_H */

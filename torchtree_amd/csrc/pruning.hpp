// Declarations shared by the pruning translation units.
#pragma once

#include "common.hpp"

namespace hx {

template <typename R>
struct PassArgs {
  const int32_t* kids;  // [2T-1][2]
  int T, N, C, K, B;
  int ntiles, tiles_per_chunk, nchunks;
  int np;                 // pattern stride of the partials: N, or 64 ntiles for S = 4
  const R* mats;          // [K][B][C][S][S]
  const uint8_t* codes;   // [T][N]
  const R* table;         // [ncodes][S]
  int ncodes;
  const R* dense;         // [T][N][S]
  R* partials;            // [I][K][N][C][S]; S = 4: [I][K][C][np][4]
  int32_t* scales;        // [I][K][N] (S != 4)
  uint8_t* codes_pad;     // S = 4: codes [T][ntiles * 64], rows 64-byte aligned (same memory)
  const R* freqs;
  int freqs_ks;           // stride between samples (0 or S)
  const R* props;
  int props_ks;           // stride between samples (0 or C)
  const double* weights;  // [N]
  double* rootsums;       // [K][nchunks][1+S+C]; S = 4: [K][ntiles][1+S+C]
  R* gpart;               // [B][K][nchunks][C][S][S]     (d loglik / d mats mode)
  const R* mats_dt;       // [K][B][C][S][S] dP/dtau; non-null selects d loglik / d tau mode
  double* gtau;           // [B][K][nchunks][C]            (d loglik / d tau mode)
  R* ppad;                // [K][B][C][2][64][64]: P and P^T zero-padded (FP32, 33..64 states);
                          // with `split`, each element holds its BF16 hi and lo halves
  int split;              // dense-state products as split BF16 on the matrix cores (FP32 only)
  // Four-state FP32 walks with tip codes: -1, or the register mode whose drop bits
  // (walk.hpp) apply: the post-order walk stores no partial of those cherries and the
  // pre-order walk rebuilds it from pcodes.
  int drop_mode;
  const int32_t* cher;    // [ncher][2] tip ids of the cherries
  int ncher;
  uint16_t* pcodes;       // [T + ncher][np]: tip codes, then each cherry's two codes (lo | hi << 8)
};

constexpr double kLn2 = 0.69314718055994530942;

// Workgroup barrier that orders LDS only. __syncthreads() also waits for every
// outstanding global load and store of the wave (vmcnt(0)); the pruning
// kernels' barriers guard LDS tiles alone, since no wave reads global data
// that another wave of the same launch writes.
__device__ __forceinline__ void lds_barrier() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
  __builtin_amdgcn_s_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
}

// Dense-state (S != 4) launchers, defined in pruning_wide.hip.
template <typename R>
int launch_post_wide(const PassArgs<R>& a, const int32_t* nodes, int width, int tips, bool root,
                     int S, hipStream_t st);
template <typename R>
int launch_pre_wide(const PassArgs<R>& a, const int32_t* nodes, int width, int tips, bool root,
                    int S, hipStream_t st);

// Four-state kernels (pruning_dna.hip): fused post-order walk and the
// pre-order adjoint one level per launch. Partials of S = 4 use the
// category-major layout [slot][K][C][np][4].
constexpr size_t kDnaLdsBudget = 64 * 1024;
constexpr size_t kDnaWalkLdsMax = 160 * 1024;  // gfx950 LDS per workgroup
template <typename R>
int launch_post_dna(const PassArgs<R>& a, const int32_t* prog, const int32_t* segtab, int nseg,
                    int nprog, int need, int tips, hipStream_t st);
template <typename R>
int launch_pre_dna(const PassArgs<R>& a, const int32_t* nodes, int width, int tips, bool root,
                   hipStream_t st);
// The whole pre-order in one launch (FP32, C <= kPreWalkMaxC; use_pre_walk).
int launch_pre_walk(const PassArgs<float>& a, const int32_t* prog, int nprog, int tips,
                    hipStream_t st);

// Dense-state kernels on MFMA (pruning_mfma.hip, FP32 and FP64); used when
// use_mfma() (layout.hpp).
template <typename R>
int launch_mfma(bool post, const PassArgs<R>& a, const int32_t* nodes, int width, int tips,
                bool root, int S, hipStream_t st);

}  // namespace hx

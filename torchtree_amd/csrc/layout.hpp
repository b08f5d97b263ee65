// Host-side shape arithmetic of the library, free of HIP: the pattern tiling of a call, the
// kernel choices that change a workspace's layout, and the workspace layouts of hx_loglik*
// and of a resident job's evaluation. The HIP sources use it through common.hpp; the
// host-only test drivers build it with g++ under AddressSanitizer and UBSan.
#pragma once

#include <stddef.h>
#include <stdint.h>

#include "hiptree.h"

namespace hx {

constexpr int kBlock = 256;      // threads per block for the pruning kernels
constexpr int kWave = 64;        // CDNA wavefront
constexpr int kMaxCodes = 256;   // tip table rows held in LDS

inline int next_pow2(int x) {
  int p = 1;
  while (p < x) p <<= 1;
  return p;
}

inline size_t align_up(size_t x, size_t a) { return (x + a - 1) / a * a; }

// The precisions a call may ask for (HX_F32_BF16X3 stores FP32, as HX_F32 does).
inline bool dtype_ok(int dtype) {
  return dtype == HX_F32 || dtype == HX_F64 || dtype == HX_F32_BF16X3;
}

// Pattern tiling shared by every pass of one call: each block owns a chunk of
// consecutive pattern tiles for one (sample, node).
struct Tiling {
  int cp;               // lanes per pattern (C rounded up to a power of two)
  int ppb;              // patterns per tile (= kBlock / cp)
  int ntiles;
  int nchunks;
  int tiles_per_chunk;
};

// Dense states run on the matrix cores (pruning_mfma.hip) at up to 4 categories, in FP32
// (with or without split BF16 products) and FP64; otherwise on the wide kernels
// (pruning_wide.hip).
inline bool use_mfma(int C, int S, int dtype) {
  return (dtype == HX_F32 || dtype == HX_F64 || dtype == HX_F32_BF16X3) && S != 4 && C >= 1 &&
         C <= 4;
}

// LDS of one wide-kernel block, and its patterns per tile.
constexpr int kWideLdsBytes = 60 * 1024;
inline int wide_nt(int C, int S, int rb) {
  if (use_mfma(C, S, rb == 8 ? HX_F64 : HX_F32)) return 16;  // MFMA tiles of 16 patterns
  const int SP = S <= 32 ? 32 : 64;
  const int G = C * SP;
  int nt = kWideLdsBytes / (7 * G * rb);
  if (nt > 8) nt = 8;
  if (nt < 1) nt = 1;
  return nt;
}

// S == 4 uses tiles of 64 patterns (a wave per category, a lane per pattern);
// other S use tiles of wide_nt() patterns handled by (child, category, state)
// lanes.
// The chunking and the kernels depend on N, S, the precision and the tile shape only, never
// on K: a sample's
// sums run in the same order whatever batch it is evaluated in, so its result is bitwise
// identical alone and in a batch.
constexpr int kTargetChunks = 128;

// The four-state FP32 pre-order walk with at most kPreWalkMaxC categories
// sums its branch statistics over the kPreWalkW tiles of a workgroup: one
// chunk per kPreWalkW tiles. It serves every such call, also those with too
// few workgroups (ceil(tiles / 4) x K) to fill the chip, where a pre-order by
// level would be faster: choosing by size would make a sample's FP32 gradient
// depend on the batch it is evaluated in.
constexpr int kPreWalkW = 4;
constexpr int kPreWalkMaxC = 4;
inline bool use_pre_walk(int S, int C, int rb, int N, int K) {
  (void)N;
  (void)K;
  return S == 4 && rb == 4 && C <= kPreWalkMaxC;
}

inline Tiling make_tiling(int N, int C, int K, int S, int rb) {
  Tiling t;
  t.cp = next_pow2(C < 1 ? 1 : C);
  int min_tiles = 1;
  if (S == 4) {
    t.ppb = kWave;  // one wave per category, one lane per pattern
  } else {
    t.ppb = wide_nt(C, S, rb);
    // FP32 at 17..64 states: chunks of at least 8 tiles, since each chunk's d/dP partials
    // (C S^2 floats per branch) and its P set-up weigh as much as several tiles of partials
    // (128 taxa, 10k patterns, 61 states: at K = 32 the step 144 -> 128 ms, at K = 1 the
    // pre-order +2 %; the 200-request serving mix, whose smaller amino-acid slices get larger
    // chunks: +14 % served throughput)
    min_tiles = (S > 16 && rb == 4) ? 8 : 4;
  }
  t.ntiles = (N + t.ppb - 1) / t.ppb;
  if (use_pre_walk(S, C, rb, N, K)) {
    t.tiles_per_chunk = kPreWalkW;
    t.nchunks = (t.ntiles + kPreWalkW - 1) / kPreWalkW;
    return t;
  }
  int cap = t.ntiles / min_tiles;
  if (cap < 1) cap = 1;
  const int nch = cap < kTargetChunks ? cap : kTargetChunks;
  t.tiles_per_chunk = (t.ntiles + nch - 1) / nch;
  t.nchunks = (t.ntiles + t.tiles_per_chunk - 1) / t.tiles_per_chunk;
  return t;
}

// FP32 gradients at 33..64 states read P and P^T from zero-padded 64 x 64 copies.
inline bool use_ppad(int S, int dtype, int grad_mode) {
  return grad_mode != 0 && dtype != HX_F64 && S > 32 && S <= 64;
}

// Workspace of one hx_loglik* call: byte offsets of its regions.
struct WsLayout {
  size_t partials, tail, scales, rootsums, gpart, ppad, pcodes, total;  // tail: padding after the partials
};

// grad_mode: 0 loglik only, 1 d loglik / d mats, 2 d loglik / d tau.
// The layout depends on the tree only through its taxon count.
inline WsLayout ws_layout_t(int taxa, int S, int N, int C, int K, int dtype, int grad_mode) {
  const size_t rb = dtype == HX_F64 ? 8 : 4;
  const Tiling t = make_tiling(N, C, K, S, (int)rb);
  const size_t I = (size_t)taxa - 1, B = 2 * (size_t)taxa - 2;
  WsLayout w;
  size_t off = 0;
  w.partials = off;
  const size_t np = S == 4 ? (size_t)t.ntiles * 64 : (size_t)N;
  // dense states: the pre-order reads up to 16 C S + 64 elements past the
  // last (node, sample) slab (rows of the last tile past N, padded states)
  w.tail = off + I * K * np * C * S * rb;
  const size_t tail = S == 4 ? 0 : (16 * (size_t)C * S + 64) * rb;
  off = align_up(w.tail + tail, 256);
  w.scales = off;  // S = 4: 64-byte-aligned copy of the tip codes [T][ntiles * 64]
  if (S == 4) off = align_up(off + (size_t)taxa * t.ntiles * 64, 256);
  else off = align_up(off + I * K * (size_t)N * 4, 256);
  w.rootsums = off;
  off = align_up(off + (size_t)K * t.ntiles * (1 + S + C) * 8, 256);
  w.gpart = off;
  if (grad_mode == 1) off = align_up(off + B * K * (size_t)t.nchunks * C * S * S * rb, 256);
  if (grad_mode == 2) off = align_up(off + B * K * (size_t)t.nchunks * C * 8, 256);
  w.ppad = off;
  if (use_ppad(S, dtype, grad_mode)) off = align_up(off + B * K * (size_t)C * 2 * 64 * 64 * rb, 256);
  // four-state FP32: 16-bit tip codes and cherry code pairs for the pre-order walk (a tree
  // has at most T / 2 cherries)
  w.pcodes = off;
  if (S == 4 && rb == 4)
    off = align_up(off + ((size_t)taxa + (size_t)taxa / 2) * t.ntiles * 64 * 2, 256);
  w.total = off;
  return w;
}

// hx_workspace_bytes for a tree of `taxa` taxa (the layout needs no plan).
inline size_t workspace_bytes_for(int taxa, int S, int N, int C, int K, int dtype, int grad_mode) {
  return ws_layout_t(taxa, S, N, C, K, dtype, grad_mode).total;
}

// Register mode of the post-order walk's drop bits (walk.hpp) for a call, or -1 when every
// internal partial is stored: cherries are dropped for four-state FP32 calls with tip codes
// whose gradient runs as the pre-order walk (use_pre_walk), the only reader that rebuilds
// them. Mode 0: the post-order walk keeps no stack depths in registers (mode 1 is the
// program's mode for a walk that keeps depths 2..4 there).
inline int walk_drop_mode(int S, int C, int N, int K, int dtype, int tip_kind) {
  if (S != 4 || dtype == HX_F64 || tip_kind != HX_TIPS_CODES || !use_pre_walk(S, C, 4, N, K))
    return -1;
  return 0;
}

// Workspace of one resident-job evaluation (hx_job_eval): P (and dP/dtau), one pattern
// slice's pruning workspace, and the per-request sums.
struct JobWs {
  size_t mats, mats_dt, prune, prune_bytes, loglik_tmp, dtau_tmp, dtau_acc, total;
};

inline JobWs job_ws_shape(int taxa, int S, int slice, int C, int K, int dtype, int grad) {
  const size_t rb = dtype == HX_F64 ? 8 : 4;
  const size_t B = 2 * (size_t)taxa - 2;
  const size_t nm = (size_t)K * B * C * S * S;
  const size_t nt = (size_t)K * B * C;
  JobWs w{};
  size_t off = 0;
  w.mats = off;
  off = align_up(off + nm * rb, 256);
  if (grad) {
    w.mats_dt = off;
    off = align_up(off + nm * rb, 256);
  }
  w.prune = off;
  w.prune_bytes = workspace_bytes_for(taxa, S, slice, C, K, dtype, grad ? HX_GRAD_TAU : HX_GRAD_NONE);
  off = align_up(off + w.prune_bytes, 256);
  w.loglik_tmp = off;
  off = align_up(off + (size_t)K * 8, 256);
  if (grad) {
    w.dtau_tmp = off;
    off = align_up(off + nt * 8, 256);
    w.dtau_acc = off;
    off = align_up(off + nt * 8, 256);
  }
  w.total = off;
  return w;
}

}  // namespace hx

// Shared helpers for the hiptree HIP sources (gfx950, wave64).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>
#include <stddef.h>

#include <string>
#include <vector>

#include "hiptree.h"

namespace hx {

// Last error detail for the calling host thread.
void set_error(const std::string& msg);
const std::string& get_error();

#define HX_HIP_TRY(expr)                                                        \
  do {                                                                          \
    hipError_t _e = (expr);                                                     \
    if (_e != hipSuccess) {                                                     \
      ::hx::set_error(std::string(#expr " -> ") + hipGetErrorString(_e));      \
      return HX_ERR_HIP;                                                        \
    }                                                                           \
  } while (0)

#define HX_REQUIRE(cond, code, msg)                                             \
  do {                                                                          \
    if (!(cond)) {                                                              \
      ::hx::set_error(msg);                                                     \
      return (code);                                                            \
    }                                                                           \
  } while (0)

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

int wide_nt(int C, int S, int rb);

// S == 4 uses tiles of 64 patterns (a wave per category, a lane per pattern);
// other S use tiles of wide_nt() patterns handled by (child, category, state)
// lanes.
// The chunking depends on N (and the tile shape) only, never on K, except
// for the choice of the four-state FP32 pre-order walk (use_pre_walk): every
// other sample's sums run in the same order whatever batch it is evaluated
// in, so its result is bitwise identical alone and in a batch.
constexpr int kTargetChunks = 128;

// The four-state FP32 pre-order walk with at most kPreWalkMaxC categories
// sums its branch statistics over the kPreWalkW tiles of a workgroup: one
// chunk per kPreWalkW tiles. It walks the tree's entries one after another,
// so it is used when it has at least kPreWalkMinGroups workgroups
// (ceil(tiles / 4) x K) to fill the chip; smaller calls run the pre-order
// by level, which is parallel over each level's nodes (measured crossover).
// The choice depends on K, so an FP32 four-state gradient may differ in its
// last bits between batch sizes on the two sides of the threshold.
constexpr int kPreWalkW = 4;
constexpr int kPreWalkMaxC = 4;
constexpr long kPreWalkMinGroups = 256;
inline bool use_pre_walk(int S, int C, int rb, int N, int K) {
  const long groups = (long)((N + 63) / 64 + kPreWalkW - 1) / kPreWalkW * K;
  return S == 4 && rb == 4 && C <= kPreWalkMaxC && groups >= kPreWalkMinGroups;
}

// hx_workspace_bytes for a tree of `taxa` taxa (the layout needs no plan);
// defined in pruning.hip.
size_t workspace_bytes_for(int taxa, int S, int N, int C, int K, int dtype, int grad_mode);

inline Tiling make_tiling(int N, int C, int K, int S, int rb) {
  Tiling t;
  t.cp = next_pow2(C < 1 ? 1 : C);
  int min_tiles = 1;
  if (S == 4) {
    t.ppb = kWave;  // one wave per category, one lane per pattern
  } else {
    t.ppb = wide_nt(C, S, rb);
    min_tiles = 4;
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

}  // namespace hx

// Opaque plan: one topology with its device-resident schedules.
struct hx_plan_s {
  int device = 0;
  int T = 0;
  int I = 0;
  int root = 0;
  int32_t* d_buf = nullptr;   // device: kids[2(2T-1)] | post_nodes[I] | pre_nodes[I] | pad |
                              //         walk[I][8] (16-byte aligned) | segtab[nseg][16] |
                              //         prew[I][16] | cher[ncher][2]
  const int32_t* d_kids = nullptr;
  const int32_t* d_post = nullptr;
  const int32_t* d_pre = nullptr;
  std::vector<int> post_off;  // host: level offsets into d_post (size levels+1)
  std::vector<int> pre_off;   // host: depth offsets into d_pre
  std::vector<int32_t> kids;  // host copy of kids
  // Post-order walk program (walk.hpp): kWalkInts ints per internal node and
  // the tip-segment table; the walk needs `walk_need` stack entries.
  const int32_t* d_walk = nullptr;
  const int32_t* d_segtab = nullptr;
  int nseg = 0;
  int walk_need = 0;
  // Pre-order walk program (walk.hpp, build_pre_walk): kPreInts ints per
  // internal node; its stack reaches `prew_depth` entries.
  const int32_t* d_prew = nullptr;
  int prew_depth = 0;
  // Cherries (walk.hpp, list_cherries): [ncher][2] tip ids.
  const int32_t* d_cher = nullptr;
  int ncher = 0;
};

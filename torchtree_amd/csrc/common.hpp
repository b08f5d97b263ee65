// Shared helpers for the hiptree HIP sources (gfx950, wave64).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>
#include <stddef.h>

#include <string>
#include <vector>

#include "hiptree.h"
#include "layout.hpp"

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
  // Cherries (walk.hpp, list_cherries): [ncher][2] tip ids; how many the post-order walk
  // drops under register mode 0 and 1.
  const int32_t* d_cher = nullptr;
  int ncher = 0;
  int ndrop[2] = {0, 0};
};

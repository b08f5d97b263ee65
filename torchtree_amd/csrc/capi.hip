// C ABI: status strings, topology plans and workspace sizing.
#include "common.hpp"
#include "walk.hpp"

#include <algorithm>
#include <string>
#include <vector>

namespace hx {

static thread_local std::string g_last_error;

void set_error(const std::string& msg) { g_last_error = msg; }
const std::string& get_error() { return g_last_error; }

}  // namespace hx

extern "C" {

const char* hx_error_string(int status) {
  switch (status) {
    case HX_OK: return "ok";
    case HX_ERR_INVALID_ARGUMENT: return "invalid argument";
    case HX_ERR_UNSUPPORTED: return "unsupported configuration";
    case HX_ERR_HIP: return "HIP runtime error";
    case HX_ERR_WORKSPACE_TOO_SMALL: return "workspace too small";
    case HX_ERR_BAD_TOPOLOGY: return "bad topology";
    case HX_ERR_OUT_OF_MEMORY: return "out of device memory";
    case HX_ERR_BUDGET: return "request exceeds the memory budget";
    case HX_ERR_NOT_FOUND: return "not found";
    default: return "unknown status";
  }
}

const char* hx_last_error(void) { return hx::get_error().c_str(); }

int hx_version(void) { return 100; }

// Builds the post-order height levels and pre-order depth levels of a rooted
// binary tree given as (node, left, right) triples in post-order, the
// convention of AbstractTreeModel.update_traversals (tree_model.py:187-195):
// tips 0..T-1, internal nodes T..2T-2, root 2T-2 last.
int hx_plan_create(int taxa, const int32_t* postorder, int device, hx_plan_t* out) {
  HX_REQUIRE(out != nullptr && postorder != nullptr, HX_ERR_INVALID_ARGUMENT, "null argument");
  HX_REQUIRE(taxa >= 2, HX_ERR_BAD_TOPOLOGY, "need at least two taxa");
  const int T = taxa, I = T - 1, nn = 2 * T - 1;
  std::vector<int32_t> kids(2 * nn, -1);
  std::vector<int> parent(nn, -1);
  std::vector<char> seen_parent(nn, 0);
  for (int r = 0; r < I; ++r) {
    int v = postorder[3 * r], a = postorder[3 * r + 1], b = postorder[3 * r + 2];
    if (v < T || v >= nn || a < 0 || a >= nn || b < 0 || b >= nn || a == b) {
      hx::set_error("triple " + std::to_string(r) + " out of range");
      return HX_ERR_BAD_TOPOLOGY;
    }
    if (seen_parent[v] || parent[a] >= 0 || parent[b] >= 0) {
      hx::set_error("node repeated in post-order triples");
      return HX_ERR_BAD_TOPOLOGY;
    }
    if ((a >= T && !seen_parent[a]) || (b >= T && !seen_parent[b])) {
      hx::set_error("child visited after its parent");
      return HX_ERR_BAD_TOPOLOGY;
    }
    seen_parent[v] = 1;
    parent[a] = v;
    parent[b] = v;
    kids[2 * v] = a;
    kids[2 * v + 1] = b;
  }
  const int root = postorder[3 * (I - 1)];
  HX_REQUIRE(root == nn - 1, HX_ERR_BAD_TOPOLOGY, "root must be node 2T-2");
  for (int v = 0; v < nn; ++v) {
    if (v != root && parent[v] < 0) {
      hx::set_error("node " + std::to_string(v) + " has no parent");
      return HX_ERR_BAD_TOPOLOGY;
    }
  }
  // heights (children first) and depths (parents first)
  std::vector<int> height(nn, 0), depth(nn, 0);
  for (int r = 0; r < I; ++r) {
    int v = postorder[3 * r];
    height[v] = 1 + std::max(height[kids[2 * v]], height[kids[2 * v + 1]]);
  }
  for (int r = I - 1; r >= 0; --r) {
    int v = postorder[3 * r];
    depth[kids[2 * v]] = depth[v] + 1;
    depth[kids[2 * v + 1]] = depth[v] + 1;
  }
  const int H = height[root];
  int D = 0;
  for (int r = 0; r < I; ++r) D = std::max(D, depth[postorder[3 * r]] + 1);
  std::vector<std::vector<int32_t>> post_lv(H), pre_lv(D);
  for (int r = 0; r < I; ++r) {
    int v = postorder[3 * r];
    post_lv[height[v] - 1].push_back(v);
    pre_lv[depth[v]].push_back(v);
  }
  std::vector<int32_t> walk, segtab;
  std::vector<uint8_t> drop;
  int walk_need = 0;
  if (!hx::build_walk(T, kids, postorder, walk, segtab, walk_need, &drop)) {
    hx::set_error("walk program construction failed");
    return HX_ERR_BAD_TOPOLOGY;
  }
  std::vector<int32_t> cher, cher_row;
  hx::list_cherries(T, kids, postorder, cher, cher_row);
  std::vector<int32_t> prew;
  int prew_depth = 0;
  if (!hx::build_pre_walk(T, kids, postorder, prew, prew_depth, &drop, &cher_row)) {
    hx::set_error("pre-order walk program construction failed");
    return HX_ERR_BAD_TOPOLOGY;
  }
  auto* plan = new hx_plan_s();
  plan->device = device;
  plan->walk_need = walk_need;
  plan->T = T;
  plan->I = I;
  plan->root = root;
  plan->kids = kids;
  const size_t walk_off = hx::align_up(2 * (size_t)nn + 2 * (size_t)I, 4);
  const size_t seg_off = walk_off + walk.size();
  const size_t prew_off = hx::align_up(seg_off + segtab.size(), 4);
  const size_t cher_off = prew_off + prew.size();
  std::vector<int32_t> host(cher_off + cher.size(), 0);
  std::copy(walk.begin(), walk.end(), host.begin() + walk_off);
  std::copy(segtab.begin(), segtab.end(), host.begin() + seg_off);
  std::copy(prew.begin(), prew.end(), host.begin() + prew_off);
  std::copy(cher.begin(), cher.end(), host.begin() + cher_off);
  std::copy(kids.begin(), kids.end(), host.begin());
  int pos = 2 * nn;
  plan->post_off.push_back(0);
  for (auto& lv : post_lv) {
    for (int v : lv) host[pos++] = v;
    plan->post_off.push_back(pos - 2 * nn);
  }
  plan->pre_off.push_back(0);
  for (auto& lv : pre_lv) {
    for (int v : lv) host[pos++] = v;
    plan->pre_off.push_back(pos - 2 * nn - I);
  }
  int prev = 0;
  hipError_t e = hipGetDevice(&prev);
  if (e == hipSuccess) e = hipSetDevice(device);
  if (e == hipSuccess) e = hipMalloc(&plan->d_buf, host.size() * sizeof(int32_t));
  if (e == hipSuccess)
    e = hipMemcpy(plan->d_buf, host.data(), host.size() * sizeof(int32_t), hipMemcpyHostToDevice);
  hipSetDevice(prev);
  if (e != hipSuccess) {
    hx::set_error(std::string("plan upload: ") + hipGetErrorString(e));
    if (plan->d_buf) hipFree(plan->d_buf);
    delete plan;
    return HX_ERR_HIP;
  }
  plan->d_kids = plan->d_buf;
  plan->d_post = plan->d_buf + 2 * nn;
  plan->d_pre = plan->d_buf + 2 * nn + I;
  plan->d_walk = plan->d_buf + walk_off;
  plan->d_segtab = plan->d_buf + seg_off;
  plan->nseg = (int)(segtab.size() / hx::kWalkSegTips);
  plan->d_prew = plan->d_buf + prew_off;
  plan->prew_depth = prew_depth;
  plan->d_cher = plan->d_buf + cher_off;
  plan->ncher = (int)(cher.size() / 2);
  for (uint8_t b : drop) {
    plan->ndrop[0] += b & 1;
    plan->ndrop[1] += (b >> 1) & 1;
  }
  *out = plan;
  return HX_OK;
}

int hx_plan_destroy(hx_plan_t plan) {
  if (!plan) return HX_OK;
  if (plan->d_buf) hipFree(plan->d_buf);
  delete plan;
  return HX_OK;
}

int hx_plan_walk_info(hx_plan_t plan, int* post_need, int* pre_depth, int* cherries,
                      int* dropped_mode0, int* dropped_mode1) {
  HX_REQUIRE(plan != nullptr, HX_ERR_INVALID_ARGUMENT, "null plan");
  if (post_need) *post_need = plan->walk_need;
  if (pre_depth) *pre_depth = plan->prew_depth;
  if (cherries) *cherries = plan->ncher;
  if (dropped_mode0) *dropped_mode0 = plan->ndrop[0];
  if (dropped_mode1) *dropped_mode1 = plan->ndrop[1];
  return HX_OK;
}

int hx_plan_info(hx_plan_t plan, int* post_levels, int* pre_levels, int* max_width) {
  HX_REQUIRE(plan != nullptr, HX_ERR_INVALID_ARGUMENT, "null plan");
  if (post_levels) *post_levels = (int)plan->post_off.size() - 1;
  if (pre_levels) *pre_levels = (int)plan->pre_off.size() - 1;
  if (max_width) {
    int w = 0;
    for (size_t i = 0; i + 1 < plan->post_off.size(); ++i)
      w = std::max(w, plan->post_off[i + 1] - plan->post_off[i]);
    for (size_t i = 0; i + 1 < plan->pre_off.size(); ++i)
      w = std::max(w, plan->pre_off[i + 1] - plan->pre_off[i]);
    *max_width = w;
  }
  return HX_OK;
}

}  // extern "C"

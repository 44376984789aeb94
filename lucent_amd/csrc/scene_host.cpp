// Host-side scene preparation: validation of the reference's packed arrays
// and their conversion into the device layout described in internal.hpp.
#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "internal.hpp"

namespace lucent {
namespace {

bool integral_field(float f, uint32_t limit) {
  return f >= 0.0f && f <= float(limit) && std::floor(f) == f;
}

}  // namespace

// Re-indexes the reference's node array (src/math/bvh.ts:6-16: children of an
// interior node are the adjacent pair leftFirst, leftFirst + 1) into one
// 64-byte record per interior node, numbered in depth-first preorder with the
// left child first so that a near-first descent walks mostly forward.
Status convert_bvh(const float* nodes, uint32_t node_count, uint32_t tri_count, HostBvhImage* out) {
  *out = HostBvhImage{};
  if (tri_count == 0) return LUCENT_OK;  // nothing to traverse: every query misses
  if (!nodes || node_count == 0) return fail(LUCENT_E_INVALID, "BVH has no nodes");
  for (uint32_t n = 0; n < node_count; ++n) {
    const float* o = nodes + size_t(n) * 8;
    if (!integral_field(o[3], 1u << 24) || !integral_field(o[7], 1u << 24))
      return fail(LUCENT_E_INVALID, "node " + std::to_string(n) + " has a non-integral leftFirst/count");
    const uint32_t first = uint32_t(o[3]), count = uint32_t(o[7]);
    if (count > 0) {
      if (uint64_t(first) + count > tri_count)
        return fail(LUCENT_E_INVALID, "leaf " + std::to_string(n) + " indexes past the triangle count");
    } else if (node_count > 1) {
      if (first <= n || uint64_t(first) + 1 >= node_count)
        return fail(LUCENT_E_INVALID, "interior node " + std::to_string(n) + " has out-of-range children");
    }
  }
  const float* root = nodes;
  if (root[7] > 0 || node_count == 1) {
    out->root_count = uint32_t(root[7]);
    out->root_link = uint32_t(root[3]);
    if (out->root_count == 0) return fail(LUCENT_E_INVALID, "single-node BVH over triangles must be a leaf");
    return LUCENT_OK;
  }
  std::vector<int64_t> pair_of(node_count, -1);
  std::vector<uint32_t> depth_of(node_count, 0);
  // Preorder numbering of interior nodes.
  std::vector<uint32_t> work{0};
  std::vector<uint32_t> preorder;
  preorder.reserve(node_count / 2 + 1);
  uint32_t max_depth = 0;
  while (!work.empty()) {
    const uint32_t n = work.back();
    work.pop_back();
    if (pair_of[n] >= 0) return fail(LUCENT_E_INVALID, "BVH node reached twice (not a tree)");
    pair_of[n] = int64_t(preorder.size());
    preorder.push_back(n);
    const uint32_t first = uint32_t(nodes[size_t(n) * 8 + 3]);
    for (int c = 1; c >= 0; --c) {
      const uint32_t child = first + uint32_t(c);
      depth_of[child] = depth_of[n] + 1;
      if (depth_of[child] > max_depth) max_depth = depth_of[child];
      if (nodes[size_t(child) * 8 + 7] == 0) work.push_back(child);
    }
  }
  out->pair_count = uint32_t(preorder.size());
  out->pairs.assign(size_t(out->pair_count) * 16, 0.0f);
  for (uint32_t p = 0; p < out->pair_count; ++p) {
    const uint32_t n = preorder[p];
    const uint32_t first = uint32_t(nodes[size_t(n) * 8 + 3]);
    float* rec = out->pairs.data() + size_t(p) * 16;
    uint32_t links[4];
    for (int c = 0; c < 2; ++c) {
      const float* ch = nodes + size_t(first + c) * 8;
      for (int a = 0; a < 3; ++a) {
        rec[a * 4 + c] = ch[a];          // lo
        rec[a * 4 + 2 + c] = ch[4 + a];  // hi
      }
      const uint32_t count = uint32_t(ch[7]);
      links[c * 2] = count > 0 ? uint32_t(ch[3]) : uint32_t(pair_of[first + c]);
      links[c * 2 + 1] = count;
    }
    std::memcpy(rec + 12, links, sizeof(links));
  }
  out->root_link = 0;
  out->root_count = 0;
  out->depth = max_depth;
  return LUCENT_OK;
}

}  // namespace lucent

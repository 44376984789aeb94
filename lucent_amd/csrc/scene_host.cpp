// Host-side scene preparation: validation of the reference's packed arrays
// and their conversion into the device layout described in internal.hpp.
#include <algorithm>
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

// Checks the reference's node array (src/math/bvh.ts:6-16: children of an
// interior node are the adjacent pair leftFirst, leftFirst + 1; leaves hold
// triangles [leftFirst, leftFirst + count) of triIndices order).
Status validate_bvh(const float* nodes, uint32_t node_count, uint32_t tri_count, uint32_t* depth) {
  *depth = 0;
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
  if (nodes[7] > 0 || node_count == 1) {
    if (nodes[7] == 0) return fail(LUCENT_E_INVALID, "single-node BVH over triangles must be a leaf");
    return LUCENT_OK;
  }
  std::vector<uint8_t> seen(node_count, 0);
  std::vector<std::pair<uint32_t, uint32_t>> work{{0u, 0u}};
  while (!work.empty()) {
    const auto [n, d] = work.back();
    work.pop_back();
    if (seen[n]) return fail(LUCENT_E_INVALID, "BVH node reached twice (not a tree)");
    seen[n] = 1;
    *depth = std::max(*depth, d);
    if (nodes[size_t(n) * 8 + 7] > 0) continue;
    const uint32_t first = uint32_t(nodes[size_t(n) * 8 + 3]);
    work.push_back({first, d + 1});
    work.push_back({first + 1, d + 1});
  }
  return LUCENT_OK;
}

// Collapses the reference's binary tree (already checked by validate_bvh)
// into 4-wide nodes for the path tracer.  Each 4-wide node replaces a binary
// interior node N: its children are N's two children, where the interior one
// with the largest surface area is repeatedly replaced by its own two
// children while fewer than 4 are collected.  Child boxes are stored as 8-bit
// codes in a per-node frame (origin = the children's minimum corner, a power
// of two scale per axis) and decoded on the device as fmaf(code, scale,
// origin); the codes are chosen with that same float expression so that every
// decoded box contains the child's true box (rounded outward).
Status convert_bvh4(const float* nodes, uint32_t node_count, uint32_t tri_count, HostBvh4Image* out) {
  *out = HostBvh4Image{};
  if (tri_count == 0) return LUCENT_OK;
  if (tri_count >= (1u << 26) - 1)
    return fail(LUCENT_E_UNSUPPORTED, "the path tracer's 4-wide nodes index at most 2^26 - 2 triangles");
  auto leaf_link = [&](uint32_t n, uint32_t* link) -> Status {
    const uint32_t first = uint32_t(nodes[size_t(n) * 8 + 3]), count = uint32_t(nodes[size_t(n) * 8 + 7]);
    if (count > 32) return fail(LUCENT_E_UNSUPPORTED, "leaves of more than 32 triangles are not supported by the path tracer");
    *link = 0x80000000u | ((count - 1u) << 26) | first;
    return LUCENT_OK;
  };
  if (nodes[7] > 0 || node_count == 1) return leaf_link(0, &out->root);
  auto area = [&](uint32_t n) {
    const float* o = nodes + size_t(n) * 8;
    const double dx = double(o[4]) - o[0], dy = double(o[5]) - o[1], dz = double(o[6]) - o[2];
    return dx * dy + dy * dz + dz * dx;
  };
  struct Item {
    uint32_t bin, node, depth;
  };
  std::vector<Item> work{{0u, 0u, 1u}};
  out->node_count = 1;
  std::vector<float> rec;
  out->nodes.clear();
  uint32_t max_depth = 1;
  while (!work.empty()) {
    const Item it = work.back();
    work.pop_back();
    uint32_t ch[4];
    int n = 0;
    const uint32_t first = uint32_t(nodes[size_t(it.bin) * 8 + 3]);
    ch[n++] = first;
    ch[n++] = first + 1;
    while (n < 4) {
      int pick = -1;
      double best = -1.0;
      for (int i = 0; i < n; ++i)
        if (nodes[size_t(ch[i]) * 8 + 7] == 0 && area(ch[i]) > best) {
          best = area(ch[i]);
          pick = i;
        }
      if (pick < 0) break;
      const uint32_t f = uint32_t(nodes[size_t(ch[pick]) * 8 + 3]);
      for (int i = n; i > pick + 1; --i) ch[i] = ch[i - 1];
      ch[pick] = f;
      ch[pick + 1] = f + 1;
      ++n;
    }
    // Frame: the children's bounding box.
    float lo[3], hi[3];
    for (int a = 0; a < 3; ++a) {
      lo[a] = nodes[size_t(ch[0]) * 8 + a];
      hi[a] = nodes[size_t(ch[0]) * 8 + 4 + a];
      for (int i = 1; i < n; ++i) {
        lo[a] = std::min(lo[a], nodes[size_t(ch[i]) * 8 + a]);
        hi[a] = std::max(hi[a], nodes[size_t(ch[i]) * 8 + 4 + a]);
      }
    }
    float scale[3];
    uint32_t qlo[3] = {0, 0, 0}, qhi[3] = {0, 0, 0};  // 4 x 8-bit codes per axis
    for (int a = 0; a < 3; ++a) {
      const double ext = double(hi[a]) - double(lo[a]);
      int e = ext > 0 ? int(std::ceil(std::log2(ext / 255.0))) : -126;
      e = std::max(-126, std::min(127, e));
      for (;; ++e) {
        if (e > 127) return fail(LUCENT_E_UNSUPPORTED, "node extent too large to quantise");
        const float sc = std::ldexp(1.0f, e);
        bool ok = true;
        uint32_t ql = 0, qh = 0;
        for (int i = 0; i < n && ok; ++i) {
          const float clo = nodes[size_t(ch[i]) * 8 + a], chi = nodes[size_t(ch[i]) * 8 + 4 + a];
          double fl = std::floor((double(clo) - lo[a]) / sc), fh = std::ceil((double(chi) - lo[a]) / sc);
          int cl = int(std::max(0.0, std::min(255.0, fl))), chh = int(std::max(0.0, std::min(255.0, fh)));
          while (cl > 0 && std::fmaf(float(cl), sc, lo[a]) > clo) --cl;
          if (std::fmaf(float(cl), sc, lo[a]) > clo) ok = false;
          while (chh < 255 && std::fmaf(float(chh), sc, lo[a]) < chi) ++chh;
          if (std::fmaf(float(chh), sc, lo[a]) < chi) ok = false;
          ql |= uint32_t(cl) << (8 * i);
          qh |= uint32_t(chh) << (8 * i);
        }
        if (ok) {
          scale[a] = sc;
          qlo[a] = ql;
          qhi[a] = qh;
          break;
        }
      }
    }
    uint32_t links[4] = {0xFFFFFFFFu, 0xFFFFFFFFu, 0xFFFFFFFFu, 0xFFFFFFFFu};
    for (int i = 0; i < n; ++i) {
      if (nodes[size_t(ch[i]) * 8 + 7] > 0) {
        Status s = leaf_link(ch[i], &links[i]);
        if (s != LUCENT_OK) return s;
      } else {
        links[i] = out->node_count++;
        work.push_back({ch[i], links[i], it.depth + 1});
        max_depth = std::max(max_depth, it.depth + 1);
      }
    }
    if (out->nodes.size() < size_t(out->node_count) * 16) out->nodes.resize(size_t(out->node_count) * 16 * 2, 0.0f);
    float* r = out->nodes.data() + size_t(it.node) * 16;
    r[0] = lo[0];
    r[1] = lo[1];
    r[2] = lo[2];
    r[3] = scale[0];
    r[4] = scale[1];
    r[5] = scale[2];
    std::memcpy(r + 6, &qlo[0], 4);
    std::memcpy(r + 7, &qhi[0], 4);
    std::memcpy(r + 8, &qlo[1], 4);
    std::memcpy(r + 9, &qhi[1], 4);
    std::memcpy(r + 10, &qlo[2], 4);
    std::memcpy(r + 11, &qhi[2], 4);
    std::memcpy(r + 12, links, sizeof(links));
    if (out->nodes_f32.size() < size_t(out->node_count) * 32)
      out->nodes_f32.resize(size_t(out->node_count) * 32 * 2, 0.0f);
    float* f = out->nodes_f32.data() + size_t(it.node) * 32;
    f[0] = lo[0];
    f[1] = lo[1];
    f[2] = lo[2];
    f[3] = 0.0f;
    for (int a = 0; a < 3; ++a) {
      const double margin = (double(hi[a]) - double(lo[a])) * 0x1p-12;
      for (int i = 0; i < 4; ++i) {
        float ol = 0.0f, oh = 0.0f;
        if (i < n) {
          const double dl = double(nodes[size_t(ch[i]) * 8 + a]) - lo[a] - margin;
          const double dh = double(nodes[size_t(ch[i]) * 8 + 4 + a]) - lo[a] + margin;
          ol = float(dl);
          if (double(ol) > dl) ol = std::nextafter(ol, -INFINITY);
          oh = float(dh);
          if (double(oh) < dh) oh = std::nextafter(oh, INFINITY);
        }
        f[4 + 8 * a + i] = ol;
        f[8 + 8 * a + i] = oh;
      }
    }
    std::memcpy(f + 28, links, sizeof(links));
  }
  out->nodes.resize(size_t(out->node_count) * 16);
  out->nodes_f32.resize(size_t(out->node_count) * 32);
  out->root = 0;
  out->depth = max_depth;
  return LUCENT_OK;
}

}  // namespace lucent

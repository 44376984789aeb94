// GPU BVH builder: a linear BVH (Karras, "Maximizing Parallelism in the
// Construction of BVHs, Octrees, and k-d Trees", HPG 2012) emitted in the
// reference's node layout (src/math/bvh.ts:6-16), so that it drops in for
// buildBVH (bvh.ts:52) wherever its output is consumed.
//
// Device: triangle centroids -> 63-bit Morton codes in the centroid box ->
// radix sort of (code, triangle) pairs -> one internal node per adjacent
// pair of sorted keys, its range and split found from the longest common
// prefixes of the keys (ties broken by position) -> bounding boxes and
// triangle counts bottom-up, each internal node finished by the second of
// its children to arrive.
// Host: the radix tree is walked from the root; a node covering at most
// max_leaf triangles becomes a leaf over its contiguous range of sorted
// triangles, and nodes are numbered as the reference numbers them (root 0,
// the two children of an interior node adjacent, left first), so
// triIndices is the sorted order itself.
//
// The tree is built by spatial order, not by surface area: its traversal
// cost is higher than the binned SAH builder's (reported by the tests and
// bench), its build is O(n log n) on the device instead of seconds on the
// host for millions of triangles.
#include <hip/hip_runtime.h>
#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <cmath>
#include <vector>

#include "internal.hpp"

namespace lucent {
namespace {

constexpr int kBlock = 256;

__device__ inline uint64_t spread21(uint64_t x) {
  x &= 0x1FFFFFu;
  x = (x | x << 32) & 0x1F00000000FFFFull;
  x = (x | x << 16) & 0x1F0000FF0000FFull;
  x = (x | x << 8) & 0x100F00F00F00F00Full;
  x = (x | x << 4) & 0x10C30C30C30C30C3ull;
  x = (x | x << 2) & 0x1249249249249249ull;
  return x;
}

// Morton code of each triangle's centroid in the centroid box.
__global__ void k_morton(const float* __restrict__ pos, uint32_t n, float3 lo, float3 scale,
                         uint64_t* __restrict__ keys, uint32_t* __restrict__ ids) {
  const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
  if (i >= n) return;
  const float* p = pos + size_t(i) * 9;
  const float cx = (p[0] + p[3] + p[6]) * (1.0f / 3.0f);
  const float cy = (p[1] + p[4] + p[7]) * (1.0f / 3.0f);
  const float cz = (p[2] + p[5] + p[8]) * (1.0f / 3.0f);
  auto q = [](float c, float l, float s) {
    const float v = (c - l) * s;
    return uint64_t(fminf(fmaxf(v, 0.0f), 2097151.0f));
  };
  keys[i] = spread21(q(cx, lo.x, scale.x)) << 2 | spread21(q(cy, lo.y, scale.y)) << 1 | spread21(q(cz, lo.z, scale.z));
  ids[i] = i;
}

// Length of the common prefix of sorted keys i and j (positions break ties);
// -1 outside [0, n).
__device__ inline int delta(const uint64_t* keys, int n, int i, int j) {
  if (j < 0 || j >= n) return -1;
  const uint64_t a = keys[i], b = keys[j];
  if (a != b) return __clzll(a ^ b);
  return 64 + __clz(uint32_t(i) ^ uint32_t(j));
}

// Internal node i of the radix tree: its children (bit 31 set = leaf index)
// and the range [first, last] of sorted triangles it covers.
__global__ void k_hierarchy(const uint64_t* __restrict__ keys, int n, uint32_t* __restrict__ child,
                            uint32_t* __restrict__ range, uint32_t* __restrict__ parent_int,
                            uint32_t* __restrict__ parent_leaf) {
  const int i = int(blockIdx.x * kBlock + threadIdx.x);
  if (i >= n - 1) return;
  const int d = (delta(keys, n, i, i + 1) - delta(keys, n, i, i - 1)) >= 0 ? 1 : -1;
  const int dmin = delta(keys, n, i, i - d);
  int lmax = 2;
  while (delta(keys, n, i, i + lmax * d) > dmin) lmax *= 2;
  int l = 0;
  for (int t = lmax / 2; t >= 1; t /= 2)
    if (delta(keys, n, i, i + (l + t) * d) > dmin) l += t;
  const int j = i + l * d;
  const int dnode = delta(keys, n, i, j);
  int s = 0;
  for (int t = (l + 1) / 2;; t = (t + 1) / 2) {
    if (delta(keys, n, i, i + (s + t) * d) > dnode) s += t;
    if (t == 1) break;
  }
  const int gamma = i + s * d + min(d, 0);
  const int first = min(i, j), last = max(i, j);
  const uint32_t left = (first == gamma) ? (0x80000000u | uint32_t(gamma)) : uint32_t(gamma);
  const uint32_t right = (last == gamma + 1) ? (0x80000000u | uint32_t(gamma + 1)) : uint32_t(gamma + 1);
  child[2 * i] = left;
  child[2 * i + 1] = right;
  range[2 * i] = uint32_t(first);
  range[2 * i + 1] = uint32_t(last);
  if (left & 0x80000000u)
    parent_leaf[left & 0x7FFFFFFFu] = uint32_t(i);
  else
    parent_int[left] = uint32_t(i);
  if (right & 0x80000000u)
    parent_leaf[right & 0x7FFFFFFFu] = uint32_t(i);
  else
    parent_int[right] = uint32_t(i);
}

// Boxes bottom-up: each leaf's box from its triangle; an internal node is
// completed by whichever child arrives second (the first only counts in).
__global__ void k_boxes(const float* __restrict__ pos, const uint32_t* __restrict__ ids, int n,
                        const uint32_t* __restrict__ child, const uint32_t* __restrict__ parent_int,
                        const uint32_t* __restrict__ parent_leaf, float* __restrict__ leaf_box,
                        float* __restrict__ int_box, uint32_t* __restrict__ arrivals) {
  const int i = int(blockIdx.x * kBlock + threadIdx.x);
  if (i >= n) return;
  const float* p = pos + size_t(ids[i]) * 9;
  float b[6];
  for (int a = 0; a < 3; ++a) {
    b[a] = fminf(fminf(p[a], p[3 + a]), p[6 + a]);
    b[3 + a] = fmaxf(fmaxf(p[a], p[3 + a]), p[6 + a]);
  }
  for (int a = 0; a < 6; ++a) leaf_box[size_t(i) * 6 + a] = b[a];
  if (n == 1) return;
  uint32_t node = parent_leaf[i];
  for (;;) {
    __threadfence();
    if (atomicAdd(&arrivals[node], 1u) == 0) return;  // the sibling finishes this node
    __threadfence();
    float u[6];
    for (int c = 0; c < 2; ++c) {
      const uint32_t ch = child[2 * node + c];
      const float* cb = (ch & 0x80000000u) ? leaf_box + size_t(ch & 0x7FFFFFFFu) * 6 : int_box + size_t(ch) * 6;
      // The sibling wrote its box before its release fence and atomic; the
      // fence above (acquire, agent scope) invalidated this CU's L1.
      float v[6];
      for (int a = 0; a < 6; ++a) v[a] = cb[a];
      if (c == 0) {
        for (int a = 0; a < 6; ++a) u[a] = v[a];
      } else {
        for (int a = 0; a < 3; ++a) {
          u[a] = fminf(u[a], v[a]);
          u[3 + a] = fmaxf(u[3 + a], v[3 + a]);
        }
      }
    }
    for (int a = 0; a < 6; ++a) int_box[size_t(node) * 6 + a] = u[a];
    if (node == 0) return;
    node = parent_int[node];
  }
}

__global__ void k_zero_u32(uint32_t* p, size_t n) {
  const size_t i = size_t(blockIdx.x) * kBlock + threadIdx.x;
  if (i < n) p[i] = 0;
}

}  // namespace

Status build_bvh_gpu(Context* ctx, const float* positions, uint32_t tri_count, uint32_t max_leaf, float* nodes_out,
                     uint32_t* node_count_out, uint32_t* tri_indices_out, double bounds_out[6]) {
  *node_count_out = 0;
  const uint32_t n = tri_count;
  if (n == 0) {
    for (int a = 0; a < 8; ++a) nodes_out[a] = 0.0f;
    *node_count_out = 1;
    if (bounds_out)
      for (int a = 0; a < 6; ++a) bounds_out[a] = 0.0;
    return LUCENT_OK;
  }
  if (max_leaf == 0) max_leaf = 4;
  // Centroid box (host, one pass) and triangle box.
  double clo[3] = {INFINITY, INFINITY, INFINITY}, chi[3] = {-INFINITY, -INFINITY, -INFINITY};
  double blo[3] = {INFINITY, INFINITY, INFINITY}, bhi[3] = {-INFINITY, -INFINITY, -INFINITY};
  for (uint32_t t = 0; t < n; ++t) {
    const float* p = positions + size_t(t) * 9;
    for (int a = 0; a < 3; ++a) {
      const double c = (double(p[a]) + p[3 + a] + p[6 + a]) / 3.0;
      clo[a] = std::min(clo[a], c);
      chi[a] = std::max(chi[a], c);
      for (int v = 0; v < 3; ++v) {
        blo[a] = std::min(blo[a], double(p[v * 3 + a]));
        bhi[a] = std::max(bhi[a], double(p[v * 3 + a]));
      }
    }
  }
  if (!(std::isfinite(blo[0]) && std::isfinite(blo[1]) && std::isfinite(blo[2]) && std::isfinite(bhi[0]) &&
        std::isfinite(bhi[1]) && std::isfinite(bhi[2])))
    return fail(LUCENT_E_INVALID, "positions must be finite");
  if (bounds_out)
    for (int a = 0; a < 3; ++a) {
      bounds_out[a] = blo[a];
      bounds_out[3 + a] = bhi[a];
    }
  float3 lo = make_float3(float(clo[0]), float(clo[1]), float(clo[2]));
  float3 scale;
  const double ext[3] = {chi[0] - clo[0], chi[1] - clo[1], chi[2] - clo[2]};
  scale.x = ext[0] > 0 ? float(2097151.0 / ext[0]) : 0.0f;
  scale.y = ext[1] > 0 ? float(2097151.0 / ext[1]) : 0.0f;
  scale.z = ext[2] > 0 ? float(2097151.0 / ext[2]) : 0.0f;

  // Device buffers.
  hipStream_t st = static_cast<hipStream_t>(ctx->stream);
  const size_t nn = n, ni = n > 1 ? n - 1 : 1;
  size_t sort_bytes = 0;
  hipError_t e = hipcub::DeviceRadixSort::SortPairs(nullptr, sort_bytes, static_cast<const uint64_t*>(nullptr),
                                                    static_cast<uint64_t*>(nullptr), static_cast<const uint32_t*>(nullptr),
                                                    static_cast<uint32_t*>(nullptr), int(n), 0, 64, st);
  if (e != hipSuccess) return fail_hip(e, "radix sort sizing");
  auto up = [](size_t b) { return (b + 255) / 256 * 256; };
  const size_t off_pos = 0, off_keys = off_pos + up(nn * 36), off_keys2 = off_keys + up(nn * 8),
               off_ids = off_keys2 + up(nn * 8), off_ids2 = off_ids + up(nn * 4), off_child = off_ids2 + up(nn * 4),
               off_range = off_child + up(ni * 8), off_pint = off_range + up(ni * 8), off_pleaf = off_pint + up(ni * 4),
               off_lbox = off_pleaf + up(nn * 4), off_ibox = off_lbox + up(nn * 24), off_arr = off_ibox + up(ni * 24),
               off_sort = off_arr + up(ni * 4), total = off_sort + up(sort_bytes);
  void* mem = nullptr;
  Status s = dev_alloc(&mem, total);
  if (s != LUCENT_OK) return s;
  char* m = static_cast<char*>(mem);
  float* d_pos = reinterpret_cast<float*>(m + off_pos);
  uint64_t* d_keys = reinterpret_cast<uint64_t*>(m + off_keys);
  uint64_t* d_keys2 = reinterpret_cast<uint64_t*>(m + off_keys2);
  uint32_t* d_ids = reinterpret_cast<uint32_t*>(m + off_ids);
  uint32_t* d_ids2 = reinterpret_cast<uint32_t*>(m + off_ids2);
  uint32_t* d_child = reinterpret_cast<uint32_t*>(m + off_child);
  uint32_t* d_range = reinterpret_cast<uint32_t*>(m + off_range);
  uint32_t* d_pint = reinterpret_cast<uint32_t*>(m + off_pint);
  uint32_t* d_pleaf = reinterpret_cast<uint32_t*>(m + off_pleaf);
  float* d_lbox = reinterpret_cast<float*>(m + off_lbox);
  float* d_ibox = reinterpret_cast<float*>(m + off_ibox);
  uint32_t* d_arr = reinterpret_cast<uint32_t*>(m + off_arr);
  void* d_sort = m + off_sort;
  auto grid = [](size_t k) { return dim3(uint32_t((k + kBlock - 1) / kBlock)); };

  s = dev_upload(d_pos, positions, nn * 36, st);
  if (s == LUCENT_OK) {
    hipLaunchKernelGGL(k_morton, grid(nn), dim3(kBlock), 0, st, d_pos, n, lo, scale, d_keys, d_ids);
    e = hipcub::DeviceRadixSort::SortPairs(d_sort, sort_bytes, d_keys, d_keys2, d_ids, d_ids2, int(n), 0, 64, st);
    if (e != hipSuccess) s = fail_hip(e, "radix sort");
  }
  if (s == LUCENT_OK) {
    if (n > 1) {
      hipLaunchKernelGGL(k_hierarchy, grid(ni), dim3(kBlock), 0, st, d_keys2, int(n), d_child, d_range, d_pint, d_pleaf);
      hipLaunchKernelGGL(k_zero_u32, grid(ni), dim3(kBlock), 0, st, d_arr, ni);
    }
    hipLaunchKernelGGL(k_boxes, grid(nn), dim3(kBlock), 0, st, d_pos, d_ids2, int(n), d_child, d_pint, d_pleaf, d_lbox,
                       d_ibox, d_arr);
    if ((e = hipGetLastError()) != hipSuccess) s = fail_hip(e, "BVH build launch");
  }
  std::vector<uint32_t> child(ni * 2), range(ni * 2), ids(nn);
  std::vector<float> lbox(nn * 6), ibox(ni * 6);
  if (s == LUCENT_OK) s = dev_download(ids.data(), d_ids2, nn * 4, st);
  if (s == LUCENT_OK) s = dev_download(lbox.data(), d_lbox, nn * 24, st);
  if (s == LUCENT_OK && n > 1) {
    if ((s = dev_download(child.data(), d_child, ni * 8, st)) == LUCENT_OK &&
        (s = dev_download(range.data(), d_range, ni * 8, st)) == LUCENT_OK)
      s = dev_download(ibox.data(), d_ibox, ni * 24, st);
  }
  if (s == LUCENT_OK) s = stream_sync(st);
  hipFree(mem);
  if (s != LUCENT_OK) return s;

  // Reference layout: root 0, children of a node adjacent (left first).
  std::copy(ids.begin(), ids.end(), tri_indices_out);
  struct Item {
    uint32_t tree;  // radix-tree reference (bit 31: leaf)
    uint32_t out;   // index in nodes_out
  };
  auto emit = [&](uint32_t out, const float* box, uint32_t left_first, uint32_t count) {
    float* o = nodes_out + size_t(out) * 8;
    o[0] = box[0];
    o[1] = box[1];
    o[2] = box[2];
    o[3] = float(left_first);
    o[4] = box[3];
    o[5] = box[4];
    o[6] = box[5];
    o[7] = float(count);
  };
  auto span = [&](uint32_t ref, uint32_t& first, uint32_t& count) {
    if (ref & 0x80000000u) {
      first = ref & 0x7FFFFFFFu;
      count = 1;
    } else {
      first = range[2 * ref];
      count = range[2 * ref + 1] - first + 1;
    }
  };
  auto box_of = [&](uint32_t ref) -> const float* {
    return (ref & 0x80000000u) ? &lbox[size_t(ref & 0x7FFFFFFFu) * 6] : &ibox[size_t(ref) * 6];
  };
  uint32_t count_nodes = 1;
  std::vector<Item> work{{n > 1 ? 0u : 0x80000000u, 0u}};
  while (!work.empty()) {
    const Item it = work.back();
    work.pop_back();
    uint32_t first, count;
    span(it.tree, first, count);
    if (count <= max_leaf || (it.tree & 0x80000000u)) {
      // A leaf over the node's contiguous range of sorted triangles; its box
      // is the node's (the union of those triangles).
      emit(it.out, box_of(it.tree), first, count);
      continue;
    }
    const uint32_t l = count_nodes;
    count_nodes += 2;
    emit(it.out, box_of(it.tree), l, 0);
    work.push_back({child[2 * it.tree + 1], l + 1});
    work.push_back({child[2 * it.tree], l});
  }
  *node_count_out = count_nodes;
  return LUCENT_OK;
}

}  // namespace lucent

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
//
// The clustered build (search radius r > 0) keeps the Morton sort and then
// merges clusters by surface area, after Meister and Bittner's parallel
// locally-ordered clustering (PLOC, TVCG 2018): the clusters, leaves at
// first, stay in Morton order; each finds the neighbour within r positions
// whose union with it has the least surface area (ties to the pair of lower
// indices, so that the globally best pair is always mutual); mutual
// nearest neighbours merge into a new node at the lower position; the list
// is compacted in order; repeat until one cluster remains (a round that
// merges few clusters is followed by one that pairs neighbours outright).  Each round is
// four kernels and a scan; a node is created after both its children, so
// subtree sizes follow in one pass over the nodes.  Its trees cost about
// what the SAH builder's do at a small part of the build time.
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

// ---------------------------------------------------------------------------
// Clustered build.  A node reference has bit 31 set for a leaf (its position
// in the sorted order) and is an internal node index otherwise.  Boxes are
// two float4 (min, max; w unused).

// Leaf boxes in sorted order; the first clusters.
__global__ void k_leaf_boxes(const float* __restrict__ pos, const uint32_t* __restrict__ ids, uint32_t n,
                             float4* __restrict__ lbox, uint32_t* __restrict__ cref, float4* __restrict__ cbox) {
  const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
  if (i >= n) return;
  const float* p = pos + size_t(ids[i]) * 9;
  const float4 lo = make_float4(fminf(fminf(p[0], p[3]), p[6]), fminf(fminf(p[1], p[4]), p[7]),
                                fminf(fminf(p[2], p[5]), p[8]), 0.0f);
  const float4 hi = make_float4(fmaxf(fmaxf(p[0], p[3]), p[6]), fmaxf(fmaxf(p[1], p[4]), p[7]),
                                fmaxf(fmaxf(p[2], p[5]), p[8]), 0.0f);
  lbox[2 * size_t(i)] = lo;
  lbox[2 * size_t(i) + 1] = hi;
  cref[i] = 0x80000000u | i;
  cbox[2 * size_t(i)] = lo;
  cbox[2 * size_t(i) + 1] = hi;
}

__device__ inline float union_area(float4 alo, float4 ahi, float4 blo, float4 bhi) {
  const float x = fmaxf(ahi.x, bhi.x) - fminf(alo.x, blo.x);
  const float y = fmaxf(ahi.y, bhi.y) - fminf(alo.y, blo.y);
  const float z = fmaxf(ahi.z, bhi.z) - fminf(alo.z, blo.z);
  return x * y + y * z + z * x;
}

constexpr int kPlocMaxRadius = 64;

// nn[i]: the cluster j != i within `radius` positions whose union with i has
// the least area.  Pairs compare by (area, distance, parity of the lower
// index, lower index), a strict order under which the best pair of the
// whole list picks itself from both ends (so every round merges), and
// equal boxes pair up as (2k, 2k+1).  With `forced` every cluster pairs
// with its neighbour i ^ 1 (a round that halves the list).  The block's
// clusters and their halo are staged in LDS.
__global__ __launch_bounds__(kBlock) void k_ploc_nearest(const float4* __restrict__ cbox, uint32_t c, int radius,
                                                         int forced, uint32_t* __restrict__ nn) {
  __shared__ float4 s_lo[kBlock + 2 * kPlocMaxRadius], s_hi[kBlock + 2 * kPlocMaxRadius];
  const int first = int(blockIdx.x) * kBlock - radius;
  const int span = kBlock + 2 * radius;
  for (int k = int(threadIdx.x); k < span; k += kBlock) {
    const int j = first + k;
    if (j >= 0 && j < int(c)) {
      s_lo[k] = cbox[2 * size_t(j)];
      s_hi[k] = cbox[2 * size_t(j) + 1];
    }
  }
  __syncthreads();
  const int i = int(blockIdx.x) * kBlock + int(threadIdx.x);
  if (i >= int(c)) return;
  if (forced) {
    nn[i] = uint32_t((i ^ 1) < int(c) ? (i ^ 1) : i);
    return;
  }
  const int li = i - first;
  const float4 lo = s_lo[li], hi = s_hi[li];
  float best = INFINITY;
  int bj = -1;
  const int j0 = max(0, i - radius), j1 = min(int(c) - 1, i + radius);
  for (int j = j0; j <= j1; ++j) {
    if (j == i) continue;
    const float a = union_area(lo, hi, s_lo[j - first], s_hi[j - first]);
    bool take = bj < 0 || a < best;
    if (bj >= 0 && a == best) {
      const int d = abs(j - i), bd = abs(bj - i);
      const int m = min(i, j), bm = min(i, bj);
      take = d != bd ? d < bd : (m & 1) != (bm & 1) ? (m & 1) < (bm & 1) : m < bm;
    }
    if (take) {
      best = a;
      bj = j;
    }
  }
  nn[i] = uint32_t(bj < 0 ? i : bj);
}

// flags[i] = survives (low word) | creates a node (high word): cluster i
// merges with its mutual nearest neighbour j; the lower of the two creates
// the node and keeps the position, the higher disappears.
__global__ void k_ploc_flags(const uint32_t* __restrict__ nn, uint32_t c, uint64_t* __restrict__ flags) {
  const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
  if (i >= c) return;
  const uint32_t j = nn[i];
  const bool mutual = j != i && nn[j] == i;
  const uint64_t survives = (mutual && j < i) ? 0u : 1u;
  const uint64_t creates = (mutual && i < j) ? 1u : 0u;
  flags[i] = survives | (creates << 32);
}

// The next cluster list (compacted, in order) and the new nodes.
__global__ void k_ploc_merge(const uint32_t* __restrict__ nn, uint32_t c, const uint64_t* __restrict__ flags,
                             const uint64_t* __restrict__ ex, uint32_t node_base, const uint32_t* __restrict__ cref,
                             const float4* __restrict__ cbox, uint32_t* __restrict__ child, float4* __restrict__ ibox,
                             uint32_t* __restrict__ cref_out, float4* __restrict__ cbox_out) {
  const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
  if (i >= c) return;
  const uint64_t f = flags[i];
  if (!(f & 0xFFFFFFFFu)) return;
  const uint32_t pos = uint32_t(ex[i] & 0xFFFFFFFFu);
  float4 lo = cbox[2 * size_t(i)], hi = cbox[2 * size_t(i) + 1];
  uint32_t ref = cref[i];
  if (f >> 32) {
    const uint32_t j = nn[i];
    const uint32_t node = node_base + uint32_t(ex[i] >> 32);
    const float4 blo = cbox[2 * size_t(j)], bhi = cbox[2 * size_t(j) + 1];
    lo = make_float4(fminf(lo.x, blo.x), fminf(lo.y, blo.y), fminf(lo.z, blo.z), 0.0f);
    hi = make_float4(fmaxf(hi.x, bhi.x), fmaxf(hi.y, bhi.y), fmaxf(hi.z, bhi.z), 0.0f);
    child[2 * size_t(node)] = ref;
    child[2 * size_t(node) + 1] = cref[j];
    ibox[2 * size_t(node)] = lo;
    ibox[2 * size_t(node) + 1] = hi;
    ref = node;
  }
  cref_out[pos] = ref;
  cbox_out[2 * size_t(pos)] = lo;
  cbox_out[2 * size_t(pos) + 1] = hi;
}

__global__ void k_zero_u32(uint32_t* p, size_t n) {
  const size_t i = size_t(blockIdx.x) * kBlock + threadIdx.x;
  if (i < n) p[i] = 0;
}

}  // namespace

// The clustered build after the sort: rounds of nearest neighbours, merges
// and compaction on the device until one cluster is left; then the host
// numbers the tree in the reference layout.  Subtrees of at most max_leaf
// triangles become leaves over consecutive slots of triIndices (their
// triangles in the subtree's left-first order).
static Status cluster_and_emit(hipStream_t st, const float* d_pos, const uint32_t* d_ids, uint32_t n,
                               uint32_t max_leaf, uint32_t radius, float* nodes_out, uint32_t* node_count_out,
                               uint32_t* tri_indices_out) {
  const size_t nn = n, ni = n > 1 ? n - 1 : 1;
  auto up = [](size_t b) { return (b + 255) / 256 * 256; };
  size_t scan_bytes = 0;
  hipError_t e = hipcub::DeviceScan::ExclusiveSum(nullptr, scan_bytes, static_cast<const uint64_t*>(nullptr),
                                                  static_cast<uint64_t*>(nullptr), int(n), st);
  if (e != hipSuccess) return fail_hip(e, "cluster scan sizing");
  const size_t off_lbox = 0, off_ibox = off_lbox + up(nn * 32), off_child = off_ibox + up(ni * 32),
               off_cref0 = off_child + up(ni * 8), off_cref1 = off_cref0 + up(nn * 4),
               off_cbox0 = off_cref1 + up(nn * 4), off_cbox1 = off_cbox0 + up(nn * 32),
               off_nn = off_cbox1 + up(nn * 32), off_flags = off_nn + up(nn * 4), off_ex = off_flags + up(nn * 8),
               off_scan = off_ex + up(nn * 8), total = off_scan + up(scan_bytes);
  void* mem = nullptr;
  Status s = dev_alloc(&mem, total);
  if (s != LUCENT_OK) return s;
  char* m = static_cast<char*>(mem);
  float4* d_lbox = reinterpret_cast<float4*>(m + off_lbox);
  float4* d_ibox = reinterpret_cast<float4*>(m + off_ibox);
  uint32_t* d_child = reinterpret_cast<uint32_t*>(m + off_child);
  uint32_t* d_cref[2] = {reinterpret_cast<uint32_t*>(m + off_cref0), reinterpret_cast<uint32_t*>(m + off_cref1)};
  float4* d_cbox[2] = {reinterpret_cast<float4*>(m + off_cbox0), reinterpret_cast<float4*>(m + off_cbox1)};
  uint32_t* d_nn = reinterpret_cast<uint32_t*>(m + off_nn);
  uint64_t* d_flags = reinterpret_cast<uint64_t*>(m + off_flags);
  uint64_t* d_ex = reinterpret_cast<uint64_t*>(m + off_ex);
  void* d_scan = m + off_scan;
  auto grid = [](size_t k) { return dim3(uint32_t((k + kBlock - 1) / kBlock)); };
  hipLaunchKernelGGL(k_leaf_boxes, grid(nn), dim3(kBlock), 0, st, d_pos, d_ids, n, d_lbox, d_cref[0], d_cbox[0]);
  uint32_t c = n, node_base = 0;
  int cur = 0;
  const int r = int(std::min<uint32_t>(radius, uint32_t(kPlocMaxRadius)));
  // A round that merges fewer than 1/64 of the clusters (long runs of equal
  // or nested boxes) is followed by one that pairs neighbours outright, so
  // the list at least halves every second round.
  bool forced = false;
  while (c > 1 && s == LUCENT_OK) {
    hipLaunchKernelGGL(k_ploc_nearest, grid(c), dim3(kBlock), 0, st, d_cbox[cur], c, r, forced ? 1 : 0, d_nn);
    hipLaunchKernelGGL(k_ploc_flags, grid(c), dim3(kBlock), 0, st, d_nn, c, d_flags);
    size_t sb = scan_bytes;
    e = hipcub::DeviceScan::ExclusiveSum(d_scan, sb, d_flags, d_ex, int(c), st);
    if (e != hipSuccess) {
      s = fail_hip(e, "cluster scan");
      break;
    }
    hipLaunchKernelGGL(k_ploc_merge, grid(c), dim3(kBlock), 0, st, d_nn, c, d_flags, d_ex, node_base, d_cref[cur],
                       d_cbox[cur], d_child, d_ibox, d_cref[1 - cur], d_cbox[1 - cur]);
    if ((e = hipGetLastError()) != hipSuccess) {
      s = fail_hip(e, "cluster round launch");
      break;
    }
    // Survivors and new nodes of the round: the last exclusive prefix plus
    // the last flags.
    uint64_t tail[2] = {0, 0};
    if ((s = dev_download(&tail[0], d_ex + (c - 1), 8, st)) != LUCENT_OK) break;
    if ((s = dev_download(&tail[1], d_flags + (c - 1), 8, st)) != LUCENT_OK) break;
    if ((s = stream_sync(st)) != LUCENT_OK) break;
    const uint64_t sum = tail[0] + tail[1];
    const uint32_t next = uint32_t(sum & 0xFFFFFFFFu), made = uint32_t(sum >> 32);
    if (made == 0 && !forced) {
      // No mutual pair (possible only when areas do not order strictly,
      // e.g. NaN coordinates): the list is unchanged; pair neighbours.
      forced = true;
      continue;
    }
    if (made == 0 || next >= c) {
      s = fail(LUCENT_E_OVERFLOW, "cluster round merged nothing");
      break;
    }
    forced = !forced && uint64_t(made) * 64 < c;
    c = next;
    node_base += made;
    cur = 1 - cur;
  }
  uint32_t root = 0x80000000u;
  std::vector<uint32_t> child(ni * 2), ids(nn);
  std::vector<float4> lbox(nn * 2), ibox(ni * 2);
  if (s == LUCENT_OK && n > 1) s = dev_download(&root, d_cref[cur], 4, st);
  if (s == LUCENT_OK) s = dev_download(ids.data(), d_ids, nn * 4, st);
  if (s == LUCENT_OK) s = dev_download(lbox.data(), d_lbox, nn * 32, st);
  if (s == LUCENT_OK && n > 1) {
    if ((s = dev_download(child.data(), d_child, ni * 8, st)) == LUCENT_OK) s = dev_download(ibox.data(), d_ibox, ni * 32, st);
  }
  if (s == LUCENT_OK) s = stream_sync(st);
  hipFree(mem);
  if (s != LUCENT_OK) return s;
  if (n > 1 && node_base != n - 1) return fail(LUCENT_E_OVERFLOW, "clustering left an incomplete tree");

  // Triangles per subtree: a node is made after both of its children.
  std::vector<uint32_t> count(ni, 0);
  auto cnt = [&](uint32_t ref) { return (ref & 0x80000000u) ? 1u : count[ref]; };
  for (uint32_t k = 0; k + 1 < n; ++k) count[k] = cnt(child[2 * k]) + cnt(child[2 * k + 1]);
  auto box_of = [&](uint32_t ref) -> const float4* {
    return (ref & 0x80000000u) ? &lbox[2 * size_t(ref & 0x7FFFFFFFu)] : &ibox[2 * size_t(ref)];
  };
  auto emit = [&](uint32_t out, uint32_t ref, uint32_t left_first, uint32_t count_) {
    const float4* b = box_of(ref);
    float* o = nodes_out + size_t(out) * 8;
    o[0] = b[0].x;
    o[1] = b[0].y;
    o[2] = b[0].z;
    o[3] = float(left_first);
    o[4] = b[1].x;
    o[5] = b[1].y;
    o[6] = b[1].z;
    o[7] = float(count_);
  };
  struct Item {
    uint32_t ref, out;
  };
  uint32_t count_nodes = 1, slot = 0;
  std::vector<Item> work{{root, 0u}};
  std::vector<uint32_t> sub;
  while (!work.empty()) {
    const Item it = work.back();
    work.pop_back();
    const uint32_t k = cnt(it.ref);
    if (k <= max_leaf) {
      // A leaf over the next k slots, filled with the subtree's triangles
      // left first.
      emit(it.out, it.ref, slot, k);
      sub.assign(1, it.ref);
      while (!sub.empty()) {
        const uint32_t ref = sub.back();
        sub.pop_back();
        if (ref & 0x80000000u) {
          tri_indices_out[slot++] = ids[ref & 0x7FFFFFFFu];
        } else {
          sub.push_back(child[2 * ref + 1]);
          sub.push_back(child[2 * ref]);
        }
      }
      continue;
    }
    const uint32_t l = count_nodes;
    count_nodes += 2;
    emit(it.out, it.ref, l, 0);
    work.push_back({child[2 * it.ref + 1], l + 1});
    work.push_back({child[2 * it.ref], l});
  }
  *node_count_out = count_nodes;
  return LUCENT_OK;
}

Status build_bvh_gpu(Context* ctx, const float* positions, uint32_t tri_count, uint32_t max_leaf, float* nodes_out,
                     uint32_t* node_count_out, uint32_t* tri_indices_out, double bounds_out[6], uint32_t radius) {
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
  if (s == LUCENT_OK && radius > 0) {
    s = cluster_and_emit(st, d_pos, d_ids2, n, max_leaf, radius, nodes_out, node_count_out, tri_indices_out);
    if (s == LUCENT_OK) s = stream_sync(st);
    hipFree(mem);
    return s;
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

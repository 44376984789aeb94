// Binned-SAH BVH builder producing Lucent's flat node layout.
//
// Behaviour follows buildBVH (reference src/math/bvh.ts:52-252) exactly, so a
// tree built here is bit-identical to one built by the TypeScript and can be
// swapped for it: 12 bins by default; all three axes are binned for ranges of
// at least 512 triangles, only the longest centroid axis below that; a range
// becomes a leaf at <= maxLeafSize triangles, or at <= 4*maxLeafSize when no
// split beats the leaf's SAH cost; unprofitable larger ranges and partitions
// that move nothing fall back to a median split; the left child of every
// split is built (and numbered) before the right one.
//
// Numerics follow JavaScript: per-triangle bounds and centroids are stored as
// float32, everything else is float64, and no operation is fused (this file is
// compiled with -ffp-contract=off and the pragma below).
#pragma clang fp contract(off)

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <thread>
#include <vector>

#include "internal.hpp"

namespace lucent {
namespace {

// Math.min / Math.max of two numbers: NaN propagates and -0 orders below +0.
inline double js_min(double a, double b) {
  if (std::isnan(a) || std::isnan(b)) return NAN;
  if (a == b) return std::signbit(a) ? a : b;
  return a < b ? a : b;
}
inline double js_max(double a, double b) {
  if (std::isnan(a) || std::isnan(b)) return NAN;
  if (a == b) return std::signbit(a) ? b : a;
  return a > b ? a : b;
}

struct Aabb {
  double lo[3], hi[3];
  void clear() {
    for (int a = 0; a < 3; ++a) {
      lo[a] = INFINITY;
      hi[a] = -INFINITY;
    }
  }
  // boxArea (bvh.ts:35-40): an inverted box has zero area.
  static double area(double x0, double y0, double z0, double x1, double y1, double z1) {
    const double dx = x1 - x0, dy = y1 - y0, dz = z1 - z0;
    return dx < 0 ? 0 : 2 * (dx * dy + dy * dz + dz * dx);
  }
};

class SahBinnedBuilder {
 public:
  SahBinnedBuilder(const float* positions, uint32_t tris, uint32_t max_leaf, uint32_t bins)
      : pos_(positions), n_(tris), max_leaf_(max_leaf), bins_(bins) {}
  // A worker over `o`'s prepared triangles and order (its own bins).
  struct Worker {};
  SahBinnedBuilder(const SahBinnedBuilder& o, Worker)
      : pos_(o.pos_), n_(o.n_), max_leaf_(o.max_leaf_), bins_(o.bins_), order_(o.order_), tlo_(o.tlo_),
        thi_(o.thi_), cen_(o.cen_) {}

  void build(float* nodes, uint32_t* node_count, uint32_t* order, double bounds[6]) {
    nodes_ = nodes;
    order_ = order;
    if (n_ == 0) {
      const double zero[6] = {0, 0, 0, 0, 0, 0};
      emit(0, zero, zero + 3, 0, 0);
      *node_count = 1;
      std::memset(bounds, 0, 6 * sizeof(double));
      return;
    }
    prepare();
    bin_lo_.assign(3 * bins_ * 3, 0.0);
    bin_hi_.assign(3 * bins_ * 3, 0.0);
    bin_n_.assign(3 * bins_, 0);
    suffix_area_.assign(bins_, 0.0);
    suffix_n_.assign(bins_, 0);

    uint32_t allocated = 1;
    std::vector<Task> todo;
    todo.push_back({0, 0, n_});
    while (!todo.empty()) {
      const Task task = todo.back();
      todo.pop_back();
      const uint32_t count = task.end - task.begin;
      bounds_of(task.begin, task.end);
      if (count <= max_leaf_) {
        emit(task.node, node_box_.lo, node_box_.hi, task.begin, count);
        continue;
      }
      Split best = best_split(task.begin, task.end, count);
      const double leaf_cost =
          count * Aabb::area(node_box_.lo[0], node_box_.lo[1], node_box_.lo[2], node_box_.hi[0], node_box_.hi[1],
                             node_box_.hi[2]);
      const bool unprofitable = best.axis < 0 || best.cost >= leaf_cost;
      if (unprofitable && count <= max_leaf_ * 4) {
        emit(task.node, node_box_.lo, node_box_.hi, task.begin, count);
        continue;
      }
      uint32_t mid = unprofitable ? (task.begin + task.end) >> 1 : partition(task.begin, task.end, best);
      if (mid == task.begin || mid == task.end) mid = (task.begin + task.end) >> 1;
      const uint32_t left = allocated++;
      const uint32_t right = allocated++;
      emit(task.node, node_box_.lo, node_box_.hi, left, 0);
      todo.push_back({right, mid, task.end});
      todo.push_back({left, task.begin, mid});
    }
    bounds_of(0, n_);
    for (int a = 0; a < 3; ++a) {
      bounds[a] = node_box_.lo[a];
      bounds[a + 3] = node_box_.hi[a];
    }
    *node_count = allocated;
  }

  // The same splits, with every range of fewer than `cutoff` triangles built
  // as a subtree of its own on a pool of `threads` threads and stitched in
  // afterwards: node numbers differ from build()'s (each subtree's nodes are
  // contiguous), the tree is the same set of splits and is deterministic.
  void build_parallel(float* nodes, uint32_t* node_count, uint32_t* order, double bounds[6], uint32_t threads,
                      uint32_t cutoff) {
    if (n_ <= cutoff || threads < 2) {
      build(nodes, node_count, order, bounds);
      return;
    }
    nodes_ = nodes;
    order_ = order;
    prepare();
    alloc_bins();
    uint32_t allocated = 1;
    std::vector<Task> todo, deferred;
    todo.push_back({0, 0, n_});
    while (!todo.empty()) {
      const Task task = todo.back();
      todo.pop_back();
      if (task.end - task.begin < cutoff) {
        deferred.push_back(task);
        continue;
      }
      const Expand e = expand(task, allocated);
      if (e.leaf) continue;
      todo.push_back({e.right, e.mid, task.end});
      todo.push_back({e.left, task.begin, e.mid});
    }
    // Subtrees: local node 0 is the subtree's root (written at its task's
    // node), local node k > 0 goes to base + k - 1.
    std::vector<std::vector<float>> local(deferred.size());
    std::vector<uint32_t> counts(deferred.size(), 0);
    std::atomic<size_t> next{0};
    auto worker = [&]() {
      SahBinnedBuilder sub(*this, Worker{});
      for (size_t i = next++; i < deferred.size(); i = next++) {
        const Task& t = deferred[i];
        local[i].assign(size_t(2 * (t.end - t.begin)) * 8, 0.0f);
        sub.nodes_ = local[i].data();
        counts[i] = sub.build_range(t.begin, t.end);
      }
    };
    std::vector<std::thread> pool;
    for (uint32_t k = 0; k + 1 < threads; ++k) pool.emplace_back(worker);
    worker();
    for (auto& th : pool) th.join();
    uint32_t base = allocated;
    for (size_t i = 0; i < deferred.size(); ++i) {
      const float* src = local[i].data();
      for (uint32_t k = 0; k < counts[i]; ++k) {
        float* dst = nodes_ + size_t(k == 0 ? deferred[i].node : base + k - 1) * 8;
        std::memcpy(dst, src + size_t(k) * 8, 8 * sizeof(float));
        if (dst[7] == 0.0f) dst[3] = float(base + uint32_t(dst[3]) - 1);  // interior: its children
      }
      base += counts[i] - 1;
    }
    bounds_of(0, n_);
    for (int a = 0; a < 3; ++a) {
      bounds[a] = node_box_.lo[a];
      bounds[a + 3] = node_box_.hi[a];
    }
    *node_count = base;
  }

 private:
  struct Task {
    uint32_t node, begin, end;
  };
  struct Split {
    int axis = -1;
    int bin = -1;
    double cost = INFINITY;
  };
  struct Expand {
    bool leaf;
    uint32_t left, right, mid;
  };
  void alloc_bins() {
    bin_lo_.assign(3 * bins_ * 3, 0.0);
    bin_hi_.assign(3 * bins_ * 3, 0.0);
    bin_n_.assign(3 * bins_, 0);
    suffix_area_.assign(bins_, 0.0);
    suffix_n_.assign(bins_, 0);
  }
  // One step of build()'s loop: the task's node becomes a leaf, or an
  // interior node whose two children are allocated.
  Expand expand(const Task& task, uint32_t& allocated) {
    const uint32_t count = task.end - task.begin;
    bounds_of(task.begin, task.end);
    if (count <= max_leaf_) {
      emit(task.node, node_box_.lo, node_box_.hi, task.begin, count);
      return {true, 0, 0, 0};
    }
    Split best = best_split(task.begin, task.end, count);
    const double leaf_cost =
        count * Aabb::area(node_box_.lo[0], node_box_.lo[1], node_box_.lo[2], node_box_.hi[0], node_box_.hi[1],
                           node_box_.hi[2]);
    const bool unprofitable = best.axis < 0 || best.cost >= leaf_cost;
    if (unprofitable && count <= max_leaf_ * 4) {
      emit(task.node, node_box_.lo, node_box_.hi, task.begin, count);
      return {true, 0, 0, 0};
    }
    uint32_t mid = unprofitable ? (task.begin + task.end) >> 1 : partition(task.begin, task.end, best);
    if (mid == task.begin || mid == task.end) mid = (task.begin + task.end) >> 1;
    const uint32_t left = allocated++;
    const uint32_t right = allocated++;
    emit(task.node, node_box_.lo, node_box_.hi, left, 0);
    return {false, left, right, mid};
  }
  // The subtree over [begin, end) with local numbering from 0 (its root);
  // returns its node count.
  uint32_t build_range(uint32_t begin, uint32_t end) {
    alloc_bins();
    uint32_t allocated = 1;
    std::vector<Task> todo;
    todo.push_back({0, begin, end});
    while (!todo.empty()) {
      const Task task = todo.back();
      todo.pop_back();
      const Expand e = expand(task, allocated);
      if (e.leaf) continue;
      todo.push_back({e.right, e.mid, task.end});
      todo.push_back({e.left, task.begin, e.mid});
    }
    return allocated;
  }


  // Per-triangle float32 bounds and centroids (bvh.ts:60-73).
  void prepare() {
    tri_lo_.resize(size_t(n_) * 3);
    tri_hi_.resize(size_t(n_) * 3);
    centroid_.resize(size_t(n_) * 3);
    tlo_ = tri_lo_.data();
    thi_ = tri_hi_.data();
    cen_ = centroid_.data();
    for (uint32_t t = 0; t < n_; ++t) {
      const float* p = pos_ + size_t(t) * 9;
      for (int a = 0; a < 3; ++a) {
        const double lo = js_min(js_min(p[a], p[3 + a]), p[6 + a]);
        const double hi = js_max(js_max(p[a], p[3 + a]), p[6 + a]);
        tri_lo_[size_t(t) * 3 + a] = float(lo);
        tri_hi_[size_t(t) * 3 + a] = float(hi);
        centroid_[size_t(t) * 3 + a] = float((lo + hi) * 0.5);
      }
    }
    for (uint32_t i = 0; i < n_; ++i) order_[i] = i;
  }

  // Triangle bounds into node_box_, centroid bounds into cen_box_.
  void bounds_of(uint32_t begin, uint32_t end) {
    node_box_.clear();
    cen_box_.clear();
    for (uint32_t i = begin; i < end; ++i) {
      const size_t t = size_t(order_[i]) * 3;
      for (int a = 0; a < 3; ++a) {
        const double lo = tlo_[t + a], hi = thi_[t + a], c = cen_[t + a];
        if (lo < node_box_.lo[a]) node_box_.lo[a] = lo;
        if (hi > node_box_.hi[a]) node_box_.hi[a] = hi;
        if (c < cen_box_.lo[a]) cen_box_.lo[a] = c;
        if (c > cen_box_.hi[a]) cen_box_.hi[a] = c;
      }
    }
  }

  double bin_index(uint32_t tri, int axis) const {
    const double f = std::floor((double(cen_[size_t(tri) * 3 + axis]) - cen_box_.lo[axis]) * scale_[axis]);
    return js_min(double(bins_ - 1), f);
  }

  Split best_split(uint32_t begin, uint32_t end, uint32_t count) {
    const double ex = cen_box_.hi[0] - cen_box_.lo[0];
    const double ey = cen_box_.hi[1] - cen_box_.lo[1];
    const double ez = cen_box_.hi[2] - cen_box_.lo[2];
    const int longest = (ex >= ey && ex >= ez) ? 0 : (ey >= ez ? 1 : 2);
    bool use[3];
    for (int a = 0; a < 3; ++a) {
      const double extent = cen_box_.hi[a] - cen_box_.lo[a];
      use[a] = extent > 1e-9 && (count >= 512 || a == longest);
      scale_[a] = double(bins_) / extent;
      if (!use[a]) continue;
      for (uint32_t k = a * bins_; k < (a + 1) * bins_; ++k) {
        for (int c = 0; c < 3; ++c) {
          bin_lo_[k * 3 + c] = INFINITY;
          bin_hi_[k * 3 + c] = -INFINITY;
        }
        bin_n_[k] = 0;
      }
    }
    for (uint32_t i = begin; i < end; ++i) {
      const uint32_t t = order_[i];
      for (int a = 0; a < 3; ++a) {
        if (!use[a]) continue;
        const uint32_t k = a * bins_ + uint32_t(bin_index(t, a));
        ++bin_n_[k];
        for (int c = 0; c < 3; ++c) {
          const double lo = tlo_[size_t(t) * 3 + c], hi = thi_[size_t(t) * 3 + c];
          if (lo < bin_lo_[k * 3 + c]) bin_lo_[k * 3 + c] = lo;
          if (hi > bin_hi_[k * 3 + c]) bin_hi_[k * 3 + c] = hi;
        }
      }
    }
    Split best;
    for (int a = 0; a < 3; ++a) {
      if (!use[a]) continue;
      // Suffix sweep (right-hand sides), then prefix sweep evaluating each plane.
      double lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
      uint32_t acc = 0;
      for (uint32_t b = bins_ - 1; b > 0; --b) {
        const uint32_t k = a * bins_ + b;
        for (int c = 0; c < 3; ++c) {
          lo[c] = js_min(lo[c], bin_lo_[k * 3 + c]);
          hi[c] = js_max(hi[c], bin_hi_[k * 3 + c]);
        }
        acc += bin_n_[k];
        suffix_area_[b] = Aabb::area(lo[0], lo[1], lo[2], hi[0], hi[1], hi[2]);
        suffix_n_[b] = acc;
      }
      for (int c = 0; c < 3; ++c) {
        lo[c] = INFINITY;
        hi[c] = -INFINITY;
      }
      uint32_t left_n = 0;
      for (uint32_t b = 0; b + 1 < bins_; ++b) {
        const uint32_t k = a * bins_ + b;
        for (int c = 0; c < 3; ++c) {
          lo[c] = js_min(lo[c], bin_lo_[k * 3 + c]);
          hi[c] = js_max(hi[c], bin_hi_[k * 3 + c]);
        }
        left_n += bin_n_[k];
        const uint32_t right_n = suffix_n_[b + 1];
        if (left_n == 0 || right_n == 0) continue;
        const double cost =
            Aabb::area(lo[0], lo[1], lo[2], hi[0], hi[1], hi[2]) * left_n + suffix_area_[b + 1] * right_n;
        if (cost < best.cost) {
          best.cost = cost;
          best.axis = a;
          best.bin = int(b);
        }
      }
    }
    return best;
  }

  // Two-index partition: keep the front while it belongs left, otherwise swap
  // it with the back (the same sequence of swaps as bvh.ts:126-140).
  uint32_t partition(uint32_t begin, uint32_t end, const Split& s) {
    int64_t front = begin, back = int64_t(end) - 1;
    while (front <= back) {
      if (bin_index(order_[front], s.axis) <= s.bin) {
        ++front;
      } else {
        const uint32_t tmp = order_[front];
        order_[front] = order_[back];
        order_[back] = tmp;
        --back;
      }
    }
    return uint32_t(front);
  }

  void emit(uint32_t idx, const double* lo, const double* hi, double first, double count) {
    float* o = nodes_ + size_t(idx) * 8;
    o[0] = float(lo[0]);
    o[1] = float(lo[1]);
    o[2] = float(lo[2]);
    o[3] = float(first);
    o[4] = float(hi[0]);
    o[5] = float(hi[1]);
    o[6] = float(hi[2]);
    o[7] = float(count);
  }

  const float* pos_;
  uint32_t n_, max_leaf_, bins_;
  float* nodes_ = nullptr;
  uint32_t* order_ = nullptr;
  std::vector<float> tri_lo_, tri_hi_, centroid_;  // owned by the builder that prepared them
  const float *tlo_ = nullptr, *thi_ = nullptr, *cen_ = nullptr;
  std::vector<double> bin_lo_, bin_hi_, suffix_area_;
  std::vector<uint32_t> bin_n_, suffix_n_;
  Aabb node_box_{}, cen_box_{};
  double scale_[3] = {0, 0, 0};
};

}  // namespace

Status build_bvh(const float* positions, uint32_t tri_count, uint32_t max_leaf, uint32_t bins, float* nodes_out,
                 uint32_t* node_count, uint32_t* tri_indices_out, double bounds[6]) {
  if (bins < 2 || bins > 4096) return fail(LUCENT_E_INVALID, "bins must be in [2, 4096]");
  if (max_leaf < 1) return fail(LUCENT_E_INVALID, "max_leaf must be >= 1");
  if (tri_count > (1u << 24)) return fail(LUCENT_E_INVALID, "more than 2^24 triangles cannot be indexed by float32 node fields");
  if (tri_count > 0 && (!positions || !tri_indices_out)) return fail(LUCENT_E_INVALID, "null array");
  if (!nodes_out || !node_count || !bounds) return fail(LUCENT_E_INVALID, "null output");
  SahBinnedBuilder b(positions, tri_count, max_leaf, bins);
  b.build(nodes_out, node_count, tri_indices_out, bounds);
  return LUCENT_OK;
}

Status build_bvh_parallel(const float* positions, uint32_t tri_count, uint32_t max_leaf, uint32_t bins,
                          float* nodes_out, uint32_t* node_count, uint32_t* tri_indices_out, double bounds[6],
                          uint32_t threads) {
  if (bins < 2 || bins > 4096) return fail(LUCENT_E_INVALID, "bins must be in [2, 4096]");
  if (max_leaf < 1) return fail(LUCENT_E_INVALID, "max_leaf must be >= 1");
  if (tri_count > (1u << 24)) return fail(LUCENT_E_INVALID, "more than 2^24 triangles cannot be indexed by float32 node fields");
  if (tri_count > 0 && (!positions || !tri_indices_out)) return fail(LUCENT_E_INVALID, "null array");
  if (!nodes_out || !node_count || !bounds) return fail(LUCENT_E_INVALID, "null output");
  // Subtrees of at most ~1/(8 x threads) of the triangles (and at least 4096)
  // keep the pool busy to the end.
  const uint32_t cutoff = std::max<uint32_t>(4096, tri_count / (8 * std::max<uint32_t>(threads, 1)));
  SahBinnedBuilder b(positions, tri_count, max_leaf, bins);
  b.build_parallel(nodes_out, node_count, tri_indices_out, bounds, threads, cutoff);
  return LUCENT_OK;
}

}  // namespace lucent

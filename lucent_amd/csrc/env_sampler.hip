// Device build of the environment importance sampler: buildEnvSampler
// (reference src/scene/environment.ts:52-85) with its Walker alias table
// built by Vose's worklists (src/math/alias.ts:17-64), in float64 on the GPU.
//
// The tables are bit-identical to the host build (lighting_host.cpp) because
// every value goes through the same sequence of IEEE float64 operations:
//   k_env_cells    one thread per grid cell: the box mean of the texels'
//                  luminance times the row's sin(theta) (the row sines come
//                  from the host helper), summed in the host loop's order;
//   k_env_total    the weights' sum in index order, a sequential float64
//                  chain run by one wave over blocks staged in LDS;
//   k_env_split    pmf = w / total, scaled = pmf * n and the small/large
//                  class of each cell, with per-block counts;
//   k_env_offsets  exclusive prefix sums of the block counts;
//   k_env_scatter  the small and large worklists in ascending index order
//                  (a stable compaction by wave ballots);
//   k_env_sweep    Vose's pairing: small and large cells taken from the
//                  tops of their lists, the current large cell's residual
//                  carried in float64 (scaled_l + scaled_s - 1 per step); a
//                  large cell whose residual falls below 1 becomes the next
//                  small one.  One wave, worklist blocks prefetched into LDS;
//   k_env_pack     the table texels (prob, alias, pmf*n, pmf[alias]*n).
// The total and the sweep are sequential chains: their rounding and the
// pairing decisions depend on every earlier step, and a reordered build
// gives a valid but different table, which would pick a different cell than
// the reference's sampler for the same random numbers.
//
// The parallel build (for callers that rebuild the sampler often, e.g. after
// every gradient step on the map) replaces both chains: the total is a tree
// sum, and the pairing is decided per cell by binary searches over FP64
// prefix sums of the small cells' deficits and the large cells' excesses
// (k_env_gather, k_env_pair).  Its table samples the same distribution
// (every cell's implied probability is its pmf up to the float32 rounding
// of the stored probabilities) with other alias pairs.
//
// Every kernel after k_env_cells returns at once when a weight was negative
// or not finite, so a failed build leaves the output table untouched.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <hipcub/hipcub.hpp>
#include <vector>

#include "render_host.hpp"

#pragma clang fp contract(off)

namespace lucent {
namespace {

constexpr uint32_t kBlock = 256;
constexpr uint32_t kStage = 2048;   // worklist entries staged in LDS per list
constexpr uint32_t kNoBad = 0xFFFFFFFFu;

// info[0]: lowest cell with a bad weight (kNoBad if none); info[1], info[2]:
// small and large list lengths.
__device__ inline bool build_failed(const uint32_t* info) { return info[0] != kNoBad; }

__device__ inline double luminance(double r, double g, double b) { return 0.2126 * r + 0.7152 * g + 0.0722 * b; }

__global__ __launch_bounds__(kBlock) void k_env_cells(const float4* __restrict__ map, uint32_t width, uint32_t height,
                                                      uint32_t gw, uint32_t gh, const double* __restrict__ row_sin,
                                                      double* __restrict__ w, uint32_t* __restrict__ info) {
  const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
  if (i >= gw * gh) return;
  const uint32_t cy = i / gw, cx = i - cy * gw;
  const uint32_t y0 = uint32_t(floor(double(cy) * height / gh));
  const uint32_t y1 = max(y0 + 1, uint32_t(floor(double(cy + 1) * height / gh)));
  const uint32_t x0 = uint32_t(floor(double(cx) * width / gw));
  const uint32_t x1 = max(x0 + 1, uint32_t(floor(double(cx + 1) * width / gw)));
  double sum = 0;
  for (uint32_t y = y0; y < y1; ++y)
    for (uint32_t x = x0; x < x1; ++x) {
      const float4 p = map[size_t(y) * width + x];
      sum += luminance(p.x, p.y, p.z);
    }
  const double v = (sum / (double(y1 - y0) * (x1 - x0))) * row_sin[cy];
  w[i] = v;
  if (!(v >= 0) || !isfinite(v)) atomicMin(info, i);
}

// The weights' sum in index order.  The next block of weights is loaded
// into registers while the current one is summed from LDS, and the sum reads
// two weights per LDS access, sixteen ahead of the chain of additions.
constexpr uint32_t kTotalPerLane = kStage / 64;
__global__ __launch_bounds__(64) void k_env_total(const double* __restrict__ w, uint32_t n,
                                                  const uint32_t* __restrict__ info, double* __restrict__ total) {
  __shared__ __align__(16) double stage[kStage];
  if (build_failed(info)) return;
  double regs[kTotalPerLane];
  auto load = [&](uint32_t base) {
#pragma unroll
    for (uint32_t j = 0; j < kTotalPerLane; ++j) {
      const uint32_t k = base + j * 64 + threadIdx.x;
      regs[j] = k < n ? w[k] : 0.0;
    }
  };
  double t = 0;
  load(0);
  for (uint32_t base = 0; base < n; base += kStage) {
#pragma unroll
    for (uint32_t j = 0; j < kTotalPerLane; ++j) stage[j * 64 + threadIdx.x] = regs[j];
    __syncthreads();
    if (base + kStage < n) load(base + kStage);
    const uint32_t m = min(kStage, n - base);
    const double2* s2 = reinterpret_cast<const double2*>(stage);
    uint32_t k = 0;
    for (; k + 16 <= m; k += 16) {
      double2 v[8];
#pragma unroll
      for (int j = 0; j < 8; ++j) v[j] = s2[k / 2 + j];
#pragma unroll
      for (int j = 0; j < 8; ++j) {
        t += v[j].x;
        t += v[j].y;
      }
    }
    for (; k < m; ++k) t += stage[k];
    __syncthreads();
  }
  if (threadIdx.x == 0) *total = t;
}

__global__ __launch_bounds__(kBlock) void k_env_split(double* __restrict__ w, double* __restrict__ scaled,
                                                      const double* __restrict__ total, uint32_t n,
                                                      const uint32_t* __restrict__ info, uint32_t* __restrict__ counts) {
  if (build_failed(info)) return;
  const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
  const double t = *total;
  const bool valid = i < n;
  bool small = false;
  if (valid) {
    const double pmf = t > 0 ? w[i] / t : 1.0 / n;
    w[i] = pmf;
    const double sc = pmf * n;
    scaled[i] = sc;
    small = sc < 1;
  }
  const int n_small = __syncthreads_count(small);
  const int n_valid = __syncthreads_count(valid);
  if (threadIdx.x == 0) {
    counts[blockIdx.x] = uint32_t(n_small);
    counts[gridDim.x + blockIdx.x] = uint32_t(n_valid - n_small);
  }
}

// Exclusive prefix sums, in place, of counts[0, nb) (small) and
// counts[nb, 2 nb) (large); the totals go to info[1] and info[2].
__global__ __launch_bounds__(1024) void k_env_offsets(uint32_t* __restrict__ counts, uint32_t nb,
                                                      uint32_t* __restrict__ info) {
  __shared__ uint32_t part[1024];
  if (build_failed(info)) return;
  const uint32_t per = (nb + 1023) / 1024;
  for (int list = 0; list < 2; ++list) {
    uint32_t* c = counts + size_t(list) * nb;
    const uint32_t lo = min(nb, threadIdx.x * per), hi = min(nb, lo + per);
    uint32_t sum = 0;
    for (uint32_t k = lo; k < hi; ++k) sum += c[k];
    part[threadIdx.x] = sum;
    __syncthreads();
    for (uint32_t d = 1; d < 1024; d <<= 1) {
      const uint32_t v = threadIdx.x >= d ? part[threadIdx.x - d] : 0;
      __syncthreads();
      part[threadIdx.x] += v;
      __syncthreads();
    }
    uint32_t run = part[threadIdx.x] - sum;  // exclusive
    for (uint32_t k = lo; k < hi; ++k) {
      const uint32_t v = c[k];
      c[k] = run;
      run += v;
    }
    if (threadIdx.x == 1023) info[1 + list] = part[1023];
    __syncthreads();
  }
}

__global__ __launch_bounds__(kBlock) void k_env_scatter(const double* __restrict__ scaled, uint32_t n,
                                                        const uint32_t* __restrict__ offsets,
                                                        const uint32_t* __restrict__ info,
                                                        uint32_t* __restrict__ small, uint32_t* __restrict__ large,
                                                        double* __restrict__ small_val, double* __restrict__ large_val) {
  __shared__ uint32_t wave_small[kBlock / 64], wave_valid[kBlock / 64];
  if (build_failed(info)) return;
  const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
  const bool valid = i < n;
  const bool is_small = valid && scaled[i] < 1;
  const uint64_t bs = __ballot(is_small), bv = __ballot(valid);
  const uint32_t lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const uint64_t below = lane ? (~0ull >> (64 - lane)) : 0ull;
  if (lane == 0) {
    wave_small[wave] = uint32_t(__popcll(bs));
    wave_valid[wave] = uint32_t(__popcll(bv));
  }
  __syncthreads();
  uint32_t ps = 0, pv = 0;
  for (uint32_t k = 0; k < wave; ++k) {
    ps += wave_small[k];
    pv += wave_valid[k];
  }
  const uint32_t rs = ps + uint32_t(__popcll(bs & below));
  const uint32_t rv = pv + uint32_t(__popcll(bv & below));
  if (!valid) return;
  const double v = scaled[i];
  if (is_small) {
    small[offsets[blockIdx.x] + rs] = i;
    small_val[offsets[blockIdx.x] + rs] = v;
  } else {
    large[offsets[gridDim.x + blockIdx.x] + (rv - rs)] = i;
    large_val[offsets[gridDim.x + blockIdx.x] + (rv - rs)] = v;
  }
}

// One worklist as the sweep reads it: entries are taken from the top
// (descending positions); blocks of kStage entries (cell index and scaled
// weight, contiguous in global memory) are staged into LDS, and the next
// entry is read from LDS one step ahead of its use, so that the LDS latency
// overlaps the step that precedes it.
struct SweepList {
  const uint32_t* idx;
  const double* val;
  uint32_t* s_idx;
  double* s_val;
  uint32_t left;        // entries not yet taken: positions [0, left)
  uint32_t base;        // first staged position
  uint32_t next_cell;   // the entry at position left - 1 (valid when left > 0)
  double next_value;

  __device__ void stage() {
    const uint32_t top = left;
    base = top > kStage ? top - kStage : 0;
    for (uint32_t k = threadIdx.x; k < top - base; k += 64) {
      s_idx[k] = idx[base + k];
      s_val[k] = val[base + k];
    }
    __syncthreads();
  }
  __device__ void peek() {
    if (left == 0) return;
    if (left - 1 < base) stage();
    next_cell = s_idx[left - 1 - base];
    next_value = s_val[left - 1 - base];
  }
  __device__ void begin(uint32_t n) {
    left = n;
    base = n;
    peek();
  }
  // Takes the top entry (left > 0) and reads the one below it.
  __device__ void take(uint32_t& cell, double& value) {
    cell = next_cell;
    value = next_value;
    --left;
    peek();
  }
};

__device__ inline void emit(float* prob, uint32_t* alias, uint32_t cell, float p, uint32_t a) {
  if (threadIdx.x == 0) {
    prob[cell] = p;
    alias[cell] = a;
  }
}

// One wave; every lane holds the same state, so all branches are uniform.
__global__ __launch_bounds__(64) void k_env_sweep(const uint32_t* __restrict__ small, const uint32_t* __restrict__ large,
                                                  const double* __restrict__ small_val,
                                                  const double* __restrict__ large_val,
                                                  const uint32_t* __restrict__ info, float* __restrict__ prob,
                                                  uint32_t* __restrict__ alias) {
  __shared__ uint32_t s_idx[kStage], l_idx[kStage];
  __shared__ double s_val[kStage], l_val[kStage];
  if (build_failed(info)) return;
  SweepList sm{small, small_val, s_idx, s_val, 0, 0, 0, 0.0};
  SweepList lg{large, large_val, l_idx, l_val, 0, 0, 0, 0.0};
  sm.begin(info[1]);
  lg.begin(info[2]);
  bool heavy = false, pending = false;  // a current large cell; a former large cell on top of the small list
  uint32_t l = 0, p = 0;
  double rl = 0, rp = 0;
  for (;;) {
    uint32_t s;
    double xs;
    if (pending) {
      s = p;
      xs = rp;
      pending = false;
    } else {
      if (sm.left == 0) break;
      sm.take(s, xs);
    }
    if (!heavy) {
      if (lg.left == 0) {
        emit(prob, alias, s, 1.0f, s);  // no large cell left: s keeps its own bucket
        break;
      }
      lg.take(l, rl);
      heavy = true;
    }
    emit(prob, alias, s, float(xs), l);
    rl = (rl + xs) - 1.0;
    if (rl < 1.0) {
      pending = true;
      p = l;
      rp = rl;
      heavy = false;
    }
  }
  if (heavy) emit(prob, alias, l, 1.0f, l);
  for (uint32_t k = threadIdx.x; k < lg.left; k += 64) {
    prob[large[k]] = 1.0f;
    alias[large[k]] = large[k];
  }
  for (uint32_t k = threadIdx.x; k < sm.left; k += 64) {
    prob[small[k]] = 1.0f;
    alias[small[k]] = small[k];
  }
}

// ---------------------------------------------------------------------------
// The parallel pairing.  Let the small cells s_1 < s_2 < ... (scaled < 1)
// have deficits d_k = 1 - scaled and the large cells l_1 < l_2 < ... have
// excesses e_j = scaled - 1, with inclusive prefix sums D_k and E_j.  A
// sweep that fills the small cells in order from the large cells in order,
// a large cell turning small (filled by the next large cell) once its
// remaining weight drops below 1, makes these choices, which prefix sums
// decide independently of one another:
//   small s_k:  prob = scaled, alias = l_j for the first j with E_j >= D_{k-1}
//               (the large cell still holding weight >= 1 when s_k's turn
//               comes);
//   large l_j:  prob = 1 - (D_k - E_j) for the first k with D_k > E_j (the
//               small cell whose fill takes l_j below 1), alias = l_{j+1};
//               prob 1 if no such k or no next large cell.
// Every cell's own share plus the shares aliased to it then add up to its
// scaled weight (exactly, in exact arithmetic), so the table samples the
// same distribution as Vose's, with other pairs.

// Deficit (small cells) or excess (large cells) of a cell, 0 otherwise.
struct Deficit {
  __host__ __device__ double operator()(double sc) const { return sc < 1 ? 1.0 - sc : 0.0; }
};
struct Excess {
  __host__ __device__ double operator()(double sc) const { return sc < 1 ? 0.0 : sc - 1.0; }
};

// dl[k] = D_k over the small list, el[j] = E_j over the large list, gathered
// from the per-cell inclusive scans.
__global__ __launch_bounds__(kBlock) void k_env_gather(const uint32_t* __restrict__ small,
                                                       const uint32_t* __restrict__ large,
                                                       const double* __restrict__ dcell, const double* __restrict__ ecell,
                                                       const uint32_t* __restrict__ info, double* __restrict__ dl,
                                                       double* __restrict__ el) {
  if (build_failed(info)) return;
  const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
  if (i < info[1]) dl[i] = dcell[small[i]];
  if (i < info[2]) el[i] = ecell[large[i]];
}

// First position p in [0, n) with a[p] >= x (kStrict: a[p] > x); n if none.
template <bool kStrict>
__device__ inline uint32_t first_reaching(const double* __restrict__ a, uint32_t n, double x) {
  uint32_t lo = 0, hi = n;
  while (lo < hi) {
    const uint32_t mid = lo + (hi - lo) / 2;
    const bool reach = kStrict ? a[mid] > x : a[mid] >= x;
    if (reach)
      hi = mid;
    else
      lo = mid + 1;
  }
  return lo;
}

// Thread i pairs small cell i and large cell i (where they exist).
__global__ __launch_bounds__(kBlock) void k_env_pair(const uint32_t* __restrict__ small,
                                                     const uint32_t* __restrict__ large,
                                                     const double* __restrict__ small_val,
                                                     const double* __restrict__ dl, const double* __restrict__ el,
                                                     const uint32_t* __restrict__ info, float* __restrict__ prob,
                                                     uint32_t* __restrict__ alias) {
  if (build_failed(info)) return;
  const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
  const uint32_t a = info[1], b = info[2];
  if (i < a) {
    const uint32_t s = small[i];
    const double before = i ? dl[i - 1] : 0.0;
    const uint32_t j = first_reaching<false>(el, b, before);
    if (j < b) {
      prob[s] = float(small_val[i]);
      alias[s] = large[j];
    } else {  // beyond the large cells' excess (rounding): its own bucket
      prob[s] = 1.0f;
      alias[s] = s;
    }
  }
  if (i < b) {
    const uint32_t l = large[i];
    const uint32_t k = first_reaching<true>(dl, a, el[i]);
    if (k < a && i + 1 < b) {
      const double keep = 1.0 - (dl[k] - el[i]);
      prob[l] = float(fmin(1.0, fmax(0.0, keep)));
      alias[l] = large[i + 1];
    } else {
      prob[l] = 1.0f;
      alias[l] = l;
    }
  }
}

__global__ __launch_bounds__(kBlock) void k_env_pack(const float* __restrict__ prob, const uint32_t* __restrict__ alias,
                                                     const double* __restrict__ pmf, uint32_t n,
                                                     const uint32_t* __restrict__ info, float4* __restrict__ table) {
  if (build_failed(info)) return;
  const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
  if (i >= n) return;
  const uint32_t a = alias[i];
  table[i] = make_float4(prob[i], float(a), float(pmf[i] * n), float(pmf[a] * n));
}

size_t align256(size_t b) { return (b + 255) & ~size_t(255); }

}  // namespace

Status build_env_sampler_gpu(Context* ctx, const float* d_rgba, uint32_t width, uint32_t height, uint32_t max_width,
                             float* d_table, uint32_t* grid_w, uint32_t* grid_h, double* total_out, void* stream,
                             float* ms_out, bool parallel) {
  if (!ctx || !d_rgba || !d_table || width == 0 || height == 0) return fail(LUCENT_E_INVALID, "empty environment map");
  if (reinterpret_cast<uintptr_t>(d_rgba) % 16 || reinterpret_cast<uintptr_t>(d_table) % 16)
    return fail(LUCENT_E_INVALID, "environment map and table must be 16-byte aligned device arrays");
  uint32_t gw, gh;
  env_grid(width, height, max_width, &gw, &gh);
  const size_t n = size_t(gw) * gh;
  if (n > (size_t(1) << 28)) return fail(LUCENT_E_INVALID, "environment sampler grid larger than 2^28 cells");
  const uint32_t nb = uint32_t((n + kBlock - 1) / kBlock);
  // Scratch: w/pmf, scaled (f64 n each), small, large (u32 n) and their
  // values (f64 n), alias (u32 n), prob (f32 n), block counts (2 nb), info
  // (4 u32), total (f64), row sines (gh).
  size_t off = 0;
  auto take = [&](size_t bytes) {
    const size_t at = off;
    off += align256(bytes);
    return at;
  };
  const size_t o_w = take(8 * n), o_sc = take(8 * n), o_small = take(4 * n), o_large = take(4 * n);
  const size_t o_small_val = take(8 * n), o_large_val = take(8 * n);
  const size_t o_alias = take(4 * n), o_prob = take(4 * n), o_counts = take(8 * size_t(nb));
  const size_t o_info = take(16), o_total = take(8), o_sin = take(8 * size_t(gh));
  // The parallel pairing's prefix sums: per cell, then per list; and the
  // library's temporary storage for the sum and the scans.
  size_t o_dcell = 0, o_ecell = 0, o_dl = 0, o_el = 0, o_temp = 0, temp_bytes = 0;
  if (parallel) {
    o_dcell = take(8 * n);
    o_ecell = take(8 * n);
    o_dl = take(8 * n);
    o_el = take(8 * n);
    size_t a = 0, b = 0;
    hipError_t qe = hipcub::DeviceReduce::Sum(nullptr, a, static_cast<const double*>(nullptr),
                                              static_cast<double*>(nullptr), int(n));
    if (qe == hipSuccess)
      qe = hipcub::DeviceScan::InclusiveSum(
          nullptr, b, hipcub::TransformInputIterator<double, Deficit, const double*>(nullptr, Deficit{}),
          static_cast<double*>(nullptr), int(n));
    if (qe != hipSuccess) return fail_hip(qe, "environment sampler scan set-up");
    temp_bytes = std::max(a, b);
    o_temp = take(temp_bytes);
  }
  Status s = ensure_scratch(ctx, off);
  if (s != LUCENT_OK) return s;
  char* base = static_cast<char*>(ctx->scratch);
  double* w = reinterpret_cast<double*>(base + o_w);
  double* scaled = reinterpret_cast<double*>(base + o_sc);
  uint32_t* small = reinterpret_cast<uint32_t*>(base + o_small);
  uint32_t* large = reinterpret_cast<uint32_t*>(base + o_large);
  double* small_val = reinterpret_cast<double*>(base + o_small_val);
  double* large_val = reinterpret_cast<double*>(base + o_large_val);
  uint32_t* alias = reinterpret_cast<uint32_t*>(base + o_alias);
  float* prob = reinterpret_cast<float*>(base + o_prob);
  uint32_t* counts = reinterpret_cast<uint32_t*>(base + o_counts);
  uint32_t* info = reinterpret_cast<uint32_t*>(base + o_info);
  double* total = reinterpret_cast<double*>(base + o_total);
  double* row_sin = reinterpret_cast<double*>(base + o_sin);

  std::vector<double> sines(gh);
  for (uint32_t cy = 0; cy < gh; ++cy) sines[cy] = env_row_sine(cy, gh);
  const uint32_t info_init[4] = {kNoBad, 0, 0, 0};
  hipStream_t st = static_cast<hipStream_t>(stream ? stream : ctx->stream);
  if ((s = dev_upload(row_sin, sines.data(), 8 * size_t(gh), st)) != LUCENT_OK) return s;
  if ((s = dev_upload(info, info_init, sizeof info_init, st)) != LUCENT_OK) return s;
  hipEvent_t ev[2] = {nullptr, nullptr};
  hipError_t e = hipSuccess;
  if (ms_out) {
    e = hipEventCreate(&ev[0]);
    if (e == hipSuccess) e = hipEventCreate(&ev[1]);
    if (e == hipSuccess) e = hipEventRecord(ev[0], st);
  }
  const uint32_t un = uint32_t(n);
  if (e == hipSuccess) {
    k_env_cells<<<nb, kBlock, 0, st>>>(reinterpret_cast<const float4*>(d_rgba), width, height, gw, gh, row_sin, w,
                                        info);
    void* temp = base + o_temp;
    if (parallel)
      e = hipcub::DeviceReduce::Sum(temp, temp_bytes, static_cast<const double*>(w), total, int(n), st);
    else
      k_env_total<<<1, 64, 0, st>>>(w, un, info, total);
    if (e == hipSuccess) {
      k_env_split<<<nb, kBlock, 0, st>>>(w, scaled, total, un, info, counts);
      k_env_offsets<<<1, 1024, 0, st>>>(counts, nb, info);
      k_env_scatter<<<nb, kBlock, 0, st>>>(scaled, un, counts, info, small, large, small_val, large_val);
    }
    if (e == hipSuccess && parallel) {
      double* dcell = reinterpret_cast<double*>(base + o_dcell);
      double* ecell = reinterpret_cast<double*>(base + o_ecell);
      double* dl = reinterpret_cast<double*>(base + o_dl);
      double* el = reinterpret_cast<double*>(base + o_el);
      size_t tb = temp_bytes;
      e = hipcub::DeviceScan::InclusiveSum(
          temp, tb, hipcub::TransformInputIterator<double, Deficit, const double*>(scaled, Deficit{}), dcell, int(n),
          st);
      tb = temp_bytes;
      if (e == hipSuccess)
        e = hipcub::DeviceScan::InclusiveSum(
            temp, tb, hipcub::TransformInputIterator<double, Excess, const double*>(scaled, Excess{}), ecell, int(n),
            st);
      if (e == hipSuccess) {
        k_env_gather<<<nb, kBlock, 0, st>>>(small, large, dcell, ecell, info, dl, el);
        k_env_pair<<<nb, kBlock, 0, st>>>(small, large, small_val, dl, el, info, prob, alias);
      }
    } else if (e == hipSuccess) {
      k_env_sweep<<<1, 64, 0, st>>>(small, large, small_val, large_val, info, prob, alias);
    }
    if (e == hipSuccess) {
      k_env_pack<<<nb, kBlock, 0, st>>>(prob, alias, w, un, info, reinterpret_cast<float4*>(d_table));
      e = hipGetLastError();
    }
  }
  if (ms_out && e == hipSuccess) e = hipEventRecord(ev[1], st);
  uint32_t info_out[4] = {0, 0, 0, 0};
  double t = 0;
  if (e == hipSuccess) e = hipMemcpyAsync(info_out, info, sizeof info_out, hipMemcpyDeviceToHost, st);
  if (e == hipSuccess) e = hipMemcpyAsync(&t, total, sizeof t, hipMemcpyDeviceToHost, st);
  if (e == hipSuccess) e = hipStreamSynchronize(st);
  if (ms_out && e == hipSuccess) e = hipEventElapsedTime(ms_out, ev[0], ev[1]);
  for (hipEvent_t x : ev)
    if (x) (void)hipEventDestroy(x);
  if (e != hipSuccess) return fail_hip(e, "environment sampler build");
  if (info_out[0] != kNoBad)
    return fail(LUCENT_E_INVALID, "weight " + std::to_string(info_out[0]) + " is negative or not finite");
  if (grid_w) *grid_w = gw;
  if (grid_h) *grid_h = gh;
  if (total_out) *total_out = t;
  return LUCENT_OK;
}

}  // namespace lucent

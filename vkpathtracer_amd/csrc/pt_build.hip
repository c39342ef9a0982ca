// pt_build.hip — the bucketed-SAH BVH build on the GPU, producing the host builder's tree exactly.
//
// ptb::build_sah (the reference's bucketed SAH, D3D12GpuPathTracer.SceneBvh.cpp:73-238) splits a node
// with min/max reductions, per-primitive bucket indices, a sequential cost scan over the buckets and a
// stable partition. Without NaNs in the primitive boxes every one of these is independent of the order
// in which the primitives are visited (min and max are exact; the only values that depend on the order
// are signs of zero, which change no split: they cancel in every difference and are absorbed by the
// 1e-4 node padding), so the tree can be built one level at a time, one workgroup per node, and equal
// the depth-first host build bit for bit. Node indices are handed out in whatever order the workgroups
// finish; the host renumbers the nodes into the host builder's pre-order at the end.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <string>
#include <vector>

#include "pt_build.h"

namespace ptg {
namespace {

constexpr float kEps = 1e-4f;
constexpr uint32_t kMaxBuckets = 16;
constexpr uint32_t kBigSegment = 2048;    // segments above this use 1024-thread workgroups
constexpr uint32_t kHugeSegment = 65536;  // segments above this are split by many workgroups
constexpr uint32_t kHugeChunk = 4096;     // references per workgroup of a huge segment's passes

// Primitive references, structure of arrays: centroid, box, original index.
struct Refs {
  float* c[3];
  float* mn[3];
  float* mx[3];
  uint32_t* orig;
};

struct Seg {
  uint32_t start, count, node, depth;
};

struct DevNode {  // ptb::Node
  float mn[3], mx[3];
  int32_t left, right, first, count;
};

enum Counter : int {
  kNextBig = 0, kNextSmall = 1, kNextHuge = 2, kNodeCount = 3, kMaxDepth = 4, kLeafCount = 5, kCounters = 8
};

// std::min / std::max as the host builder writes them
__device__ __forceinline__ float smin(float a, float b) { return (b < a) ? b : a; }
__device__ __forceinline__ float smax(float a, float b) { return (a < b) ? b : a; }

// Order-preserving float <-> uint mapping for LDS min/max atomics (NaN-free inputs).
__device__ __forceinline__ uint32_t f2u(float f) {
  const uint32_t b = __float_as_uint(f);
  return (b & 0x80000000u) ? ~b : (b | 0x80000000u);
}
__device__ __forceinline__ float u2f(uint32_t u) { return __uint_as_float((u & 0x80000000u) ? (u & 0x7fffffffu) : ~u); }

__device__ __forceinline__ float bvh_sa(const float* mn, const float* mx) {
  float dx = mx[0] - mn[0], dy = mx[1] - mn[1], dz = mx[2] - mn[2];
  if (dx < 0) dx = 0;
  if (dy < 0) dy = 0;
  if (dz < 0) dz = 0;
  return 2.0f * (dx * dy + dy * dz + dz * dx);
}

// build_sah's reference records: box = (min3, max3) of the three vertices, centroid = (mn + mx) * 0.5f.
__global__ void k_build_refs(const float* raw, uint32_t n, Refs r) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float* t = raw + 9u * i;
  for (int k = 0; k < 3; ++k) {
    float m = t[k], M = t[k];
    if (t[3 + k] < m) m = t[3 + k];
    if (t[6 + k] < m) m = t[6 + k];
    if (M < t[3 + k]) M = t[3 + k];
    if (M < t[6 + k]) M = t[6 + k];
    r.mn[k][i] = m;
    r.mx[k][i] = M;
    r.c[k][i] = (m + M) * 0.5f;
  }
  r.orig[i] = i;
}

template <uint32_t NT>
__device__ __forceinline__ void block_minmax(float* v, bool* is_max, float (*red)[NT / 64u], uint32_t count) {
  const uint32_t w = threadIdx.x >> 6, lane = threadIdx.x & 63u;
  for (uint32_t k = 0; k < count; ++k) {
    float x = v[k];
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
      const float y = __shfl_xor(x, off);
      x = is_max[k] ? fmaxf(x, y) : fminf(x, y);
    }
    if (lane == 0u) red[k][w] = x;
  }
  __syncthreads();
  for (uint32_t k = 0; k < count; ++k) {
    float x = red[k][0];
    for (uint32_t j = 1; j < NT / 64u; ++j) x = is_max[k] ? fmaxf(x, red[k][j]) : fminf(x, red[k][j]);
    v[k] = x;
  }
  __syncthreads();
}

template <uint32_t NT>
__device__ __forceinline__ uint32_t block_sum(uint32_t x, uint32_t* red) {
  const uint32_t w = threadIdx.x >> 6, lane = threadIdx.x & 63u;
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) x += __shfl_xor(x, off);
  if (lane == 0u) red[w] = x;
  __syncthreads();
  uint32_t s = 0;
  for (uint32_t j = 0; j < NT / 64u; ++j) s += red[j];
  __syncthreads();
  return s;
}

// One node per workgroup: its bounds, then a leaf, or the bucketed SAH split (binning of all three
// axes into LDS, the host builder's cost scan by one thread) and the stable partition of its
// references from A into B, and its two children appended to the next level.
template <uint32_t NT>
__global__ __launch_bounds__(NT) void k_split(const Seg* segs, Refs A, Refs B, DevNode* nodes, uint32_t* order,
                                               Seg* next_big, Seg* next_small, uint32_t* counters, uint32_t leaf,
                                               uint32_t nb) {
  __shared__ float red[12][NT / 64u];
  __shared__ uint32_t ured[NT / 64u];
  __shared__ uint32_t bin_n[3][kMaxBuckets];
  __shared__ uint32_t bin_b[3][kMaxBuckets][6];  // mn xyz, mx xyz as ordered uints
  __shared__ int s_axis, s_found;
  __shared__ float s_split;
  const Seg sg = segs[blockIdx.x];
  const uint32_t tid = threadIdx.x;

  // bounds of the boxes and of the centroids (the host starts from +-1e30)
  float v[12];
  bool is_max[12];
  for (int k = 0; k < 3; ++k) {
    v[k] = 1e30f; is_max[k] = false;            // box min
    v[3 + k] = -1e30f; is_max[3 + k] = true;    // box max
    v[6 + k] = 1e30f; is_max[6 + k] = false;    // centroid min
    v[9 + k] = -1e30f; is_max[9 + k] = true;    // centroid max
  }
  for (uint32_t i = tid; i < sg.count; i += NT) {
    const uint32_t p = sg.start + i;
    for (int k = 0; k < 3; ++k) {
      v[k] = smin(v[k], A.mn[k][p]);
      v[3 + k] = smax(v[3 + k], A.mx[k][p]);
      v[6 + k] = smin(v[6 + k], A.c[k][p]);
      v[9 + k] = smax(v[9 + k], A.c[k][p]);
    }
  }
  block_minmax<NT>(v, is_max, red, 12);
  const float bmn[3] = {v[0], v[1], v[2]}, bmx[3] = {v[3], v[4], v[5]};
  DevNode node;
  for (int k = 0; k < 3; ++k) {
    node.mn[k] = bmn[k] - kEps;
    node.mx[k] = bmx[k] + kEps;
  }
  if (sg.count <= leaf) {  // a leaf: its references in order are its slice of the leaf order
    for (uint32_t i = tid; i < sg.count; i += NT) order[sg.start + i] = A.orig[sg.start + i];
    if (tid == 0u) {
      node.left = node.right = -1;
      node.first = static_cast<int32_t>(sg.start);
      node.count = static_cast<int32_t>(sg.count);
      nodes[sg.node] = node;
      atomicMax(counters + kMaxDepth, sg.depth);
      atomicAdd(counters + kLeafCount, 1u);
    }
    return;
  }
  const float fB = static_cast<float>(nb);
  float cmin[3], extent[3];
  for (int k = 0; k < 3; ++k) {
    cmin[k] = v[6 + k];
    extent[k] = v[9 + k] - v[6 + k];
  }
  for (uint32_t i = tid; i < 3u * kMaxBuckets; i += NT) {
    bin_n[i / kMaxBuckets][i % kMaxBuckets] = 0u;
    for (int k = 0; k < 3; ++k) {
      bin_b[i / kMaxBuckets][i % kMaxBuckets][k] = f2u(1e30f);
      bin_b[i / kMaxBuckets][i % kMaxBuckets][3 + k] = f2u(-1e30f);
    }
  }
  __syncthreads();
  for (uint32_t i = tid; i < sg.count; i += NT) {
    const uint32_t p = sg.start + i;
    uint32_t bm[6];
    for (int k = 0; k < 3; ++k) {
      bm[k] = f2u(A.mn[k][p]);
      bm[3 + k] = f2u(A.mx[k][p]);
    }
    for (int axis = 0; axis < 3; ++axis) {
      if (extent[axis] < 1e-6f) continue;
      // the host builder's bucket index, NaN and infinity included (not reached: such scenes build on
      // the host), so the LDS index is in range whatever the input
      float x = (A.c[axis][p] - cmin[axis]) / extent[axis] * fB;
      if (!(x >= 0.0f)) x = 0.0f;
      int b = x < fB ? static_cast<int>(x) : static_cast<int>(nb) - 1;
      if (static_cast<int>(nb) - 1 < b) b = static_cast<int>(nb) - 1;
      atomicAdd(&bin_n[axis][b], 1u);
      for (int k = 0; k < 3; ++k) {
        atomicMin(&bin_b[axis][b][k], bm[k]);
        atomicMax(&bin_b[axis][b][3 + k], bm[3 + k]);
      }
    }
  }
  __syncthreads();
  if (tid == 0u) {  // the host builder's cost scan, in its order and float expressions
    const float node_sa = smax(1e-9f, bvh_sa(bmn, bmx));
    int best_axis = 0;
    float best_split = 0.0f, best_cost = 1e30f;
    int found = 0;
    for (int axis = 0; axis < 3; ++axis) {
      if (extent[axis] < 1e-6f) continue;
      for (uint32_t s = 1; s < nb; ++s) {
        float lmn[3] = {1e30f, 1e30f, 1e30f}, lmx[3] = {-1e30f, -1e30f, -1e30f};
        float rmn[3] = {1e30f, 1e30f, 1e30f}, rmx[3] = {-1e30f, -1e30f, -1e30f};
        uint32_t lc = 0, rc = 0;
        for (uint32_t b = 0; b < s; ++b) {
          if (!bin_n[axis][b]) continue;
          lc += bin_n[axis][b];
          for (int k = 0; k < 3; ++k) {
            lmn[k] = smin(lmn[k], u2f(bin_b[axis][b][k]));
            lmx[k] = smax(lmx[k], u2f(bin_b[axis][b][3 + k]));
          }
        }
        for (uint32_t b = s; b < nb; ++b) {
          if (!bin_n[axis][b]) continue;
          rc += bin_n[axis][b];
          for (int k = 0; k < 3; ++k) {
            rmn[k] = smin(rmn[k], u2f(bin_b[axis][b][k]));
            rmx[k] = smax(rmx[k], u2f(bin_b[axis][b][3 + k]));
          }
        }
        if (!lc || !rc) continue;
        const float cost = (bvh_sa(lmn, lmx) * static_cast<float>(lc) + bvh_sa(rmn, rmx) * static_cast<float>(rc)) / node_sa;
        if (cost < best_cost) {
          best_cost = cost;
          best_axis = axis;
          best_split = cmin[axis] + static_cast<float>(s) / fB * extent[axis];
          found = 1;
        }
      }
    }
    s_axis = best_axis;
    s_split = best_split;
    s_found = found;
  }
  __syncthreads();
  const int axis = s_axis;
  const float split = s_split;
  const bool found = s_found != 0;
  // stable partition into B: the left side (centroid below the split) keeps its order, then the right
  uint32_t nl = 0;
  if (found) {
    uint32_t mine = 0;
    for (uint32_t i = tid; i < sg.count; i += NT) mine += A.c[axis][sg.start + i] < split ? 1u : 0u;
    nl = block_sum<NT>(mine, ured);
  }
  uint32_t lbase = 0, rbase = 0;
  for (uint32_t t0 = 0; t0 < sg.count; t0 += NT) {
    const uint32_t i = t0 + tid;
    const bool in = i < sg.count;
    const uint32_t p = sg.start + i;
    const bool left = in && found && A.c[axis][p] < split;
    // exclusive rank of this thread among the tile's left (and right) threads
    const uint64_t lm = __ballot(left), am = __ballot(in);
    const uint32_t lane = tid & 63u, w = tid >> 6;
    const uint64_t below = lane ? (~0ull >> (64u - lane)) : 0ull;
    const uint32_t lrank_w = static_cast<uint32_t>(__popcll(lm & below));
    const uint32_t rrank_w = static_cast<uint32_t>(__popcll(am & ~lm & below));
    __shared__ uint32_t wl[NT / 64u], wr[NT / 64u];
    if (lane == 0u) {
      wl[w] = static_cast<uint32_t>(__popcll(lm));
      wr[w] = static_cast<uint32_t>(__popcll(am & ~lm));
    }
    __syncthreads();
    uint32_t lpre = 0, rpre = 0, ltot = 0, rtot = 0;
    for (uint32_t j = 0; j < NT / 64u; ++j) {
      if (j < w) {
        lpre += wl[j];
        rpre += wr[j];
      }
      ltot += wl[j];
      rtot += wr[j];
    }
    __syncthreads();
    if (in) {
      const uint32_t dst = sg.start + (left ? lbase + lpre + lrank_w : nl + rbase + rpre + rrank_w);
      for (int k = 0; k < 3; ++k) {
        B.c[k][dst] = A.c[k][p];
        B.mn[k][dst] = A.mn[k][p];
        B.mx[k][dst] = A.mx[k][p];
      }
      B.orig[dst] = A.orig[p];
    }
    lbase += ltot;
    rbase += rtot;
  }
  if (tid == 0u) {
    uint32_t mid = found ? nl : sg.count / 2u;
    if (mid == 0u || mid >= sg.count) mid = sg.count / 2u;
    const uint32_t base = atomicAdd(counters + kNodeCount, 2u);
    node.left = static_cast<int32_t>(base);
    node.right = static_cast<int32_t>(base + 1u);
    node.first = -1;
    node.count = 0;
    nodes[sg.node] = node;
    const Seg kids[2] = {{sg.start, mid, base, sg.depth + 1u}, {sg.start + mid, sg.count - mid, base + 1u, sg.depth + 1u}};
    for (int c = 0; c < 2; ++c) {
      if (kids[c].count > kBigSegment) next_big[atomicAdd(counters + kNextBig, 1u)] = kids[c];
      else next_small[atomicAdd(counters + kNextSmall, 1u)] = kids[c];
    }
  }
}

// A huge segment (more references than kHugeSegment) is split by the same steps spread over many
// workgroups of kHugeChunk references each; one launch per step serves every huge segment of a level
// (blockIdx.y is the segment). Bounds and bins are combined with global atomics (min, max and integer
// sums: the same values in any order), the split is chosen by one thread as in k_split, and the stable
// partition places each chunk by the left counts of the chunks before it.
struct HugeAcc {
  uint32_t bounds[12];  // ordered: box min xyz, box max xyz, centroid min xyz, centroid max xyz
  uint32_t bin_n[3][kMaxBuckets];
  uint32_t bin_b[3][kMaxBuckets][6];
  int found, axis;
  float split;
  uint32_t nl;
};

__device__ __forceinline__ bool huge_chunk(const Seg* huge, Seg& sg, uint32_t& c0, uint32_t& c1) {
  sg = huge[blockIdx.y];
  c0 = blockIdx.x * kHugeChunk;
  c1 = min(sg.count, c0 + kHugeChunk);
  return c0 < sg.count;
}

__global__ void k_huge_init(HugeAcc* accs) {
  HugeAcc* acc = accs + blockIdx.x;
  for (uint32_t i = threadIdx.x; i < 12u; i += blockDim.x) acc->bounds[i] = f2u((i % 6u) < 3u ? 1e30f : -1e30f);
  for (uint32_t i = threadIdx.x; i < 3u * kMaxBuckets; i += blockDim.x) {
    acc->bin_n[i / kMaxBuckets][i % kMaxBuckets] = 0u;
    for (int k = 0; k < 3; ++k) {
      acc->bin_b[i / kMaxBuckets][i % kMaxBuckets][k] = f2u(1e30f);
      acc->bin_b[i / kMaxBuckets][i % kMaxBuckets][3 + k] = f2u(-1e30f);
    }
  }
}

__global__ __launch_bounds__(256) void k_huge_bounds(const Seg* huge, Refs A, HugeAcc* accs) {
  __shared__ float red[12][4];
  Seg sg;
  uint32_t c0, c1;
  if (!huge_chunk(huge, sg, c0, c1)) return;
  float v[12];
  bool is_max[12];
  for (int k = 0; k < 3; ++k) {
    v[k] = 1e30f; is_max[k] = false;
    v[3 + k] = -1e30f; is_max[3 + k] = true;
    v[6 + k] = 1e30f; is_max[6 + k] = false;
    v[9 + k] = -1e30f; is_max[9 + k] = true;
  }
  for (uint32_t i = c0 + threadIdx.x; i < c1; i += 256u) {
    const uint32_t p = sg.start + i;
    for (int k = 0; k < 3; ++k) {
      v[k] = smin(v[k], A.mn[k][p]);
      v[3 + k] = smax(v[3 + k], A.mx[k][p]);
      v[6 + k] = smin(v[6 + k], A.c[k][p]);
      v[9 + k] = smax(v[9 + k], A.c[k][p]);
    }
  }
  block_minmax<256>(v, is_max, red, 12);
  if (threadIdx.x < 12u) {
    HugeAcc* acc = accs + blockIdx.y;
    const uint32_t u = f2u(v[threadIdx.x]);
    if (is_max[threadIdx.x]) atomicMax(&acc->bounds[threadIdx.x], u);
    else atomicMin(&acc->bounds[threadIdx.x], u);
  }
}

__global__ __launch_bounds__(256) void k_huge_bins(const Seg* huge, Refs A, HugeAcc* accs, uint32_t nb) {
  __shared__ uint32_t bin_n[3][kMaxBuckets];
  __shared__ uint32_t bin_b[3][kMaxBuckets][6];
  Seg sg;
  uint32_t c0, c1;
  if (!huge_chunk(huge, sg, c0, c1)) return;
  HugeAcc* acc = accs + blockIdx.y;
  const float fB = static_cast<float>(nb);
  float cmin[3], extent[3];
  for (int k = 0; k < 3; ++k) {
    cmin[k] = u2f(acc->bounds[6 + k]);
    extent[k] = u2f(acc->bounds[9 + k]) - cmin[k];
  }
  for (uint32_t i = threadIdx.x; i < 3u * kMaxBuckets; i += 256u) {
    bin_n[i / kMaxBuckets][i % kMaxBuckets] = 0u;
    for (int k = 0; k < 3; ++k) {
      bin_b[i / kMaxBuckets][i % kMaxBuckets][k] = f2u(1e30f);
      bin_b[i / kMaxBuckets][i % kMaxBuckets][3 + k] = f2u(-1e30f);
    }
  }
  __syncthreads();
  for (uint32_t i = c0 + threadIdx.x; i < c1; i += 256u) {
    const uint32_t p = sg.start + i;
    uint32_t bm[6];
    for (int k = 0; k < 3; ++k) {
      bm[k] = f2u(A.mn[k][p]);
      bm[3 + k] = f2u(A.mx[k][p]);
    }
    for (int axis = 0; axis < 3; ++axis) {
      if (extent[axis] < 1e-6f) continue;
      float x = (A.c[axis][p] - cmin[axis]) / extent[axis] * fB;  // k_split's bucket index
      if (!(x >= 0.0f)) x = 0.0f;
      int b = x < fB ? static_cast<int>(x) : static_cast<int>(nb) - 1;
      if (static_cast<int>(nb) - 1 < b) b = static_cast<int>(nb) - 1;
      atomicAdd(&bin_n[axis][b], 1u);
      for (int k = 0; k < 3; ++k) {
        atomicMin(&bin_b[axis][b][k], bm[k]);
        atomicMax(&bin_b[axis][b][3 + k], bm[3 + k]);
      }
    }
  }
  __syncthreads();
  for (uint32_t i = threadIdx.x; i < 3u * kMaxBuckets; i += 256u) {
    const uint32_t a = i / kMaxBuckets, b = i % kMaxBuckets;
    if (!bin_n[a][b]) continue;
    atomicAdd(&acc->bin_n[a][b], bin_n[a][b]);
    for (int k = 0; k < 3; ++k) {
      atomicMin(&acc->bin_b[a][b][k], bin_b[a][b][k]);
      atomicMax(&acc->bin_b[a][b][3 + k], bin_b[a][b][3 + k]);
    }
  }
}

// k_split's cost scan over the combined bins, one thread per segment.
__global__ void k_huge_decide(HugeAcc* accs, uint32_t nb) {
  if (threadIdx.x != 0u) return;
  HugeAcc* acc = accs + blockIdx.x;
  const float fB = static_cast<float>(nb);
  float bmn[3], bmx[3], cmin[3], extent[3];
  for (int k = 0; k < 3; ++k) {
    bmn[k] = u2f(acc->bounds[k]);
    bmx[k] = u2f(acc->bounds[3 + k]);
    cmin[k] = u2f(acc->bounds[6 + k]);
    extent[k] = u2f(acc->bounds[9 + k]) - cmin[k];
  }
  const float node_sa = smax(1e-9f, bvh_sa(bmn, bmx));
  int best_axis = 0;
  float best_split = 0.0f, best_cost = 1e30f;
  int found = 0;
  for (int axis = 0; axis < 3; ++axis) {
    if (extent[axis] < 1e-6f) continue;
    for (uint32_t s = 1; s < nb; ++s) {
      float lmn[3] = {1e30f, 1e30f, 1e30f}, lmx[3] = {-1e30f, -1e30f, -1e30f};
      float rmn[3] = {1e30f, 1e30f, 1e30f}, rmx[3] = {-1e30f, -1e30f, -1e30f};
      uint32_t lc = 0, rc = 0;
      for (uint32_t b = 0; b < s; ++b) {
        if (!acc->bin_n[axis][b]) continue;
        lc += acc->bin_n[axis][b];
        for (int k = 0; k < 3; ++k) {
          lmn[k] = smin(lmn[k], u2f(acc->bin_b[axis][b][k]));
          lmx[k] = smax(lmx[k], u2f(acc->bin_b[axis][b][3 + k]));
        }
      }
      for (uint32_t b = s; b < nb; ++b) {
        if (!acc->bin_n[axis][b]) continue;
        rc += acc->bin_n[axis][b];
        for (int k = 0; k < 3; ++k) {
          rmn[k] = smin(rmn[k], u2f(acc->bin_b[axis][b][k]));
          rmx[k] = smax(rmx[k], u2f(acc->bin_b[axis][b][3 + k]));
        }
      }
      if (!lc || !rc) continue;
      const float cost = (bvh_sa(lmn, lmx) * static_cast<float>(lc) + bvh_sa(rmn, rmx) * static_cast<float>(rc)) / node_sa;
      if (cost < best_cost) {
        best_cost = cost;
        best_axis = axis;
        best_split = cmin[axis] + static_cast<float>(s) / fB * extent[axis];
        found = 1;
      }
    }
  }
  acc->found = found;
  acc->axis = best_axis;
  acc->split = best_split;
}

__global__ __launch_bounds__(256) void k_huge_count(const Seg* huge, Refs A, const HugeAcc* accs, uint32_t* chunk_left,
                                                    uint32_t max_chunks) {
  __shared__ uint32_t ured[4];
  Seg sg;
  uint32_t c0, c1;
  if (!huge_chunk(huge, sg, c0, c1)) return;
  const HugeAcc* acc = accs + blockIdx.y;
  uint32_t mine = 0;
  if (acc->found) {
    const int axis = acc->axis;
    const float split = acc->split;
    for (uint32_t i = c0 + threadIdx.x; i < c1; i += 256u) mine += A.c[axis][sg.start + i] < split ? 1u : 0u;
  }
  const uint32_t s = block_sum<256>(mine, ured);
  if (threadIdx.x == 0u) chunk_left[blockIdx.y * max_chunks + blockIdx.x] = s;
}

// Per segment (one thread): the chunks' left counts become exclusive offsets in place, then the node
// record and the two children, as k_split writes them.
__global__ void k_huge_finish(const Seg* huge, HugeAcc* accs, uint32_t* chunk_left, uint32_t max_chunks,
                              DevNode* nodes, Seg* next_huge, Seg* next_big, Seg* next_small, uint32_t* counters) {
  if (threadIdx.x != 0u) return;
  const Seg sg = huge[blockIdx.x];
  HugeAcc* acc = accs + blockIdx.x;
  uint32_t* cl = chunk_left + blockIdx.x * max_chunks;
  const uint32_t chunks = (sg.count + kHugeChunk - 1u) / kHugeChunk;
  uint32_t nl = 0;
  for (uint32_t b = 0; b < chunks; ++b) {
    const uint32_t c = cl[b];
    cl[b] = nl;
    nl += c;
  }
  acc->nl = nl;
  uint32_t mid = acc->found ? nl : sg.count / 2u;
  if (mid == 0u || mid >= sg.count) mid = sg.count / 2u;
  DevNode node;
  for (int k = 0; k < 3; ++k) {
    node.mn[k] = u2f(acc->bounds[k]) - kEps;
    node.mx[k] = u2f(acc->bounds[3 + k]) + kEps;
  }
  const uint32_t base = atomicAdd(counters + kNodeCount, 2u);
  node.left = static_cast<int32_t>(base);
  node.right = static_cast<int32_t>(base + 1u);
  node.first = -1;
  node.count = 0;
  nodes[sg.node] = node;
  const Seg kids[2] = {{sg.start, mid, base, sg.depth + 1u}, {sg.start + mid, sg.count - mid, base + 1u, sg.depth + 1u}};
  for (int c = 0; c < 2; ++c) {
    if (kids[c].count > kHugeSegment) next_huge[atomicAdd(counters + kNextHuge, 1u)] = kids[c];
    else if (kids[c].count > kBigSegment) next_big[atomicAdd(counters + kNextBig, 1u)] = kids[c];
    else next_small[atomicAdd(counters + kNextSmall, 1u)] = kids[c];
  }
}

// k_split's stable partition, each chunk starting from the offsets of the chunks before it (every
// earlier chunk is full, so its right-side count is its size minus its left count).
__global__ __launch_bounds__(256) void k_huge_scatter(const Seg* huge, Refs A, Refs B, const HugeAcc* accs,
                                                      const uint32_t* chunk_left, uint32_t max_chunks) {
  __shared__ uint32_t wl[4], wr[4];
  Seg sg;
  uint32_t c0, c1;
  if (!huge_chunk(huge, sg, c0, c1)) return;
  const HugeAcc* acc = accs + blockIdx.y;
  const bool found = acc->found != 0;
  const int axis = acc->axis;
  const float split = acc->split;
  const uint32_t nl = acc->nl;
  const uint32_t first_left = chunk_left[blockIdx.y * max_chunks + blockIdx.x];
  uint32_t lbase = first_left, rbase = c0 - first_left;
  const uint32_t lane = threadIdx.x & 63u, w = threadIdx.x >> 6;
  const uint64_t below = lane ? (~0ull >> (64u - lane)) : 0ull;
  for (uint32_t t0 = c0; t0 < c1; t0 += 256u) {
    const uint32_t i = t0 + threadIdx.x;
    const bool in = i < c1;
    const uint32_t p = sg.start + i;
    const bool left = in && found && A.c[axis][p] < split;
    const uint64_t lm = __ballot(left), am = __ballot(in);
    if (lane == 0u) {
      wl[w] = static_cast<uint32_t>(__popcll(lm));
      wr[w] = static_cast<uint32_t>(__popcll(am & ~lm));
    }
    __syncthreads();
    uint32_t lpre = 0, rpre = 0, ltot = 0, rtot = 0;
    for (uint32_t j = 0; j < 4u; ++j) {
      if (j < w) {
        lpre += wl[j];
        rpre += wr[j];
      }
      ltot += wl[j];
      rtot += wr[j];
    }
    __syncthreads();
    if (in) {
      const uint32_t dst = sg.start + (left ? lbase + lpre + static_cast<uint32_t>(__popcll(lm & below))
                                            : nl + rbase + rpre + static_cast<uint32_t>(__popcll(am & ~lm & below)));
      for (int k = 0; k < 3; ++k) {
        B.c[k][dst] = A.c[k][p];
        B.mn[k][dst] = A.mn[k][p];
        B.mx[k][dst] = A.mx[k][p];
      }
      B.orig[dst] = A.orig[p];
    }
    lbase += ltot;
    rbase += rtot;
  }
}

// Renumbering into the host builder's pre-order. The nodes created while level L is split are level
// L + 1, a contiguous range of indices, so each pass is one launch per level: subtree sizes from the
// deepest level up, pre-order positions from the root down, then the scatter.
__global__ void k_subtree_sizes(const DevNode* nodes, uint32_t* size, uint32_t lo, uint32_t hi) {
  const uint32_t b = lo + blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= hi) return;
  const DevNode& d = nodes[b];
  size[b] = d.count > 0 ? 1u : 1u + size[d.left] + size[d.right];
}

__global__ void k_preorder(const DevNode* nodes, const uint32_t* size, uint32_t* pre, uint32_t lo, uint32_t hi) {
  const uint32_t b = lo + blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= hi) return;
  const DevNode& d = nodes[b];
  if (d.count > 0) return;
  pre[d.left] = pre[b] + 1u;
  pre[d.right] = pre[b] + 1u + size[d.left];
}

__global__ void k_scatter_preorder(const DevNode* nodes, const uint32_t* pre, DevNode* out, uint32_t n) {
  const uint32_t b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= n) return;
  DevNode d = nodes[b];
  if (d.count == 0) {
    d.left = static_cast<int32_t>(pre[d.left]);
    d.right = static_cast<int32_t>(pre[d.right]);
  }
  out[pre[b]] = d;
}

// Stream-ordered scratch: allocation and release follow the build's stream, so builds of several
// scenes on their own streams (the batch builds them from parallel host threads) do not synchronise
// the device as hipFree would.
template <typename T>
struct DevBuf {
  T* p = nullptr;
  hipStream_t s = nullptr;
  ~DevBuf() {
    if (p) (void)hipFreeAsync(p, s);
  }
  hipError_t alloc(size_t n, hipStream_t stream) {
    s = stream;
    return hipMallocAsync(reinterpret_cast<void**>(&p), std::max<size_t>(n, 1) * sizeof(T), stream);
  }
};

}  // namespace

int build_sah_device(hipStream_t stream, ptb::Bvh& bvh, const std::vector<float>& raw, uint32_t leaf_size,
                     uint32_t buckets, std::string& err) {
  const uint32_t n = static_cast<uint32_t>(bvh.prims.size());
  bvh.nodes.clear();
  bvh.order.clear();
  bvh.max_depth = 0;
  bvh.leaf_count = 0;
  if (n == 0) return 0;
  // NaN boxes make min/max order-dependent, and coordinates beyond the builder's 1e30 sentinels can
  // overflow a centroid: the host builds such scenes
  for (float x : raw)
    if (!(std::fabs(x) <= 1e30f)) return 1;
  const uint32_t leaf = leaf_size ? leaf_size : 4u;
  const uint32_t nb = std::clamp(buckets ? buckets : 8u, 2u, kMaxBuckets);
  DevBuf<float> d_raw, d_ref[2];
  DevBuf<uint32_t> d_orig[2], d_order, d_counters;
  DevBuf<Seg> d_seg[4];   // big and small lists of the current and the next level
  DevBuf<Seg> d_huge[2];  // huge lists of the current and the next level
  DevBuf<HugeAcc> d_acc;
  DevBuf<uint32_t> d_chunk_left;
  DevBuf<DevNode> d_nodes, d_pre_nodes;
  auto check = [&](hipError_t e, const char* what) {
    if (e != hipSuccess) err = std::string("device BVH build: ") + what + ": " + hipGetErrorString(e);
    return e == hipSuccess;
  };
  const size_t n9 = 9u * static_cast<size_t>(n), n2 = 2u * static_cast<size_t>(n);
  if (!check(d_raw.alloc(n9, stream), "allocate") || !check(d_ref[0].alloc(n9, stream), "allocate") ||
      !check(d_ref[1].alloc(n9, stream), "allocate") || !check(d_orig[0].alloc(n, stream), "allocate") ||
      !check(d_orig[1].alloc(n, stream), "allocate") || !check(d_order.alloc(n, stream), "allocate") ||
      !check(d_counters.alloc(kCounters, stream), "allocate") || !check(d_nodes.alloc(n2, stream), "allocate") ||
      !check(d_pre_nodes.alloc(n2, stream), "allocate"))
    return -1;
  for (auto& s : d_seg)
    if (!check(s.alloc(n, stream), "allocate")) return -1;
  // huge segments are disjoint and each holds more than kHugeSegment references
  const uint32_t max_huge = n / kHugeSegment + 1u, max_chunks = (n + kHugeChunk - 1u) / kHugeChunk;
  for (auto& s : d_huge)
    if (!check(s.alloc(max_huge, stream), "allocate")) return -1;
  if (!check(d_acc.alloc(max_huge, stream), "allocate") ||
      !check(d_chunk_left.alloc(static_cast<size_t>(max_huge) * max_chunks, stream), "allocate"))
    return -1;
  Refs R[2];
  for (int b = 0; b < 2; ++b) {
    for (int k = 0; k < 3; ++k) {
      R[b].c[k] = d_ref[b].p + static_cast<size_t>(k) * n;
      R[b].mn[k] = d_ref[b].p + static_cast<size_t>(3 + k) * n;
      R[b].mx[k] = d_ref[b].p + static_cast<size_t>(6 + k) * n;
    }
    R[b].orig = d_orig[b].p;
  }
  if (!check(hipMemcpyAsync(d_raw.p, raw.data(), raw.size() * sizeof(float), hipMemcpyHostToDevice, stream), "upload"))
    return -1;
  hipLaunchKernelGGL(k_build_refs, dim3((n + 255u) / 256u), dim3(256), 0, stream, d_raw.p, n, R[0]);
  uint32_t counters[kCounters] = {};
  counters[kNodeCount] = 1u;  // the root is node 0
  const Seg root{0u, n, 0u, 0u};
  uint32_t nhuge = n > kHugeSegment ? 1u : 0u, nbig = !nhuge && n > kBigSegment ? 1u : 0u;
  uint32_t nsmall = nhuge + nbig ? 0u : 1u;
  Seg* root_list = nhuge ? d_huge[0].p : d_seg[nbig ? 0 : 1].p;
  std::vector<Seg> huge(max_huge);
  if (nhuge) huge[0] = root;
  if (!check(hipMemcpyAsync(root_list, &root, sizeof(Seg), hipMemcpyHostToDevice, stream), "upload") ||
      !check(hipMemcpyAsync(d_counters.p, counters, sizeof(counters), hipMemcpyHostToDevice, stream), "upload"))
    return -1;
  int cur = 0;  // lists 2 * cur (big), 2 * cur + 1 (small), d_huge[cur]; references in R[level % 2]
  std::vector<uint32_t> level_start{0u, 1u};  // nodes of level L: [level_start[L], level_start[L + 1])
  for (uint32_t level = 0; nhuge + nbig + nsmall > 0u; ++level) {
    const uint32_t zero3[3] = {0u, 0u, 0u};  // kNextBig, kNextSmall, kNextHuge
    if (!check(hipMemcpyAsync(d_counters.p, zero3, sizeof(zero3), hipMemcpyHostToDevice, stream), "reset"))
      return -1;
    const Refs& A = R[level % 2u];
    const Refs& B = R[(level + 1u) % 2u];
    Seg* nb_list = d_seg[2 * (1 - cur)].p;
    Seg* ns_list = d_seg[2 * (1 - cur) + 1].p;
    if (nhuge) {
      uint32_t grid_x = 0;
      for (uint32_t h = 0; h < nhuge; ++h) grid_x = std::max(grid_x, (huge[h].count + kHugeChunk - 1u) / kHugeChunk);
      const Seg* hl = d_huge[cur].p;
      const dim3 grid(grid_x, nhuge);
      hipLaunchKernelGGL(k_huge_init, dim3(nhuge), dim3(256), 0, stream, d_acc.p);
      hipLaunchKernelGGL(k_huge_bounds, grid, dim3(256), 0, stream, hl, A, d_acc.p);
      hipLaunchKernelGGL(k_huge_bins, grid, dim3(256), 0, stream, hl, A, d_acc.p, nb);
      hipLaunchKernelGGL(k_huge_decide, dim3(nhuge), dim3(64), 0, stream, d_acc.p, nb);
      hipLaunchKernelGGL(k_huge_count, grid, dim3(256), 0, stream, hl, A, d_acc.p, d_chunk_left.p, max_chunks);
      hipLaunchKernelGGL(k_huge_finish, dim3(nhuge), dim3(64), 0, stream, hl, d_acc.p, d_chunk_left.p, max_chunks,
                         d_nodes.p, d_huge[1 - cur].p, nb_list, ns_list, d_counters.p);
      hipLaunchKernelGGL(k_huge_scatter, grid, dim3(256), 0, stream, hl, A, B, d_acc.p, d_chunk_left.p, max_chunks);
    }
    if (nbig)
      hipLaunchKernelGGL(k_split<1024>, dim3(nbig), dim3(1024), 0, stream, d_seg[2 * cur].p, A, B, d_nodes.p,
                         d_order.p, nb_list, ns_list, d_counters.p, leaf, nb);
    if (nsmall)
      hipLaunchKernelGGL(k_split<64>, dim3(nsmall), dim3(64), 0, stream, d_seg[2 * cur + 1].p, A, B, d_nodes.p,
                         d_order.p, nb_list, ns_list, d_counters.p, leaf, nb);
    // the next huge list travels with the counters (it is short: at most n / kHugeSegment entries)
    if (!check(hipGetLastError(), "launch") ||
        !check(hipMemcpyAsync(counters, d_counters.p, sizeof(counters), hipMemcpyDeviceToHost, stream), "read") ||
        (nhuge && !check(hipMemcpyAsync(huge.data(), d_huge[1 - cur].p, max_huge * sizeof(Seg), hipMemcpyDeviceToHost,
                                        stream), "read")) ||
        !check(hipStreamSynchronize(stream), "synchronize"))
      return -1;
    nhuge = counters[kNextHuge];
    nbig = counters[kNextBig];
    nsmall = counters[kNextSmall];
    level_start.push_back(counters[kNodeCount]);
    cur = 1 - cur;
  }
  const uint32_t node_count = counters[kNodeCount];
  // pre-order renumbering on the device (the arrays of the finished lists are reused)
  uint32_t* size = reinterpret_cast<uint32_t*>(d_seg[0].p);  // node_count <= 2n - 1 < 4n words
  uint32_t* pre = reinterpret_cast<uint32_t*>(d_seg[2].p);
  DevNode* out = d_pre_nodes.p;
  static_assert(sizeof(Seg) == 16 && sizeof(DevNode) == 40 && sizeof(ptb::Node) == 40, "record sizes");
  const uint32_t levels = static_cast<uint32_t>(level_start.size()) - 1u;
  for (uint32_t L = levels; L-- > 0;) {
    const uint32_t lo = level_start[L], hi = level_start[L + 1];
    if (hi > lo)
      hipLaunchKernelGGL(k_subtree_sizes, dim3((hi - lo + 255u) / 256u), dim3(256), 0, stream, d_nodes.p, size, lo, hi);
  }
  const uint32_t zero = 0u;
  if (!check(hipMemcpyAsync(pre, &zero, sizeof(zero), hipMemcpyHostToDevice, stream), "upload")) return -1;
  for (uint32_t L = 0; L < levels; ++L) {
    const uint32_t lo = level_start[L], hi = level_start[L + 1];
    if (hi > lo)
      hipLaunchKernelGGL(k_preorder, dim3((hi - lo + 255u) / 256u), dim3(256), 0, stream, d_nodes.p, size, pre, lo, hi);
  }
  hipLaunchKernelGGL(k_scatter_preorder, dim3((node_count + 255u) / 256u), dim3(256), 0, stream, d_nodes.p, pre, out,
                     node_count);
  bvh.nodes.resize(node_count);
  bvh.order.resize(n);
  if (!check(hipGetLastError(), "launch") ||
      !check(hipMemcpyAsync(bvh.nodes.data(), out, node_count * sizeof(DevNode), hipMemcpyDeviceToHost, stream),
             "download") ||
      !check(hipMemcpyAsync(bvh.order.data(), d_order.p, n * sizeof(uint32_t), hipMemcpyDeviceToHost, stream),
             "download") ||
      !check(hipStreamSynchronize(stream), "synchronize"))
    return -1;
  bvh.max_depth = counters[kMaxDepth];
  bvh.leaf_count = counters[kLeafCount];
  return 0;
}

}  // namespace ptg

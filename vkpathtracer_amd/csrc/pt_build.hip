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
#include <cstring>
#include <memory>
#include <string>
#include <vector>

#include "pt_build.h"
#include "pt_scan.h"

namespace ptg {
namespace {

constexpr float kEps = 1e-4f;
constexpr uint32_t kMaxBuckets = 16;
constexpr uint32_t kBigSegment = 2048;    // segments above this use 1024-thread workgroups
constexpr uint32_t kHugeSegment = 65536;  // segments above this are split by many workgroups
constexpr uint32_t kHugeChunk = 4096;     // references per workgroup of a huge segment's passes
constexpr uint64_t kMaxForestRefs = 64ull << 20;  // references per level-synchronous build of several scenes

// Primitive references, structure of arrays: centroid, box, original index.
struct Refs {
  float* c[3];
  float* mn[3];
  float* mx[3];
  uint32_t* orig;
};

// A node still to be split: its references [start, start + count), its node index, depth and scene.
struct Seg {
  uint32_t start, count, node, depth, scene;
};

// Per scene of a build: its leaf size and bucket count, and where its references start in the
// concatenated reference arrays.
struct SceneTab {
  uint32_t leaf, nb, ref_base, pad;
};

// Per scene, written by the build: the host builder's max_depth and leaf_count, and where the scene's
// nodes start (and how many there are) in the renumbered node array.
struct SceneOut {
  uint32_t max_depth, leaf_count, node_base, node_count;
};

struct DevNode {  // ptb::Node
  float mn[3], mx[3];
  int32_t left, right, first, count;
};

enum Counter : int { kNextBig = 0, kNextHuge = 1, kInner = 2, kCounters = 4 };

// Node numbering without shared counters. A level's segments are listed in node order (the segment at
// position i is node level_base + i); an exclusive prefix count of its inner segments (those above
// their scene's leaf size) gives the segment at i its rank r among them, its children the nodes
// next_base + 2r and next_base + 2r + 1, and their places 2r and 2r + 1 in the next level's list.

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

// The host builder's cost scan over the bins, one candidate split (axis, bucket boundary s) per lane of
// one wave. Each lane evaluates its candidate with the scan's own float expressions; the scan keeps the
// first candidate, in axis-major order, whose cost is below both 1e30 and every earlier cost, which is the
// least cost with ties going to the earliest candidate. Costs are non-negative, so their bit patterns
// order like the values and one 64-bit minimum (cost bits, candidate index) finds it. All 64 lanes of the
// wave must call this; bin_n is [3][kMaxBuckets], bin_b [3][kMaxBuckets][6] (ordered uints), in LDS
// or global memory.
__device__ __forceinline__ void scan_splits(const uint32_t* bin_n, const uint32_t* bin_b, uint32_t nb,
                                            const float* bmn, const float* bmx, const float* cmin,
                                            const float* extent, int& found, int& best_axis, float& best_split) {
  const uint32_t lane = threadIdx.x & 63u, per_axis = nb - 1u;
  const float fB = static_cast<float>(nb);
  const float node_sa = smax(1e-9f, bvh_sa(bmn, bmx));
  uint64_t key = ~0ull;
  if (lane < 3u * per_axis) {
    const uint32_t axis = lane / per_axis, s = lane % per_axis + 1u;
    if (!(extent[axis] < 1e-6f)) {
      float lmn[3] = {1e30f, 1e30f, 1e30f}, lmx[3] = {-1e30f, -1e30f, -1e30f};
      float rmn[3] = {1e30f, 1e30f, 1e30f}, rmx[3] = {-1e30f, -1e30f, -1e30f};
      uint32_t lc = 0, rc = 0;
      for (uint32_t b = 0; b < nb; ++b) {
        const uint32_t c = bin_n[axis * kMaxBuckets + b];
        if (!c) continue;
        const uint32_t* q = bin_b + (axis * kMaxBuckets + b) * 6u;
        if (b < s) {
          lc += c;
          for (int k = 0; k < 3; ++k) {
            lmn[k] = smin(lmn[k], u2f(q[k]));
            lmx[k] = smax(lmx[k], u2f(q[3 + k]));
          }
        } else {
          rc += c;
          for (int k = 0; k < 3; ++k) {
            rmn[k] = smin(rmn[k], u2f(q[k]));
            rmx[k] = smax(rmx[k], u2f(q[3 + k]));
          }
        }
      }
      if (lc && rc) {
        const float cost =
            (bvh_sa(lmn, lmx) * static_cast<float>(lc) + bvh_sa(rmn, rmx) * static_cast<float>(rc)) / node_sa;
        if (cost < 1e30f) key = (static_cast<uint64_t>(__float_as_uint(cost)) << 32) | lane;
      }
    }
  }
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) {
    const uint64_t other = __shfl_xor(key, off);
    key = other < key ? other : key;
  }
  found = key != ~0ull ? 1 : 0;
  best_axis = 0;
  best_split = 0.0f;
  if (found) {
    const uint32_t c = static_cast<uint32_t>(key & 0xffffffffu);
    best_axis = static_cast<int>(c / per_axis);
    best_split = cmin[best_axis] + static_cast<float>(c % per_axis + 1u) / fB * extent[best_axis];
  }
}

// Chunk-local exclusive ranks of the inner segments and each chunk's inner count.
__global__ __launch_bounds__(256) void k_seg_rank(const Seg* segs, uint32_t m, const SceneTab* tab,
                                                  uint32_t* rank_local, uint32_t* chunk_sum) {
  chunk_rank(
      m,
      [&](uint32_t i) {
        const Seg sg = segs[i];
        return sg.count > tab[sg.scene].leaf ? 1u : 0u;
      },
      rank_local, chunk_sum);
}

// Where an inner segment's children go (ranks as above).
struct Numbering {
  const uint32_t* rank_local;
  const uint32_t* chunk_off;
  uint32_t level_base, next_base;
  Seg* next_all;  // the next level's segments, in node order
  Seg* next_big;  // and, appended, those of them above kBigSegment (and at most kHugeSegment)
  Seg* next_huge; // those above kHugeSegment
  uint32_t* counters;
};

// Writes the inner node sg (its box in `node`) and its children [start, start + mid) and the rest.
__device__ void emit_inner(const Numbering& N, const Seg& sg, DevNode node, uint32_t mid, DevNode* nodes,
                           uint32_t* node_scene) {
  const uint32_t r = rank_of(N.rank_local, N.chunk_off, sg.node - N.level_base);
  const uint32_t base = N.next_base + 2u * r;
  node.left = static_cast<int32_t>(base);
  node.right = static_cast<int32_t>(base + 1u);
  node.first = -1;
  node.count = 0;
  nodes[sg.node] = node;
  node_scene[sg.node] = sg.scene;
  const Seg kids[2] = {{sg.start, mid, base, sg.depth + 1u, sg.scene},
                       {sg.start + mid, sg.count - mid, base + 1u, sg.depth + 1u, sg.scene}};
  for (uint32_t c = 0; c < 2u; ++c) {
    N.next_all[2u * r + c] = kids[c];
    if (kids[c].count > kHugeSegment) N.next_huge[atomicAdd(N.counters + kNextHuge, 1u)] = kids[c];
    else if (kids[c].count > kBigSegment) N.next_big[atomicAdd(N.counters + kNextBig, 1u)] = kids[c];
  }
}

// One node per workgroup: its bounds, then a leaf, or the bucketed SAH split (binning of all three
// axes into LDS, the host builder's cost scan on one wave) and the stable partition of its
// references from A into B, and its two children. The 64-thread form runs over a level's whole list
// and leaves the segments above kBigSegment to the 1024-thread form (run over the big list) and the
// huge-segment passes.
template <uint32_t NT>
__global__ __launch_bounds__(NT) void k_split(const Seg* segs, Refs A, Refs B, DevNode* nodes, uint32_t* node_scene,
                                               uint32_t* order, Numbering N, const SceneTab* tab) {
  __shared__ float red[12][NT / 64u];
  __shared__ uint32_t ured[NT / 64u];
  __shared__ uint32_t bin_n[3][kMaxBuckets];
  __shared__ uint32_t bin_b[3][kMaxBuckets][6];  // mn xyz, mx xyz as ordered uints
  __shared__ int s_axis, s_found;
  __shared__ float s_split;
  const Seg sg = segs[blockIdx.x];
  if (NT == 64u && sg.count > kBigSegment) return;
  const SceneTab P = tab[sg.scene];
  const uint32_t leaf = P.leaf, nb = P.nb;
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
  if (sg.count <= leaf) {  // a leaf: its references in order are its slice of the scene's leaf order
    for (uint32_t i = tid; i < sg.count; i += NT) order[sg.start + i] = A.orig[sg.start + i] - P.ref_base;
    if (tid == 0u) {
      node.left = node.right = -1;
      node.first = static_cast<int32_t>(sg.start - P.ref_base);
      node.count = static_cast<int32_t>(sg.count);
      nodes[sg.node] = node;
      node_scene[sg.node] = sg.scene;
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
  if (tid < 64u) {  // the host builder's cost scan, on the first wave
    int found, best_axis;
    float best_split;
    scan_splits(&bin_n[0][0], &bin_b[0][0][0], nb, bmn, bmx, cmin, extent, found, best_axis, best_split);
    if (tid == 0u) {
      s_axis = best_axis;
      s_split = best_split;
      s_found = found;
    }
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
    emit_inner(N, sg, node, mid, nodes, node_scene);
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

__global__ __launch_bounds__(256) void k_huge_bins(const Seg* huge, Refs A, HugeAcc* accs, const SceneTab* tab) {
  __shared__ uint32_t bin_n[3][kMaxBuckets];
  __shared__ uint32_t bin_b[3][kMaxBuckets][6];
  Seg sg;
  uint32_t c0, c1;
  if (!huge_chunk(huge, sg, c0, c1)) return;
  HugeAcc* acc = accs + blockIdx.y;
  const uint32_t nb = tab[sg.scene].nb;
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

// k_split's cost scan over the combined bins, one wave per segment.
__global__ __launch_bounds__(64) void k_huge_decide(const Seg* huge, HugeAcc* accs, const SceneTab* tab) {
  HugeAcc* acc = accs + blockIdx.x;
  const uint32_t nb = tab[huge[blockIdx.x].scene].nb;
  float bmn[3], bmx[3], cmin[3], extent[3];
  for (int k = 0; k < 3; ++k) {
    bmn[k] = u2f(acc->bounds[k]);
    bmx[k] = u2f(acc->bounds[3 + k]);
    cmin[k] = u2f(acc->bounds[6 + k]);
    extent[k] = u2f(acc->bounds[9 + k]) - cmin[k];
  }
  int found, best_axis;
  float best_split;
  scan_splits(&acc->bin_n[0][0], &acc->bin_b[0][0][0], nb, bmn, bmx, cmin, extent, found, best_axis, best_split);
  if (threadIdx.x == 0u) {
    acc->found = found;
    acc->axis = best_axis;
    acc->split = best_split;
  }
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
                              DevNode* nodes, uint32_t* node_scene, Numbering N) {
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
  emit_inner(N, sg, node, mid, nodes, node_scene);
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

// Renumbering into the host builder's pre-order, per scene. The nodes created while level L is split
// are level L + 1, a contiguous range of indices, so each pass is one launch per level: subtree sizes
// from the deepest level up, pre-order positions from the roots down (scene s's root is node s), then
// the scatter of each scene's nodes to its own range of the output.
// Subtree sizes and heights (a root's height is its scene's max_depth: the host builder's deepest node).
__global__ void k_subtree_sizes(const DevNode* nodes, uint32_t* size, uint32_t* height, uint32_t lo, uint32_t hi) {
  const uint32_t b = lo + blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= hi) return;
  const DevNode& d = nodes[b];
  size[b] = d.count > 0 ? 1u : 1u + size[d.left] + size[d.right];
  height[b] = d.count > 0 ? 0u : 1u + max(height[d.left], height[d.right]);
}

__global__ void k_preorder(const DevNode* nodes, const uint32_t* size, uint32_t* pre, uint32_t lo, uint32_t hi) {
  const uint32_t b = lo + blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= hi) return;
  const DevNode& d = nodes[b];
  if (d.count > 0) return;
  pre[d.left] = pre[b] + 1u;
  pre[d.right] = pre[b] + 1u + size[d.left];
}

// The scenes' node ranges of the output, depths and leaf counts (one thread; scene s's root is node s,
// and a tree whose inner nodes all have two children has (nodes + 1) / 2 leaves).
__global__ void k_scene_bases(const uint32_t* size, const uint32_t* height, SceneOut* sout, uint32_t scenes) {
  if (blockIdx.x != 0u || threadIdx.x != 0u) return;
  uint32_t base = 0;
  for (uint32_t s = 0; s < scenes; ++s) {
    sout[s].node_base = base;
    sout[s].node_count = size[s];
    sout[s].max_depth = height[s];
    sout[s].leaf_count = (size[s] + 1u) / 2u;
    base += size[s];
  }
}

__global__ void k_scatter_preorder(const DevNode* nodes, const uint32_t* node_scene, const uint32_t* pre,
                                   const SceneOut* sout, DevNode* out, uint32_t n) {
  const uint32_t b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= n) return;
  DevNode d = nodes[b];
  if (d.count == 0) {
    d.left = static_cast<int32_t>(pre[d.left]);
    d.right = static_cast<int32_t>(pre[d.right]);
  }
  out[sout[node_scene[b]].node_base + pre[b]] = d;
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
  // hands the buffer to a shared owner that releases it on the same stream
  std::shared_ptr<void> share() {
    T* q = p;
    p = nullptr;
    const hipStream_t st = s;
    return std::shared_ptr<void>(q, [st](void* x) { (void)hipFreeAsync(x, st); });
  }
};

// Each scene's root record (its node 0 in the renumbered order).
__global__ void k_gather_roots(const DevNode* out, const SceneOut* sout, DevNode* roots, uint32_t scenes) {
  const uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s < scenes) roots[s] = out[sout[s].node_base];
}

// One level-synchronous build of several scenes (a forest): their references are concatenated and
// every scene's root is a segment of the first level, so the scenes share each level's launches and
// its one synchronisation.
int build_forest(hipStream_t stream, std::vector<DeviceBuild>& jobs, const std::vector<uint32_t>& sel,
                 std::string& err, std::vector<DeviceForest>* keep) {
  const uint32_t S = static_cast<uint32_t>(sel.size());
  std::vector<SceneTab> tab(S);
  uint32_t n = 0;
  for (uint32_t s = 0; s < S; ++s) {
    const DeviceBuild& d = jobs[sel[s]];
    tab[s].leaf = d.leaf_size ? d.leaf_size : 4u;
    tab[s].nb = std::clamp(d.buckets ? d.buckets : 8u, 2u, kMaxBuckets);
    tab[s].ref_base = n;
    tab[s].pad = 0u;
    n += static_cast<uint32_t>(d.bvh->prims.size());
  }
  DevBuf<float> d_raw, d_ref[2];
  DevBuf<uint32_t> d_orig[2], d_order, d_counters, d_node_scene, d_rank, d_chunk;
  DevBuf<Seg> d_all[2];   // a level's segments in node order: this level's and the next
  DevBuf<Seg> d_big[2];   // of those, the ones above kBigSegment (and at most kHugeSegment)
  DevBuf<Seg> d_huge[2];  // and the ones above kHugeSegment
  DevBuf<HugeAcc> d_acc;
  DevBuf<uint32_t> d_chunk_left;
  DevBuf<DevNode> d_nodes, d_pre_nodes;
  DevBuf<SceneTab> d_tab;
  DevBuf<SceneOut> d_sout;
  DevBuf<uint32_t> d_mat;   // kept trees: the primitives' instance materials
  DevBuf<DevNode> d_roots;  // kept trees: each scene's root record
  auto check = [&](hipError_t e, const char* what) {
    if (e != hipSuccess) err = std::string("device BVH build: ") + what + ": " + hipGetErrorString(e);
    return e == hipSuccess;
  };
  if (keep)
    for (uint32_t s = 0; s < S; ++s)
      if (!jobs[sel[s]].material || jobs[sel[s]].material->size() != jobs[sel[s]].bvh->prims.size()) {
        err = "device BVH build: a kept tree needs its primitives' materials";
        return -1;
      }
  // segments of one level are disjoint: at most n of them, n / kBigSegment above kBigSegment and
  // n / kHugeSegment above kHugeSegment
  const uint32_t max_big = n / kBigSegment + 1u, max_huge = n / kHugeSegment + 1u;
  const uint32_t max_chunks = (n + kHugeChunk - 1u) / kHugeChunk, max_scan = (n + kScanChunk - 1u) / kScanChunk;
  const size_t n9 = 9u * static_cast<size_t>(n), n2 = 2u * static_cast<size_t>(n);
  if (!check(d_raw.alloc(n9, stream), "allocate") || !check(d_ref[0].alloc(n9, stream), "allocate") ||
      !check(d_ref[1].alloc(n9, stream), "allocate") || !check(d_orig[0].alloc(n, stream), "allocate") ||
      !check(d_orig[1].alloc(n, stream), "allocate") || !check(d_order.alloc(n, stream), "allocate") ||
      !check(d_counters.alloc(kCounters, stream), "allocate") || !check(d_nodes.alloc(n2, stream), "allocate") ||
      !check(d_node_scene.alloc(n2, stream), "allocate") || !check(d_pre_nodes.alloc(n2, stream), "allocate") ||
      !check(d_rank.alloc(n, stream), "allocate") || !check(d_chunk.alloc(max_scan, stream), "allocate") ||
      !check(d_acc.alloc(max_huge, stream), "allocate") ||
      !check(d_chunk_left.alloc(static_cast<size_t>(max_huge) * max_chunks, stream), "allocate") ||
      !check(d_tab.alloc(S, stream), "allocate") || !check(d_sout.alloc(S, stream), "allocate"))
    return -1;
  for (int k = 0; k < 2; ++k)
    if (!check(d_all[k].alloc(n, stream), "allocate") || !check(d_big[k].alloc(max_big, stream), "allocate") ||
        !check(d_huge[k].alloc(max_huge, stream), "allocate"))
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
  if (keep && (!check(d_mat.alloc(n, stream), "allocate") || !check(d_roots.alloc(S, stream), "allocate")))
    return -1;
  for (uint32_t s = 0; s < S; ++s) {
    const ptb::HostVec<float>& raw = *jobs[sel[s]].raw;
    if (!check(hipMemcpyAsync(d_raw.p + 9u * static_cast<size_t>(tab[s].ref_base), raw.data(),
                              raw.size() * sizeof(float), hipMemcpyHostToDevice, stream), "upload"))
      return -1;
    if (keep) {
      const ptb::HostVec<uint32_t>& mat = *jobs[sel[s]].material;
      if (!check(hipMemcpyAsync(d_mat.p + tab[s].ref_base, mat.data(), mat.size() * sizeof(uint32_t),
                                hipMemcpyHostToDevice, stream), "upload"))
        return -1;
    }
  }
  hipLaunchKernelGGL(k_build_refs, dim3((n + 255u) / 256u), dim3(256), 0, stream, d_raw.p, n, R[0]);
  // level 0: the roots (scene s is node s)
  std::vector<Seg> roots(S), big0, huge;
  for (uint32_t s = 0; s < S; ++s) {
    const uint32_t c = static_cast<uint32_t>(jobs[sel[s]].bvh->prims.size());
    roots[s] = Seg{tab[s].ref_base, c, s, 0u, s};
    if (c > kHugeSegment) huge.push_back(roots[s]);
    else if (c > kBigSegment) big0.push_back(roots[s]);
  }
  uint32_t nhuge = static_cast<uint32_t>(huge.size()), nbig = static_cast<uint32_t>(big0.size());
  huge.resize(max_huge);
  if (!check(hipMemcpyAsync(d_all[0].p, roots.data(), S * sizeof(Seg), hipMemcpyHostToDevice, stream), "upload") ||
      (nhuge && !check(hipMemcpyAsync(d_huge[0].p, huge.data(), nhuge * sizeof(Seg), hipMemcpyHostToDevice, stream),
                       "upload")) ||
      (nbig && !check(hipMemcpyAsync(d_big[0].p, big0.data(), nbig * sizeof(Seg), hipMemcpyHostToDevice, stream),
                      "upload")) ||
      !check(hipMemcpyAsync(d_tab.p, tab.data(), S * sizeof(SceneTab), hipMemcpyHostToDevice, stream), "upload"))
    return -1;
  int cur = 0;  // the level's lists are d_all/d_big/d_huge[cur]; its references are in R[level % 2]
  std::vector<uint32_t> level_start{0u};  // nodes of level L: [level_start[L], level_start[L + 1])
  uint32_t level_base = 0, m = S;
  uint32_t counters[kCounters] = {};
  for (uint32_t level = 0; m > 0u; ++level) {
    const uint32_t zero2[2] = {0u, 0u};  // kNextBig, kNextHuge
    if (!check(hipMemcpyAsync(d_counters.p, zero2, sizeof(zero2), hipMemcpyHostToDevice, stream), "reset"))
      return -1;
    const Refs& A = R[level % 2u];
    const Refs& B = R[(level + 1u) % 2u];
    const uint32_t scan_chunks = (m + kScanChunk - 1u) / kScanChunk;
    hipLaunchKernelGGL(k_seg_rank, dim3(scan_chunks), dim3(256), 0, stream, d_all[cur].p, m, d_tab.p, d_rank.p,
                       d_chunk.p);
    hipLaunchKernelGGL(k_chunk_offsets, dim3(1), dim3(1024), 0, stream, d_chunk.p, scan_chunks,
                       d_counters.p + kInner);
    const Numbering N{d_rank.p, d_chunk.p, level_base, level_base + m, d_all[1 - cur].p, d_big[1 - cur].p,
                      d_huge[1 - cur].p, d_counters.p};
    if (nhuge) {
      uint32_t grid_x = 0;
      for (uint32_t h = 0; h < nhuge; ++h) grid_x = std::max(grid_x, (huge[h].count + kHugeChunk - 1u) / kHugeChunk);
      const Seg* hl = d_huge[cur].p;
      const dim3 grid(grid_x, nhuge);
      hipLaunchKernelGGL(k_huge_init, dim3(nhuge), dim3(256), 0, stream, d_acc.p);
      hipLaunchKernelGGL(k_huge_bounds, grid, dim3(256), 0, stream, hl, A, d_acc.p);
      hipLaunchKernelGGL(k_huge_bins, grid, dim3(256), 0, stream, hl, A, d_acc.p, d_tab.p);
      hipLaunchKernelGGL(k_huge_decide, dim3(nhuge), dim3(64), 0, stream, hl, d_acc.p, d_tab.p);
      hipLaunchKernelGGL(k_huge_count, grid, dim3(256), 0, stream, hl, A, d_acc.p, d_chunk_left.p, max_chunks);
      hipLaunchKernelGGL(k_huge_finish, dim3(nhuge), dim3(64), 0, stream, hl, d_acc.p, d_chunk_left.p, max_chunks,
                         d_nodes.p, d_node_scene.p, N);
      hipLaunchKernelGGL(k_huge_scatter, grid, dim3(256), 0, stream, hl, A, B, d_acc.p, d_chunk_left.p, max_chunks);
    }
    if (nbig)
      hipLaunchKernelGGL(k_split<1024>, dim3(nbig), dim3(1024), 0, stream, d_big[cur].p, A, B, d_nodes.p,
                         d_node_scene.p, d_order.p, N, d_tab.p);
    hipLaunchKernelGGL(k_split<64>, dim3(m), dim3(64), 0, stream, d_all[cur].p, A, B, d_nodes.p, d_node_scene.p,
                       d_order.p, N, d_tab.p);
    // the next huge list travels with the counters (it is short: at most n / kHugeSegment entries)
    if (!check(hipGetLastError(), "launch") ||
        !check(hipMemcpyAsync(counters, d_counters.p, sizeof(counters), hipMemcpyDeviceToHost, stream), "read") ||
        (nhuge && !check(hipMemcpyAsync(huge.data(), d_huge[1 - cur].p, max_huge * sizeof(Seg), hipMemcpyDeviceToHost,
                                        stream), "read")) ||
        !check(hipStreamSynchronize(stream), "synchronize"))
      return -1;
    nhuge = counters[kNextHuge];
    nbig = counters[kNextBig];
    level_base += m;
    level_start.push_back(level_base);
    m = 2u * counters[kInner];
    cur = 1 - cur;
  }
  const uint32_t node_count = level_base;
  // pre-order renumbering on the device (the arrays of the finished lists are reused: each holds 5n
  // words, and node_count < 2n)
  static_assert(sizeof(Seg) == 20 && sizeof(DevNode) == 40 && sizeof(ptb::Node) == 40, "record sizes");
  uint32_t* size = reinterpret_cast<uint32_t*>(d_all[0].p);
  uint32_t* height = size + n2;
  uint32_t* pre = reinterpret_cast<uint32_t*>(d_all[1].p);
  DevNode* out = d_pre_nodes.p;
  const uint32_t levels = static_cast<uint32_t>(level_start.size()) - 1u;
  for (uint32_t L = levels; L-- > 0;) {
    const uint32_t lo = level_start[L], hi = level_start[L + 1];
    hipLaunchKernelGGL(k_subtree_sizes, dim3((hi - lo + 255u) / 256u), dim3(256), 0, stream, d_nodes.p, size, height,
                       lo, hi);
  }
  if (!check(hipMemsetAsync(pre, 0, S * sizeof(uint32_t), stream), "clear")) return -1;  // every root is 0
  for (uint32_t L = 0; L < levels; ++L) {
    const uint32_t lo = level_start[L], hi = level_start[L + 1];
    hipLaunchKernelGGL(k_preorder, dim3((hi - lo + 255u) / 256u), dim3(256), 0, stream, d_nodes.p, size, pre, lo, hi);
  }
  hipLaunchKernelGGL(k_scene_bases, dim3(1), dim3(64), 0, stream, size, height, d_sout.p, S);
  hipLaunchKernelGGL(k_scatter_preorder, dim3((node_count + 255u) / 256u), dim3(256), 0, stream, d_nodes.p,
                     d_node_scene.p, pre, d_sout.p, out, node_count);
  std::vector<SceneOut> sout(S);
  std::vector<DevNode> root_nodes(keep ? S : 0u);
  if (keep) hipLaunchKernelGGL(k_gather_roots, dim3((S + 63u) / 64u), dim3(64), 0, stream, out, d_sout.p, d_roots.p, S);
  if (!check(hipGetLastError(), "launch") ||
      !check(hipMemcpyAsync(sout.data(), d_sout.p, S * sizeof(SceneOut), hipMemcpyDeviceToHost, stream), "read") ||
      (keep && !check(hipMemcpyAsync(root_nodes.data(), d_roots.p, S * sizeof(DevNode), hipMemcpyDeviceToHost, stream),
                      "read")) ||
      !check(hipStreamSynchronize(stream), "synchronize"))
    return -1;
  if (keep) {  // the trees stay where they are
    DeviceForest f;
    f.nodes = reinterpret_cast<const ptb::Node*>(d_pre_nodes.p);
    f.order = d_order.p;
    f.raw = d_raw.p;
    f.material = d_mat.p;
    f.node_total = node_count;
    f.ref_total = n;
    f.storage = {d_pre_nodes.share(), d_order.share(), d_raw.share(), d_mat.share()};
    for (uint32_t s = 0; s < S; ++s) {
      DeviceBuild& d = jobs[sel[s]];
      ForestScene fs{sout[s].node_base, sout[s].node_count, tab[s].ref_base,
                     static_cast<uint32_t>(d.bvh->prims.size()), ptb::Node{}};
      std::memcpy(&fs.root, &root_nodes[s], sizeof(ptb::Node));
      f.scenes.push_back(fs);
      d.forest = static_cast<int>(keep->size());
      d.scene = s;
      d.bvh->max_depth = sout[s].max_depth;
      d.bvh->leaf_count = sout[s].leaf_count;
    }
    keep->push_back(std::move(f));
    return 0;
  }
  for (uint32_t s = 0; s < S; ++s) {
    ptb::Bvh& bvh = *jobs[sel[s]].bvh;
    const size_t prims = bvh.prims.size();
    bvh.nodes.resize(sout[s].node_count);
    bvh.order.resize(prims);
    if (!check(hipMemcpyAsync(bvh.nodes.data(), out + sout[s].node_base, sout[s].node_count * sizeof(DevNode),
                              hipMemcpyDeviceToHost, stream), "download") ||
        !check(hipMemcpyAsync(bvh.order.data(), d_order.p + tab[s].ref_base, prims * sizeof(uint32_t),
                              hipMemcpyDeviceToHost, stream), "download"))
      return -1;
    bvh.max_depth = sout[s].max_depth;
    bvh.leaf_count = sout[s].leaf_count;
  }
  return check(hipStreamSynchronize(stream), "synchronize") ? 0 : -1;
}

}  // namespace

int build_sah_device_many(hipStream_t stream, std::vector<DeviceBuild>& jobs, std::string& err,
                          std::vector<DeviceForest>* keep) {
  std::vector<uint32_t> sel;
  uint64_t refs = 0;
  for (uint32_t j = 0; j < jobs.size(); ++j) {
    DeviceBuild& d = jobs[j];
    ptb::Bvh& bvh = *d.bvh;
    bvh.nodes.clear();
    bvh.order.clear();
    bvh.max_depth = 0;
    bvh.leaf_count = 0;
    d.rc = 0;
    if (bvh.prims.empty()) continue;
    // NaN boxes make min/max order-dependent, and coordinates beyond the builder's 1e30 sentinels can
    // overflow a centroid: the host builds such scenes
    bool finite = d.bounded == 1;
    if (d.bounded < 0) {
      finite = true;
      for (float x : *d.raw)
        if (!(std::fabs(x) <= 1e30f)) {
          finite = false;
          break;
        }
    }
    if (!finite) {
      d.rc = 1;
      continue;
    }
    // one forest holds at most kMaxForestRefs references (device scratch of about 370 bytes each)
    if (!sel.empty() && refs + bvh.prims.size() > kMaxForestRefs) {
      if (build_forest(stream, jobs, sel, err, keep) < 0) return -1;
      sel.clear();
      refs = 0;
    }
    sel.push_back(j);
    refs += bvh.prims.size();
  }
  if (!sel.empty() && build_forest(stream, jobs, sel, err, keep) < 0) return -1;
  return 0;
}

int build_sah_device(hipStream_t stream, ptb::Bvh& bvh, const ptb::HostVec<float>& raw, uint32_t leaf_size,
                     uint32_t buckets, std::string& err) {
  std::vector<DeviceBuild> jobs{DeviceBuild{&bvh, &raw, leaf_size, buckets, 0}};
  return build_sah_device_many(stream, jobs, err) < 0 ? -1 : jobs[0].rc;
}

}  // namespace ptg

// pt_pool.hip — the device scene pool packed on the device from trees the GPU build kept there.
//
// The host path (ptb::pack_device_layout, then the pool assembly in pt_capi.cpp) downloads a tree,
// reformats it and uploads it again; here the same records are written in place by three launches
// over a forest (nodes, triangles, primitives), with the same float operations so that every word
// equals the host layout's: e1 = v1 - v0, e2 = v2 - v0, the flattening's normalize(cross(e1, e2)) with
// correctly rounded square root and division, and the reference's leaf and child encodings.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>
#include <string>
#include <vector>

#include "pt_build.h"
#include "pt_scan.h"

namespace ptg {
namespace {

constexpr uint32_t kLeafBits = static_cast<uint32_t>(ptb::kLeafCountBits);

// One forest scene and its place in the pool.
struct PackScene {
  uint32_t f_node_base, ref_base, prims;
  uint32_t node_base, tri_base, prim_base, mat_base, last_mat, with_uv;
};

// The scene holding index i: the last scene whose range starts at or before i (ranges ascend).
template <typename Start>
__device__ __forceinline__ uint32_t scene_of(uint32_t i, uint32_t scenes, Start start) {
  uint32_t lo = 0, hi = scenes;  // start(lo) <= i < start(hi)
  while (hi - lo > 1u) {
    const uint32_t mid = (lo + hi) / 2u;
    if (start(mid) <= i) lo = mid;
    else hi = mid;
  }
  return lo;
}

__global__ __launch_bounds__(256) void k_inner_rank(const ptb::Node* nodes, uint32_t m, uint32_t* rank_local,
                                                    uint32_t* chunk_sum) {
  chunk_rank(m, [&](uint32_t i) { return nodes[i].count > 0 ? 0u : 1u; }, rank_local, chunk_sum);
}

// ptb::pack_device_layout's node record (both children's boxes, then the two references), rebased.
__global__ __launch_bounds__(256) void k_pack_nodes(const ptb::Node* nodes, uint32_t total, const uint32_t* rank_local,
                                                    const uint32_t* chunk_off, const PackScene* tab, uint32_t scenes,
                                                    float4* out) {
  const uint32_t b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= total) return;
  const ptb::Node nd = nodes[b];
  if (nd.count > 0) return;
  const PackScene P = tab[scene_of(b, scenes, [&](uint32_t s) { return tab[s].f_node_base; })];
  const uint32_t rank0 = rank_of(rank_local, chunk_off, P.f_node_base);
  const uint32_t kids[2] = {P.f_node_base + static_cast<uint32_t>(nd.left),
                            P.f_node_base + static_cast<uint32_t>(nd.right)};
  ptb::Node c[2];
  int32_t ref[2];
  for (int k = 0; k < 2; ++k) {
    c[k] = nodes[kids[k]];
    if (c[k].count > 0) {
      const uint32_t code = ((static_cast<uint32_t>(c[k].first) + P.tri_base) << kLeafBits) |
                            static_cast<uint32_t>(c[k].count);
      ref[k] = ~static_cast<int32_t>(code);
    } else {
      ref[k] = static_cast<int32_t>(P.node_base + rank_of(rank_local, chunk_off, kids[k]) - rank0);
    }
  }
  float4* g = out + 4u * (P.node_base + rank_of(rank_local, chunk_off, b) - rank0);
  g[0] = make_float4(c[0].mn[0], c[0].mn[1], c[0].mn[2], c[0].mx[0]);
  g[1] = make_float4(c[0].mx[1], c[0].mx[2], c[1].mn[0], c[1].mn[1]);
  g[2] = make_float4(c[1].mn[2], c[1].mx[0], c[1].mx[1], c[1].mx[2]);
  g[3] = make_float4(__int_as_float(ref[0]), __int_as_float(ref[1]), 0.0f, 0.0f);
}

// Triangle records in leaf order ({v0, pool primitive id}, {e1, 0}, {e2, 0}) and, when asked, each
// primitive's record index.
__global__ __launch_bounds__(256) void k_pack_tris(const uint32_t* order, const float* raw, uint32_t total,
                                                   const PackScene* tab, uint32_t scenes, float4* tris,
                                                   uint32_t* prim_tri) {
  const uint32_t j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= total) return;
  const PackScene P = tab[scene_of(j, scenes, [&](uint32_t s) { return tab[s].ref_base; })];
  const uint32_t jl = j - P.ref_base, pid = order[j];
  const float* v = raw + 9u * (static_cast<size_t>(P.ref_base) + pid);
  const float a0 = v[0], a1 = v[1], a2 = v[2];
  float4* t = tris + 3u * (static_cast<size_t>(P.tri_base) + jl);
  t[0] = make_float4(a0, a1, a2, __uint_as_float(pid + P.prim_base));
  t[1] = make_float4(v[3] - a0, v[4] - a1, v[5] - a2, 0.0f);
  t[2] = make_float4(v[6] - a0, v[7] - a1, v[8] - a2, 0.0f);
  if (prim_tri && P.with_uv) prim_tri[P.prim_base + pid] = P.tri_base + jl;
}

// Primitive records {face normal, pool material}: the flattening's normalize_vec3(cross(e1, e2))
// (CpuBvhAccelerator.cpp:76-82) and the scene's last-material clamp.
__global__ __launch_bounds__(256) void k_pack_prims(const float* raw, const uint32_t* material, uint32_t total,
                                                    const PackScene* tab, uint32_t scenes, float4* prims) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= total) return;
  const PackScene P = tab[scene_of(i, scenes, [&](uint32_t s) { return tab[s].ref_base; })];
  const float* v = raw + 9u * static_cast<size_t>(i);
  const float e1x = v[3] - v[0], e1y = v[4] - v[1], e1z = v[5] - v[2];
  const float e2x = v[6] - v[0], e2y = v[7] - v[1], e2z = v[8] - v[2];
  float nx = e1y * e2z - e1z * e2y, ny = e1z * e2x - e1x * e2z, nz = e1x * e2y - e1y * e2x;
  const float len_sq = nx * nx + ny * ny + nz * nz;
  constexpr float kEps = 1e-4f;
  if (len_sq <= kEps * kEps) {
    nx = 0.0f;
    ny = 1.0f;
    nz = 0.0f;
  } else {
    const float inv = 1.0f / sqrtf(len_sq);
    nx *= inv;
    ny *= inv;
    nz *= inv;
  }
  const uint32_t m = min(material[i], P.last_mat) + P.mat_base;
  prims[P.prim_base + (i - P.ref_base)] = make_float4(nx, ny, nz, __uint_as_float(m));
}

// The hot-copy selection (one thread): per tree a frontier ordered by (box area, node index), largest
// first, from the root down; non-finite areas rank last.
__device__ __forceinline__ float box_area(float x0, float y0, float z0, float x1, float y1, float z1) {
  const float ex = x1 - x0, ey = y1 - y0, ez = z1 - z0;
  const float a = ex * ey + ey * ez + ez * ex;
  return isfinite(a) ? a : -1.0f;
}

constexpr uint32_t kFrontier = 2u * 256u + 2u;  // the frontier grows by at most one per copy taken

__global__ void k_hot_select(const float4* nodes, const int32_t* roots, const float* root_boxes, uint32_t trees,
                             uint32_t per_tree, uint32_t max_total, uint32_t* hot_src, uint32_t* count) {
  __shared__ float fa[kFrontier];
  __shared__ int32_t fr[kFrontier];
  if (threadIdx.x != 0u) return;
  uint32_t total = 0;
  for (uint32_t t = 0; t < trees; ++t) {
    const uint32_t quota = min(per_tree, max_total - total);
    if (quota == 0u) break;
    const float* rb = root_boxes + 6u * t;
    uint32_t fn = 1, taken = 0;
    fa[0] = box_area(rb[0], rb[1], rb[2], rb[3], rb[4], rb[5]);
    fr[0] = roots[t];
    while (fn > 0u && taken < quota) {
      uint32_t best = 0;
      for (uint32_t k = 1; k < fn; ++k)
        if (fa[k] > fa[best] || (fa[k] == fa[best] && fr[k] > fr[best])) best = k;
      const int32_t ni = fr[best];
      --fn;
      fa[best] = fa[fn];
      fr[best] = fr[fn];
      hot_src[total++] = static_cast<uint32_t>(ni);
      ++taken;
      const float4* f = nodes + 4u * static_cast<uint32_t>(ni);
      const float4 q0 = f[0], q1 = f[1], q2 = f[2], q3 = f[3];
      const int32_t r0 = __float_as_int(q3.x), r1 = __float_as_int(q3.y);
      if (r0 >= 0 && fn < kFrontier) {
        fa[fn] = box_area(q0.x, q0.y, q0.z, q0.w, q1.x, q1.y);
        fr[fn++] = r0;
      }
      if (r1 >= 0 && fn < kFrontier) {
        fa[fn] = box_area(q1.z, q1.w, q2.x, q2.y, q2.z, q2.w);
        fr[fn++] = r1;
      }
    }
  }
  *count = total;
}

// Copy h of hot_src[h] goes to hot_base + h; children that are hot point at their copies.
__global__ void k_hot_copy(float4* nodes, const uint32_t* hot_src, uint32_t hot_base, uint32_t count) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= 4u * count) return;
  const uint32_t h = i / 4u, k = i % 4u;
  float4 q = nodes[4u * hot_src[h] + k];
  if (k == 3u) {
    int32_t r[2] = {__float_as_int(q.x), __float_as_int(q.y)};
    for (int c = 0; c < 2; ++c)
      if (r[c] >= 0)
        for (uint32_t j = 0; j < count; ++j)
          if (hot_src[j] == static_cast<uint32_t>(r[c])) {
            r[c] = static_cast<int32_t>(hot_base + j);
            break;
          }
    q.x = __int_as_float(r[0]);
    q.y = __int_as_float(r[1]);
  }
  nodes[4u * (hot_base + h) + k] = q;
}

__global__ void k_leaf_positions(const float4* tris, uint32_t count, uint32_t* leaf_pos) {
  const uint32_t j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j < count) leaf_pos[__float_as_uint(tris[3u * j].w)] = j;
}

template <typename T>
struct Scratch {  // stream-ordered scratch of one call
  T* p = nullptr;
  hipStream_t s = nullptr;
  ~Scratch() {
    if (p) (void)hipFreeAsync(p, s);
  }
  hipError_t alloc(size_t n, hipStream_t stream) {
    s = stream;
    return hipMallocAsync(reinterpret_cast<void**>(&p), std::max<size_t>(n, 1) * sizeof(T), stream);
  }
};

}  // namespace

int pack_pool(hipStream_t stream, const DeviceForest& f, const std::vector<PoolDest>& dest, const PoolArrays& out,
              std::string& err) {
  auto check = [&](hipError_t e, const char* what) {
    if (e != hipSuccess) err = std::string("device pool packing: ") + what + ": " + hipGetErrorString(e);
    return e == hipSuccess;
  };
  const uint32_t S = static_cast<uint32_t>(f.scenes.size());
  if (dest.size() != S) {
    err = "device pool packing: one destination per forest scene";
    return -1;
  }
  if (S == 0u) return 0;
  std::vector<PackScene> tab(S);
  for (uint32_t s = 0; s < S; ++s) {
    const ForestScene& fs = f.scenes[s];
    const PoolDest& d = dest[s];
    tab[s] = PackScene{fs.node_base, fs.ref_base, fs.prims, d.node_base, d.tri_base, d.prim_base, d.mat_base,
                       d.last_mat, d.with_uv};
  }
  const uint32_t N = f.node_total, chunks = (N + kScanChunk - 1u) / kScanChunk;
  Scratch<PackScene> d_tab;
  Scratch<uint32_t> d_rank, d_chunk;
  if (!check(d_tab.alloc(S, stream), "allocate") || !check(d_rank.alloc(N, stream), "allocate") ||
      !check(d_chunk.alloc(chunks + 1u, stream), "allocate") ||
      !check(hipMemcpyAsync(d_tab.p, tab.data(), S * sizeof(PackScene), hipMemcpyHostToDevice, stream), "upload"))
    return -1;
  if (N) {
    hipLaunchKernelGGL(k_inner_rank, dim3(chunks), dim3(256), 0, stream, f.nodes, N, d_rank.p, d_chunk.p);
    hipLaunchKernelGGL(k_chunk_offsets, dim3(1), dim3(1024), 0, stream, d_chunk.p, chunks, d_chunk.p + chunks);
    hipLaunchKernelGGL(k_pack_nodes, dim3((N + 255u) / 256u), dim3(256), 0, stream, f.nodes, N, d_rank.p, d_chunk.p,
                       d_tab.p, S, out.nodes);
  }
  const uint32_t R = f.ref_total;
  if (R) {
    hipLaunchKernelGGL(k_pack_tris, dim3((R + 255u) / 256u), dim3(256), 0, stream, f.order, f.raw, R, d_tab.p, S,
                       out.tris, out.prim_tri);
    hipLaunchKernelGGL(k_pack_prims, dim3((R + 255u) / 256u), dim3(256), 0, stream, f.raw, f.material, R, d_tab.p,
                       S, out.prims);
  }
  return check(hipGetLastError(), "launch") ? 0 : -1;
}

int select_hot_nodes(hipStream_t stream, float4* nodes, uint32_t inner_total, const std::vector<int32_t>& roots,
                     const std::vector<float>& root_boxes, uint32_t per_tree, uint32_t max_total, uint32_t* hot_src,
                     std::vector<uint32_t>& hot_host, std::string& err) {
  auto check = [&](hipError_t e, const char* what) {
    if (e != hipSuccess) err = std::string("hot node selection: ") + what + ": " + hipGetErrorString(e);
    return e == hipSuccess;
  };
  hot_host.clear();
  const uint32_t trees = static_cast<uint32_t>(roots.size());
  if (trees == 0u || max_total == 0u) return 0;
  if (max_total > 256u || root_boxes.size() != 6u * roots.size()) {
    err = "hot node selection: at most 256 copies, six box floats per tree";
    return -1;
  }
  Scratch<int32_t> d_roots;
  Scratch<float> d_boxes;
  Scratch<uint32_t> d_count;
  uint32_t count = 0;
  if (!check(d_roots.alloc(trees, stream), "allocate") || !check(d_boxes.alloc(6u * trees, stream), "allocate") ||
      !check(d_count.alloc(1, stream), "allocate") ||
      !check(hipMemcpyAsync(d_roots.p, roots.data(), trees * sizeof(int32_t), hipMemcpyHostToDevice, stream),
             "upload") ||
      !check(hipMemcpyAsync(d_boxes.p, root_boxes.data(), root_boxes.size() * sizeof(float), hipMemcpyHostToDevice,
                            stream), "upload"))
    return -1;
  hipLaunchKernelGGL(k_hot_select, dim3(1), dim3(64), 0, stream, nodes, d_roots.p, d_boxes.p, trees, per_tree,
                     max_total, hot_src, d_count.p);
  if (!check(hipGetLastError(), "launch") ||
      !check(hipMemcpyAsync(&count, d_count.p, sizeof(count), hipMemcpyDeviceToHost, stream), "read") ||
      !check(hipStreamSynchronize(stream), "synchronize"))
    return -1;
  hot_host.resize(count);
  if (count) {
    if (!check(hipMemcpyAsync(hot_host.data(), hot_src, count * sizeof(uint32_t), hipMemcpyDeviceToHost, stream),
               "read"))
      return -1;
    hipLaunchKernelGGL(k_hot_copy, dim3((4u * count + 255u) / 256u), dim3(256), 0, stream, nodes, hot_src, inner_total,
                       count);
    if (!check(hipGetLastError(), "launch") || !check(hipStreamSynchronize(stream), "synchronize")) return -1;
  }
  return static_cast<int>(count);
}

void launch_leaf_positions(hipStream_t stream, const float4* tris, uint32_t count, uint32_t* leaf_pos) {
  if (count == 0u) return;
  hipLaunchKernelGGL(k_leaf_positions, dim3((count + 255u) / 256u), dim3(256), 0, stream, tris, count, leaf_pos);
}

}  // namespace ptg

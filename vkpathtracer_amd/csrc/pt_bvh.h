// pt_bvh.h — host-side scene flattening, BVH builders and the device BVH layout.
#pragma once

#include <cstdint>
#include <memory>
#include <string>
#include <thread>
#include <vector>

#include "vkpt_hip.h"

namespace ptb {

// An allocator whose resize leaves new elements of trivial types unwritten: the large host arrays of a
// build are written once, element by element, by several threads, so their pages are first touched
// in parallel rather than zeroed by one thread.
template <typename T>
struct NoInitAlloc : std::allocator<T> {
  template <typename U>
  struct rebind {
    using other = NoInitAlloc<U>;
  };
  NoInitAlloc() = default;
  template <typename U>
  NoInitAlloc(const NoInitAlloc<U>&) noexcept {}
  template <typename U>
  void construct(U*) noexcept {}
  template <typename U, typename... A>
  void construct(U* p, A&&... a) {
    ::new (static_cast<void*>(p)) U(std::forward<A>(a)...);
  }
};
template <typename T>
using HostVec = std::vector<T, NoInitAlloc<T>>;

// fn(begin, end) over [0, n) split between up to `threads` threads (one below 64k elements).
template <typename F>
void parallel_for(size_t n, unsigned threads, F fn) {
  if (threads <= 1u || n < 65536u) {
    fn(size_t{0}, n);
    return;
  }
  const size_t chunk = (n + threads - 1u) / threads;
  std::vector<std::thread> pool;
  for (unsigned t = 0; t < threads; ++t) {
    const size_t a = t * chunk, b = std::min(n, a + chunk);
    if (a < b) pool.emplace_back(fn, a, b);
  }
  for (std::thread& th : pool) th.join();
}

// One flattened triangle (CpuBvhAccelerator::Primitive, CpuBvhAccelerator.cpp:497-505):
// primitive_index is its position in the flattened array.
struct Prim {
  uint32_t material;
  float v0[3], e1[3], e2[3], n[3];
  float uv[6];  // texture coordinates of v0, v1, v2 (zero without texcoords)
};

struct Node {
  float mn[3], mx[3];  // padded by 1e-4 (make_triangle_aabb, CpuBvhAccelerator.cpp:84-95)
  int32_t left, right, first, count;
};

struct Bvh {
  HostVec<Prim> prims;
  HostVec<Node> nodes;      // node 0 is the root
  HostVec<uint32_t> order;  // leaf-ordered primitive indices
  uint32_t max_depth = 0;
  uint32_t leaf_count = 0;
};

// Flatten instances into primitives; raw receives the 3 unpadded vertices of each primitive.
// texcoords (2 floats per vertex of scene.vertices, may be null) give the primitives of indexed
// geometry their texture coordinates. inst_first (may be null) receives each instance's first
// primitive and, last, the primitive count (CpuBvhAccelerator's instance ranges). Returns false (with
// err set) on a malformed snapshot. `threads` host threads share the triangles. materials (may be null)
// receives each primitive's instance material; bounded (may be null) whether every vertex coordinate is
// finite and within 1e30 (what the device build needs).
bool flatten(const PtSceneView& scene, HostVec<Prim>& prims, HostVec<float>& raw, std::string& err,
             const float* texcoords = nullptr, std::vector<uint32_t>* inst_first = nullptr, unsigned threads = 1,
             HostVec<uint32_t>* materials = nullptr, bool* bounded = nullptr);

// Bucketed SAH (D3D12GpuPathTracer.SceneBvh.cpp:73-238) and midpoint (ParallelBvhBuilder.cpp:61-163).
void build_sah(Bvh& bvh, const HostVec<float>& raw, uint32_t leaf_size, uint32_t buckets);
void build_midpoint(Bvh& bvh, const HostVec<float>& raw, uint32_t leaf_size);

// Device layout.
//   node (64 B):  float4 {c0.min.xyz, c0.max.x}, {c0.max.yz, c1.min.xy}, {c1.min.z, c1.max.xyz},
//                 int4  {ref0, ref1, 0, 0}
//   ref >= 0 : inner node index;  ref < 0 : leaf, ~ref = (first << 5) | count
//   tri  (48 B):  float4 {v0.xyz, prim id bits}, {e1.xyz, 0}, {e2.xyz, 0} in leaf order
//   prim (16 B):  float4 {normal.xyz, material bits} by primitive index
struct DeviceLayout {
  HostVec<float> nodes;       // 16 floats per inner node
  HostVec<float> tris;        // 12 floats per leaf-ordered triangle
  HostVec<float> prims;       // 4 floats per primitive
  HostVec<float> prim_uv;     // 6 floats per primitive (texture coordinates), empty without texcoords
  HostVec<uint32_t> prim_tri; // per primitive: its position in leaf order (with prim_uv)
  float root_box[6] = {0, 0, 0, 0, 0, 0};
  int32_t root_ref = 0;
  uint32_t inner_count = 0;
};

constexpr int kLeafCountBits = 5;
constexpr uint32_t kMaxLeafSize = (1u << kLeafCountBits) - 1u;
constexpr uint32_t kMaxPrimitives = 1u << (31 - kLeafCountBits);

// threads: host threads for the element-wise passes (1: the calling thread only).
void pack_device_layout(const Bvh& bvh, DeviceLayout& out, bool with_uv = false, unsigned threads = 1);

}  // namespace ptb

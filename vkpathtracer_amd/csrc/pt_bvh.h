// pt_bvh.h — host-side scene flattening, BVH builders and the device BVH layout.
#pragma once

#include <cstdint>
#include <string>
#include <vector>

#include "vkpt_hip.h"

namespace ptb {

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
  std::vector<Prim> prims;
  std::vector<Node> nodes;     // node 0 is the root
  std::vector<uint32_t> order; // leaf-ordered primitive indices
  uint32_t max_depth = 0;
  uint32_t leaf_count = 0;
};

// Flatten instances into primitives; raw receives the 3 unpadded vertices of each primitive.
// texcoords (2 floats per vertex of scene.vertices, may be null) give the primitives of indexed
// geometry their texture coordinates. inst_first (may be null) receives each instance's first
// primitive and, last, the primitive count (CpuBvhAccelerator's instance ranges). Returns false (with
// err set) on a malformed snapshot.
bool flatten(const PtSceneView& scene, std::vector<Prim>& prims, std::vector<float>& raw, std::string& err,
             const float* texcoords = nullptr, std::vector<uint32_t>* inst_first = nullptr);

// Bucketed SAH (D3D12GpuPathTracer.SceneBvh.cpp:73-238) and midpoint (ParallelBvhBuilder.cpp:61-163).
void build_sah(Bvh& bvh, const std::vector<float>& raw, uint32_t leaf_size, uint32_t buckets);
void build_midpoint(Bvh& bvh, const std::vector<float>& raw, uint32_t leaf_size);

// Device layout.
//   node (64 B):  float4 {c0.min.xyz, c0.max.x}, {c0.max.yz, c1.min.xy}, {c1.min.z, c1.max.xyz},
//                 int4  {ref0, ref1, 0, 0}
//   ref >= 0 : inner node index;  ref < 0 : leaf, ~ref = (first << 5) | count
//   tri  (48 B):  float4 {v0.xyz, prim id bits}, {e1.xyz, 0}, {e2.xyz, 0} in leaf order
//   prim (16 B):  float4 {normal.xyz, material bits} by primitive index
struct DeviceLayout {
  std::vector<float> nodes;   // 16 floats per inner node
  std::vector<float> tris;    // 12 floats per leaf-ordered triangle
  std::vector<float> prims;   // 4 floats per primitive
  std::vector<float> prim_uv; // 6 floats per primitive (texture coordinates), empty without texcoords
  std::vector<uint32_t> prim_tri; // per primitive: its position in leaf order (with prim_uv)
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

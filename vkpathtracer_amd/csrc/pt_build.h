// pt_build.h — the GPU bucketed-SAH build (pt_build.hip) and the scene pool packed from it on the
// device (pt_pool.hip).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <memory>
#include <string>
#include <vector>

#include "pt_bvh.h"

namespace ptg {

// ptb::build_sah on the device: the same tree (nodes in the host builder's pre-order, leaf order,
// depth, leaf count) from bvh.prims and their raw vertices, on `stream`. Returns 0 when built, 1 when
// the scene has a non-finite vertex (its min/max reductions would depend on the visiting order; the
// caller builds on the host), -1 on a HIP error (err says which).
int build_sah_device(hipStream_t stream, ptb::Bvh& bvh, const ptb::HostVec<float>& raw, uint32_t leaf_size,
                     uint32_t buckets, std::string& err);

// One scene of build_sah_device_many: its tree (bvh.prims set; nodes and order written unless the
// build keeps its trees on the device), raw vertices, the primitives' instance materials (read when
// the trees stay on the device), parameters, and on return rc: 0 built, 1 not built (a non-finite
// vertex; build it on the host), and where a kept tree is: forest `forest`, its scene `scene`.
struct DeviceBuild {
  ptb::Bvh* bvh;
  const ptb::HostVec<float>* raw;
  uint32_t leaf_size, buckets;
  int rc;
  const ptb::HostVec<uint32_t>* material = nullptr;
  int forest = -1;
  uint32_t scene = 0;
  int bounded = -1;  // 1: every coordinate finite and within 1e30 (ptb::flatten's check), 0: not, -1: unknown
};

// Trees kept on the device after a build of several scenes, in the host builder's form: each scene's
// nodes (ptb::Node records, the host builder's pre-order) from node_base, and from ref_base its leaf
// order (scene-local primitive indices), its primitives' vertices (9 floats each, flattening order)
// and instance materials. The storage is released on the build's stream when the last copy goes.
struct ForestScene {
  uint32_t node_base, node_count, ref_base, prims;
  ptb::Node root;
};
struct DeviceForest {
  std::vector<std::shared_ptr<void>> storage;
  const ptb::Node* nodes = nullptr;
  const uint32_t* order = nullptr;
  const float* raw = nullptr;
  const uint32_t* material = nullptr;
  uint32_t node_total = 0, ref_total = 0;
  std::vector<ForestScene> scenes;
};

// build_sah_device for several scenes at once: they share every level's launches, so many small scenes
// cost about as many synchronisations as the deepest one. With `keep`, the trees stay on the device
// (appended to *keep; bvh.nodes and bvh.order are left empty; max_depth and leaf_count are set).
// Returns 0, or -1 on a HIP error.
int build_sah_device_many(hipStream_t stream, std::vector<DeviceBuild>& jobs, std::string& err,
                          std::vector<DeviceForest>* keep = nullptr);

// Where a kept scene goes in the device scene pool, and the pool's arrays (the device layout that
// ptb::pack_device_layout makes on the host, with every reference rebased into the pool).
struct PoolDest {
  uint32_t node_base, tri_base, prim_base, mat_base, last_mat;
  uint32_t with_uv;  // the scene has texture coordinates: its primitives' triangle records go to prim_tri
};
struct PoolArrays {
  float4* nodes;       // 4 per inner node
  float4* tris;        // 3 per triangle, leaf order
  float4* prims;       // 1 per primitive
  uint32_t* prim_tri;  // per primitive its triangle record (may be null)
};

// Writes a forest's scenes (dest[i] for forest scene i) into the pool on `stream`: each scene's inner
// nodes numbered in pre-order, triangles in leaf order with e1 = v1 - v0 and e2 = v2 - v0, primitive
// records with the face normal (the flattening's normalize(cross(e1, e2))) and the scene's material
// clamp. The same values as the host layout rebased into the pool, bit for bit. Returns 0 or -1.
int pack_pool(hipStream_t stream, const DeviceForest& forest, const std::vector<PoolDest>& dest,
              const PoolArrays& out, std::string& err);

// Hot copies for the traversal kernels' LDS: for each tree (root reference and root box), the inner
// nodes of largest box surface area taken from the root down, up to `per_tree` each and `max_total` in
// all, written to hot_src (pool node indices) and copied after the pool's `inner_total` nodes with their
// hot children re-pointed at the copies. Returns the number of copies (read back), or -1.
int select_hot_nodes(hipStream_t stream, float4* nodes, uint32_t inner_total, const std::vector<int32_t>& roots,
                     const std::vector<float>& root_boxes, uint32_t per_tree, uint32_t max_total,
                     uint32_t* hot_src, std::vector<uint32_t>& hot_host, std::string& err);

// Per primitive its position in leaf order, from the triangle records of a one-scene pool.
void launch_leaf_positions(hipStream_t stream, const float4* tris, uint32_t count, uint32_t* leaf_pos);

}  // namespace ptg

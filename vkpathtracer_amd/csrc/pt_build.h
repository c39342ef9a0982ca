// pt_build.h — the GPU bucketed-SAH build (pt_build.hip).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
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

// One scene of build_sah_device_many: its tree (bvh.prims set, nodes and order written), raw vertices,
// parameters, and on return rc: 0 built, 1 not built (a non-finite vertex; build it on the host).
struct DeviceBuild {
  ptb::Bvh* bvh;
  const ptb::HostVec<float>* raw;
  uint32_t leaf_size, buckets;
  int rc;
};

// build_sah_device for several scenes at once: they share every level's launches, so many small scenes
// cost about as many synchronisations as the deepest one. Returns 0, or -1 on a HIP error.
int build_sah_device_many(hipStream_t stream, std::vector<DeviceBuild>& jobs, std::string& err);

}  // namespace ptg

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
int build_sah_device(hipStream_t stream, ptb::Bvh& bvh, const std::vector<float>& raw, uint32_t leaf_size,
                     uint32_t buckets, std::string& err);

}  // namespace ptg

"""Host-side BVH construction through the C ABI (pt_bvh_build_host): no device needed."""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import abi


def build_host(scene, builder: str = "sah", leaf_size: int = 4, bucket_count: int = 8) -> dict:
    """Flatten `scene` and build its BVH exactly as pt_hip_build_accel does, returning
    bounds (N, 6) f32, meta (N, 4) i32 [left, right, first, count], order (P,) u32,
    geo (P, 12) f32 [v0, e1, e2, normal] and max_depth."""
    lib = abi.load_library()
    view = abi.SceneView(scene)
    p = abi.PtBvhParams()
    p.builder = {"sah": 0, "midpoint": 1}[builder]
    p.leaf_size, p.bucket_count = leaf_size, bucket_count
    nn, npr, md = C.c_uint32(), C.c_uint32(), C.c_uint32()
    rc = lib.pt_bvh_build_host(C.byref(view.view), C.byref(p), C.addressof(nn), C.addressof(npr), C.addressof(md),
                               None, None, None, None)
    if rc != 0:
        raise RuntimeError(f"pt_bvh_build_host failed ({abi.STATUS_NAMES[rc]})")
    bounds = np.zeros((nn.value, 6), np.float32)
    meta = np.zeros((nn.value, 4), np.int32)
    order = np.zeros(npr.value, np.uint32)
    geo = np.zeros((npr.value, 12), np.float32)
    rc = lib.pt_bvh_build_host(C.byref(view.view), C.byref(p), None, None, None, bounds.ctypes.data, meta.ctypes.data,
                               order.ctypes.data, geo.ctypes.data)
    if rc != 0:
        raise RuntimeError(f"pt_bvh_build_host failed ({abi.STATUS_NAMES[rc]})")
    return dict(bounds=bounds, meta=meta, order=order, geo=geo, max_depth=md.value)

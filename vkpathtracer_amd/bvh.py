"""BVH construction through the C ABI: on the host (pt_bvh_build_host, no device needed) or the SAH
tree on the GPU (pt_bvh_build_device), both exported in the host builder's numbering."""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import abi


def build_host(scene, builder: str = "sah", leaf_size: int = 4, bucket_count: int = 8) -> dict:
    """Flatten `scene` and build its BVH exactly as pt_hip_build_accel does, returning
    bounds (N, 6) f32, meta (N, 4) i32 [left, right, first, count], order (P,) u32,
    geo (P, 12) f32 [v0, e1, e2, normal] and max_depth."""
    return _build(scene, builder, leaf_size, bucket_count, device=False)


def build_device(scene, builder: str = "sah", leaf_size: int = 4, bucket_count: int = 8) -> dict:
    """build_host's tree built by pt_bvh_build_device (the SAH tree on the current GPU); the result
    also carries built_on_device."""
    return _build(scene, builder, leaf_size, bucket_count, device=True)


def _build(scene, builder, leaf_size, bucket_count, device):
    lib = abi.load_library()
    view = abi.SceneView(scene)
    p = abi.PtBvhParams()
    p.builder = {"sah": 0, "midpoint": 1}[builder]
    p.leaf_size, p.bucket_count = leaf_size, bucket_count
    nn, npr, md, dev = C.c_uint32(), C.c_uint32(), C.c_uint32(), C.c_uint32()
    fn = lib.pt_bvh_build_device if device else lib.pt_bvh_build_host
    extra = (C.addressof(dev),) if device else ()
    rc = fn(C.byref(view.view), C.byref(p), C.addressof(nn), C.addressof(npr), C.addressof(md), None, None, None, None,
            *extra)
    if rc != 0:
        raise RuntimeError(f"{fn.__name__} failed ({abi.STATUS_NAMES[rc]})")
    bounds = np.zeros((nn.value, 6), np.float32)
    meta = np.zeros((nn.value, 4), np.int32)
    order = np.zeros(npr.value, np.uint32)
    geo = np.zeros((npr.value, 12), np.float32)
    rc = fn(C.byref(view.view), C.byref(p), None, None, None, bounds.ctypes.data, meta.ctypes.data, order.ctypes.data,
            geo.ctypes.data, *((None,) if device else ()))
    if rc != 0:
        raise RuntimeError(f"{fn.__name__} failed ({abi.STATUS_NAMES[rc]})")
    out = dict(bounds=bounds, meta=meta, order=order, geo=geo, max_depth=md.value)
    if device:
        out["built_on_device"] = bool(dev.value)
    return out

"""ctypes mirrors of the records in include/vkpt_hip.h and the loader for libvkpt_hip.so."""
from __future__ import annotations

import ctypes as C
import os
import pathlib

import numpy as np

from . import scene as _scene

PKG_DIR = pathlib.Path(__file__).resolve().parent
LIB_PATH = PKG_DIR / "lib" / "libvkpt_hip.so"

STATUS_NAMES = ["ok", "unsupported", "invalid_argument", "not_ready", "busy", "alloc_failed", "timeout",
                "cancelled", "internal_error"]  # vkpt::core::StatusCodeName, Result.h:69-91


class PtCamera(C.Structure):
    _fields_ = [("position", C.c_float * 3), ("target", C.c_float * 3), ("up", C.c_float * 3),
                ("fov_deg", C.c_float), ("aperture_radius", C.c_float), ("focus_distance", C.c_float),
                ("iris_blade_count", C.c_uint32), ("iris_rotation_degrees", C.c_float),
                ("iris_roundness", C.c_float), ("anamorphic_squeeze", C.c_float)]


class PtSceneView(C.Structure):
    _fields_ = [("vertices", C.c_void_p), ("indices", C.c_void_p), ("local_vertices", C.c_void_p),
                ("local_indices", C.c_void_p), ("instances", C.c_void_p), ("materials", C.c_void_p),
                ("lights", C.c_void_p), ("environment_map", C.c_void_p),
                ("vertex_count", C.c_uint32), ("index_count", C.c_uint32), ("local_vertex_count", C.c_uint32),
                ("local_index_count", C.c_uint32), ("instance_count", C.c_uint32), ("material_count", C.c_uint32),
                ("light_count", C.c_uint32), ("environment_map_width", C.c_uint32),
                ("environment_map_height", C.c_uint32), ("environment_color", C.c_float * 3),
                ("environment_map_scale", C.c_float * 3), ("camera", PtCamera)]


class PtRenderSettings(C.Structure):
    _fields_ = [("seed", C.c_uint64), ("width", C.c_uint32), ("height", C.c_uint32), ("spp", C.c_uint32),
                ("max_depth", C.c_uint32), ("enable_nee", C.c_uint32), ("enable_mis", C.c_uint32),
                ("deterministic", C.c_uint32), ("rr_start_depth", C.c_uint32), ("rr_min_survival", C.c_float),
                ("rr_max_survival", C.c_float), ("camera_aperture_radius", C.c_float),
                ("camera_focus_distance", C.c_float)]


class PtRenderTile(C.Structure):
    _fields_ = [(n, C.c_uint32) for n in ("x", "y", "width", "height", "sample_index", "tile_id", "gpu_id")]


COUNTER_FIELDS = ("samples", "rays", "triangle_tests", "sdf_tests", "sdf_steps", "triangle_hits", "sdf_hits",
                  "sdf_misses", "bvh_node_visits", "bvh_leaf_visits", "shadow_tests")


class PtSampleCounters(C.Structure):
    _fields_ = [(n, C.c_uint64) for n in COUNTER_FIELDS]

    def as_dict(self) -> dict:
        return {n: int(getattr(self, n)) for n in COUNTER_FIELDS}


PT_BVH_FLAG_HOST_BUILD = 1  # PtBvhParams.flags: build the SAH tree on the host


class PtBvhParams(C.Structure):
    _fields_ = [("builder", C.c_uint32), ("leaf_size", C.c_uint32), ("bucket_count", C.c_uint32),
                ("flags", C.c_uint32)]


class PtBuildInfo(C.Structure):
    _fields_ = [("build_ms", C.c_double), ("device_bytes", C.c_uint64), ("built", C.c_uint32),
                ("primitive_count", C.c_uint32), ("node_count", C.c_uint32), ("leaf_count", C.c_uint32),
                ("max_depth", C.c_uint32), ("deterministic", C.c_uint32), ("device_built", C.c_uint32),
                ("pad0", C.c_uint32)]

    def as_dict(self) -> dict:
        return {n: getattr(self, n) for n, _ in self._fields_}


class PtTracerStatus(C.Structure):
    _fields_ = [("total_samples", C.c_uint64), ("accumulation_gen", C.c_uint64), ("lifecycle", C.c_uint32),
                ("scene_loaded", C.c_uint32), ("accel_valid", C.c_uint32), ("ready_to_render", C.c_uint32),
                ("current_sample", C.c_uint32), ("pad0", C.c_uint32)]


class PtTuning(C.Structure):
    _fields_ = [(n, C.c_uint32) for n in ("wavefront_capacity", "refill_min", "traversal_blocks", "shade_blocks")]

    def as_dict(self) -> dict:
        return {n: int(getattr(self, n)) for n, _ in self._fields_}


class PtTextureDesc(C.Structure):
    _fields_ = [(n, C.c_uint32) for n in ("offset", "width", "height", "pad")]


class PtFilmResolve(C.Structure):
    _fields_ = [("exposure", C.c_float), ("white_balance_kelvin", C.c_float), ("tone_map", C.c_uint32),
                ("output_transform", C.c_uint32), ("gamma", C.c_float), ("clamp_output", C.c_uint32)]


class PtFilmReadbackToken(C.Structure):
    _fields_ = [("id", C.c_uint64), ("fence_value", C.c_uint64), ("width", C.c_uint32), ("height", C.c_uint32)]


READBACK_STATES = ("invalid", "pending", "ready", "failed", "unsupported")  # FilmReadbackState


class PtInstanceTransformUpdate(C.Structure):
    _fields_ = [("entity_id", C.c_uint64), ("instance_index", C.c_uint32), ("flags", C.c_uint32),
                ("transform_revision", C.c_uint32), ("translation", C.c_float * 3), ("rotation", C.c_float * 4),
                ("scale", C.c_float * 3), ("pad0", C.c_uint32)]


class PtInstanceTransformResult(C.Structure):
    _fields_ = [("status", C.c_uint32), ("requested_count", C.c_uint32), ("matched_count", C.c_uint32),
                ("applied_count", C.c_uint32), ("refit_ms", C.c_double)]


# InstanceTransformUpdateStatus (PathTracer.h:565-575)
TRANSFORM_STATUS = ("applied_metadata_only", "applied_instance_buffer_only", "applied_dynamic_accel_update",
                    "applied_full_static_accel_rebuild", "applied_full_scene_reload",
                    "blocked_needs_full_static_accel_rebuild", "blocked_needs_full_scene_reload", "unsupported",
                    "failed")


def make_transform_updates(updates) -> C.Array:
    """PtInstanceTransformUpdate array of scene.InstanceTransformUpdate records."""
    arr = (PtInstanceTransformUpdate * max(1, len(updates)))()
    for a, u in zip(arr, updates):
        a.entity_id = u.entity_id
        a.instance_index = u.instance_index
        a.flags = u.flags
        a.transform_revision = u.transform_revision
        a.translation[:] = list(u.translation)
        a.rotation[:] = list(u.rotation)
        a.scale[:] = list(u.scale)
    return arr


def make_resolve(r: "_scene.FilmResolve | None") -> PtFilmResolve:
    """PtFilmResolve of a FilmResolve (None: FilmResolveSettings defaults)."""
    r = r or _scene.FilmResolve()
    if r.tone_map not in _scene.TONE_MAPS or r.output_transform not in _scene.OUTPUT_TRANSFORMS:
        raise ValueError(f"unknown tone map {r.tone_map!r} or output transform {r.output_transform!r}")
    out = PtFilmResolve()
    out.exposure = r.exposure
    out.white_balance_kelvin = r.white_balance_kelvin
    out.tone_map = _scene.TONE_MAPS.index(r.tone_map)
    out.output_transform = _scene.OUTPUT_TRANSFORMS.index(r.output_transform)
    out.gamma = r.gamma
    out.clamp_output = 1 if r.clamp_output else 0
    return out


# PtRayQueryHit (RayQueryHit, PathTracer.h:976-983): 40 bytes.
RAY_QUERY_HIT_DTYPE = np.dtype([("hit", np.uint32), ("t", np.float32), ("position", np.float32, 3),
                                ("normal", np.float32, 3), ("material_index", np.uint32),
                                ("primitive_index", np.uint32)])
# RTMaterialUpdate / RTLightUpdate / RTInstanceMaterialUpdate (PathTracer.h:438-452).
MATERIAL_UPDATE_DTYPE = np.dtype([("material_index", np.uint32), ("pad", np.uint32),
                                  ("material", _scene.MATERIAL_DTYPE)])
LIGHT_UPDATE_DTYPE = np.dtype([("light_index", np.uint32), ("light", _scene.LIGHT_DTYPE)])
INSTANCE_MATERIAL_UPDATE_DTYPE = np.dtype([("entity_id", np.uint64), ("instance_index", np.uint32),
                                           ("material_index", np.uint32)])
INVALID_INSTANCE_INDEX = 0xFFFFFFFF  # kInvalidRTInstanceIndex


class PtSceneDeltaUpdate(C.Structure):
    _fields_ = [("materials", C.c_void_p), ("lights", C.c_void_p), ("instance_materials", C.c_void_p),
                ("material_count", C.c_uint32), ("light_count", C.c_uint32),
                ("instance_material_count", C.c_uint32), ("environment_color_changed", C.c_uint32),
                ("environment_color", C.c_float * 3), ("pad", C.c_uint32)]


KERNELS = ("raygen", "extend", "shade", "shadow", "film")


class PtKernelStats(C.Structure):
    _fields_ = [("total_ms", C.c_double * 5), ("launches", C.c_uint64 * 5), ("items", C.c_uint64 * 5),
                ("inner_node_fetches", C.c_uint64), ("shadow_node_visits", C.c_uint64),
                ("shadow_triangle_tests", C.c_uint64), ("shadow_inner_node_fetches", C.c_uint64),
                ("shadow_leaf_visits", C.c_uint64), ("shadow_rays", C.c_uint64)]


# Every symbol include/vkpt_hip.h declares, with its ctypes signature.
_P = C.c_void_p
SIGNATURES = {
    "pt_hip_create": (C.c_int, [C.c_int, C.POINTER(C.c_void_p)]),
    "pt_hip_destroy": (None, [_P]),
    "pt_hip_configure": (C.c_int, [_P, C.POINTER(PtRenderSettings)]),
    "pt_hip_load_scene": (C.c_int, [_P, C.POINTER(PtSceneView)]),
    "pt_hip_build_accel": (C.c_int, [_P, C.POINTER(PtBvhParams), C.POINTER(PtBuildInfo)]),
    "pt_hip_reset_accumulation": (C.c_int, [_P]),
    "pt_hip_render_tile": (C.c_int, [_P, C.POINTER(PtRenderTile), C.c_uint64]),
    "pt_hip_render_tiles": (C.c_int, [_P, C.POINTER(PtRenderTile), C.c_uint32, C.c_uint64]),
    "pt_hip_read_film": (C.c_int, [_P, _P, _P]),
    "pt_hip_resolve_hdr": (C.c_int, [_P, _P]),
    "pt_hip_resolve_ldr": (C.c_int, [_P, C.POINTER(PtFilmResolve), _P]),
    "pt_hip_replace_film": (C.c_int, [_P, _P, _P]),
    "pt_hip_request_film_readback": (C.c_int, [_P, C.POINTER(PtFilmResolve), C.POINTER(PtFilmReadbackToken)]),
    "pt_hip_poll_film": (C.c_int, [_P, C.POINTER(PtFilmReadbackToken), _P, C.POINTER(C.c_uint32)]),
    "pt_hip_update_camera": (C.c_int, [_P, C.POINTER(PtCamera)]),
    "pt_hip_load_sdf_primitives": (C.c_int, [_P, _P, C.c_uint32]),
    "pt_hip_plan_instance_transform_update": (C.c_int, [_P, _P, C.c_uint32, C.POINTER(PtInstanceTransformResult)]),
    "pt_hip_apply_instance_transform_update": (C.c_int, [_P, _P, C.c_uint32, C.POINTER(PtInstanceTransformResult)]),
    "pt_hip_update_render_settings": (C.c_int, [_P, C.POINTER(PtRenderSettings)]),
    "pt_hip_read_counters": (C.c_int, [_P, C.POINTER(PtSampleCounters)]),
    "pt_hip_status": (C.c_int, [_P, C.POINTER(PtTracerStatus)]),
    "pt_hip_shutdown": (C.c_int, [_P]),
    "pt_hip_last_error": (C.c_char_p, [_P]),
    "pt_hip_trace_rays": (C.c_int, [_P, C.c_int, _P, _P, C.c_uint32, _P, _P, C.POINTER(PtSampleCounters)]),
    "pt_hip_trace_rays_noalloc": (C.c_int, [_P, C.c_int, _P, _P, C.c_uint32, _P, _P, C.POINTER(PtSampleCounters)]),
    "pt_hip_intersect": (C.c_int, [_P, C.c_int, _P, _P, C.c_uint32, _P, C.POINTER(PtSampleCounters)]),
    "pt_hip_apply_scene_delta": (C.c_int, [_P, C.POINTER(PtSceneDeltaUpdate)]),
    "pt_hip_synchronize": (C.c_int, [_P]),
    "pt_hip_set_wavefront_capacity": (C.c_int, [_P, C.c_uint32]),
    "pt_hip_set_profiling": (C.c_int, [_P, C.c_int]),
    "pt_hip_set_tuning": (C.c_int, [_P, C.POINTER(PtTuning)]),
    "pt_hip_get_tuning": (C.c_int, [_P, C.POINTER(PtTuning)]),
    "pt_hip_read_kernel_stats": (C.c_int, [_P, C.POINTER(PtKernelStats)]),
    "pt_hip_reset_kernel_stats": (C.c_int, [_P]),
    "pt_bvh_build_host": (C.c_int, [C.POINTER(PtSceneView), C.POINTER(PtBvhParams), _P, _P, _P, _P, _P, _P, _P]),
    "pt_bvh_build_device": (C.c_int, [C.POINTER(PtSceneView), C.POINTER(PtBvhParams), _P, _P, _P, _P, _P, _P, _P, _P]),
    "pt_hip_update_scene_delta": (C.c_int, [_P, _P, C.c_uint32, _P, C.c_uint32]),
    "pt_hip_render_adjoint": (C.c_int, [_P, C.POINTER(PtRenderTile), C.c_uint32, C.c_uint64, _P, _P, _P, _P, _P,
                                        _P]),
    "pt_batch_render_adjoint": (C.c_int, [_P, C.c_uint64, _P, _P, _P, _P, _P, _P]),
    "pt_batch_update_scene_delta": (C.c_int, [_P, _P, C.c_uint32, _P, C.c_uint32]),
    "pt_batch_update_texels": (C.c_int, [_P, _P, C.c_uint64]),
    "pt_hip_load_textures": (C.c_int, [_P, _P, C.c_uint32, _P, C.c_uint64, _P, C.c_uint32]),
    "pt_hip_update_texels": (C.c_int, [_P, _P, C.c_uint64]),
    "pt_batch_create": (C.c_int, [C.c_int, C.POINTER(PtRenderSettings), C.POINTER(C.c_void_p)]),
    "pt_batch_destroy": (None, [_P]),
    "pt_batch_last_error": (C.c_char_p, [_P]),
    "pt_batch_add_scene": (C.c_int, [_P, C.POINTER(PtSceneView), C.c_uint32, C.c_uint32, C.c_uint32,
                                     C.POINTER(PtBvhParams), C.POINTER(C.c_uint32)]),
    "pt_batch_add_view": (C.c_int, [_P, C.c_uint32, C.POINTER(PtCamera), C.c_uint32, C.c_uint32, C.c_uint32,
                                    C.POINTER(C.c_uint32)]),
    "pt_batch_load_textures": (C.c_int, [_P, C.c_uint32, _P, C.c_uint32, _P, C.c_uint64, _P, C.c_uint32]),
    "pt_batch_build": (C.c_int, [_P, C.POINTER(PtBuildInfo)]),
    "pt_batch_reset_accumulation": (C.c_int, [_P]),
    "pt_batch_render": (C.c_int, [_P, C.c_uint64, C.c_int]),
    "pt_batch_read_film": (C.c_int, [_P, C.c_uint32, _P, _P]),
    "pt_batch_resolve_hdr": (C.c_int, [_P, C.c_uint32, _P]),
    "pt_batch_update_camera": (C.c_int, [_P, C.c_uint32, C.POINTER(PtCamera)]),
    "pt_batch_load_sdf_primitives": (C.c_int, [_P, C.c_uint32, _P, C.c_uint32]),
    "pt_batch_resolve_ldr": (C.c_int, [_P, C.c_uint32, C.POINTER(PtFilmResolve), _P]),
    "pt_batch_read_counters": (C.c_int, [_P, C.POINTER(PtSampleCounters)]),
    "pt_batch_set_tuning": (C.c_int, [_P, C.POINTER(PtTuning)]),
    "pt_batch_get_tuning": (C.c_int, [_P, C.POINTER(PtTuning)]),
    "pt_batch_set_profiling": (C.c_int, [_P, C.c_int]),
    "pt_batch_read_kernel_stats": (C.c_int, [_P, C.POINTER(PtKernelStats)]),
    "pt_batch_reset_kernel_stats": (C.c_int, [_P]),
    "pt_batch_synchronize": (C.c_int, [_P]),
    "pt_hip_sizeof": (C.c_uint32, [C.c_int]),
    "pt_hip_version": (C.c_char_p, []),
}

_lib = None


class NativeLibraryMissing(RuntimeError):
    pass


def load_library(path: os.PathLike | None = None) -> C.CDLL:
    """Load libvkpt_hip.so (built in-tree by __graft_entry__.build()). Raises if it is absent."""
    global _lib
    if _lib is not None and path is None:
        return _lib
    p = pathlib.Path(path) if path else LIB_PATH
    # torch ships its own HIP runtime with the same sonames (libamdhip64.so.7, libhsa-runtime64.so.1).
    # Loaded first, it satisfies this library's dependencies and the process has one runtime; loaded
    # after this library, it would bring a second HSA runtime whose initialisation finds no GPU.
    try:
        import torch  # noqa: F401
    except ImportError:
        pass
    if not p.exists():
        raise NativeLibraryMissing(f"{p} is missing: build the HIP library first (python -c 'import __graft_entry__ as g; g.build()')")
    lib = C.CDLL(str(p))
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    if path is None:
        _lib = lib
    return lib


def _ptr(a: np.ndarray | None):
    if a is None or a.size == 0:
        return None
    return a.ctypes.data


class SceneView:
    """PtSceneView over a Scene's arrays; keeps the arrays alive while the view exists."""

    def __init__(self, sc: "_scene.Scene"):
        self.scene = sc
        self.keep = [sc.vertices, sc.indices, sc.local_vertices, sc.local_indices, sc.instances, sc.materials,
                     sc.lights]
        v = PtSceneView()
        v.vertices = _ptr(sc.vertices)
        v.indices = _ptr(sc.indices)
        v.local_vertices = _ptr(sc.local_vertices)
        v.local_indices = _ptr(sc.local_indices)
        v.instances = _ptr(sc.instances)
        v.materials = _ptr(sc.materials)
        v.lights = _ptr(sc.lights)
        env = sc.environment_map
        v.environment_map = _ptr(env)
        if env is not None:
            self.keep.append(env)
            v.environment_map_height, v.environment_map_width = env.shape[0], env.shape[1]
        v.environment_map_scale[:] = list(sc.environment_map_scale)
        v.vertex_count = len(sc.vertices)
        v.index_count = len(sc.indices)
        v.local_vertex_count = len(sc.local_vertices)
        v.local_index_count = len(sc.local_indices)
        v.instance_count = len(sc.instances)
        v.material_count = len(sc.materials)
        v.light_count = len(sc.lights)
        v.environment_color[:] = list(sc.environment_color)
        v.camera = make_camera(sc.camera)
        self.view = v
        # texture arrays (pt_hip_load_textures / pt_batch_load_textures)
        self.texcoords = sc.texcoords
        self.texels, self.texture_descs = sc.texture_arrays()

    def texture_args(self) -> tuple:
        """(texcoords, texcoord_count, texels, texel_count, descs, desc_count) for *_load_textures."""
        tc = self.texcoords
        return (_ptr(tc), 0 if tc is None else len(tc), _ptr(self.texels), len(self.texels),
                _ptr(self.texture_descs), len(self.texture_descs))

    @property
    def has_textures(self) -> bool:
        return self.texcoords is not None or len(self.texture_descs) > 0


def make_camera(c: "_scene.Camera") -> PtCamera:
    cam = PtCamera()
    cam.position[:] = list(c.position)
    cam.target[:] = list(c.target)
    cam.up[:] = list(c.up)
    cam.fov_deg = c.fov_deg
    cam.aperture_radius = c.aperture_radius
    cam.focus_distance = c.focus_distance
    cam.iris_blade_count = c.iris_blade_count
    cam.iris_rotation_degrees = c.iris_rotation_degrees
    cam.iris_roundness = c.iris_roundness
    cam.anamorphic_squeeze = c.anamorphic_squeeze
    return cam


def make_settings(s: "_scene.RenderSettings") -> PtRenderSettings:
    out = PtRenderSettings()
    out.seed = s.seed & 0xFFFFFFFFFFFFFFFF
    out.width, out.height, out.spp, out.max_depth = s.width, s.height, s.spp, s.max_depth
    out.enable_nee, out.enable_mis, out.deterministic = int(s.enable_nee), int(s.enable_mis), int(s.deterministic)
    out.rr_start_depth = s.rr_start_depth
    out.rr_min_survival, out.rr_max_survival = s.rr_min_survival, s.rr_max_survival
    out.camera_aperture_radius, out.camera_focus_distance = s.camera_aperture_radius, s.camera_focus_distance
    return out

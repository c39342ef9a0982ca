"""HipPathTracer — the host-side mirror of the reference's IPathTracer over the C ABI.

Method names, argument meaning and error behaviour follow
``vkpt::pathtracer::IPathTracer`` (src/pathtracer/PathTracer.h:1069-1184):
``configure`` / ``load_scene_snapshot`` / ``build_or_update_acceleration`` return a
:class:`Status`; ``render_tile`` / ``render_tiles`` / ``reset_accumulation`` return a
bool and leave the message in ``status().last_error``. Every call goes to
libvkpt_hip.so; there is no CPU fallback — construction raises if the library or a
GPU is missing.
"""
from __future__ import annotations

import ctypes as C
import dataclasses

import numpy as np

from . import abi
from .scene import Camera as _Camera
from .scene import RenderSettings, Scene


@dataclasses.dataclass
class Status:
    """vkpt::core::Status (src/core/contracts/Result.h:92-119)."""
    code: int = 0
    message: str = ""

    @property
    def code_name(self) -> str:
        return abi.STATUS_NAMES[self.code] if 0 <= self.code < len(abi.STATUS_NAMES) else "unknown"

    def is_ok(self) -> bool:
        return self.code == 0

    def __bool__(self) -> bool:
        return self.is_ok()


class PathTracerError(RuntimeError):
    def __init__(self, status: Status, what: str):
        super().__init__(f"{what}: {status.code_name}: {status.message}")
        self.status = status


@dataclasses.dataclass
class TracerStatus:
    """PathTracerStatus subset (PathTracer.h:854-865)."""
    backend: str
    lifecycle: str
    scene_loaded: bool
    accel_valid: bool
    ready_to_render: bool
    current_sample: int
    total_samples: int
    accumulation_gen: int
    last_error: str


_LIFECYCLE = ["Uninitialized", "Configured", "SceneLoaded", "Ready"]


def _dptr(t):
    return None if t is None or t.numel() == 0 else t.data_ptr()


class HipPathTracer:
    """IPathTracer implemented by the gfx950 wavefront kernels."""

    backend_name = "hip-wavefront"

    def __init__(self, device: int = 0, library=None):
        self._lib = abi.load_library(library)
        h = C.c_void_p()
        rc = self._lib.pt_hip_create(device, C.byref(h))
        if rc != 0:
            raise PathTracerError(Status(rc, "no HIP device available for the path tracer"), "pt_hip_create")
        self._h = h
        self._settings: RenderSettings | None = None
        self._view: abi.SceneView | None = None
        self._camera = None
        self.device = device

    # -- lifetime
    def close(self):
        if getattr(self, "_h", None):
            self._lib.pt_hip_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def _status(self, rc: int) -> Status:
        msg = self._lib.pt_hip_last_error(self._h) if rc else b""
        return Status(rc, (msg or b"").decode())

    def _check(self, rc: int, what: str):
        if rc != 0:
            raise PathTracerError(self._status(rc), what)

    # -- IPathTracer
    def configure(self, settings: RenderSettings) -> Status:
        self._settings = settings
        s = abi.make_settings(settings)
        return self._status(self._lib.pt_hip_configure(self._h, C.byref(s)))

    def load_scene_snapshot(self, scene: Scene) -> Status:
        view = abi.SceneView(scene)
        st = self._status(self._lib.pt_hip_load_scene(self._h, C.byref(view.view)))
        if st and view.has_textures:
            st = self._status(self._lib.pt_hip_load_textures(self._h, *view.texture_args()))
        sdf = getattr(scene, "sdf_primitives", None)
        if st and sdf is not None and len(sdf):
            sdf = np.ascontiguousarray(sdf)
            view.keep.append(sdf)
            st = self._status(self._lib.pt_hip_load_sdf_primitives(self._h, sdf.ctypes.data, len(sdf)))
        if st:
            self._view = view
            self._camera = scene.camera
        return st

    def update_render_settings(self, settings: RenderSettings) -> bool:
        """IPathTracer::update_render_settings: False when the backend needs configure (a new size)."""
        s = abi.make_settings(settings)
        ok = self._lib.pt_hip_update_render_settings(self._h, C.byref(s)) == 0
        if ok:
            self._settings = settings
        return ok

    def update_camera_state(self, camera) -> bool:
        """IPathTracer::update_camera_state: a new camera (scene.Camera) without re-upload or rebuild."""
        c = abi.make_camera(camera)
        ok = self._lib.pt_hip_update_camera(self._h, C.byref(c)) == 0
        if ok:
            self._camera = camera
        return ok

    def update_camera(self, position, target, up, fov_deg: float) -> bool:
        """IPathTracer::update_camera: pose and field of view; the lens fields stay."""
        cam = dataclasses.replace(self._camera or _Camera(), position=tuple(position), target=tuple(target),
                                  up=tuple(up), fov_deg=fov_deg)
        return self.update_camera_state(cam)

    def plan_instance_transform_update(self, updates) -> tuple[str, int]:
        """plan_instance_transform_update: (InstanceTransformUpdateStatus name, matched count); a pure
        query."""
        r = abi.PtInstanceTransformResult()
        arr = abi.make_transform_updates(updates)
        self._check(self._lib.pt_hip_plan_instance_transform_update(self._h, arr, len(updates), C.byref(r)),
                    "plan_instance_transform_update")
        return abi.TRANSFORM_STATUS[r.status], r.matched_count

    def apply_instance_transform_update(self, updates) -> tuple[str, int, float]:
        """apply_instance_transform_update: the updated dynamic instances move and the BVH is refit on
        the GPU. (status name, applied count, refit milliseconds)."""
        r = abi.PtInstanceTransformResult()
        arr = abi.make_transform_updates(updates)
        self._check(self._lib.pt_hip_apply_instance_transform_update(self._h, arr, len(updates), C.byref(r)),
                    "apply_instance_transform_update")
        return abi.TRANSFORM_STATUS[r.status], r.applied_count, r.refit_ms

    def update_instance_transforms(self, updates) -> bool:
        """IPathTracer::update_instance_transforms: True when the transforms were applied."""
        return self.apply_instance_transform_update(updates)[0].startswith("applied")

    def request_film_readback(self, settings=None) -> "abi.PtFilmReadbackToken":
        """IPathTracer::request_film_readback: the LDR resolve of the film as queued so far, copied
        to host memory while later rendering proceeds. Poll it with poll_film."""
        tok = abi.PtFilmReadbackToken()
        rs = abi.make_resolve(settings)
        self._check(self._lib.pt_hip_request_film_readback(self._h, C.byref(rs), C.byref(tok)),
                    "request_film_readback")
        return tok

    def poll_film(self, token) -> tuple[str, np.ndarray | None]:
        """IPathTracer::poll_film: ("pending" | "ready" | "invalid" | "failed", (H, W, 4) u8 when ready)."""
        state = C.c_uint32()
        img = np.zeros((max(token.height, 1), max(token.width, 1), 4), np.uint8)
        self._check(self._lib.pt_hip_poll_film(self._h, C.byref(token), img.ctypes.data, C.byref(state)), "poll_film")
        name = abi.READBACK_STATES[state.value]
        return name, (img if name == "ready" else None)

    def update_texels(self, texels) -> None:
        """Replace every texel of the built scene (same count): a (N, 3) float32 numpy array or a
        contiguous float32 device tensor on this tracer's device."""
        if isinstance(texels, np.ndarray):
            t = np.ascontiguousarray(texels, np.float32).reshape(-1, 3)
            self._check(self._lib.pt_hip_update_texels(self._h, t.ctypes.data, len(t)), "update_texels")
        else:
            import torch

            torch.cuda.synchronize(self.device)
            self._check(self._lib.pt_hip_update_texels(self._h, texels.data_ptr(), texels.numel() // 3),
                        "update_texels")

    def build_or_update_acceleration(self, builder: str = "sah", leaf_size: int = 4, bucket_count: int = 8,
                                     host_build: bool = False) -> Status:
        """host_build: build the SAH tree on the host (PT_BVH_FLAG_HOST_BUILD) instead of the GPU."""
        p = abi.PtBvhParams()
        p.builder = {"sah": 0, "midpoint": 1}[builder]
        p.leaf_size = leaf_size
        p.bucket_count = bucket_count
        p.flags = abi.PT_BVH_FLAG_HOST_BUILD if host_build else 0
        info = abi.PtBuildInfo()
        st = self._status(self._lib.pt_hip_build_accel(self._h, C.byref(p), C.byref(info)))
        self.build_info = info.as_dict()
        return st

    def reset_accumulation(self) -> bool:
        return self._lib.pt_hip_reset_accumulation(self._h) == 0

    def render_tile(self, tile: dict | tuple, frame_index: int = 0) -> bool:
        return self.render_tiles([tile], frame_index)

    def render_tiles(self, tiles, frame_index: int = 0) -> bool:
        arr = (abi.PtRenderTile * max(1, len(tiles)))()
        for i, t in enumerate(tiles):
            if isinstance(t, dict):
                t = (t.get("x", 0), t.get("y", 0), t["width"], t["height"], t.get("sample_index", 0),
                     t.get("tile_id", 0), t.get("gpu_id", 0))
            arr[i] = abi.PtRenderTile(*[int(v) for v in (tuple(t) + (0, 0, 0, 0, 0, 0, 0))[:7]])
        return self._lib.pt_hip_render_tiles(self._h, arr, len(tiles), frame_index) == 0

    @staticmethod
    def _tile_array(tiles):
        arr = (abi.PtRenderTile * max(1, len(tiles)))()
        for i, t in enumerate(tiles):
            if isinstance(t, dict):
                t = (t.get("x", 0), t.get("y", 0), t["width"], t["height"], t.get("sample_index", 0),
                     t.get("tile_id", 0), t.get("gpu_id", 0))
            arr[i] = abi.PtRenderTile(*[int(v) for v in (tuple(t) + (0, 0, 0, 0, 0, 0, 0))[:7]])
        return arr

    def update_scene_delta(self, materials: np.ndarray | None = None, lights: np.ndarray | None = None) -> bool:
        """IPathTracer::update_scene_delta: replace material and/or light records (same counts)."""
        from .scene import LIGHT_DTYPE, MATERIAL_DTYPE

        m = None if materials is None else np.ascontiguousarray(materials, MATERIAL_DTYPE)
        l = None if lights is None else np.ascontiguousarray(lights, LIGHT_DTYPE)
        rc = self._lib.pt_hip_update_scene_delta(self._h, None if m is None else m.ctypes.data, 0 if m is None else len(m),
                                                 None if l is None else l.ctypes.data, 0 if l is None else len(l))
        return rc == 0

    def apply_scene_delta(self, materials=None, lights=None, instance_materials=None,
                          environment_color=None) -> bool:
        """IPathTracer::update_scene_delta with an RTSceneDeltaUpdate (PathTracer.h:455-461):
        materials = [(index, record)], lights = [(index, record)], instance_materials =
        [(instance_index, material_index)] or [(entity_id, INVALID_INSTANCE_INDEX, material_index)],
        environment_color = rgb or None. False (status message kept) when an index is out of range or
        nothing changed, as ScalarCpuPathTracer::update_scene_delta returns false."""
        mu = np.zeros(len(materials or ()), abi.MATERIAL_UPDATE_DTYPE)
        for i, (k, rec) in enumerate(materials or ()):
            mu[i]["material_index"] = k
            mu[i]["material"] = rec
        lu = np.zeros(len(lights or ()), abi.LIGHT_UPDATE_DTYPE)
        for i, (k, rec) in enumerate(lights or ()):
            lu[i]["light_index"] = k
            lu[i]["light"] = rec
        iu = np.zeros(len(instance_materials or ()), abi.INSTANCE_MATERIAL_UPDATE_DTYPE)
        for i, e in enumerate(instance_materials or ()):
            if len(e) == 2:
                iu[i]["instance_index"], iu[i]["material_index"] = e
            else:
                iu[i]["entity_id"], iu[i]["instance_index"], iu[i]["material_index"] = e
        u = abi.PtSceneDeltaUpdate()
        u.materials = mu.ctypes.data if len(mu) else None
        u.lights = lu.ctypes.data if len(lu) else None
        u.instance_materials = iu.ctypes.data if len(iu) else None
        u.material_count, u.light_count, u.instance_material_count = len(mu), len(lu), len(iu)
        if environment_color is not None:
            u.environment_color_changed = 1
            u.environment_color[:] = [float(c) for c in environment_color]
        return self._lib.pt_hip_apply_scene_delta(self._h, C.byref(u)) == 0

    def render_adjoint(self, tiles, film_grad, grad_albedo, grad_emission, grad_light, frame_index: int = 0,
                       grad_roughness=None, grad_texels=None):
        """Path-replay adjoint of render_tiles(tiles): adds d(sum film_grad * film)/d(parameters) into the
        device tensors grad_albedo / grad_emission (M, 3), grad_light (L, 3) and, when given,
        grad_roughness (M,) and grad_texels (N, 3) over the scene's textures. film_grad: (H, W, 3)
        device tensor. Torch tensors must be contiguous float32 on this tracer's device."""
        import torch

        torch.cuda.synchronize(self.device)
        arr = self._tile_array(tiles)
        rc = self._lib.pt_hip_render_adjoint(self._h, arr, len(tiles), frame_index, film_grad.data_ptr(),
                                             grad_albedo.data_ptr(), grad_emission.data_ptr(),
                                             grad_light.data_ptr() if grad_light.numel() else None,
                                             _dptr(grad_roughness), _dptr(grad_texels))
        self._check(rc, "render_adjoint")
        self.synchronize()

    def render_frame_samples(self, sample_begin: int, sample_count: int, frame_index: int = 0) -> bool:
        """Full-frame tiles for samples [begin, begin+count) in one batched call (render_tiles)."""
        s = self._settings
        tiles = [(0, 0, s.width, s.height, sample_begin + i, 0, 0) for i in range(sample_count)]
        return self.render_tiles(tiles, frame_index)

    def synchronize(self):
        self._check(self._lib.pt_hip_synchronize(self._h), "synchronize")

    def film(self) -> tuple[np.ndarray, np.ndarray]:
        """(rgb sum (H, W, 3) f32, sample counts (H, W) u32) — FilmBuffer::raw / sample_counts."""
        s = self._settings
        rgb = np.zeros((s.height, s.width, 3), np.float32)
        cnt = np.zeros((s.height, s.width), np.uint32)
        self._check(self._lib.pt_hip_read_film(self._h, rgb.ctypes.data, cnt.ctypes.data), "read_film")
        return rgb, cnt

    def replace_film_history(self, rgb_sum, counts) -> None:
        """IPathTracer::replace_film_history: the film becomes (rgb sums (H, W, 3), counts (H, W))."""
        s = self._settings
        rgb = np.ascontiguousarray(rgb_sum, np.float32).reshape(s.height, s.width, 3)
        cnt = np.ascontiguousarray(counts, np.uint32).reshape(s.height, s.width)
        self._check(self._lib.pt_hip_replace_film(self._h, rgb.ctypes.data, cnt.ctypes.data), "replace_film_history")

    def resolve_hdr(self, out=None):
        """FilmBuffer::resolve_hdr on the GPU: (H, W, 3) f32 mean radiance. out: an optional device
        tensor of H*W*3 floats that receives the image instead of a new numpy array."""
        s = self._settings
        rgb = out if out is not None else np.zeros((s.height, s.width, 3), np.float32)
        ptr = _dptr(out) if out is not None else rgb.ctypes.data
        self._check(self._lib.pt_hip_resolve_hdr(self._h, ptr), "resolve_hdr")
        return rgb

    def resolve_ldr(self, settings=None, out=None):
        """IPathTracer::resolve_ldr on the GPU (ApplyFilmResolve): (H, W, 4) u8 rgba. settings: a
        FilmResolve (None: defaults); out: an optional device tensor of H*W*4 bytes."""
        s = self._settings
        rs = abi.make_resolve(settings)
        img = out if out is not None else np.zeros((s.height, s.width, 4), np.uint8)
        ptr = _dptr(out) if out is not None else img.ctypes.data
        self._check(self._lib.pt_hip_resolve_ldr(self._h, C.byref(rs), ptr), "resolve_ldr")
        return img

    def read_counters(self) -> dict:
        c = abi.PtSampleCounters()
        self._check(self._lib.pt_hip_read_counters(self._h, C.byref(c)), "read_counters")
        return c.as_dict()

    def status(self) -> TracerStatus:
        st = abi.PtTracerStatus()
        self._lib.pt_hip_status(self._h, C.byref(st))
        return TracerStatus(self.backend_name, _LIFECYCLE[st.lifecycle], bool(st.scene_loaded), bool(st.accel_valid),
                            bool(st.ready_to_render), int(st.current_sample), int(st.total_samples),
                            int(st.accumulation_gen), self._lib.pt_hip_last_error(self._h).decode())

    def shutdown(self):
        self._lib.pt_hip_shutdown(self._h)

    # -- tuning / measurement
    def set_wavefront_capacity(self, paths: int):
        self._check(self._lib.pt_hip_set_wavefront_capacity(self._h, int(paths)), "set_wavefront_capacity")

    def set_tuning(self, wavefront_capacity: int = 0, refill_min: int = 0, traversal_blocks: int = 0,
                   shade_blocks: int = 0):
        """Launch shape (PtTuning): zero capacity / refill_min keep the current value, zero block
        counts restore the occupancy-derived defaults."""
        tu = abi.PtTuning(int(wavefront_capacity), int(refill_min), int(traversal_blocks), int(shade_blocks))
        self._check(self._lib.pt_hip_set_tuning(self._h, C.byref(tu)), "set_tuning")

    def tuning(self) -> dict:
        tu = abi.PtTuning()
        self._check(self._lib.pt_hip_get_tuning(self._h, C.byref(tu)), "get_tuning")
        return tu.as_dict()

    def set_profiling(self, enabled: bool):
        self._check(self._lib.pt_hip_set_profiling(self._h, int(enabled)), "set_profiling")

    def reset_kernel_stats(self):
        self._check(self._lib.pt_hip_reset_kernel_stats(self._h), "reset_kernel_stats")

    def kernel_stats(self) -> dict:
        k = abi.PtKernelStats()
        self._check(self._lib.pt_hip_read_kernel_stats(self._h, C.byref(k)), "read_kernel_stats")
        out = {name: {"total_ms": k.total_ms[i], "launches": int(k.launches[i]), "items": int(k.items[i])}
               for i, name in enumerate(abi.KERNELS)}
        for name in ("inner_node_fetches", "shadow_node_visits", "shadow_triangle_tests", "shadow_inner_node_fetches",
                     "shadow_leaf_visits", "shadow_rays"):
            out[name] = int(getattr(k, name))
        return out

    # -- IRayAccelerator::intersect, batched
    def trace_rays(self, rays, any_hit: bool = False, tmax=None, noalloc: bool = False):
        """Closest-hit (or any-hit below tmax) for an (N, 6) array of rays.

        Accepts numpy arrays or torch device tensors; returns (t, primitive_index, counters)
        as numpy arrays (t = inf and primitive_index = -1 on a miss). noalloc: through
        pt_hip_trace_rays_noalloc (the library's static scratch instead of its own allocations)."""
        import torch

        dev = torch.device("cuda", self.device)
        r = torch.as_tensor(np.ascontiguousarray(rays, np.float32) if isinstance(rays, np.ndarray) else rays)
        r = r.to(dev, torch.float32).contiguous()
        n = r.shape[0]
        tm = None
        if tmax is not None:
            tm = torch.as_tensor(np.ascontiguousarray(tmax, np.float32) if isinstance(tmax, np.ndarray) else tmax)
            tm = tm.to(dev, torch.float32).contiguous()
        out_t = torch.empty(n, dtype=torch.float32, device=dev)
        out_p = torch.empty(n, dtype=torch.int32, device=dev)
        torch.cuda.synchronize(dev)
        c = abi.PtSampleCounters()
        fn = self._lib.pt_hip_trace_rays_noalloc if noalloc else self._lib.pt_hip_trace_rays
        rc = fn(self._h, 1 if any_hit else 0, r.data_ptr(), tm.data_ptr() if tm is not None else None, n,
                out_t.data_ptr(), out_p.data_ptr(), C.byref(c))
        self._check(rc, "trace_rays_noalloc" if noalloc else "trace_rays")
        return out_t.cpu().numpy(), out_p.cpu().numpy(), c.as_dict()

    def intersect(self, rays, any_hit: bool = False, tmax=None):
        """IRayAccelerator::intersect over an (N, 6) numpy array of rays (host memory, staged by the
        library): (RAY_QUERY_HIT_DTYPE records, RayQueryStats as a dict)."""
        r = np.ascontiguousarray(rays, np.float32).reshape(-1, 6)
        tm = None if tmax is None else np.ascontiguousarray(tmax, np.float32)
        hits = np.zeros(len(r), abi.RAY_QUERY_HIT_DTYPE)
        c = abi.PtSampleCounters()
        rc = self._lib.pt_hip_intersect(self._h, 1 if any_hit else 0, r.ctypes.data,
                                        None if tm is None else tm.ctypes.data, len(r), hits.ctypes.data, C.byref(c))
        self._check(rc, "intersect")
        d = c.as_dict()
        return hits, {k: d[k] for k in ("triangle_tests", "triangle_hits", "bvh_node_visits", "bvh_leaf_visits")}


class HipBatch:
    """Many scenes rendered through one scene pool and one wavefront launch sequence per frame
    (pt_batch_*; no counterpart in the reference). Integrator settings are shared; each scene has
    its own image size, spp and BVH parameters."""

    def __init__(self, settings: RenderSettings, device: int = 0, library=None):
        self._lib = abi.load_library(library)
        h = C.c_void_p()
        s = abi.make_settings(settings)
        rc = self._lib.pt_batch_create(device, C.byref(s), C.byref(h))
        if rc != 0:
            raise PathTracerError(Status(rc, "no HIP device available for the batch"), "pt_batch_create")
        self._h = h
        self.settings = settings
        self.scenes: list[tuple[int, int, int]] = []  # (width, height, spp)
        self.device = device

    def close(self):
        if getattr(self, "_h", None):
            self._lib.pt_batch_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def _check(self, rc: int, what: str):
        if rc != 0:
            msg = (self._lib.pt_batch_last_error(self._h) or b"").decode()
            raise PathTracerError(Status(rc, msg), what)

    def add_scene(self, scene: Scene, width: int, height: int, spp: int, builder: str = "sah", leaf_size: int = 4,
                  bucket_count: int = 8, host_build: bool = False) -> int:
        view = abi.SceneView(scene)
        p = abi.PtBvhParams({"sah": 0, "midpoint": 1}[builder], leaf_size, bucket_count,
                            abi.PT_BVH_FLAG_HOST_BUILD if host_build else 0)
        idx = C.c_uint32()
        self._check(self._lib.pt_batch_add_scene(self._h, C.byref(view.view), width, height, spp, C.byref(p),
                                                 C.byref(idx)), "add_scene")
        if view.has_textures:
            self._check(self._lib.pt_batch_load_textures(self._h, idx.value, *view.texture_args()), "load_textures")
        sdf = getattr(scene, "sdf_primitives", None)
        if sdf is not None and len(sdf):
            sdf = np.ascontiguousarray(sdf)
            self._check(self._lib.pt_batch_load_sdf_primitives(self._h, idx.value, sdf.ctypes.data, len(sdf)),
                        "load_sdf_primitives")
        self.scenes.append((width, height, spp))
        return int(idx.value)

    def update_camera(self, scene: int, camera) -> None:
        """The camera of scene or view `scene` (scene.Camera); graph replays see it."""
        c = abi.make_camera(camera)
        self._check(self._lib.pt_batch_update_camera(self._h, scene, C.byref(c)), "update_camera")

    def add_view(self, scene: int, camera, width: int, height: int, spp: int) -> int:
        """Another camera on batch scene `scene`: shares its geometry, materials and textures, has its own
        image and film."""
        cam = abi.make_camera(camera)
        idx = C.c_uint32()
        self._check(self._lib.pt_batch_add_view(self._h, scene, C.byref(cam), width, height, spp, C.byref(idx)),
                    "add_view")
        self.scenes.append((width, height, spp))
        return int(idx.value)

    def build(self) -> dict:
        info = abi.PtBuildInfo()
        self._check(self._lib.pt_batch_build(self._h, C.byref(info)), "build")
        self.build_info = info.as_dict()
        return self.build_info

    @property
    def paths_per_frame(self) -> int:
        return sum(w * h * s for w, h, s in self.scenes)

    def reset_accumulation(self):
        self._check(self._lib.pt_batch_reset_accumulation(self._h), "reset_accumulation")

    def render(self, frame_index: int = 0, use_graph: bool = False):
        self._check(self._lib.pt_batch_render(self._h, frame_index, int(use_graph)), "render")

    def synchronize(self):
        self._check(self._lib.pt_batch_synchronize(self._h), "synchronize")

    def update_scene_delta(self, materials: np.ndarray | None = None, lights: np.ndarray | None = None):
        """Replace the pool's materials and/or lights (every scene's, in the order they were added)."""
        from .scene import LIGHT_DTYPE, MATERIAL_DTYPE

        m = None if materials is None else np.ascontiguousarray(materials, MATERIAL_DTYPE)
        l = None if lights is None else np.ascontiguousarray(lights, LIGHT_DTYPE)
        self._check(self._lib.pt_batch_update_scene_delta(self._h, None if m is None else m.ctypes.data,
                                                          0 if m is None else len(m),
                                                          None if l is None else l.ctypes.data,
                                                          0 if l is None else len(l)), "update_scene_delta")

    def update_texels(self, texels) -> None:
        """Replace every texel of the pool: (N, 3) float32 numpy array or contiguous float32 device tensor."""
        if isinstance(texels, np.ndarray):
            t = np.ascontiguousarray(texels, np.float32).reshape(-1, 3)
            self._check(self._lib.pt_batch_update_texels(self._h, t.ctypes.data, len(t)), "update_texels")
        else:
            import torch

            torch.cuda.synchronize(self.device)
            self._check(self._lib.pt_batch_update_texels(self._h, texels.data_ptr(), texels.numel() // 3),
                        "update_texels")

    def render_adjoint(self, film_grad, grad_albedo, grad_emission, grad_light, frame_index: int = 0,
                       grad_roughness=None, grad_texels=None):
        """Path-replay adjoint of render(frame_index): film_grad is a device tensor of every film's pixels
        (films in batch order, 3 floats each); gradients are pool-ordered device tensors (added to)."""
        import torch

        torch.cuda.synchronize(self.device)
        rc = self._lib.pt_batch_render_adjoint(self._h, frame_index, film_grad.data_ptr(), grad_albedo.data_ptr(),
                                               grad_emission.data_ptr(), _dptr(grad_light), _dptr(grad_roughness),
                                               _dptr(grad_texels))
        self._check(rc, "render_adjoint")
        self.synchronize()

    def films(self) -> list:
        """Every film (rgb sum, count), in batch order."""
        return [self.film(i) for i in range(len(self.scenes))]

    def film(self, scene: int) -> tuple[np.ndarray, np.ndarray]:
        w, h, _ = self.scenes[scene]
        rgb = np.zeros((h, w, 3), np.float32)
        cnt = np.zeros((h, w), np.uint32)
        self._check(self._lib.pt_batch_read_film(self._h, scene, rgb.ctypes.data, cnt.ctypes.data), "read_film")
        return rgb, cnt

    def resolve_hdr(self, scene: int) -> np.ndarray:
        w, h, _ = self.scenes[scene]
        rgb = np.zeros((h, w, 3), np.float32)
        self._check(self._lib.pt_batch_resolve_hdr(self._h, scene, rgb.ctypes.data), "resolve_hdr")
        return rgb

    def resolve_ldr(self, scene: int, settings=None) -> np.ndarray:
        w, h, _ = self.scenes[scene]
        img = np.zeros((h, w, 4), np.uint8)
        rs = abi.make_resolve(settings)
        self._check(self._lib.pt_batch_resolve_ldr(self._h, scene, C.byref(rs), img.ctypes.data), "resolve_ldr")
        return img

    def read_counters(self) -> dict:
        c = abi.PtSampleCounters()
        self._check(self._lib.pt_batch_read_counters(self._h, C.byref(c)), "read_counters")
        return c.as_dict()

    def set_tuning(self, wavefront_capacity: int = 0, refill_min: int = 0, traversal_blocks: int = 0,
                   shade_blocks: int = 0):
        tu = abi.PtTuning(int(wavefront_capacity), int(refill_min), int(traversal_blocks), int(shade_blocks))
        self._check(self._lib.pt_batch_set_tuning(self._h, C.byref(tu)), "set_tuning")

    def tuning(self) -> dict:
        tu = abi.PtTuning()
        self._check(self._lib.pt_batch_get_tuning(self._h, C.byref(tu)), "get_tuning")
        return tu.as_dict()

    def set_profiling(self, enabled: bool):
        self._check(self._lib.pt_batch_set_profiling(self._h, int(enabled)), "set_profiling")

    def reset_kernel_stats(self):
        self._check(self._lib.pt_batch_reset_kernel_stats(self._h), "reset_kernel_stats")

    def kernel_stats(self) -> dict:
        k = abi.PtKernelStats()
        self._check(self._lib.pt_batch_read_kernel_stats(self._h, C.byref(k)), "read_kernel_stats")
        out = {name: {"total_ms": k.total_ms[i], "launches": int(k.launches[i]), "items": int(k.items[i])}
               for i, name in enumerate(abi.KERNELS)}
        for name in ("inner_node_fetches", "shadow_node_visits", "shadow_triangle_tests", "shadow_inner_node_fetches",
                     "shadow_leaf_visits", "shadow_rays"):
            out[name] = int(getattr(k, name))
        return out

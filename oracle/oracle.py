"""ORACLE — test infrastructure only.

ctypes front end of the CPU restatement in pto.c (see pto.h). Only tests/,
__graft_entry__.smoke() and bench.py's cpu_baseline leg import this module, and
only to check or time against the product; nothing in vkpathtracer_amd/ imports it.

Variants (oracle/Makefile):
  "f32"       portable transcendentals — bit-level parity anchor for the GPU path
  "f32_libm"  glibc sinf/cosf/powf, as the reference is built
  "f64"       the same estimator evaluated in double precision
"""
from __future__ import annotations

import ctypes as C
import pathlib
import subprocess

import numpy as np

HERE = pathlib.Path(__file__).resolve().parent
LIB_DIR = HERE / "lib"
VARIANTS = ("f32", "f32_libm", "f64", "f32_avx2", "f32_avx2_libm")


class PtoCamera(C.Structure):
    _fields_ = [("position", C.c_float * 3), ("target", C.c_float * 3), ("up", C.c_float * 3),
                ("fov_deg", C.c_float), ("aperture_radius", C.c_float), ("focus_distance", C.c_float),
                ("iris_blade_count", C.c_uint32), ("iris_rotation_degrees", C.c_float),
                ("iris_roundness", C.c_float), ("anamorphic_squeeze", C.c_float)]


class PtoScene(C.Structure):
    _fields_ = [("vertices", C.c_void_p), ("indices", C.c_void_p), ("local_vertices", C.c_void_p),
                ("local_indices", C.c_void_p), ("instances", C.c_void_p), ("materials", C.c_void_p),
                ("lights", C.c_void_p), ("environment_map", C.c_void_p),
                ("vertex_count", C.c_uint32), ("index_count", C.c_uint32), ("local_vertex_count", C.c_uint32),
                ("local_index_count", C.c_uint32), ("instance_count", C.c_uint32), ("material_count", C.c_uint32),
                ("light_count", C.c_uint32), ("environment_map_width", C.c_uint32),
                ("environment_map_height", C.c_uint32), ("environment_color", C.c_float * 3),
                ("environment_map_scale", C.c_float * 3), ("camera", PtoCamera)]


class PtoSettings(C.Structure):
    _fields_ = [("seed", C.c_uint64), ("width", C.c_uint32), ("height", C.c_uint32), ("spp", C.c_uint32),
                ("max_depth", C.c_uint32), ("enable_nee", C.c_uint32), ("enable_mis", C.c_uint32),
                ("deterministic", C.c_uint32), ("rr_start_depth", C.c_uint32), ("rr_min_survival", C.c_float),
                ("rr_max_survival", C.c_float), ("camera_aperture_radius", C.c_float),
                ("camera_focus_distance", C.c_float)]


COUNTER_FIELDS = ("samples", "rays", "triangle_tests", "sdf_tests", "sdf_steps", "triangle_hits", "sdf_hits",
                  "sdf_misses", "bvh_node_visits", "bvh_leaf_visits", "shadow_tests")


class PtoCounters(C.Structure):
    _fields_ = [(n, C.c_uint64) for n in COUNTER_FIELDS]

    def as_dict(self):
        return {n: int(getattr(self, n)) for n in COUNTER_FIELDS}


SHADOW_CLOSEST, SHADOW_ANYHIT = 0, 1
BUILDER = {"midpoint": 0, "sah": 1}

_P = C.c_void_p
_SIGS = {
    "pto_real_bytes": (C.c_uint32, []),
    "pto_uses_libm": (C.c_uint32, []),
    "pto_sizeof": (C.c_uint32, [C.c_int]),
    "pto_sinf": (C.c_float, [C.c_float]),
    "pto_cosf": (C.c_float, [C.c_float]),
    "pto_powf": (C.c_float, [C.c_float, C.c_float]),
    "pto_atan2f": (C.c_float, [C.c_float, C.c_float]),
    "pto_acosf": (C.c_float, [C.c_float]),
    "pto_splitmix64": (C.c_uint64, [C.c_uint64]),
    "pto_rng_seed": (C.c_uint64, [C.c_uint64] * 7),
    "pto_rng_draws": (None, [C.c_uint64, C.c_uint32, _P, _P]),
    "pto_accel_build": (_P, [C.POINTER(PtoScene), C.c_int, C.c_uint32, C.c_uint32]),
    "pto_accel_free": (None, [_P]),
    "pto_accel_set_textures": (C.c_int, [_P, _P, C.c_uint32, _P, C.c_uint64, _P, C.c_uint32]),
    "pto_texture_sample": (C.c_int, [_P, C.c_uint32, C.c_uint32, C.c_float, C.c_float, _P, _P, _P]),
    "pto_accel_node_count": (C.c_uint32, [_P]),
    "pto_accel_prim_count": (C.c_uint32, [_P]),
    "pto_accel_max_depth": (C.c_uint32, [_P]),
    "pto_accel_export": (C.c_int, [_P, _P, _P, _P, _P, _P]),
    "pto_trace_rays": (C.c_int, [_P, _P, C.c_uint32, C.c_int, _P, _P, _P, C.POINTER(PtoCounters)]),
    "pto_brute_force_rays": (C.c_int, [_P, _P, C.c_uint32, _P, _P]),
    "pto_render_pixels": (C.c_int, [C.POINTER(PtoScene), _P, C.POINTER(PtoSettings), _P, C.c_uint32, C.c_uint32,
                                    C.c_uint64, C.c_int, C.c_int, _P, _P, C.POINTER(PtoCounters)]),
    "pto_render_tile": (C.c_int, [C.POINTER(PtoScene), _P, C.POINTER(PtoSettings), C.c_uint32, C.c_uint32, C.c_uint32,
                                  C.c_uint32, C.c_uint32, C.c_uint32, C.c_uint64, C.c_int, C.c_int, _P, _P,
                                  C.POINTER(PtoCounters)]),
    "pto_accel_refit": (C.c_int, [_P, C.POINTER(PtoScene)]),
    "pto_accel_set_sdfs": (C.c_int, [_P, _P, C.c_uint32]),
    "pto_intersect_scene_rays": (C.c_int, [C.POINTER(PtoScene), _P, _P, C.c_uint32, _P, _P, _P, _P,
                                           C.POINTER(PtoCounters)]),
    "pto_color_temperature_rgb": (None, [C.c_float, _P]),
    "pto_white_balance_scale": (None, [C.c_float, _P]),
    "pto_resolve_hdr": (None, [_P, _P, C.c_uint64, _P]),
}



class PtoResolve(C.Structure):
    _fields_ = [("exposure", C.c_float), ("white_balance", C.c_float * 3), ("inv_gamma", C.c_float),
                ("tone_map", C.c_uint32), ("output_transform", C.c_uint32), ("clamp_output", C.c_uint32)]


_SIGS["pto_resolve_ldr"] = (None, [_P, _P, C.c_uint64, C.POINTER(PtoResolve), _P])

_SIGS["pto_camera_rays"] = (C.c_int, [C.POINTER(PtoScene), C.POINTER(PtoSettings), _P, C.c_uint32, C.c_uint32,
                                       C.c_uint64, _P, _P])
_SIGS["pto_camera_dir"] = (C.c_int, [C.POINTER(PtoScene), C.POINTER(PtoSettings), C.c_uint32, C.c_uint32, C.c_float,
                                     C.c_float, _P])
_SIGS["pto_hemisphere_samples"] = (C.c_int, [_P, C.c_uint64, C.c_uint32, _P])
_SIGS["pto_light_pdf"] = (C.c_double, [C.POINTER(PtoScene), _P, _P])

_SIGS["pto_det_rejections"] = (C.c_int, [_P, _P, C.c_uint32, C.c_int, _P])
_SIGS["pto_render_bands"] = (C.c_int, [C.POINTER(PtoScene), _P, C.POINTER(PtoSettings), C.c_uint32, C.c_uint32,
                                       C.c_uint32, C.c_uint32, C.c_uint32, C.c_uint64, C.c_int, C.c_int, _P, _P,
                                       C.POINTER(PtoCounters)])

_libs: dict[str, C.CDLL] = {}


def build(quiet: bool = True) -> None:
    """Compile the three oracle variants with gcc (make -C oracle)."""
    subprocess.run(["make", "-s", "-C", str(HERE), "-j3"], check=True,
                   stdout=subprocess.DEVNULL if quiet else None)


def lib(variant: str = "f32") -> C.CDLL:
    if variant not in _libs:
        path = LIB_DIR / f"liboracle_{variant}.so"
        if not path.exists():
            build()
        L = C.CDLL(str(path))
        for name, (res, args) in _SIGS.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _libs[variant] = L
    return _libs[variant]


def _ptr(a):
    return None if a is None or a.size == 0 else a.ctypes.data


class OracleScene:
    """PtoScene over a scene object exposing the snapshot arrays (vkpathtracer_amd.scene.Scene layout)."""

    def __init__(self, sc):
        self.sc = sc
        s = PtoScene()
        s.vertices, s.indices = _ptr(sc.vertices), _ptr(sc.indices)
        s.local_vertices, s.local_indices = _ptr(sc.local_vertices), _ptr(sc.local_indices)
        s.instances, s.materials, s.lights = _ptr(sc.instances), _ptr(sc.materials), _ptr(sc.lights)
        env = getattr(sc, "environment_map", None)
        s.environment_map = _ptr(env)
        if env is not None:
            s.environment_map_height, s.environment_map_width = env.shape[0], env.shape[1]
        s.environment_map_scale[:] = list(getattr(sc, "environment_map_scale", (0.0, 0.0, 0.0)))
        s.vertex_count, s.index_count = len(sc.vertices), len(sc.indices)
        s.local_vertex_count, s.local_index_count = len(sc.local_vertices), len(sc.local_indices)
        s.instance_count, s.material_count, s.light_count = len(sc.instances), len(sc.materials), len(sc.lights)
        s.environment_color[:] = list(sc.environment_color)
        c = sc.camera
        cam = PtoCamera()
        cam.position[:], cam.target[:], cam.up[:] = list(c.position), list(c.target), list(c.up)
        cam.fov_deg, cam.aperture_radius, cam.focus_distance = c.fov_deg, c.aperture_radius, c.focus_distance
        cam.iris_blade_count, cam.iris_rotation_degrees = c.iris_blade_count, c.iris_rotation_degrees
        cam.iris_roundness, cam.anamorphic_squeeze = c.iris_roundness, c.anamorphic_squeeze
        s.camera = cam
        self.s = s


def settings_struct(st) -> PtoSettings:
    o = PtoSettings()
    o.seed = st.seed & 0xFFFFFFFFFFFFFFFF
    o.width, o.height, o.spp, o.max_depth = st.width, st.height, st.spp, st.max_depth
    o.enable_nee, o.enable_mis, o.deterministic = int(st.enable_nee), int(st.enable_mis), int(st.deterministic)
    o.rr_start_depth, o.rr_min_survival, o.rr_max_survival = st.rr_start_depth, st.rr_min_survival, st.rr_max_survival
    o.camera_aperture_radius, o.camera_focus_distance = st.camera_aperture_radius, st.camera_focus_distance
    return o


class Oracle:
    """One scene + acceleration structure in one oracle variant."""

    def __init__(self, scene, builder: str = "sah", leaf_size: int = 4, bucket_count: int = 8, variant: str = "f32"):
        self.L = lib(variant)
        self.variant = variant
        self.real = np.float64 if self.L.pto_real_bytes() == 8 else np.float32
        self.scene = OracleScene(scene)
        self.accel = self.L.pto_accel_build(C.byref(self.scene.s), BUILDER[builder], leaf_size, bucket_count)
        if not self.accel:
            raise RuntimeError("oracle accel build failed")
        self.texel_count = 0
        sdf = getattr(scene, "sdf_primitives", None)
        if sdf is not None and len(sdf):
            self._sdf = np.ascontiguousarray(sdf)
            if self.L.pto_accel_set_sdfs(self.accel, self._sdf.ctypes.data, len(self._sdf)) != 0:
                raise RuntimeError("oracle SDF primitives rejected")
        tc = getattr(scene, "texcoords", None)
        if tc is not None:
            texels, descs = scene.texture_arrays()
            tc = np.ascontiguousarray(tc, np.float32)
            rc = self.L.pto_accel_set_textures(self.accel, _ptr(tc), len(tc), _ptr(texels), len(texels), _ptr(descs),
                                               len(descs))
            if rc != 0:
                raise RuntimeError("oracle textures rejected")
            self.texel_count = len(texels)

    def __del__(self):
        if getattr(self, "accel", None):
            self.L.pto_accel_free(self.accel)
            self.accel = None

    def intersect_scene(self, rays):
        """intersect_scene for (N, 6) rays: (t, normal (N, 3), material, is_sdf, counters)."""
        rays = np.ascontiguousarray(rays, np.float32)
        n = len(rays)
        t = np.zeros(n, np.float32)
        nrm = np.zeros((n, 3), np.float32)
        mat = np.zeros(n, np.uint32)
        sdf = np.zeros(n, np.uint8)
        c = PtoCounters()
        self.L.pto_intersect_scene_rays(C.byref(self.scene.s), self.accel, rays.ctypes.data, n, t.ctypes.data,
                                        nrm.ctypes.data, mat.ctypes.data, sdf.ctypes.data, C.byref(c))
        return t, nrm, mat, sdf.astype(bool), c.as_dict()

    def refit(self, scene) -> None:
        """CpuBvhAccelerator's dynamic-instance refit (pto_accel_refit): `scene` is this scene with
        updated instance transforms; primitives and node boxes follow them, the topology stays."""
        self.scene = OracleScene(scene)
        rc = self.L.pto_accel_refit(self.accel, C.byref(self.scene.s))
        if rc != 0:
            raise RuntimeError(f"oracle refit failed ({rc})")

    def texture_sample(self, tex: int, prim: int, bu: float, bv: float):
        """pto_texture_sample: (rgb (3,), texel indices (4,), weights (4,)) of one bilinear tap."""
        rgb = np.zeros(3, np.float64)
        idx = np.zeros(4, np.uint32)
        w = np.zeros(4, np.float64)
        if self.L.pto_texture_sample(self.accel, tex, prim, bu, bv, rgb.ctypes.data, idx.ctypes.data, w.ctypes.data):
            raise ValueError("no such texture or primitive")
        return rgb, idx, w

    @property
    def prim_count(self):
        return self.L.pto_accel_prim_count(self.accel)

    @property
    def node_count(self):
        return self.L.pto_accel_node_count(self.accel)

    @property
    def max_depth(self):
        return self.L.pto_accel_max_depth(self.accel)

    def export(self):
        n, p = self.node_count, self.prim_count
        bounds = np.zeros((n, 6), self.real)
        meta = np.zeros((n, 4), np.int32)
        order = np.zeros(p, np.uint32)
        geo = np.zeros((p, 12), self.real)
        mat = np.zeros(p, np.uint32)
        self.L.pto_accel_export(self.accel, bounds.ctypes.data, meta.ctypes.data, order.ctypes.data,
                                geo.ctypes.data, mat.ctypes.data)
        return dict(bounds=bounds, meta=meta, order=order, geo=geo, material=mat)

    def trace_rays(self, rays: np.ndarray, any_hit: bool = False, tmax: np.ndarray | None = None):
        rays = np.ascontiguousarray(rays, np.float32)
        n = len(rays)
        t = np.zeros(n, np.float32)
        p = np.zeros(n, np.int32)
        c = PtoCounters()
        tm = None if tmax is None else np.ascontiguousarray(tmax, np.float32)
        self.L.pto_trace_rays(self.accel, rays.ctypes.data, n, int(any_hit), None if tm is None else tm.ctypes.data,
                              t.ctypes.data, p.ctypes.data, C.byref(c))
        return t, p, c.as_dict()

    def brute_force(self, rays: np.ndarray):
        rays = np.ascontiguousarray(rays, np.float32)
        n = len(rays)
        t = np.zeros(n, np.float32)
        p = np.zeros(n, np.int32)
        self.L.pto_brute_force_rays(self.accel, rays.ctypes.data, n, t.ctypes.data, p.ctypes.data)
        return t, p

    def det_rejections(self, rays: np.ndarray, any_hit: bool = False) -> tuple[int, int]:
        """(triangle tests, tests rejected by |det| < 1e-4) of closest- (or any-) hit queries."""
        rays = np.ascontiguousarray(rays, np.float32)
        out = np.zeros(2, np.uint64)
        self.L.pto_det_rejections(self.accel, rays.ctypes.data, len(rays), int(any_hit), out.ctypes.data)
        return int(out[0]), int(out[1])

    def render_bands(self, settings, band_rows: int = 16, band_first: int = 0, band_stride: int = 1,
                     sample_begin: int = 0, sample_count: int | None = None, frame_index: int = 0, threads: int = 0,
                     shadow: int = SHADOW_ANYHIT):
        """render_tile over whole-width bands of band_rows rows (bands band_first + k * band_stride), one
        band per job per sample: (film (H, W, 3), counts (H, W), counters)."""
        W, H = settings.width, settings.height
        film = np.zeros((H, W, 3), self.real)
        cnt = np.zeros((H, W), np.uint32)
        c = PtoCounters()
        st = settings_struct(settings)
        count = settings.spp if sample_count is None else sample_count
        rc = self.L.pto_render_bands(C.byref(self.scene.s), self.accel, C.byref(st), band_rows, band_first, band_stride,
                                     sample_begin, count, frame_index, threads, shadow, film.ctypes.data,
                                     cnt.ctypes.data, C.byref(c))
        if rc != 0:
            raise RuntimeError(f"oracle render_bands failed ({rc})")
        return film, cnt, c.as_dict()

    def camera_rays(self, settings, pixels: np.ndarray, sample_index: int, frame_index: int = 0):
        """The primary rays shade_pixel makes for one sample of each pixel: (rays (N, 6), jitter (N, 2))."""
        px = np.ascontiguousarray(pixels, np.uint32)
        rays = np.zeros((len(px), 6), np.float32)
        jit = np.zeros((len(px), 2), np.float32)
        st = settings_struct(settings)
        if self.L.pto_camera_rays(C.byref(self.scene.s), C.byref(st), px.ctypes.data, len(px), sample_index,
                                  frame_index, rays.ctypes.data, jit.ctypes.data) != 0:
            raise RuntimeError("oracle camera_rays failed")
        return rays, jit

    def camera_dir(self, settings, x: int, y: int, u: float, v: float) -> np.ndarray:
        """The pinhole direction through pixel (x, y) at jitter (u, v)."""
        out = np.zeros(3, np.float32)
        st = settings_struct(settings)
        if self.L.pto_camera_dir(C.byref(self.scene.s), C.byref(st), x, y, u, v, out.ctypes.data) != 0:
            raise RuntimeError("oracle camera_dir failed")
        return out

    def light_pdf(self, pos, direction) -> float:
        """light_pdf_for_direction at pos for a direction."""
        p = np.ascontiguousarray(pos, np.float32)
        d = np.ascontiguousarray(direction, np.float32)
        return float(self.L.pto_light_pdf(C.byref(self.scene.s), p.ctypes.data, d.ctypes.data))

    def render_pixels(self, settings, pixels: np.ndarray, sample_index: int, frame_index: int = 0, threads: int = 0,
                      shadow: int = SHADOW_ANYHIT):
        px = np.ascontiguousarray(pixels, np.uint32)
        n = len(px)
        rgb = np.zeros((n, 3), self.real)
        valid = np.zeros(n, np.uint8)
        c = PtoCounters()
        st = settings_struct(settings)
        rc = self.L.pto_render_pixels(C.byref(self.scene.s), self.accel, C.byref(st), px.ctypes.data, n, sample_index,
                                      frame_index, threads, shadow, rgb.ctypes.data, valid.ctypes.data, C.byref(c))
        if rc != 0:
            raise RuntimeError(f"oracle render_pixels failed ({rc})")
        return rgb, valid.astype(bool), c.as_dict()

    def render(self, settings, sample_begin: int = 0, sample_count: int | None = None, frame_index: int = 0,
               threads: int = 0, shadow: int = SHADOW_ANYHIT, tile=None):
        """Film (sum, count) after render_tile for each sample in order. tile = (x, y, w, h) or full frame."""
        W, H = settings.width, settings.height
        film = np.zeros((H, W, 3), self.real)
        cnt = np.zeros((H, W), np.uint32)
        c = PtoCounters()
        st = settings_struct(settings)
        x, y, w, h = tile if tile is not None else (0, 0, W, H)
        count = settings.spp if sample_count is None else sample_count
        rc = self.L.pto_render_tile(C.byref(self.scene.s), self.accel, C.byref(st), x, y, w, h, sample_begin, count,
                                    frame_index, threads, shadow, film.ctypes.data, cnt.ctypes.data, C.byref(c))
        if rc != 0:
            raise RuntimeError(f"oracle render failed ({rc})")
        return film, cnt, c.as_dict()


_SIGS_ADJ = (C.c_int, [C.POINTER(PtoScene), _P, C.POINTER(PtoSettings), _P, C.c_uint32, C.c_uint32, C.c_uint64, C.c_int,
                       C.c_int, _P, _P, _P, _P, _P, _P, _P])


def adjoint_gradients(orc: Oracle, settings, pixels, sample_index, delta, frame_index=0, threads=0,
                      shadow=SHADOW_ANYHIT) -> dict:
    """Path-replay adjoint (pto_adjoint_pixels) of one sample per listed pixel: a dict of fp64
    gradients albedo (M,3), emission (M,3), light (Lt,3), roughness (M,), texture (N,3) per texel of the
    scene's textures, and the forward rgb."""
    fn = orc.L.pto_adjoint_pixels
    fn.restype, fn.argtypes = _SIGS_ADJ
    px = np.ascontiguousarray(pixels, np.uint32)
    d = np.ascontiguousarray(delta, np.float32).reshape(-1, 3)
    nm, nl = orc.scene.s.material_count, orc.scene.s.light_count
    ga = np.zeros((nm, 3), np.float64)
    ge = np.zeros((nm, 3), np.float64)
    gl = np.zeros((max(nl, 1), 3), np.float64)
    gr = np.zeros(max(nm, 1), np.float64)
    gt = np.zeros((max(orc.texel_count, 1), 3), np.float64)
    rgb = np.zeros((len(px), 3), orc.real)
    st = settings_struct(settings)
    rc = fn(C.byref(orc.scene.s), orc.accel, C.byref(st), px.ctypes.data, len(px), sample_index, frame_index, threads,
            shadow, d.ctypes.data, ga.ctypes.data, ge.ctypes.data, gl.ctypes.data, gr.ctypes.data,
            gt.ctypes.data if orc.texel_count else None, rgb.ctypes.data)
    if rc != 0:
        raise RuntimeError(f"oracle adjoint failed ({rc})")
    return {"albedo": ga, "emission": ge, "light": gl[:nl], "roughness": gr[:nm], "texture": gt[:orc.texel_count],
            "rgb": rgb}


def adjoint_pixels(orc: Oracle, settings, pixels, sample_index, delta, frame_index=0, threads=0,
                   shadow=SHADOW_ANYHIT):
    """(g_albedo (M,3), g_emission (M,3), g_light (Lt,3), rgb) of adjoint_gradients."""
    g = adjoint_gradients(orc, settings, pixels, sample_index, delta, frame_index, threads, shadow)
    return g["albedo"], g["emission"], g["light"], g["rgb"]


def phong_samples(exponent: float, refl, normal, ps: float, state: int, count: int, variant: str = "f64"):
    """pto_phong_samples: (directions (count, 3) float32, score (count,) float64)."""
    L = lib(variant)
    fn = L.pto_phong_samples
    fn.restype = C.c_int
    fn.argtypes = [C.c_float, _P, _P, C.c_float, C.c_uint64, C.c_uint32, _P, _P]
    r = np.ascontiguousarray(refl, np.float32)
    n = np.ascontiguousarray(normal, np.float32)
    d = np.zeros((count, 3), np.float32)
    sc = np.zeros(count, np.float64)
    fn(exponent, r.ctypes.data, n.ctypes.data, ps, state & 0xFFFFFFFFFFFFFFFF, count, d.ctypes.data, sc.ctypes.data)
    return d, sc


def hemisphere_samples(normal, state: int, count: int, variant: str = "f32") -> np.ndarray:
    """pto_hemisphere_samples: sample_hemisphere directions (count, 3) about `normal`."""
    n = np.ascontiguousarray(normal, np.float32)
    d = np.zeros((count, 3), np.float32)
    lib(variant).pto_hemisphere_samples(n.ctypes.data, state & 0xFFFFFFFFFFFFFFFF, count, d.ctypes.data)
    return d


# ------------------------------------------------------------------ film resolve
TONE_MAPS = ("linear", "reinhard", "filmic_approx", "aces_approx")


def color_temperature_rgb(kelvin: float, variant: str = "f32") -> np.ndarray:
    out = np.zeros(3, np.float32)
    lib(variant).pto_color_temperature_rgb(kelvin, out.ctypes.data)
    return out


def white_balance_scale(kelvin: float, variant: str = "f32") -> np.ndarray:
    out = np.zeros(3, np.float32)
    lib(variant).pto_white_balance_scale(kelvin, out.ctypes.data)
    return out


def resolve_hdr(rgb_sum, count, variant: str = "f32") -> np.ndarray:
    """FilmBuffer::resolve_hdr of a film (rgb sums (H, W, 3), counts (H, W))."""
    s = np.ascontiguousarray(rgb_sum, np.float32)
    c = np.ascontiguousarray(count, np.uint32)
    out = np.zeros(s.shape, np.float32)
    lib(variant).pto_resolve_hdr(s.ctypes.data, c.ctypes.data, c.size, out.ctypes.data)
    return out


def resolve_ldr(rgb_sum, count, exposure=1.0, white_balance_kelvin=6500.0, tone_map="linear",
                output_transform="gamma", gamma=2.2, clamp_output=True, variant: str = "f32") -> np.ndarray:
    """ApplyFilmResolve of resolve_hdr's image: (H, W, 4) u8. The settings are FilmResolveSettings'
    fields; the white-balance scale and inverse gamma are formed here as the reference's host does."""
    s = np.ascontiguousarray(rgb_sum, np.float32)
    c = np.ascontiguousarray(count, np.uint32)
    rs = PtoResolve()
    rs.exposure = exposure
    rs.white_balance[:] = list(white_balance_scale(white_balance_kelvin, variant))
    rs.inv_gamma = float(np.float32(1.0) / max(np.float32(0.01), np.float32(gamma)))
    rs.tone_map = TONE_MAPS.index(tone_map)
    rs.output_transform = 0 if output_transform == "gamma" else 1
    rs.clamp_output = 1 if clamp_output else 0
    out = np.zeros(c.shape + (4,), np.uint8)
    lib(variant).pto_resolve_ldr(s.ctypes.data, c.ctypes.data, c.size, C.byref(rs), out.ctypes.data)
    return out

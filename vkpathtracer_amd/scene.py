"""Scene snapshots and the procedural workloads the path is measured on.

A :class:`Scene` holds the flattened arrays of the reference's
``PathTracerSceneSnapshot`` (src/pathtracer/PathTracer.h:358-409) as numpy
arrays whose record layouts match ``include/vkpt_hip.h`` byte for byte.

Generators (all seeded, all fp32):

* :func:`cornell_box` — the 34-triangle box with two rotated blocks and a
  0.25-side ceiling light (the small CPU-runnable workload).
* :func:`procedural_scene` — a displaced icosphere plus instanced UV spheres
  and quad lights (the 1M-triangle workload at its default sizes).
"""
from __future__ import annotations

import dataclasses
import math

import numpy as np

MATERIAL_DTYPE = np.dtype([
    ("albedo", "<f4", (3,)), ("emissive", "<f4", (3,)),
    ("roughness", "<f4"), ("metallic", "<f4"), ("ior", "<f4"), ("transmission", "<f4"),
    ("clearcoat", "<f4"), ("sheen", "<f4"), ("anisotropy", "<f4"), ("alpha", "<f4"),
    ("material_model", "<u4"), ("material_effect", "<u4"), ("material_flags", "<u4"),
    ("base_color_texture_index", "<u4"), ("normal_texture_index", "<u4"), ("pad0", "<u4"),
])
LIGHT_DTYPE = np.dtype([
    ("position", "<f4", (3,)), ("color", "<f4", (3,)), ("intensity", "<f4"), ("radius", "<f4"),
    ("direction", "<f4", (3,)), ("spot_inner_cos", "<f4"), ("spot_outer_cos", "<f4"), ("pad", "<f4", (3,)),
])
INSTANCE_DTYPE = np.dtype([
    ("entity_id", "<u8"), ("geometry_id", "<u4"), ("first_triangle", "<u4"), ("triangle_count", "<u4"),
    ("material_index", "<u4"), ("flags", "<u4"), ("transform_revision", "<u4"),
    ("local_first_vertex", "<u4"), ("local_vertex_count", "<u4"), ("local_first_index", "<u4"),
    ("local_index_count", "<u4"), ("translation", "<f4", (3,)), ("rotation", "<f4", (4,)),
    ("scale", "<f4", (3,)), ("pad", "<u4", (2,)),
])
# RTSdfPrimitive (PathTracer.h:330-339); shape: SdfShape (PathTracer.h:232-240)
SDF_DTYPE = np.dtype([
    ("shape", "<u4"), ("position", "<f4", (3,)), ("scale", "<f4", (3,)), ("rotation", "<f4", (3,)),
    ("material_index", "<u4"), ("radius", "<f4"), ("param_a", "<f4"), ("param_b", "<f4"),
])
SDF_SHAPES = ("sphere", "box", "rounded_box", "plane", "torus", "capsule", "unknown")
assert MATERIAL_DTYPE.itemsize == 80 and LIGHT_DTYPE.itemsize == 64 and INSTANCE_DTYPE.itemsize == 96
assert SDF_DTYPE.itemsize == 56

# RTInstance flag (PathTracer.h:266): vertices come from local_* arrays through the TRS transform.
INSTANCE_FLAG_DYNAMIC_TRANSFORM = 1

# material_model values used on this path (SceneConversion.cpp:518-560).
MODEL_LAMBERT, MODEL_EMISSIVE, MODEL_MIRROR, MODEL_GLOSSY = 0, 1, 2, 3


def default_material() -> np.ndarray:
    """RTMaterial{} defaults (PathTracer.h:242-261)."""
    m = np.zeros((), MATERIAL_DTYPE)
    m["albedo"] = 1.0
    m["roughness"] = 1.0
    m["ior"] = 1.5
    m["alpha"] = 1.0
    m["base_color_texture_index"] = 0xFFFFFFFF
    m["normal_texture_index"] = 0xFFFFFFFF
    return m


def make_materials(n: int) -> np.ndarray:
    out = np.zeros(n, MATERIAL_DTYPE)
    out[:] = default_material()
    return out


def make_sdf(shape: str, position, scale=(1.0, 1.0, 1.0), rotation=(0.0, 0.0, 0.0), material_index: int = 0,
             radius: float = 0.5, param_a: float = 0.0, param_b: float = 0.0) -> np.ndarray:
    """An RTSdfPrimitive record (PathTracer.h:330-339); rotation: Euler angles in radians."""
    p = np.zeros((), SDF_DTYPE)
    p["shape"] = SDF_SHAPES.index(shape)
    p["position"], p["scale"], p["rotation"] = position, scale, rotation
    p["material_index"], p["radius"], p["param_a"], p["param_b"] = material_index, radius, param_a, param_b
    return p


def make_light(position, color, intensity, radius=0.0) -> np.ndarray:
    """RTHitLight with the reference's defaults for direction and spot cone (PathTracer.h:341-349)."""
    l = np.zeros((), LIGHT_DTYPE)
    l["position"] = position
    l["color"] = color
    l["intensity"] = intensity
    l["radius"] = radius
    l["direction"] = (0.0, -1.0, 0.0)
    l["spot_inner_cos"] = -1.0
    l["spot_outer_cos"] = -1.0
    return l


@dataclasses.dataclass
class Camera:
    position: tuple = (0.0, 0.0, 0.0)
    target: tuple = (0.0, 0.0, -1.0)
    up: tuple = (0.0, 1.0, 0.0)
    fov_deg: float = 60.0
    aperture_radius: float = 0.0
    focus_distance: float = 0.0
    iris_blade_count: int = 0
    iris_rotation_degrees: float = 0.0
    iris_roundness: float = 1.0
    anamorphic_squeeze: float = 1.0


@dataclasses.dataclass
class Scene:
    vertices: np.ndarray                      # (V, 3) f32
    indices: np.ndarray                       # (3T,) u32
    instances: np.ndarray                     # INSTANCE_DTYPE
    materials: np.ndarray                     # MATERIAL_DTYPE
    lights: np.ndarray                        # LIGHT_DTYPE
    camera: Camera
    local_vertices: np.ndarray = dataclasses.field(default_factory=lambda: np.zeros((0, 3), np.float32))
    local_indices: np.ndarray = dataclasses.field(default_factory=lambda: np.zeros(0, np.uint32))
    environment_color: tuple = (0.0, 0.0, 0.0)
    name: str = "scene"
    environment_map: np.ndarray | None = None  # (H, W, 3) f32 equirectangular, or None
    environment_map_scale: tuple = (0.0, 0.0, 0.0)
    # albedo textures (no counterpart in the reference's CPU path): per-vertex (u, v) of `vertices`
    # and (H, W, 3) f32 images, rows in increasing v; materials name them by base_color_texture_index
    texcoords: np.ndarray | None = None
    textures: list = dataclasses.field(default_factory=list)
    # SDF primitives intersect_scene tests after the triangles (SDF_DTYPE)
    sdf_primitives: np.ndarray = dataclasses.field(default_factory=lambda: np.zeros(0, SDF_DTYPE))

    def __post_init__(self):
        if self.environment_map is not None:
            self.environment_map = np.ascontiguousarray(self.environment_map, np.float32).reshape(
                self.environment_map.shape[0], self.environment_map.shape[1], 3)
        self.vertices = np.ascontiguousarray(self.vertices, np.float32).reshape(-1, 3)
        self.indices = np.ascontiguousarray(self.indices, np.uint32).reshape(-1)
        self.local_vertices = np.ascontiguousarray(self.local_vertices, np.float32).reshape(-1, 3)
        self.local_indices = np.ascontiguousarray(self.local_indices, np.uint32).reshape(-1)
        self.instances = np.ascontiguousarray(self.instances, INSTANCE_DTYPE).reshape(-1)
        self.materials = np.ascontiguousarray(self.materials, MATERIAL_DTYPE).reshape(-1)
        self.lights = np.ascontiguousarray(self.lights, LIGHT_DTYPE).reshape(-1)
        if self.texcoords is not None:
            self.texcoords = np.ascontiguousarray(self.texcoords, np.float32).reshape(-1, 2)
        self.textures = [np.ascontiguousarray(t, np.float32).reshape(t.shape[0], t.shape[1], 3) for t in self.textures]
        self.sdf_primitives = np.ascontiguousarray(self.sdf_primitives, SDF_DTYPE).reshape(-1)

    def texture_arrays(self) -> tuple[np.ndarray, np.ndarray]:
        """All textures as one (N, 3) texel array and (K, 4) u32 records (offset, width, height, 0)."""
        descs = np.zeros((len(self.textures), 4), np.uint32)
        off = 0
        for i, t in enumerate(self.textures):
            descs[i] = (off, t.shape[1], t.shape[0], 0)
            off += t.shape[0] * t.shape[1]
        texels = (np.concatenate([t.reshape(-1, 3) for t in self.textures]) if self.textures
                  else np.zeros((0, 3), np.float32))
        return np.ascontiguousarray(texels, np.float32), descs

    def save_snapshot(self, path) -> None:
        """Write the scene as a binary snapshot ("VKPTSNAP"), the file format the C++ host
        (vkpathtracer_amd/host, pt_hip_bench --scene) reads: magic, a header of counts, environment
        colour and scale, the camera record, then every array in the C ABI's record layouts. A scene
        with SDF primitives is written as version 2: version 1 followed by a u32 count and the
        PtSdfPrimitive records; without them the file is version 1."""
        import struct

        env = self.environment_map
        ew, eh = (0, 0) if env is None else (env.shape[1], env.shape[0])
        texels, descs = self.texture_arrays()
        tc = self.texcoords if self.texcoords is not None else np.zeros((0, 2), np.float32)
        sdf = self.sdf_primitives if self.sdf_primitives is not None else np.zeros(0, SDF_DTYPE)
        version = 2 if len(sdf) else 1
        c = self.camera
        cam = struct.pack("<12fI3f", *c.position, *c.target, *c.up, c.fov_deg, c.aperture_radius, c.focus_distance,
                          c.iris_blade_count, c.iris_rotation_degrees, c.iris_roundness, c.anamorphic_squeeze)
        head = struct.pack("<12IQ6f", version, len(self.vertices), len(self.indices), len(self.local_vertices),
                           len(self.local_indices), len(self.instances), len(self.materials), len(self.lights), ew, eh,
                           len(tc), len(descs), len(texels), *self.environment_color, *self.environment_map_scale)
        with open(path, "wb") as f:
            f.write(b"VKPTSNAP" + head + cam)
            for a in (self.vertices, self.indices, self.local_vertices, self.local_indices, self.instances,
                      self.materials, self.lights, env if env is not None else np.zeros(0, np.float32), tc, texels,
                      descs):
                f.write(np.ascontiguousarray(a).tobytes())
            if version == 2:
                f.write(struct.pack("<I", len(sdf)) + np.ascontiguousarray(sdf, SDF_DTYPE).tobytes())

    @property
    def triangle_count(self) -> int:
        return int(self.instances["triangle_count"].astype(np.int64).sum())


@dataclasses.dataclass
class RenderSettings:
    """RenderSettings fields the hot path reads (PathTracer.h:134-178); ptbench defaults for seed."""
    width: int = 256
    height: int = 256
    spp: int = 8
    max_depth: int = 6
    seed: int = 0xBADC0FFEE
    enable_nee: bool = True
    enable_mis: bool = True
    deterministic: bool = True
    rr_start_depth: int = 3
    rr_min_survival: float = 0.1
    rr_max_survival: float = 0.99
    camera_aperture_radius: float = 0.0
    camera_focus_distance: float = 0.0


@dataclasses.dataclass
class InstanceTransformUpdate:
    """RTInstanceTransformUpdate (PathTracer.h:298-306): a new TRS for the instance at instance_index,
    or, when that index is out of range, for the instance whose entity_id matches."""
    translation: tuple = (0.0, 0.0, 0.0)
    rotation: tuple = (0.0, 0.0, 0.0, 1.0)
    scale: tuple = (1.0, 1.0, 1.0)
    instance_index: int = 0xFFFFFFFF
    entity_id: int = 0
    flags: int = 0
    transform_revision: int = 0


def resolve_instance_index(scene: "Scene", u: InstanceTransformUpdate) -> int | None:
    """resolve_instance_index (CpuBvhAccelerator.cpp:507-525)."""
    i = u.instance_index
    if i >= len(scene.instances) and u.entity_id != 0:
        hits = np.nonzero(scene.instances["entity_id"] == np.uint64(u.entity_id))[0]
        if len(hits):
            i = int(hits[0])
    return i if i < len(scene.instances) else None


def with_instance_transforms(scene: "Scene", updates) -> "Scene":
    """ApplyInstanceTransformUpdates in its metadata mode (PathTracer.cpp:364-390): a copy of the scene
    whose instance records carry the updates' TRS, flags (or-ed in) and revisions. Dynamic instances
    take their triangles from local geometry through these transforms."""
    inst = scene.instances.copy()
    probe = dataclasses.replace(scene, instances=inst)
    for u in updates:
        i = resolve_instance_index(probe, u)
        if i is None:
            continue
        inst[i]["translation"] = u.translation
        inst[i]["rotation"] = u.rotation
        inst[i]["scale"] = u.scale
        inst[i]["flags"] |= np.uint32(u.flags)
        if u.transform_revision:
            inst[i]["transform_revision"] = u.transform_revision
    return dataclasses.replace(scene, instances=inst)


TONE_MAPS = ("linear", "reinhard", "filmic_approx", "aces_approx")  # ToneMapMode, PathTracer.h:55-60
OUTPUT_TRANSFORMS = ("gamma", "linear")                              # OutputTransformMode, PathTracer.h:62-65


@dataclasses.dataclass
class FilmResolve:
    """FilmResolveSettings (PathTracer.h:67-74): the LDR resolve's exposure, white balance, tone map,
    output transform and clamp."""
    exposure: float = 1.0
    white_balance_kelvin: float = 6500.0
    tone_map: str = "linear"
    output_transform: str = "gamma"
    gamma: float = 2.2
    clamp_output: bool = True


def camera_adjusted_resolve(base: FilmResolve, f_stop: float = 0.0, shutter_seconds: float = 0.0166666675,
                            iso: float = 100.0, exposure_compensation: float = 0.0,
                            white_balance_kelvin: float = 6500.0) -> FilmResolve:
    """CameraAdjustedFilmResolveSettings (FilmBuffer.cpp:226-253) in float32: the camera's f-stop,
    shutter and ISO relative to f/2.8, 1/60 s, ISO 100 scale the exposure (when all three are positive),
    exposure compensation multiplies it by 2^EV, and an in-range white balance replaces the base's. The
    defaults are PathTracerSceneSnapshot's (PathTracer.h:400-404)."""
    f32 = np.float32
    exposure = f32(base.exposure)
    fs, sh, iso_ = f32(f_stop), f32(shutter_seconds), f32(iso)
    if all(math.isfinite(v) for v in (fs, sh, iso_)) and fs > 0 and sh > 0 and iso_ > 0:
        reference = (f32(1.0) / f32(60.0) * f32(100.0)) / (f32(2.8) * f32(2.8))
        physical = (sh * iso_) / (fs * fs)
        ratio = physical / max(f32(1.0e-6), reference)
        exposure = f32(exposure * min(f32(1.0e6), max(f32(1.0e-6), ratio)))
    ec = f32(exposure_compensation)
    if math.isfinite(ec):
        exposure = f32(exposure * np.power(f32(2.0), min(f32(32.0), max(f32(-32.0), ec))))
    kelvin = base.white_balance_kelvin
    wb = f32(white_balance_kelvin)
    if math.isfinite(wb) and 1000.0 <= wb <= 40000.0:
        kelvin = float(wb)
    return dataclasses.replace(base, exposure=float(exposure), white_balance_kelvin=kelvin)


# ---------------------------------------------------------------------- geometry helpers
def _quad(p0, p1, p2, p3):
    return np.array([p0, p1, p2, p3], np.float64), np.array([[0, 1, 2], [0, 2, 3]], np.int64)


def _box(size, center, yaw_deg):
    sx, sy, sz = (s * 0.5 for s in size)
    corners = np.array([[x, y, z] for x in (-sx, sx) for y in (-sy, sy) for z in (-sz, sz)], np.float64)
    a = math.radians(yaw_deg)
    rot = np.array([[math.cos(a), 0, math.sin(a)], [0, 1, 0], [-math.sin(a), 0, math.cos(a)]])
    corners = corners @ rot.T + np.asarray(center, np.float64)
    # corner index = 4*ix + 2*iy + iz
    faces = [(0, 1, 3, 2), (4, 6, 7, 5), (0, 4, 5, 1), (2, 3, 7, 6), (0, 2, 6, 4), (1, 5, 7, 3)]
    tris = []
    for f in faces:
        tris += [[f[0], f[1], f[2]], [f[0], f[2], f[3]]]
    return corners, np.array(tris, np.int64)


def cornell_box() -> Scene:
    """Closed box of 5 axis-aligned walls, two blocks rotated +-17 degrees and a ceiling light.

    Geometry: box [-1,1]x[0,2]x[-1,1] (as assets/scenes/cornell_native.json); 10 wall + 24 block
    triangles, each its own instance with its own Lambert material (34 per-triangle albedos), plus a
    0.25 x 0.25 emissive quad at y = 1.98 (2 triangles sharing one emissive material, emission
    (17, 12, 4)). NEE samples one RTHitLight: colour (17, 12, 4), intensity = quad area (0.0625),
    radius 0.005, centred 1 cm below the quad so the sampled sphere stays under it.
    Camera (0,1,3) -> (0,1,0), fov 55.
    """
    grey, red, green = (0.73, 0.73, 0.73), (0.65, 0.05, 0.05), (0.12, 0.45, 0.15)
    parts = [
        (_quad((-1, 0, -1), (1, 0, -1), (1, 0, 1), (-1, 0, 1)), grey),      # floor
        (_quad((-1, 2, -1), (-1, 2, 1), (1, 2, 1), (1, 2, -1)), grey),      # ceiling
        (_quad((-1, 0, -1), (-1, 2, -1), (1, 2, -1), (1, 0, -1)), grey),    # back
        (_quad((-1, 0, -1), (-1, 0, 1), (-1, 2, 1), (-1, 2, -1)), red),     # left
        (_quad((1, 0, -1), (1, 2, -1), (1, 2, 1), (1, 0, 1)), green),       # right
        (_box((0.6, 0.6, 0.6), (0.35, 0.3, 0.3), -17.0), grey),             # short block
        (_box((0.6, 1.2, 0.6), (-0.35, 0.6, -0.35), 17.0), grey),           # tall block
    ]
    verts, tris, albedos = [], [], []
    for (v, t), alb in parts:
        base = sum(len(x) for x in verts)
        verts.append(v)
        tris.append(t + base)
        albedos += [alb] * len(t)
    h = 0.125
    lv, lt = _quad((-h, 1.98, -h), (-h, 1.98, h), (h, 1.98, h), (h, 1.98, -h))
    base = sum(len(x) for x in verts)
    verts.append(lv)
    tris.append(lt + base)
    vertices = np.concatenate(verts).astype(np.float32)
    indices = np.concatenate(tris).astype(np.uint32).reshape(-1)
    ntri = len(indices) // 3
    assert ntri == 36
    materials = make_materials(35)
    for i, alb in enumerate(albedos):
        materials[i]["albedo"] = alb
    materials[34]["albedo"] = (0.78, 0.78, 0.78)
    materials[34]["emissive"] = (17.0, 12.0, 4.0)
    materials[34]["material_model"] = MODEL_EMISSIVE
    instances = np.zeros(ntri, INSTANCE_DTYPE)
    instances["first_triangle"] = np.arange(ntri)
    instances["triangle_count"] = 1
    instances["material_index"] = np.minimum(np.arange(ntri), 34)
    instances["entity_id"] = np.arange(ntri) + 1
    instances["rotation"] = (0.0, 0.0, 0.0, 1.0)
    instances["scale"] = 1.0
    lights = np.array([make_light((0.0, 1.97, 0.0), (17.0, 12.0, 4.0), 0.0625, 0.005)], LIGHT_DTYPE)
    cam = Camera(position=(0.0, 1.0, 3.0), target=(0.0, 1.0, 0.0), up=(0.0, 1.0, 0.0), fov_deg=55.0)
    return Scene(vertices, indices, instances, materials, lights, cam, name="cornell")


def feature_scene(seed: int = 5) -> Scene:
    """The Cornell box with every material class and procedural effect of the integrator switched on:
    per-triangle materials cycling through material_effect 0-19 and the models lambert / glossy /
    mirror / metal / transmissive / sheen / clearcoat / toon, a spot light beside the area light, a
    thin-lens camera with a six-blade iris and anamorphic squeeze, and a seeded 32x16 equirectangular
    environment map seen through the open front of the box."""
    rng = np.random.default_rng(seed)
    sc = cornell_box()
    mats = sc.materials.copy()
    models = [MODEL_LAMBERT, MODEL_GLOSSY, MODEL_MIRROR, 4, 5, 6, 7, 8]
    for i in range(34):
        m = mats[i]
        m["material_effect"] = i % 20
        model = models[i % len(models)]
        m["material_model"] = model
        m["roughness"] = {MODEL_LAMBERT: 1.0, MODEL_GLOSSY: 0.3, MODEL_MIRROR: 0.0}.get(model, 0.5)
        if model == 4:
            m["metallic"] = 0.9
        if model == 5:
            m["transmission"] = 0.8
            m["alpha"] = 0.7
        if model == 6:
            m["sheen"] = 0.6
        if model == 7:
            m["clearcoat"] = 0.5
        m["albedo"] = rng.uniform(0.2, 0.9, 3)
    spot = make_light((0.5, 1.9, 0.5), (3.0, 3.0, 2.5), 0.5, 0.02)
    spot["direction"] = (0.0, -1.0, 0.0)
    spot["spot_inner_cos"] = 0.95
    spot["spot_outer_cos"] = 0.8
    lights = np.concatenate([sc.lights, np.array([spot], LIGHT_DTYPE)])
    cam = Camera(position=(0.0, 1.0, 3.0), target=(0.0, 1.0, 0.0), up=(0.0, 1.0, 0.0), fov_deg=55.0,
                 aperture_radius=0.05, focus_distance=3.0, iris_blade_count=6, iris_rotation_degrees=10.0,
                 iris_roundness=0.3, anamorphic_squeeze=1.3)
    env = rng.uniform(0.0, 2.0, (16, 32, 3)).astype(np.float32)
    return dataclasses.replace(sc, materials=mats, lights=lights, camera=cam, environment_map=env,
                               environment_map_scale=(1.0, 0.9, 0.8), environment_color=(0.1, 0.1, 0.1),
                               name="features")


def perturbed_albedo_scene(scene: Scene, sigma: float = 0.1, seed: int = 7) -> Scene:
    """Copy of `scene` whose non-emissive albedos are perturbed by N(0, sigma), clipped to [0, 1]."""
    rng = np.random.default_rng(seed)
    mats = scene.materials.copy()
    for i in range(len(mats)):
        if mats[i]["material_model"] == MODEL_EMISSIVE:
            continue
        mats[i]["albedo"] = np.clip(mats[i]["albedo"] + rng.normal(0.0, sigma, 3), 0.0, 1.0)
    return dataclasses.replace(scene, materials=mats, name=scene.name + "-perturbed")


def icosphere(subdivisions: int):
    """Unit icosphere: 20 * 4**subdivisions triangles (float64 vertices, int64 faces)."""
    t = (1.0 + 5.0 ** 0.5) / 2.0
    v = np.array([[-1, t, 0], [1, t, 0], [-1, -t, 0], [1, -t, 0], [0, -1, t], [0, 1, t], [0, -1, -t], [0, 1, -t],
                  [t, 0, -1], [t, 0, 1], [-t, 0, -1], [-t, 0, 1]], np.float64)
    v /= np.linalg.norm(v, axis=1, keepdims=True)
    f = np.array([[0, 11, 5], [0, 5, 1], [0, 1, 7], [0, 7, 10], [0, 10, 11], [1, 5, 9], [5, 11, 4], [11, 10, 2],
                  [10, 7, 6], [7, 1, 8], [3, 9, 4], [3, 4, 2], [3, 2, 6], [3, 6, 8], [3, 8, 9], [4, 9, 5],
                  [2, 4, 11], [6, 2, 10], [8, 6, 7], [9, 8, 1]], np.int64)
    for _ in range(subdivisions):
        v, f = _subdivide(v, f)
    return v, f


def _subdivide(v, f):
    edges = np.concatenate([f[:, [0, 1]], f[:, [1, 2]], f[:, [2, 0]]])
    edges.sort(axis=1)
    uniq, inv = np.unique(edges, axis=0, return_inverse=True)
    inv = inv.reshape(-1)
    mid = v[uniq[:, 0]] + v[uniq[:, 1]]
    mid /= np.linalg.norm(mid, axis=1, keepdims=True)
    nv = len(v)
    v2 = np.concatenate([v, mid])
    nf = len(f)
    m01, m12, m20 = inv[:nf] + nv, inv[nf:2 * nf] + nv, inv[2 * nf:] + nv
    a, b, c = f[:, 0], f[:, 1], f[:, 2]
    f2 = np.concatenate([np.stack([a, m01, m20], 1), np.stack([b, m12, m01], 1), np.stack([c, m20, m12], 1),
                         np.stack([m01, m12, m20], 1)])
    return v2, f2


def _bisect(v, f):
    """Split every triangle in two at the midpoint of its first edge (projected to the sphere)."""
    e = np.sort(f[:, [0, 1]], axis=1)
    uniq, inv = np.unique(e, axis=0, return_inverse=True)
    mid = v[uniq[:, 0]] + v[uniq[:, 1]]
    mid /= np.linalg.norm(mid, axis=1, keepdims=True)
    m = inv.reshape(-1) + len(v)
    v2 = np.concatenate([v, mid])
    f2 = np.concatenate([np.stack([f[:, 0], m, f[:, 2]], 1), np.stack([m, f[:, 1], f[:, 2]], 1)])
    return v2, f2


def perlin3(p: np.ndarray, seed: int) -> np.ndarray:
    """Improved gradient noise (Perlin 2002) at points p (N, 3), seeded permutation; range about [-1, 1]."""
    rng = np.random.default_rng(seed)
    perm = rng.permutation(256)
    perm = np.concatenate([perm, perm]).astype(np.int64)
    grads = np.array([[1, 1, 0], [-1, 1, 0], [1, -1, 0], [-1, -1, 0], [1, 0, 1], [-1, 0, 1], [1, 0, -1], [-1, 0, -1],
                      [0, 1, 1], [0, -1, 1], [0, 1, -1], [0, -1, -1], [1, 1, 0], [0, -1, 1], [-1, 1, 0], [0, -1, -1]],
                     np.float64)
    pi = np.floor(p).astype(np.int64)
    pf = p - pi
    pi &= 255
    fade = pf * pf * pf * (pf * (pf * 6 - 15) + 10)

    def grad(ix, iy, iz, fx, fy, fz):
        h = perm[perm[perm[ix] + iy] + iz] & 15
        g = grads[h]
        return g[:, 0] * fx + g[:, 1] * fy + g[:, 2] * fz

    x, y, z = pi[:, 0], pi[:, 1], pi[:, 2]
    fx, fy, fz = pf[:, 0], pf[:, 1], pf[:, 2]
    n = {}
    for dx in (0, 1):
        for dy in (0, 1):
            for dz in (0, 1):
                n[dx, dy, dz] = grad((x + dx) & 255, (y + dy) & 255, (z + dz) & 255, fx - dx, fy - dy, fz - dz)
    u, v, w = fade[:, 0], fade[:, 1], fade[:, 2]
    lerp = lambda a, b, t: a + t * (b - a)
    x00 = lerp(n[0, 0, 0], n[1, 0, 0], u)
    x10 = lerp(n[0, 1, 0], n[1, 1, 0], u)
    x01 = lerp(n[0, 0, 1], n[1, 0, 1], u)
    x11 = lerp(n[0, 1, 1], n[1, 1, 1], u)
    return lerp(lerp(x00, x10, v), lerp(x01, x11, v), w)


def uv_sphere(slices: int = 12, stacks: int = 8):
    """Unit UV sphere with slices*(stacks-2)*2 + 2*slices triangles (168 for 12 x 8)."""
    verts = [[0.0, 1.0, 0.0]]
    for i in range(1, stacks):
        th = math.pi * i / stacks
        for j in range(slices):
            ph = 2.0 * math.pi * j / slices
            verts.append([math.sin(th) * math.cos(ph), math.cos(th), math.sin(th) * math.sin(ph)])
    verts.append([0.0, -1.0, 0.0])
    bottom = len(verts) - 1
    tris = []
    for j in range(slices):
        tris.append([0, 1 + (j + 1) % slices, 1 + j])
    for i in range(stacks - 2):
        r0, r1 = 1 + i * slices, 1 + (i + 1) * slices
        for j in range(slices):
            a, b = r0 + j, r0 + (j + 1) % slices
            c, d = r1 + j, r1 + (j + 1) % slices
            tris += [[a, b, d], [a, d, c]]
    r = 1 + (stacks - 2) * slices
    for j in range(slices):
        tris.append([bottom, r + j, r + (j + 1) % slices])
    return np.array(verts, np.float64), np.array(tris, np.int64)


def _random_quaternions(rng, n):
    q = rng.normal(size=(n, 4))
    return q / np.linalg.norm(q, axis=1, keepdims=True)


def procedural_scene(seed: int = 1, ico_subdivisions: int = 7, bisect: bool = True, sphere_count: int = 2048,
                     light_count: int = 16, radius: float = 5.0, box: float = 20.0) -> Scene:
    """Displaced icosphere + instanced UV spheres + quad lights (default: 999,424 + 32 triangles).

    * Icosphere: 20*4^7 faces, each bisected once -> 655,360 triangles, radius 5, displaced along the
      normal by 0.05 * radius * (3-octave Perlin fBm), seeded.
    * `sphere_count` 168-triangle UV spheres as dynamic-transform instances of one local mesh
      (flag kRTInstanceFlagDynamicTransform): centres uniform in the box, radii log-uniform
      [0.05, 0.5], random rotations.
    * `light_count` downward-facing emissive quads (side 1.2) in the upper part of the box, emission
      e*(1,1,1) with e ~ U[2, 20]; each has an RTHitLight (colour e, intensity = quad area,
      radius 0.05) 0.1 below its centre.
    * Materials per instance: 60% Lambert, 30% glossy (model 3, roughness U[0.05, 1]), 10% mirror.
    Camera inside the box looking at the icosphere; black environment.
    """
    rng = np.random.default_rng(seed)
    half = box * 0.5
    v, f = icosphere(ico_subdivisions)
    if bisect:
        v, f = _bisect(v, f)
    noise = np.zeros(len(v))
    amp = 1.0
    for octave in range(3):
        noise += amp * perlin3(v * (1.5 * 2 ** octave) + 17.0, seed * 131 + octave)
        amp *= 0.5
    v = v * (radius * (1.0 + 0.05 * noise))[:, None]
    vertices = [v]
    tris = [f]
    ico_tris = len(f)

    # light quads baked into the global mesh
    light_pos = np.stack([rng.uniform(-half * 0.8, half * 0.8, light_count),
                          rng.uniform(half * 0.55, half * 0.9, light_count),
                          rng.uniform(-half * 0.8, half * 0.8, light_count)], 1)
    side = 1.2
    base = len(v)
    lv, lt = [], []
    for i, c in enumerate(light_pos):
        h = side / 2
        q = np.array([[c[0] - h, c[1], c[2] - h], [c[0] - h, c[1], c[2] + h], [c[0] + h, c[1], c[2] + h],
                      [c[0] + h, c[1], c[2] - h]])
        lv.append(q)
        lt.append(np.array([[0, 1, 2], [0, 2, 3]]) + base + 4 * i)
    vertices.append(np.concatenate(lv))
    tris.append(np.concatenate(lt))
    vertices = np.concatenate(vertices).astype(np.float32)
    indices = np.concatenate(tris).astype(np.uint32).reshape(-1)

    sv, sf = uv_sphere(12, 8)
    n_inst = 1 + light_count + sphere_count
    instances = np.zeros(n_inst, INSTANCE_DTYPE)
    instances["rotation"] = (0.0, 0.0, 0.0, 1.0)
    instances["scale"] = 1.0
    instances["entity_id"] = np.arange(n_inst) + 1
    instances[0]["first_triangle"] = 0
    instances[0]["triangle_count"] = ico_tris
    for i in range(light_count):
        instances[1 + i]["first_triangle"] = ico_tris + 2 * i
        instances[1 + i]["triangle_count"] = 2
    s0 = 1 + light_count
    centres = rng.uniform(-half, half, (sphere_count, 3))
    radii = np.exp(rng.uniform(math.log(0.05), math.log(0.5), sphere_count))
    quats = _random_quaternions(rng, sphere_count)
    ins = instances[s0:]
    ins["flags"] = INSTANCE_FLAG_DYNAMIC_TRANSFORM
    ins["triangle_count"] = len(sf)
    ins["first_triangle"] = 0
    ins["local_first_vertex"] = 0
    ins["local_vertex_count"] = len(sv)
    ins["local_first_index"] = 0
    ins["local_index_count"] = sf.size
    ins["translation"] = centres
    ins["rotation"] = quats
    ins["scale"] = np.repeat(radii[:, None], 3, 1)
    instances[s0:] = ins

    # materials: one per surface instance + one per light
    n_surf = 1 + sphere_count
    materials = make_materials(n_surf + light_count)
    kind = rng.uniform(size=n_surf)
    alb = rng.uniform(0.2, 0.9, (n_surf, 3))
    rough = rng.uniform(0.05, 1.0, n_surf)
    for i in range(n_surf):
        materials[i]["albedo"] = alb[i]
        if kind[i] < 0.6:
            materials[i]["material_model"] = MODEL_LAMBERT
            materials[i]["roughness"] = 1.0
        elif kind[i] < 0.9:
            materials[i]["material_model"] = MODEL_GLOSSY
            materials[i]["roughness"] = rough[i]
        else:
            materials[i]["material_model"] = MODEL_MIRROR
            materials[i]["roughness"] = 0.0
    instances[0]["material_index"] = 0
    instances[s0:]["material_index"] = 1 + np.arange(sphere_count)
    emission = rng.uniform(2.0, 20.0, light_count)
    lights = np.zeros(light_count, LIGHT_DTYPE)
    for i in range(light_count):
        mi = n_surf + i
        materials[mi]["albedo"] = (0.8, 0.8, 0.8)
        materials[mi]["emissive"] = (emission[i],) * 3
        materials[mi]["material_model"] = MODEL_EMISSIVE
        instances[1 + i]["material_index"] = mi
        c = light_pos[i]
        lights[i] = make_light((c[0], c[1] - 0.1, c[2]), (emission[i],) * 3, side * side, 0.05)
    cam = Camera(position=(0.0, 2.0, 9.0), target=(0.0, 0.0, 0.0), up=(0.0, 1.0, 0.0), fov_deg=55.0)
    return Scene(vertices, indices, instances, materials, lights, cam, local_vertices=sv.astype(np.float32),
                 local_indices=sf.astype(np.uint32).reshape(-1), name=f"procedural-{seed}")


def random_rays(scene: Scene, n: int, seed: int = 0) -> np.ndarray:
    """(n, 6) f32 rays: origins uniform in the scene's bounding box (slightly enlarged),
    directions uniform on the sphere, with some exactly axis-aligned components."""
    rng = np.random.default_rng(seed)
    v = scene.vertices
    if scene.local_vertices.size:
        v = np.concatenate([v, scene.instances["translation"]])
    lo, hi = v.min(0), v.max(0)
    pad = (hi - lo) * 0.1 + 1e-3
    o = rng.uniform(lo - pad, hi + pad, (n, 3))
    d = rng.normal(size=(n, 3))
    k = rng.integers(0, 8, n)
    d[k == 0, 0] = 0.0          # parallel-axis cases of the slab test
    d[k == 1, 1] = 0.0
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    return np.concatenate([o, d], 1).astype(np.float32)


def batch_scenes(seed: int = 7, count: int = 64, min_tris: int = 1_000, max_tris: int = 500_000,
                 resolutions=(128, 256, 512, 1024), spps=(4, 16, 64)):
    """`count` independent seeded scenes for batched rendering: (scene, width, height, spp) each.

    Triangle counts are log-uniform in [min_tris, max_tris]: a displaced icosphere takes the largest
    20*4^L (bisected: 40*4^L) face count not above 2/3 of the target, 168-triangle instanced spheres
    and 8 quad lights fill the rest (the recipe of `procedural_scene`). Square resolutions and spp are
    uniform over the given sets.
    """
    rng = np.random.default_rng(seed)
    out = []
    for i in range(count):
        target = int(round(math.exp(rng.uniform(math.log(min_tris), math.log(max_tris)))))
        level, bisect = 0, False
        for lv in range(8):
            for bi in (False, True):
                n = 20 * 4 ** lv * (2 if bi else 1)
                if n <= (2 * target) // 3:
                    level, bisect = lv, bi
        ico = 20 * 4 ** level * (2 if bisect else 1)
        lights = 8
        spheres = max(0, (target - ico - 2 * lights) // 168)
        sc = procedural_scene(seed=seed * 1000 + i, ico_subdivisions=level, bisect=bisect, sphere_count=spheres,
                              light_count=lights)
        res = int(rng.choice(resolutions))
        spp = int(rng.choice(spps))
        out.append((dataclasses.replace(sc, name=f"batch-{seed}-{i}"), res, res, spp))
    return out


def geodesic_sphere(frequency: int):
    """Unit geodesic sphere: every icosahedron face split into frequency^2 triangles on a barycentric
    grid and projected to the sphere (20 * frequency^2 triangles). Returns the grid points (P, 3)
    float64 and triangles (20 * frequency^2, 3) int64 indexing them; grid points on the icosahedron's
    edges appear once per face."""
    n = int(frequency)
    v0, f0 = icosphere(0)
    ii, jj = [], []
    for i in range(n + 1):
        for j in range(n + 1 - i):
            ii.append(i)
            jj.append(j)
    ii = np.array(ii)
    jj = np.array(jj)
    idx = -np.ones((n + 1, n + 1), np.int64)
    idx[ii, jj] = np.arange(len(ii))
    up = [(idx[i, j], idx[i + 1, j], idx[i, j + 1]) for i in range(n) for j in range(n - i)]
    down = [(idx[i + 1, j], idx[i + 1, j + 1], idx[i, j + 1]) for i in range(n) for j in range(n - i - 1)]
    local = np.array(up + down, np.int64)
    a, b, c = v0[f0[:, 0]], v0[f0[:, 1]], v0[f0[:, 2]]
    wa = ((n - ii - jj) / n)[None, :, None]
    pts = a[:, None] * wa + b[:, None] * (ii / n)[None, :, None] + c[:, None] * (jj / n)[None, :, None]
    pts /= np.linalg.norm(pts, axis=2, keepdims=True)
    npf = len(ii)
    faces = (local[None] + (np.arange(20) * npf)[:, None, None]).reshape(-1, 3)
    return pts.reshape(-1, 3), faces


def value_noise_texture(size: int = 1024, seed: int = 3, cells: int = 16, octaves: int = 4, lo: float = 0.15,
                        hi: float = 0.95) -> np.ndarray:
    """Tileable seeded value noise, (size, size, 3) float32 in [lo, hi]: per octave a random lattice of
    cells * 2^o cells per side (wrapping) interpolated with smoothstep weights, amplitudes halving."""
    rng = np.random.default_rng(seed)
    out = np.zeros((size, size, 3), np.float64)
    amp, total = 1.0, 0.0
    for o in range(octaves):
        g = cells * 2 ** o
        lat = rng.uniform(0.0, 1.0, (g, g, 3))
        x = (np.arange(size) + 0.5) / size * g
        i0 = np.floor(x).astype(np.int64) % g
        i1 = (i0 + 1) % g
        f = x - np.floor(x)
        f = f * f * (3.0 - 2.0 * f)
        rows = lat[i0] * (1.0 - f)[:, None, None] + lat[i1] * f[:, None, None]          # (size, g, 3)
        out += amp * (rows[:, i0] * (1.0 - f)[None, :, None] + rows[:, i1] * f[None, :, None])
        total += amp
        amp *= 0.5
    out /= total
    mn, mx = out.min(), out.max()
    return (lo + (hi - lo) * (out - mn) / max(mx - mn, 1e-12)).astype(np.float32)


def ring_cameras(count: int = 16, radius: float = 13.0, height: float = 3.0, fov_deg: float = 42.0) -> list:
    """`count` pinhole cameras evenly spaced on a horizontal ring, all looking at the origin."""
    cams = []
    for k in range(count):
        a = 2.0 * math.pi * k / count
        cams.append(Camera(position=(radius * math.cos(a), height, radius * math.sin(a)), target=(0.0, 0.0, 0.0),
                           up=(0.0, 1.0, 0.0), fov_deg=fov_deg))
    return cams


def textured_sphere_scene(seed: int = 11, frequency: int = 100, texture_size: int = 1024, radius: float = 5.0,
                          material_count: int = 8, light_count: int = 6, texture=None) -> Scene:
    """Displaced geodesic sphere with a UV-mapped albedo texture (default 200,000 triangles and a
    1024^2 rgb fp32 value-noise texture), on a ground quad, lit by emissive quads.

    * Sphere: 20 * frequency^2 triangles, radius 5, displaced along the normal by 0.05 * radius *
      (3-octave Perlin fBm), seeded. Vertices are per triangle corner, with equirectangular texture
      coordinates u = 0.5 + atan2(z, x) / 2pi, v = acos(y) / pi of the undisplaced direction (a
      triangle across the u seam has its small u values moved up by 1; the texture repeats).
    * `material_count` glossy materials (model 3), one per longitude sector, albedo 1, roughness
      U[0.15, 0.9], all using texture 0.
    * Ground: a 40 x 40 Lambert quad (albedo 0.6) at y = -1.3 * radius, untextured.
    * `light_count` downward emissive quads (side 1.5) on a ring of radius 6 at y = 9, emission
      U[4, 12], each with an RTHitLight (intensity = quad area, radius 0.05) 0.1 below its centre.
    Camera: the first of `ring_cameras()`.
    """
    rng = np.random.default_rng(seed)
    pts, faces = geodesic_sphere(frequency)
    noise = np.zeros(len(pts))
    amp = 1.0
    for octave in range(3):
        noise += amp * perlin3(pts * (1.5 * 2 ** octave) + 17.0, seed * 131 + octave)
        amp *= 0.5
    disp = pts * (radius * (1.0 + 0.05 * noise))[:, None]
    u = 0.5 + np.arctan2(pts[:, 2], pts[:, 0]) / (2.0 * math.pi)
    v = np.arccos(np.clip(pts[:, 1], -1.0, 1.0)) / math.pi
    tu = u[faces]                                              # (T, 3)
    seam = (tu.max(1) - tu.min(1)) > 0.5
    tu = np.where(seam[:, None] & (tu < 0.5), tu + 1.0, tu)
    tv = v[faces]
    # sectors by centroid longitude; triangles grouped by sector for contiguous instance ranges
    sector = np.minimum((np.mod(tu.mean(1), 1.0) * material_count).astype(np.int64), material_count - 1)
    order = np.argsort(sector, kind="stable")
    faces, tu, tv, sector = faces[order], tu[order], tv[order], sector[order]
    n_tri = len(faces)
    sph_v = disp[faces].reshape(-1, 3)
    sph_uv = np.stack([tu.reshape(-1), tv.reshape(-1)], 1)

    g = 1.3 * radius
    ground_v = np.array([[-20.0, -g, -20.0], [-20.0, -g, 20.0], [20.0, -g, 20.0], [20.0, -g, -20.0]])
    side = 1.5
    lv = []
    light_pos = []
    for k in range(light_count):
        a = 2.0 * math.pi * (k + 0.5) / light_count
        c = np.array([6.0 * math.cos(a), 9.0, 6.0 * math.sin(a)])
        light_pos.append(c)
        h = side / 2
        lv.append(np.array([[c[0] - h, c[1], c[2] - h], [c[0] - h, c[1], c[2] + h], [c[0] + h, c[1], c[2] + h],
                            [c[0] + h, c[1], c[2] - h]]))
    quad = np.array([[0, 1, 2], [0, 2, 3]])
    vertices = np.concatenate([sph_v, ground_v] + lv).astype(np.float32)
    base = len(sph_v)
    indices = np.concatenate([np.arange(3 * n_tri).reshape(-1, 3), quad + base] +
                             [quad + base + 4 + 4 * k for k in range(light_count)]).astype(np.uint32).reshape(-1)
    texcoords = np.zeros((len(vertices), 2), np.float32)
    texcoords[:len(sph_uv)] = sph_uv

    counts = np.bincount(sector, minlength=material_count)
    n_inst = material_count + 1 + light_count
    instances = np.zeros(n_inst, INSTANCE_DTYPE)
    instances["rotation"] = (0.0, 0.0, 0.0, 1.0)
    instances["scale"] = 1.0
    instances["entity_id"] = np.arange(n_inst) + 1
    first = 0
    for k in range(material_count):
        instances[k]["first_triangle"] = first
        instances[k]["triangle_count"] = counts[k]
        instances[k]["material_index"] = k
        first += int(counts[k])
    instances[material_count]["first_triangle"] = n_tri
    instances[material_count]["triangle_count"] = 2
    instances[material_count]["material_index"] = material_count
    materials = make_materials(material_count + 1 + light_count)
    rough = rng.uniform(0.15, 0.9, material_count)
    for k in range(material_count):
        materials[k]["albedo"] = (1.0, 1.0, 1.0)
        materials[k]["material_model"] = MODEL_GLOSSY
        materials[k]["roughness"] = rough[k]
        materials[k]["base_color_texture_index"] = 0
    materials[material_count]["albedo"] = (0.6, 0.6, 0.6)
    materials[material_count]["material_model"] = MODEL_LAMBERT
    emission = rng.uniform(4.0, 12.0, light_count)
    lights = np.zeros(light_count, LIGHT_DTYPE)
    for k in range(light_count):
        i = material_count + 1 + k
        instances[i]["first_triangle"] = n_tri + 2 + 2 * k
        instances[i]["triangle_count"] = 2
        instances[i]["material_index"] = i
        materials[i]["albedo"] = (0.8, 0.8, 0.8)
        materials[i]["emissive"] = (emission[k],) * 3
        materials[i]["material_model"] = MODEL_EMISSIVE
        c = light_pos[k]
        lights[k] = make_light((c[0], c[1] - 0.1, c[2]), (emission[k],) * 3, side * side, 0.05)
    tex = value_noise_texture(texture_size, seed=seed * 7 + 1) if texture is None else texture
    return Scene(vertices, indices, instances, materials, lights, ring_cameras()[0], name=f"textured-{seed}",
                 texcoords=texcoords, textures=[tex])


def perturbed_texture(texture: np.ndarray, sigma: float = 0.1, seed: int = 9) -> np.ndarray:
    """texture + N(0, sigma) per texel and channel, clipped to [0, 1]."""
    rng = np.random.default_rng(seed)
    return np.clip(texture + rng.normal(0.0, sigma, texture.shape), 0.0, 1.0).astype(np.float32)

"""Scene documents: the reference's JSON scene fixtures (tests/golden/scenes, copied from
assets/scenes) converted to scenes, checked against what the fixtures state, closed forms for
transforms (scipy's rotations as an independent implementation), the conversion's rules for
lights, cameras, culling and validation, and a render of the converted Cornell box by the oracle."""
import copy
import json
import math
import pathlib
import struct

import numpy as np
import pytest

GOLDEN = pathlib.Path(__file__).resolve().parent / "golden" / "scenes"


def _load(name):
    return json.loads((GOLDEN / name).read_text())


def _quad_doc(entities, **extra):
    """A minimal valid document: one unit quad geometry (id 1) and one material (id 1)."""
    doc = {"metadata": {"schema": "1.0"},
           "materials": [{"id": 1, "name": "m", "albedo": [0.5, 0.5, 0.5]}],
           "geometry": [{"id": 1, "primitive": "triangle", "material_id": 1,
                         "vertices": [[-0.5, 0, -0.5], [0.5, 0, -0.5], [0.5, 0, 0.5], [-0.5, 0, 0.5]],
                         "indices": [0, 2, 1, 0, 3, 2]}],
           "entities": entities}
    doc.update(extra)
    return doc


def test_cornell_document_instances():
    """cornell_native.json: 6 quads + 3 cuboids = 48 triangles in 9 instances. Every mesh entity has a
    transform, so every instance is a dynamic-transform instance with local geometry (what the
    reference's transform_drag_throughput.cpp:132-153 relies on); the three cuboids carry dynamic
    physics bodies."""
    from vkpathtracer_amd import scene as S
    from vkpathtracer_amd import scene_document as SD

    sc = SD.document_to_scene(_load("cornell_native.json"))
    assert sc.triangle_count == 48 and len(sc.indices) == 144
    assert len(sc.instances) == 9 and len(sc.materials) == 7 and len(sc.lights) == 1
    ins = sc.instances
    assert list(ins["entity_id"]) == [3, 4, 5, 6, 7, 8, 9, 10, 11]
    assert list(ins["geometry_id"]) == list(range(1, 10))
    assert list(ins["material_index"]) == [0, 0, 0, 1, 2, 3, 4, 5, 6]
    assert list(ins["triangle_count"]) == [2] * 6 + [12] * 3
    assert list(ins["first_triangle"]) == [0, 2, 4, 6, 8, 10, 12, 24, 36]
    dyn = S.INSTANCE_FLAG_DYNAMIC_TRANSFORM | SD.INSTANCE_FLAG_TRANSFORM_DIRTY
    assert list(ins["flags"]) == [dyn] * 6 + [dyn | SD.INSTANCE_FLAG_PHYSICS] * 3
    assert list(ins["local_vertex_count"]) == [4] * 6 + [8] * 3
    assert list(ins["local_index_count"]) == [6] * 6 + [36] * 3
    assert list(ins["transform_revision"]) == list(range(1, 10))
    assert len(sc.local_vertices) == 48 and len(sc.local_indices) == 144
    # local indices are relative to the instance's local vertex block
    assert sc.local_indices.max() < 8


def test_cornell_document_world_vertices():
    """World vertices are translation + scale * local vertex (identity rotations); the cuboids' boxes
    follow from their transforms in the fixture."""
    from vkpathtracer_amd import scene_document as SD

    doc = _load("cornell_native.json")
    sc = SD.document_to_scene(doc)
    geom = {g["id"]: np.asarray(g["vertices"], np.float64) for g in doc["geometry"]}
    ents = [e for e in doc["entities"] if "mesh" in e]
    off = 0
    for e in ents:
        v = geom[e["mesh"]["mesh_id"]]
        t = np.asarray(e["transform"]["translation"], np.float64)
        s = np.asarray(e["transform"]["scale"], np.float64)
        np.testing.assert_allclose(sc.vertices[off:off + len(v)], t + s * v, atol=1e-6)
        off += len(v)
    # glossy cuboid: translation (0.575, 0.225, 0.325), scale (0.55, 0.45, 0.45) on the unit cube
    i = list(sc.instances["entity_id"]).index(11)
    tri = sc.indices[3 * sc.instances["first_triangle"][i]:3 * (sc.instances["first_triangle"][i] + 12)]
    box = sc.vertices[tri]
    np.testing.assert_allclose(box.min(0), [0.3, 0.0, 0.1], atol=1e-6)
    np.testing.assert_allclose(box.max(0), [0.85, 0.45, 0.55], atol=1e-6)


def test_cornell_document_materials_camera_light():
    """Materials without a family are "diffuse": the loader's Override preset, then the explicit
    fields, then the conversion's FillGenericDefaults, which turns the generic roughness 1.0 into the
    diffuse family's 0.85 and leaves authored values alone; emission is colour times intensity."""
    from vkpathtracer_amd import scene_document as SD

    sc = SD.document_to_scene(_load("cornell_native.json"))
    m = sc.materials
    np.testing.assert_allclose(m["roughness"], [0.85, 0.85, 0.85, 0.85, 0.85, 0.0, 0.3], atol=0)
    np.testing.assert_allclose(m["albedo"][1], np.float32([0.65, 0.05, 0.05]))
    np.testing.assert_allclose(m["emissive"][3], [12, 12, 12])
    assert not m["emissive"][[0, 1, 2, 4, 5, 6]].any()
    assert (m["material_model"] == 0).all() and (m["material_effect"] == 0).all()
    assert (m["ior"] == np.float32(1.5)).all() and (m["alpha"] == 1).all() and (m["metallic"] == 0).all()
    assert (m["base_color_texture_index"] == 0xFFFFFFFF).all()
    c = sc.camera
    assert c.position == (0.0, 1.0, 3.0) and c.target == (0.0, 1.0, 2.0) and c.up == (0.0, 1.0, 0.0)
    assert c.fov_deg == 55.0
    l = sc.lights[0]
    np.testing.assert_array_equal(l["position"], np.float32([0, 1.85, 0]))
    np.testing.assert_array_equal(l["color"], np.float32([1.0, 0.98, 0.95]))
    assert l["intensity"] == 14 and l["radius"] == 0
    np.testing.assert_array_equal(l["direction"], [0, -1, 0])
    assert l["spot_inner_cos"] == -1 and l["spot_outer_cos"] == -1
    assert sc.environment_color == (0.0, 0.0, 0.0)


def test_family_presets():
    """Authored families get their presets; FillGenericDefaults leaves authored values and only
    replaces fields at their generic defaults (the values are those of Scene.cpp's table)."""
    from vkpathtracer_amd import scene_document as SD

    mats = [{"id": 1, "family": "Mirror"}, {"id": 2, "family": "dielectric glass", "ior": 1.33},
            {"id": 3, "family": "emissive", "albedo": [0.1, 0.9, 0.5]}, {"id": 4, "family": "car-paint"},
            {"id": 5, "name": "Brushed Metal", "family": ""}, {"id": 6, "family": "velvet", "roughness": 0.4},
            {"id": 7, "family": "blackbody_emission", "emission": [1, 0.5, 0.25]}]
    doc = _quad_doc([], materials=mats)
    m = SD.document_to_scene(doc).materials
    assert m["roughness"][0] == 0 and m["metallic"][0] == 1 and m["material_model"][0] == 2
    assert m["ior"][1] == np.float32(1.33) and m["transmission"][1] == 1 and m["alpha"][1] == np.float32(0.35)
    assert m["material_model"][1] == 5 and m["roughness"][1] == np.float32(0.02)
    # emissive: the loader's Override preset sets emission = max(0.2, albedo) before the authored
    # albedo is read, so from the default albedo (1, 1, 1); intensity 3
    np.testing.assert_array_equal(m["emissive"][2], [3, 3, 3])
    assert m["material_model"][2] == 1
    assert m["clearcoat"][3] == 1 and m["metallic"][3] == np.float32(0.25) and m["material_model"][3] == 7
    # an empty family takes the normalised name
    assert m["material_model"][4] == 4 and m["anisotropy"][4] == np.float32(0.65) and m["metallic"][4] == 1
    assert m["roughness"][5] == np.float32(0.4) and m["sheen"][5] == np.float32(0.85)
    assert m["material_effect"][5] == 1 and m["material_model"][5] == 6
    np.testing.assert_allclose(m["emissive"][6], np.float32([5, 2.5, 1.25]))
    assert m["material_effect"][6] == 10


def _quat(axis, deg):
    from scipy.spatial.transform import Rotation

    return Rotation.from_rotvec(np.asarray(axis, np.float64) / np.linalg.norm(axis) * math.radians(deg)).as_quat()


@pytest.mark.parametrize("seed", range(6))
def test_hierarchy_world_transform_matches_scipy(seed):
    """A three-level hierarchy with random rotations (unnormalised quaternions), translations and
    uniform parent scales: world vertices equal the product of the three TRS transforms computed in
    float64 with scipy's rotations."""
    from scipy.spatial.transform import Rotation

    from vkpathtracer_amd import scene_document as SD

    rng = np.random.default_rng(seed)
    trs = []
    for level in range(3):
        q = rng.normal(size=4) * 1.7
        t = rng.uniform(-3, 3, 3)
        s = np.full(3, rng.uniform(0.5, 2.0)) if level < 2 else rng.uniform(0.5, 2.0, 3)
        trs.append((q, t, s))
    ents = [{"id": 1, "transform": {"translation": list(trs[0][1]), "rotation": list(trs[0][0]),
                                    "scale": list(trs[0][2])}},
            {"id": 2, "hierarchy": {"parent": 1},
             "transform": {"translation": list(trs[1][1]), "rotation": list(trs[1][0]), "scale": list(trs[1][2])}},
            {"id": 3, "hierarchy": {"parent": 2}, "mesh": {"mesh_id": 1, "material_id": 1},
             "transform": {"translation": list(trs[2][1]), "rotation": list(trs[2][0]), "scale": list(trs[2][2])}}]
    doc = _quad_doc(ents)
    sc = SD.document_to_scene(doc)
    v = np.asarray(doc["geometry"][0]["vertices"], np.float64)
    for q, t, s in reversed(trs):
        v = Rotation.from_quat(q.astype(np.float32).astype(np.float64)).apply(v * s) + t
    np.testing.assert_allclose(sc.vertices, v, atol=2e-5 * max(1.0, np.abs(v).max()))
    # the instance keeps the world TRS: composed rotation, product of scales, matrix translation
    ins = sc.instances[0]
    qw = (Rotation.from_quat(trs[0][0]) * Rotation.from_quat(trs[1][0]) * Rotation.from_quat(trs[2][0])).as_quat()
    qi = ins["rotation"].astype(np.float64)
    assert min(np.abs(qi - qw).max(), np.abs(qi + qw).max()) < 1e-5
    np.testing.assert_allclose(ins["scale"], trs[0][2] * trs[1][2] * trs[2][2], rtol=1e-6)
    local = Rotation.from_quat(qi).apply(sc.local_vertices.astype(np.float64) * ins["scale"]) + ins["translation"]
    np.testing.assert_allclose(local, sc.vertices, atol=2e-5 * max(1.0, np.abs(v).max()))


def test_transforms_array_and_parent_without_transform():
    """The document's "transforms" array overrides an entity's transform (last entry wins) and its
    parent; a parent without a transform contributes the identity and ends the chain."""
    from vkpathtracer_amd import scene_document as SD

    ents = [{"id": 1, "transform": {"translation": [10, 0, 0]}},
            {"id": 2, "hierarchy": {"parent": 1}},
            {"id": 3, "hierarchy": {"parent": 2}, "mesh": {"mesh_id": 1}, "transform": {"translation": [0, 1, 0]}},
            {"id": 4, "mesh": {"mesh_id": 1}, "transform": {"translation": [0, 0, 0]}}]
    doc = _quad_doc(ents, transforms=[{"id": 4, "transform": {"translation": [0, 0, 1]}},
                                      {"id": 4, "parent": 1, "transform": {"translation": [0, 0, 2]}}])
    sc = SD.document_to_scene(doc)
    quad = np.asarray(doc["geometry"][0]["vertices"], np.float32)
    np.testing.assert_array_equal(sc.vertices[:4], quad + [0, 1, 0])
    np.testing.assert_array_equal(sc.vertices[4:], quad + [10, 0, 2])


def test_camera_from_rotation_and_auto_placement():
    """A camera looks down its rotated -Z with its rotated +Y up; a camera entity without a transform
    is placed from the scene bounds (center + (0, max(0.6, 0.4 r), max(2, 2.2 r)))."""
    from scipy.spatial.transform import Rotation

    from vkpathtracer_amd import scene_document as SD

    q = _quat([1, 2, 0.5], 70)
    ents = [{"id": 1, "camera": {"fov": 40}, "transform": {"translation": [1, 2, 3], "rotation": list(q)}},
            {"id": 2, "mesh": {"mesh_id": 1}}]
    c = SD.document_to_scene(_quad_doc(ents)).camera
    r = Rotation.from_quat(q)
    np.testing.assert_allclose(np.subtract(c.target, c.position), r.apply([0, 0, -1]), atol=1e-6)
    np.testing.assert_allclose(c.up, r.apply([0, 1, 0]), atol=1e-6)
    assert c.position == (1.0, 2.0, 3.0) and c.fov_deg == 40.0

    ents = [{"id": 1, "camera": {}}, {"id": 2, "mesh": {"mesh_id": 1}, "transform": {"scale": [6, 1, 2]}}]
    c = SD.document_to_scene(_quad_doc(ents)).camera
    # bounds x in [-3, 3], z in [-1, 1], y = 0: radius = max(0.5, 0.5 * 6) = 3
    np.testing.assert_allclose(c.target, [0, 0, 0])
    np.testing.assert_allclose(c.position, [0, 1.2, 6.6], rtol=1e-6)
    assert c.fov_deg == 60.0


def test_lights_spot_environment_visibility():
    """Spot cones are cos(clamp(beam / 2, 1, 89)) and cos(clamp(inner (1 + blend), inner + 0.25, 89.5));
    a zero direction falls back to the rotated -Z; environment lights add colour x intensity to the
    environment; a light under an invisible parent keeps its record with intensity 0; a top-level
    light whose id has no light component is added at its entity's position."""
    from scipy.spatial.transform import Rotation

    from vkpathtracer_amd import scene_document as SD

    q = _quat([0, 1, 0], 90)
    ents = [{"id": 1, "transform": {"translation": [1, 2, 3], "rotation": list(q)},
             "light": {"type": "Spot Light", "beam_angle": 40, "blend": 0.5, "direction": [0, 0, 0], "intensity": 5}},
            {"id": 2, "light": {"type": "spot", "beam_angle": 200, "blend": 0.0}},
            {"id": 3, "light": {"type": "sky", "color": [0.5, 0.25, 1], "intensity": 2}},
            {"id": 4, "light": {"type": "open-sky", "color": [1, 1, 1], "intensity": 0.5}},
            {"id": 5, "visible": False},
            {"id": 6, "hierarchy": {"parent": 5}, "light": {"intensity": 7}},
            {"id": 7, "transform": {"translation": [4, 5, 6]}}]
    sc = SD.document_to_scene(_quad_doc(ents, lights=[{"id": 7, "type": "point", "intensity": 3, "radius": 0.2}]))
    L = sc.lights
    assert len(L) == 4
    assert abs(L[0]["spot_inner_cos"] - math.cos(math.radians(20))) < 1e-6
    assert abs(L[0]["spot_outer_cos"] - math.cos(math.radians(30))) < 1e-6
    np.testing.assert_allclose(L[0]["direction"], Rotation.from_quat(q).apply([0, 0, -1]), atol=1e-6)
    np.testing.assert_array_equal(L[0]["position"], [1, 2, 3])
    assert L[0]["intensity"] == 5
    assert abs(L[1]["spot_inner_cos"] - math.cos(math.radians(89))) < 1e-6
    assert abs(L[1]["spot_outer_cos"] - math.cos(math.radians(89.25))) < 1e-6
    np.testing.assert_array_equal(L[1]["position"], np.float32([0, 1.8, 0]))
    np.testing.assert_array_equal(L[1]["direction"], [0, -1, 0])
    assert L[2]["intensity"] == 0
    np.testing.assert_array_equal(L[3]["position"], [4, 5, 6])
    assert L[3]["intensity"] == 3 and L[3]["radius"] == np.float32(0.2)
    np.testing.assert_allclose(sc.environment_color, [1.5, 1.0, 2.5])


def test_visibility_scripts_and_static_meshes():
    """A mesh under an invisible ancestor is dropped; a mesh without a transform, physics body or
    script in its ancestry is static (world vertices only); a scripted ancestor makes it dynamic."""
    from vkpathtracer_amd import scene_document as SD

    ents = [{"id": 1, "visible": False}, {"id": 2, "hierarchy": {"parent": 1}, "mesh": {"mesh_id": 1}},
            {"id": 3, "mesh": {"mesh_id": 1}},
            {"id": 4, "script": {"source": "x.lua"}}, {"id": 5, "hierarchy": {"parent": 4}, "mesh": {"mesh_id": 1}},
            {"id": 6, "script": {"source": "x.lua", "enabled": False}},
            {"id": 7, "hierarchy": {"parent": 6}, "mesh": {"mesh_id": 1}}]
    sc = SD.document_to_scene(_quad_doc(ents))
    assert list(sc.instances["entity_id"]) == [3, 5, 7]
    assert list(sc.instances["flags"]) == [0, 1, 0]
    assert list(sc.instances["local_vertex_count"]) == [0, 4, 0]


def test_texture_indices_assets_first():
    """Texture indices are dense in the order assets, then each material's textures; URIs are
    normalised to forward slashes and only image extensions count."""
    from vkpathtracer_amd import scene_document as SD

    mats = [{"id": 1, "base_color_texture": "b.png", "normal_texture": "tex\\a.JPG"},
            {"id": 2, "base_color_texture": "tex/a.JPG", "normal_texture": "n.txt"}]
    assets = [{"id": 1, "type": "texture", "uri": "c.tga"}, {"id": 2, "type": "environment/hdri", "uri": "e.hdr"},
              {"id": 3, "type": "audio", "uri": "s.ogg"}]
    m = SD.document_to_scene(_quad_doc([], materials=mats, assets=assets)).materials
    assert list(m["base_color_texture_index"]) == [1, 2]
    assert list(m["normal_texture_index"]) == [2, 0xFFFFFFFF]


def test_performance_culling():
    """With performance culling on, a mesh behind the camera or beyond max_distance is dropped, one in
    view is kept, and a dynamic mesh survives when cull_dynamic is off."""
    from vkpathtracer_amd import scene_document as SD

    def doc(cull_dynamic=True):
        ents = [{"id": 1, "camera": {"fov": 60}, "transform": {}},
                {"id": 2, "mesh": {"mesh_id": 1}, "transform": {"translation": [0, 0, -5]}},
                {"id": 3, "mesh": {"mesh_id": 1}, "transform": {"translation": [0, 0, 5]}},
                {"id": 4, "mesh": {"mesh_id": 1}, "transform": {"translation": [0, 0, -500]}},
                {"id": 5, "mesh": {"mesh_id": 1}, "transform": {"translation": [40, 0, -5]}},
                {"id": 6, "mesh": {"mesh_id": 1}, "transform": {"translation": [0, 0, 5]},
                 "physics": {"body_type": "dynamic"}}]
        return _quad_doc(ents, performance_culling={"enabled": True, "max_distance": 100,
                                                    "cull_dynamic": cull_dynamic})

    assert list(SD.document_to_scene(doc()).instances["entity_id"]) == [2]
    assert list(SD.document_to_scene(doc(False)).instances["entity_id"]) == [2, 6]
    off = doc()
    off["performance_culling"]["enabled"] = False
    assert len(SD.document_to_scene(off).instances) == 5


@pytest.mark.parametrize("yaw", [0, 25, 180])
def test_showcase_culling_matches_float64_frustum(yaw):
    """material_shader_physics_showcase.json enables culling. With its camera turned by `yaw` degrees
    about +Y, the conversion keeps exactly the instances whose bounding sphere (the geometry's sphere
    moved by the instance transform, radius scaled by the largest |scale|) meets the padded view
    frustum within the distance limit, decided here in float64 (spheres within 1e-3 of a plane are
    not judged)."""
    from scipy.spatial.transform import Rotation

    from vkpathtracer_amd import scene_document as SD

    doc = _load("material_shader_physics_showcase.json")
    cam_e = next(e for e in doc["entities"] if "camera" in e)
    q = Rotation.from_rotvec([0, math.radians(yaw), 0]) * Rotation.from_quat(cam_e["transform"]["rotation"])
    cam_e["transform"]["rotation"] = list(q.as_quat())
    culled = SD.document_to_scene(doc)
    off = copy.deepcopy(doc)
    off["performance_culling"]["enabled"] = False
    full = SD.document_to_scene(off)
    kept = set(culled.instances["entity_id"].tolist())
    s = doc["performance_culling"]
    cam = full.camera
    p = np.asarray(cam.position, np.float64)
    f = np.subtract(cam.target, p)
    f /= np.linalg.norm(f)
    r = np.cross(f, cam.up)
    r /= np.linalg.norm(r)
    u = np.cross(r, f)
    near = cam_e["camera"]["near_plane"]
    ty = math.tan(math.radians(cam.fov_deg) / 2) * s["frustum_padding"]
    tx = ty * s["aspect_ratio"]
    judged = 0
    for ins in full.instances:
        a, n = ins["local_first_vertex"], ins["local_vertex_count"]
        if n == 0:  # the one static mesh (no transform) has no local copy of its geometry
            continue
        v = full.local_vertices[a:a + n].astype(np.float64)
        c0 = (v.min(0) + v.max(0)) / 2
        rad = np.linalg.norm(v - c0, axis=1).max() * max(0.001, np.abs(ins["scale"]).max())
        rad = max(s["min_instance_radius"], rad)
        c = Rotation.from_quat(ins["rotation"].astype(np.float64)).apply(c0 * ins["scale"]) + ins["translation"]
        rel = c - p
        z = rel @ f
        pz = max(z, near)
        margins = [s["max_distance"] + rad - np.linalg.norm(rel), z + rad - near,
                   pz * tx + rad - abs(rel @ r), pz * ty + rad - abs(rel @ u)]
        if min(abs(m) for m in margins) < 1e-3:
            continue
        judged += 1
        assert (ins["entity_id"] in kept) == all(m > 0 for m in margins), ins["entity_id"]
    assert judged > 60
    if yaw == 180:
        assert not kept


@pytest.mark.parametrize("mutate, message", [
    (lambda d: d["metadata"].pop("schema"), "schema is empty"),
    (lambda d: d["entities"].append({"id": 1}), "duplicate entity id 1"),
    (lambda d: d["entities"].extend([{"id": 8, "hierarchy": {"parent": 9}}, {"id": 9, "hierarchy": {"parent": 8}}]),
     "hierarchy cycle"),
    (lambda d: d["entities"].append({"id": 8, "hierarchy": {"parent": 99}}), "missing parent 99"),
    (lambda d: d["entities"].append({"id": 8, "light": {"intensity": -1}}), "light has invalid values"),
    (lambda d: d["geometry"][0]["indices"].extend([0, 1, 9]), "index out of range"),
    (lambda d: d.update(transforms=[{"id": 42}]), "transform references missing entity 42"),
    (lambda d: d["entities"].append({"id": 8, "transform": {"scale": [1, 0, 1]}}), "invalid transform"),
    (lambda d: d["materials"][0].update(roughness=1.5), "roughness out of range"),
    (lambda d: d.update(entities={}), "wrong JSON kind"),
])
def test_validation_rejects(mutate, message):
    """Documents the reference's loader refuses (its validate() gate) raise SceneDocumentError."""
    from vkpathtracer_amd import scene_document as SD

    d = _quad_doc([{"id": 1, "mesh": {"mesh_id": 1}}])
    SD.document_to_scene(copy.deepcopy(d))
    mutate(d)
    with pytest.raises(SD.SceneDocumentError, match=message):
        SD.document_to_scene(d)


@pytest.mark.parametrize("name", sorted(p.name for p in GOLDEN.glob("*.json")))
def test_fixture_triangle_counts(name):
    """Each fixture converts; its triangle count is the index count / 3 of the visible mesh entities
    (all fixture triangles are valid, culling aside)."""
    from vkpathtracer_amd import scene_document as SD

    doc = _load(name)
    if doc.get("performance_culling", {}).get("enabled"):
        doc["performance_culling"]["enabled"] = False
    sc = SD.document_to_scene(doc)
    geom = {g["id"]: len(g["indices"]) // 3 for g in doc["geometry"]}
    ents = {e["id"]: e for e in doc["entities"]}

    def visible(e):
        while e is not None:
            if e.get("visible", True) is False:
                return False
            e = ents.get(e.get("hierarchy", {}).get("parent", 0))
        return True

    want = sum(geom.get(e["mesh"]["mesh_id"], 0) for e in doc["entities"] if "mesh" in e and visible(e))
    assert sc.triangle_count == want
    assert (len(sc.vertices) or len(sc.sdf_primitives)) and np.isfinite(sc.vertices).all()
    assert (sc.instances["material_index"] < len(sc.materials)).all()


def test_snapshot_file_of_a_document(tmp_path):
    """The command-line conversion writes a VKPTSNAP file whose header counts match the scene."""
    from vkpathtracer_amd import scene_document as SD

    out = tmp_path / "cornell.ptsnap"
    SD.main([str(GOLDEN / "cornell_native.json"), "--out", str(out)])
    raw = out.read_bytes()
    assert raw[:8] == b"VKPTSNAP"
    head = struct.unpack_from("<12IQ6f", raw, 8)
    assert head[0] == 1 and head[1:8] == (48, 144, 48, 144, 9, 7, 1)


def test_oracle_renders_cornell_document():
    """The oracle path-traces the converted Cornell box: a lit image whose brightest pixels are the
    ceiling light (emission 12) and whose mean is finite and positive."""
    from oracle import oracle as O
    from vkpathtracer_amd import scene as S
    from vkpathtracer_amd import scene_document as SD

    sc = SD.document_to_scene(_load("cornell_native.json"))
    st = S.RenderSettings(width=32, height=32, spp=4, max_depth=5)
    film, cnt, counters = O.Oracle(sc).render(st)
    img = film / np.maximum(cnt[..., None], 1)
    assert np.isfinite(img).all() and img.mean() > 0.05
    assert counters["rays"] > 32 * 32 * 4


def test_host_reads_converted_snapshot(tmp_path):
    """The C++ host's snapshot reader (pt_hip_bench --check-snapshot) reads a converted document with
    every count intact."""
    import subprocess

    from vkpathtracer_amd import scene_document as SD

    bench = GOLDEN.parent.parent.parent / "vkpathtracer_amd" / "bin" / "pt_hip_bench"
    if not bench.exists():
        pytest.skip("pt_hip_bench is not built")
    sc = SD.document_to_scene(_load("material_shader_physics_showcase.json"))
    path = tmp_path / "showcase.ptsnap"
    sc.save_snapshot(path)
    got = json.loads(subprocess.run([str(bench), "--check-snapshot", str(path)], check=True, capture_output=True,
                                    text=True, timeout=60).stdout)
    assert got["triangles"] == sc.triangle_count and got["instances"] == len(sc.instances)
    assert got["local_vertices"] == len(sc.local_vertices) and got["materials"] == len(sc.materials)
    vsum = float(np.sum(sc.vertices, dtype=np.float64))
    assert abs(got["vertex_sum"] - vsum) <= 1e-6 * max(1.0, abs(vsum))


def test_sdf_primitives_conversion():
    """SDF entities become primitives with their world translation and scale, the world quaternion's
    x, y, z as Euler angles (parse_shape_rotation), radius at least 0.01 and the entity's material override;
    document-level entries add shapes unless an entity has their id; unknown shapes are dropped."""
    from vkpathtracer_amd import scene as S
    from vkpathtracer_amd import scene_document as SD

    sc = SD.document_to_scene(_load("sdf_complexity.json"))
    p = sc.sdf_primitives
    assert [S.SDF_SHAPES[k] for k in p["shape"]] == ["sphere", "box"]
    np.testing.assert_array_equal(p["position"], np.float32([[0, 1, 0], [1.2, 0.5, 0]]))
    np.testing.assert_array_equal(p["scale"], np.float32([[0.8, 0.8, 0.8], [0.4, 0.6, 0.4]]))
    assert p["radius"][0] == np.float32(0.8) and p["param_a"][1] == np.float32(0.4)
    ents = [{"id": 1, "sdf_primitive": {"shape": "torus", "radius": 0.001, "param_a": 2.0},
             "transform": {"translation": [1, 2, 3], "rotation": [0.1, 0.2, 0.3, 0.9]}, "material": {"id": 2}},
            {"id": 2, "sdf_primitive": {"shape": "cone"}}]
    doc = _quad_doc(ents, materials=[{"id": 1}, {"id": 2, "albedo": [1, 0, 0]}],
                    sdf_primitives=[{"id": 1, "shape": "sphere"}, {"id": 7, "shape": "capsule",
                                     "transform": {"translation": [0, 5, 0]}, "primitive": {"radius": 0.3}}])
    p = SD.document_to_scene(doc).sdf_primitives
    assert [S.SDF_SHAPES[k] for k in p["shape"]] == ["torus", "capsule"]
    assert p["radius"][0] == np.float32(0.01) and p["material_index"][0] == 1
    # the world transform's (normalised) quaternion x, y, z
    np.testing.assert_allclose(p["rotation"][0], np.array([0.1, 0.2, 0.3]) / math.sqrt(0.95), rtol=1e-6)
    np.testing.assert_array_equal(p["position"][1], [0, 5, 0])
    assert p["radius"][1] == np.float32(0.3) and p["material_index"][1] == 0
    # a document-level entry names its shape at the top level; without one the loader rejects it
    doc["sdf_primitives"].append({"id": 8, "primitive": {"shape": "sphere"}})
    with pytest.raises(SD.SceneDocumentError, match="sdf primitive is invalid 8"):
        SD.document_to_scene(doc)


def test_snapshot_version_2_carries_sdf_primitives(tmp_path):
    """A scene with SDF primitives is written as snapshot version 2 (version 1 plus the SDF records),
    which the C++ host reads back field for field; a scene without them stays version 1."""
    import struct
    import subprocess

    from vkpathtracer_amd import scene as S
    from vkpathtracer_amd import scene_document as SD

    bench = GOLDEN.parent.parent.parent / "vkpathtracer_amd" / "bin" / "pt_hip_bench"
    if not bench.exists():
        pytest.skip("pt_hip_bench is not built")
    sc = SD.document_to_scene(_load("sdf_complexity.json"))
    assert len(sc.sdf_primitives) > 0
    path = tmp_path / "sdf.ptsnap"
    sc.save_snapshot(path)
    raw = path.read_bytes()
    assert raw[:8] == b"VKPTSNAP" and struct.unpack("<I", raw[8:12])[0] == 2
    n = len(sc.sdf_primitives)
    assert struct.unpack("<I", raw[-(4 + 56 * n):-56 * n])[0] == n
    assert raw[-56 * n:] == np.ascontiguousarray(sc.sdf_primitives, S.SDF_DTYPE).tobytes()
    got = json.loads(subprocess.run([str(bench), "--check-snapshot", str(path)], check=True, capture_output=True,
                                    text=True, timeout=60).stdout)
    p = sc.sdf_primitives
    want = float(sum(np.sum(p[k], dtype=np.float64) for k in ("shape", "material_index", "radius", "param_a",
                                                                  "param_b", "position", "scale", "rotation")))
    assert got["sdf_primitives"] == n and abs(got["sdf_sum"] - want) <= 1e-9 * max(1.0, abs(want))
    plain = tmp_path / "plain.ptsnap"
    SD.document_to_scene(_load("cornell_native.json")).save_snapshot(plain)
    assert struct.unpack("<I", plain.read_bytes()[8:12])[0] == 1
    got = json.loads(subprocess.run([str(bench), "--check-snapshot", str(plain)], check=True, capture_output=True,
                                    text=True, timeout=60).stdout)
    assert got["sdf_primitives"] == 0

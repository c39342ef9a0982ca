"""The GPU bucketed-SAH build (pt_bvh_build_device) against the host builder (pt_bvh_build_host, itself
pinned node for node to the oracle's restatement): node boxes, child links, leaf ranges, the leaf
order, the depth, for leaf sizes 1 to 31 and 2 to 16 buckets, on the benchmark scenes and on scenes
made to hit the builder's corner cases (every centroid equal, so no axis has extent; duplicated
triangles; a flat scene; a single triangle; a non-finite vertex, which the build leaves to the host).
The tracer's own build (pt_hip_build_accel) reports that it ran on the device."""
import numpy as np
import pytest

from vkpathtracer_amd import bvh, scene as S

pytestmark = pytest.mark.gpu


def _soup(tris, seed=0):
    """A scene of the given (T, 3, 3) triangles, one instance."""
    tris = np.asarray(tris, np.float32)
    v = tris.reshape(-1, 3)
    inst = np.zeros(1, S.INSTANCE_DTYPE)
    inst["triangle_count"] = len(tris)
    inst["rotation"] = (0, 0, 0, 1)
    inst["scale"] = 1
    inst["entity_id"] = 1
    return S.Scene(v, np.arange(len(v), dtype=np.uint32), inst, S.make_materials(1), np.zeros(0, S.LIGHT_DTYPE),
                   S.Camera(), name=f"soup{seed}")


def _corner_scenes():
    rng = np.random.default_rng(11)
    one = rng.uniform(-1, 1, (1, 3, 3))
    yield "single", _soup(one)
    same = np.repeat(rng.uniform(-1, 1, (1, 3, 3)), 700, axis=0)  # every centroid equal: no split axis
    yield "identical", _soup(same)
    base = rng.uniform(-1, 1, (300, 3, 3))
    yield "duplicated", _soup(np.concatenate([base, base, base[:50]]))
    flat = rng.uniform(-1, 1, (2000, 3, 3))
    flat[:, :, 1] = 0.25  # all in one plane: zero extent on y
    yield "flat", _soup(flat)
    tiny = rng.uniform(0, 1e-7, (500, 3, 3)) + np.array([3.0, -2.0, 1.0])  # extents below 1e-6
    yield "tiny", _soup(tiny)
    yield "random", _soup(rng.normal(0, 10, (20000, 3, 3)) * rng.uniform(0.01, 1, (20000, 1, 1)))


def _assert_same(a, d, what):
    assert d["built_on_device"], what
    assert a["max_depth"] == d["max_depth"], what
    assert np.array_equal(a["meta"], d["meta"]), what
    assert np.array_equal(a["order"], d["order"]), what
    assert np.array_equal(a["bounds"].view(np.uint32), d["bounds"].view(np.uint32)), what
    assert np.array_equal(a["geo"], d["geo"]), what


@pytest.mark.parametrize("leaf,buckets", [(4, 8), (1, 2), (2, 16), (8, 5), (31, 8)])
def test_device_sah_equals_host_on_corner_cases(leaf, buckets):
    for name, sc in _corner_scenes():
        _assert_same(bvh.build_host(sc, "sah", leaf, buckets), bvh.build_device(sc, "sah", leaf, buckets),
                     (name, leaf, buckets))


def test_device_sah_equals_host_on_huge_segments():
    """Segments above 65536 references are split by many workgroups at once: the same tree, also when
    no split is found (identical centroids) and when one axis is flat."""
    rng = np.random.default_rng(23)
    same = np.repeat(rng.uniform(-1, 1, (1, 3, 3)), 70001, axis=0)
    flat = rng.uniform(-1, 1, (150000, 3, 3))
    flat[:, :, 2] = -0.5
    clustered = np.concatenate([rng.normal(0, 0.01, (90000, 3, 3)), rng.normal(5, 3, (110000, 3, 3))])
    for name, t in (("identical", same), ("flat", flat), ("clustered", clustered)):
        sc = _soup(t)
        for leaf, buckets in ((4, 8), (2, 13)):
            _assert_same(bvh.build_host(sc, "sah", leaf, buckets), bvh.build_device(sc, "sah", leaf, buckets),
                         (name, leaf, buckets))


def test_device_sah_equals_host_on_benchmark_scenes():
    scenes = [("cornell", S.cornell_box(), 4), ("features", S.feature_scene(), 4),
              ("procedural 1M", S.procedural_scene(seed=1), 4)]
    scenes += [(f"batch {i}", sc, 2) for i, (sc, *_) in enumerate(S.batch_scenes(seed=7, count=64)) if i % 16 == 0]
    for name, sc, leaf in scenes:
        a, d = bvh.build_host(sc, "sah", leaf, 8), bvh.build_device(sc, "sah", leaf, 8)
        _assert_same(a, d, name)
        print(f"{name}: {len(a['meta'])} nodes, depth {a['max_depth']}: device tree = host tree", flush=True)


def test_non_finite_vertex_builds_on_host():
    rng = np.random.default_rng(5)
    t = rng.uniform(-1, 1, (100, 3, 3))
    t[17, 1, 2] = np.inf
    sc = _soup(t)
    d = bvh.build_device(sc, "sah", 4, 8)
    assert not d["built_on_device"]
    a = bvh.build_host(sc, "sah", 4, 8)
    assert np.array_equal(a["meta"], d["meta"]) and np.array_equal(a["order"], d["order"])


def test_tracer_build_runs_on_device(tracer_factory):
    sc = S.procedural_scene(seed=3, ico_subdivisions=3, bisect=True, sphere_count=40, light_count=3)
    pt = tracer_factory(sc, S.RenderSettings(width=64, height=48, spp=2, max_depth=4))
    info = pt.build_info
    assert info["device_built"] == 1 and info["built"] == 1
    host = bvh.build_host(sc, "sah", 4, 8)
    assert info["node_count"] == len(host["meta"]) and info["max_depth"] == host["max_depth"]

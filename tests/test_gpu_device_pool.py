"""The scene pool packed on the device from trees the GPU build keeps there, against the pool the host
assembles from a host-built tree (PT_BVH_FLAG_HOST_BUILD: the same SAH tree, node for node, through
ptb::pack_device_layout). Every pool word the kernels read is equal when films and all counters are:
node boxes and child references decide the traversal counters, triangle records the hits, primitive
records the shading normal and material, the texture map the textured albedo. Single tracers (an
opaque procedural scene, the feature scene with transmission and an environment map, a textured
sphere, a one-triangle scene whose root is a leaf) and a batch mixing device- and host-built trees."""
import dataclasses

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

COUNTERS = ("samples", "rays", "shadow_tests", "bvh_node_visits", "bvh_leaf_visits", "triangle_tests",
            "triangle_hits")


def _render(scene, settings, host_build, frames=(0,)):
    from vkpathtracer_amd.tracer import HipPathTracer

    with HipPathTracer(0) as pt:
        assert pt.configure(settings) and pt.load_scene_snapshot(scene)
        st = pt.build_or_update_acceleration("sah", 4, 8, host_build=host_build)
        assert st, st.message
        info = dict(pt.build_info)
        tiles = [(0, 0, settings.width, settings.height, s, 0, 0) for s in range(settings.spp)]
        for f in frames:
            assert pt.render_tiles(tiles, f)
        film, cnt = pt.film()
        return film, cnt, pt.read_counters(), info


def _one_triangle():
    from vkpathtracer_amd import scene as S

    sc = S.cornell_box()
    keep = 1  # the first triangle alone: the root is a leaf
    inst = sc.instances[:1].copy()
    inst["triangle_count"] = keep
    return dataclasses.replace(sc, instances=inst, name="one-triangle")


def _scenes():
    from vkpathtracer_amd import scene as S

    yield "procedural", S.procedural_scene(seed=5, ico_subdivisions=4, bisect=True, sphere_count=300, light_count=4)
    yield "features", S.feature_scene()
    yield "textured", S.textured_sphere_scene(seed=11, frequency=10, texture_size=64)
    yield "one triangle", _one_triangle()


@pytest.mark.parametrize("name", ["procedural", "features", "textured", "one triangle"])
def test_device_packed_pool_renders_as_the_host_layout(name):
    from vkpathtracer_amd import scene as S

    sc = dict(_scenes())[name]
    st = S.RenderSettings(width=96, height=64, spp=3, max_depth=6)
    fd, cd, kd, idev = _render(sc, st, host_build=False, frames=(0, 1))
    fh, ch, kh, ihost = _render(sc, st, host_build=True, frames=(0, 1))
    assert idev["device_built"] == 1 and ihost["device_built"] == 0
    for k in ("node_count", "leaf_count", "max_depth", "primitive_count", "device_bytes"):
        assert idev[k] == ihost[k], (name, k, idev[k], ihost[k])
    assert np.array_equal(cd, ch), name
    assert np.array_equal(fd, fh), (name, float(np.abs(fd - fh).max()))
    for k in COUNTERS:
        assert kd[k] == kh[k], (name, k, kd[k], kh[k])
    assert kd["bvh_node_visits"] > 0


def test_batch_pool_mixes_device_and_host_trees():
    """Scenes 0 and 2 built on the host, 1 and 3 on the device (one forest), in one pool: films equal
    the all-host batch's."""
    from vkpathtracer_amd import scene as S
    from vkpathtracer_amd.tracer import HipBatch

    named = dict(_scenes())
    jobs = [(named["features"], 40, 30, 2), (named["procedural"], 48, 32, 3), (named["textured"], 32, 32, 2),
            (S.cornell_box(), 24, 24, 2)]
    settings = S.RenderSettings(width=1, height=1, spp=1, max_depth=6)

    def run(host_of):
        with HipBatch(settings) as b:
            for i, (sc, w, h, spp) in enumerate(jobs):
                b.add_scene(sc, w, h, spp, host_build=host_of(i))
            info = b.build()
            b.render(0)
            b.render(1, use_graph=True)
            return [b.film(i) for i in range(len(jobs))], b.read_counters(), info

    mixed, cm, im = run(lambda i: i % 2 == 0)
    host, ch, ih = run(lambda i: True)
    assert im["device_built"] == 0 and ih["device_built"] == 0  # the pool reports the AND over its scenes
    assert im["node_count"] == ih["node_count"] and im["device_bytes"] == ih["device_bytes"]
    for i, ((fm, nm), (fh, nh)) in enumerate(zip(mixed, host)):
        assert np.array_equal(nm, nh), i
        assert np.array_equal(fm, fh), (i, float(np.abs(fm - fh).max()))
    for k in COUNTERS:
        assert cm[k] == ch[k], k

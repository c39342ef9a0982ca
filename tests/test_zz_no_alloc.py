"""pt_hip_trace_rays_noalloc, the ray query that allocates nothing (its scratch is a static device array
per GPU), against pt_hip_trace_rays on the same tracer: t, primitive ids and counters identical, on
consecutive calls with shrinking inputs (the shared scratch keeps the previous call's contents), for
closest and any hit, and a tree deeper than the static spill area refused before any launch."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _chain_scene(levels: int):
    """Triangles in planes x = 2^k: the midpoint builder splits the widest one off per level, so the
    tree is about `levels` deep."""
    from vkpathtracer_amd import scene as S

    n = levels + 8
    v = []
    for k in range(n):
        x = float(2.0 ** k)
        v += [(x, 0.0, 0.0), (x, 1.0, 0.0), (x, 0.0, 1.0)]
    verts = np.array(v, np.float32)
    idx = np.arange(3 * n, dtype=np.uint32)
    inst = np.zeros(1, S.INSTANCE_DTYPE)
    inst["triangle_count"] = n
    inst["rotation"] = (0.0, 0.0, 0.0, 1.0)
    inst["scale"] = 1.0
    cam = S.Camera(position=(-3.0, 0.5, 0.5), target=(0.0, 0.5, 0.5))
    return S.Scene(verts, idx, inst, S.make_materials(1), np.zeros(0, S.LIGHT_DTYPE), cam, name="chain")


@pytest.mark.parametrize("builder", ["sah", "midpoint"])
def test_noalloc_trace_rays_equals_trace_rays(small_procedural, tracer_factory, builder):
    from vkpathtracer_amd import scene as S

    pt = tracer_factory(small_procedural, S.RenderSettings(width=8, height=8), builder=builder)
    rays = S.random_rays(small_procedural, 150_000, seed=17)
    tmax = np.random.default_rng(3).uniform(0.5, 30.0, len(rays)).astype(np.float32)
    for n in (150_000, 2_500, 0, 63):  # each call smaller than a previous one; an empty call between
        for any_hit in (False, True):
            tm = tmax[:n] if any_hit else None
            print(f"noalloc {builder} n={n} any_hit={any_hit}", flush=True)
            t0, p0, c0 = pt.trace_rays(rays[:n], any_hit=any_hit, tmax=tm)
            t1, p1, c1 = pt.trace_rays(rays[:n], any_hit=any_hit, tmax=tm, noalloc=True)
            assert np.array_equal(p1, p0)
            assert np.array_equal(t1, t0)
            assert c1 == c0
            if n:
                assert c1["bvh_node_visits"] > 0


def test_noalloc_refuses_trees_deeper_than_its_spill_area(tracer_factory):
    from vkpathtracer_amd import scene as S
    from vkpathtracer_amd.tracer import PathTracerError

    pt = tracer_factory(_chain_scene(48), S.RenderSettings(width=8, height=8), builder="midpoint")
    assert pt.build_info["max_depth"] > 32  # PT_NOALLOC_MAX_TREE_DEPTH
    rays = np.array([[-1.0, 0.25, 0.25, 1.0, 0.0, 0.0]] * 5, np.float32)
    _, p0, _ = pt.trace_rays(rays)
    assert (p0 >= 0).all()
    with pytest.raises(PathTracerError) as e:
        pt.trace_rays(rays, noalloc=True)
    assert "deeper" in str(e.value) and "unsupported" in str(e.value)

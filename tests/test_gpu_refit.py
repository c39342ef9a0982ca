"""Dynamic-instance refit on the GPU (pt_hip_apply_instance_transform_update) against the oracle's
restatement of CpuBvhAccelerator's refit: after each round of instance moves, ray queries (t, primitive,
traversal counters) and rendered films (sums, counts, counters) are identical to the oracle's on its
refit tree; plans are pure queries with the reference's statuses; a full rebuild afterwards sees the
moved instances."""
import dataclasses

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

COUNTERS = ("samples", "rays", "shadow_tests", "bvh_node_visits", "bvh_leaf_visits", "triangle_tests",
            "triangle_hits")


def _scene():
    from vkpathtracer_amd import scene as S

    return S.procedural_scene(seed=4, ico_subdivisions=4, bisect=False, sphere_count=48, light_count=3)


def _updates(sc, seed, k=8, by_entity=False):
    from vkpathtracer_amd import scene as S

    rng = np.random.default_rng(seed)
    dyn = np.nonzero(sc.instances["flags"] & S.INSTANCE_FLAG_DYNAMIC_TRANSFORM)[0]
    out = []
    for i in rng.choice(dyn, k, replace=False):
        u = S.InstanceTransformUpdate(translation=tuple(rng.uniform(-5, 5, 3)), rotation=tuple(rng.normal(size=4)),
                                      scale=tuple(rng.uniform(0.4, 1.4, 3)), instance_index=int(i))
        if by_entity:
            u = dataclasses.replace(u, instance_index=0xFFFFFFFF, entity_id=int(sc.instances["entity_id"][i]))
        out.append(u)
    return out


def _tracer(sc, st, builder="sah"):
    from vkpathtracer_amd.tracer import HipPathTracer

    pt = HipPathTracer(0)
    assert pt.configure(st) and pt.load_scene_snapshot(sc) and pt.build_or_update_acceleration(builder, 4, 8)
    return pt


@pytest.mark.parametrize("builder", ["sah", "midpoint"])
def test_refit_rounds_match_oracle(builder):
    from oracle import oracle as O
    from vkpathtracer_amd import scene as S

    sc = _scene()
    st = S.RenderSettings(width=48, height=32, spp=2, max_depth=5)
    pt = _tracer(sc, st, builder)
    orc = O.Oracle(sc, builder=builder)
    try:
        cur = sc
        rounds = [_updates(sc, 1), _updates(sc, 2, by_entity=True), _updates(sc, 3, k=3) + _updates(sc, 3, k=3)]
        for r, ups in enumerate(rounds):
            print(f"refit round {r} ({builder})", flush=True)
            status, applied, ms = pt.apply_instance_transform_update(ups)
            assert status == "applied_dynamic_accel_update" and applied == len(ups), status
            cur = S.with_instance_transforms(cur, ups)
            orc.refit(cur)
            rays = S.random_rays(cur, 100_000, seed=20 + r)
            t, p, c = pt.trace_rays(rays)
            ot, op, oc = orc.trace_rays(rays)
            assert np.array_equal(t, ot) and np.array_equal(p, op)
            for k in ("bvh_node_visits", "bvh_leaf_visits", "triangle_tests", "triangle_hits"):
                assert c[k] == oc[k], (k, c[k], oc[k])
            assert pt.reset_accumulation() and pt.render_frame_samples(0, st.spp)
            film, cnt = pt.film()
            counters = pt.read_counters()
            rf, rc, rcnt = orc.render(st)
            assert np.array_equal(cnt, rc) and np.array_equal(film, rf), float(np.abs(film - rf).max())
            for k in COUNTERS:
                assert counters[k] == rcnt[k], (k, counters[k], rcnt[k])
        # a full rebuild sees the moved instances
        assert pt.build_or_update_acceleration(builder, 4, 8)
        assert pt.reset_accumulation() and pt.render_frame_samples(0, st.spp)
        film, _ = pt.film()
        rf, _, _ = O.Oracle(cur, builder=builder).render(st)
        assert np.array_equal(film, rf)
    finally:
        pt.close()


def test_plan_statuses_and_purity():
    from vkpathtracer_amd import scene as S

    sc = _scene()
    st = S.RenderSettings(width=16, height=16, spp=1, max_depth=4)
    pt = _tracer(sc, st)
    try:
        ups = _updates(sc, 5, k=4)
        assert pt.plan_instance_transform_update(ups) == ("applied_dynamic_accel_update", 4)
        assert pt.plan_instance_transform_update(ups) == ("applied_dynamic_accel_update", 4)
        static = int(np.nonzero((sc.instances["flags"] & S.INSTANCE_FLAG_DYNAMIC_TRANSFORM) == 0)[0][0])
        blocked = ups[:2] + [S.InstanceTransformUpdate(instance_index=static)]
        assert pt.plan_instance_transform_update(blocked) == ("blocked_needs_full_static_accel_rebuild", 2)
        assert pt.apply_instance_transform_update(blocked)[:2] == ("blocked_needs_full_static_accel_rebuild", 0)
        unknown = [S.InstanceTransformUpdate(instance_index=10**6, entity_id=987654321)]
        assert pt.plan_instance_transform_update(unknown) == ("failed", 0)
        rays = S.random_rays(sc, 20000, seed=3)
        t0, p0, _ = pt.trace_rays(rays)
        pt.plan_instance_transform_update(ups)
        t1, p1, _ = pt.trace_rays(rays)
        assert np.array_equal(t0, t1) and np.array_equal(p0, p1)
        assert not pt.update_instance_transforms(blocked)
        assert pt.update_instance_transforms(ups)
    finally:
        pt.close()


def test_refit_cornell_document_cuboid():
    """Move the glossy cuboid of the reference's Cornell document (a physics-driven dynamic instance):
    the render equals the oracle's after its refit."""
    import pathlib

    from oracle import oracle as O
    from vkpathtracer_amd import scene as S
    from vkpathtracer_amd import scene_document as SD

    sc = SD.load_scene_document(pathlib.Path(__file__).parent / "golden" / "scenes" / "cornell_native.json")
    st = S.RenderSettings(width=40, height=40, spp=3, max_depth=6)
    pt = _tracer(sc, st)
    orc = O.Oracle(sc)
    try:
        u = S.InstanceTransformUpdate(translation=(0.2, 0.5, 0.1), rotation=(0.0, 0.3826834, 0.0, 0.9238795),
                                      scale=(0.55, 0.45, 0.45), entity_id=11)
        assert pt.update_instance_transforms([u])
        moved = S.with_instance_transforms(sc, [u])
        orc.refit(moved)
        assert pt.reset_accumulation() and pt.render_frame_samples(0, st.spp)
        film, cnt = pt.film()
        rf, rc, _ = orc.render(st)
        assert np.array_equal(film, rf) and np.array_equal(cnt, rc)
    finally:
        pt.close()

"""Progressive and interactive entry points on the GPU: camera and render-settings updates without a
scene re-upload render exactly what a freshly loaded tracer renders, for a tracer and for a batch
(eager and graph replays); the split film readback returns the LDR resolve of the film as it stood at
the request while later rendering proceeds, and only the latest token is valid."""
import dataclasses
import time

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _fresh(scene, st):
    from vkpathtracer_amd.tracer import HipPathTracer

    pt = HipPathTracer(0)
    assert pt.configure(st) and pt.load_scene_snapshot(scene) and pt.build_or_update_acceleration("sah", 4, 8)
    return pt


def _render(pt, spp, frame=0):
    assert pt.reset_accumulation()
    assert pt.render_frame_samples(0, spp, frame)
    return pt.film()


def test_update_camera_equals_reload(cornell):
    from vkpathtracer_amd import scene as S

    st = S.RenderSettings(width=40, height=32, spp=3, max_depth=6)
    cam2 = dataclasses.replace(cornell.camera, position=(0.3, 1.1, 3.4), target=(0.0, 0.9, 0.0), fov_deg=47.0,
                               aperture_radius=0.02, focus_distance=3.0)
    pt = _fresh(cornell, st)
    ref = _fresh(dataclasses.replace(cornell, camera=cam2), st)
    try:
        a0 = _render(pt, st.spp)
        assert pt.update_camera_state(cam2)
        a1 = _render(pt, st.spp)
        b1 = _render(ref, st.spp)
        assert np.array_equal(a1[0], b1[0]) and np.array_equal(a1[1], b1[1])
        assert not np.array_equal(a0[0], a1[0])
        assert pt.update_camera(cornell.camera.position, cornell.camera.target, cornell.camera.up,
                                cornell.camera.fov_deg)
        # pose and fov back, lens of cam2 kept
        back = _fresh(dataclasses.replace(cornell, camera=dataclasses.replace(
            cam2, position=cornell.camera.position, target=cornell.camera.target, up=cornell.camera.up,
            fov_deg=cornell.camera.fov_deg)), st)
        try:
            assert np.array_equal(_render(pt, st.spp)[0], _render(back, st.spp)[0])
        finally:
            back.close()
    finally:
        pt.close()
        ref.close()


def test_update_render_settings(cornell):
    from vkpathtracer_amd import scene as S

    st = S.RenderSettings(width=32, height=24, spp=2, max_depth=6)
    st2 = dataclasses.replace(st, seed=12345, max_depth=3, spp=3, enable_mis=False, camera_aperture_radius=0.03,
                              camera_focus_distance=2.5)
    pt = _fresh(cornell, st)
    ref = _fresh(cornell, st2)
    try:
        _render(pt, st.spp)
        assert pt.update_render_settings(st2)
        assert np.array_equal(_render(pt, st2.spp)[0], _render(ref, st2.spp)[0])
        assert not pt.update_render_settings(dataclasses.replace(st2, width=64))
        assert pt.status().ready_to_render
    finally:
        pt.close()
        ref.close()


@pytest.mark.parametrize("use_graph", [False, True])
def test_batch_update_camera(cornell, use_graph):
    from vkpathtracer_amd import scene as S
    from vkpathtracer_amd.tracer import HipBatch

    st = S.RenderSettings(width=1, height=1, spp=1, max_depth=6)
    cam2 = dataclasses.replace(cornell.camera, position=(-0.4, 1.2, 3.2), fov_deg=50.0)
    with HipBatch(st) as b:
        b.add_scene(cornell, 32, 24, 2)
        b.add_view(0, cornell.camera, 16, 16, 1)
        b.build()
        b.render(0, use_graph=use_graph)
        b.update_camera(1, cam2)
        b.update_camera(0, cam2)
        b.reset_accumulation()
        b.render(0, use_graph=use_graph)
        for i, (w, h, spp) in enumerate([(32, 24, 2), (16, 16, 1)]):
            ref = _fresh(dataclasses.replace(cornell, camera=cam2), dataclasses.replace(st, width=w, height=h, spp=spp))
            try:
                want = _render(ref, spp)
            finally:
                ref.close()
            film, cnt = b.film(i)
            assert np.array_equal(film, want[0]) and np.array_equal(cnt, want[1]), i


def _poll(pt, tok, timeout=10.0):
    t0 = time.time()
    while True:
        state, img = pt.poll_film(tok)
        if state != "pending" or time.time() - t0 > timeout:
            return state, img
        time.sleep(0.001)


def test_split_film_readback(cornell):
    from vkpathtracer_amd import scene as S

    st = S.RenderSettings(width=48, height=36, spp=4, max_depth=6)
    r = S.FilmResolve(tone_map="reinhard", exposure=1.3)
    pt = _fresh(cornell, st)
    try:
        _render(pt, st.spp)
        want = pt.resolve_ldr(r)
        tok = pt.request_film_readback(r)
        assert tok.id != 0 and (tok.width, tok.height) == (48, 36)
        # more samples queued after the request do not reach the requested image
        assert pt.render_frame_samples(4, 4)
        state, img = _poll(pt, tok)
        assert state == "ready" and np.array_equal(img, want)
        assert not np.array_equal(pt.resolve_ldr(r), want)
        tok2 = pt.request_film_readback(r)
        assert pt.poll_film(tok)[0] == "invalid"
        state, img = _poll(pt, tok2)
        assert state == "ready" and np.array_equal(img, pt.resolve_ldr(r))
        from vkpathtracer_amd import abi
        assert pt.poll_film(abi.PtFilmReadbackToken())[0] == "invalid"
    finally:
        pt.close()


def test_scene_delta_update_follows_the_reference(cornell):
    """pt_hip_apply_scene_delta restates ScalarCpuPathTracer::update_scene_delta (ApplySceneDeltaUpdate,
    PathTracer.cpp:427-482): the environment colour and per-index material and light records reach the
    next render; instance materials are baked into the accelerator's primitives, so they reach renders
    only after the next build (CpuBvhAccelerator.cpp:232-238); an out-of-range index changes nothing and
    an update that changes nothing is refused. Every film equals the oracle's render of the snapshot the
    reference would be rendering."""
    from oracle import oracle as O
    from vkpathtracer_amd import abi, scene as S

    st = S.RenderSettings(width=40, height=32, spp=2, max_depth=6)

    def oracle_film(sc):
        return O.Oracle(sc).render(st)[0]

    pt = _fresh(cornell, st)
    try:
        env = (0.3, 0.2, 0.1)
        sc_env = dataclasses.replace(cornell, environment_color=env)
        assert pt.apply_scene_delta(environment_color=env)
        assert np.array_equal(_render(pt, st.spp)[0], oracle_film(sc_env))

        inst = sc_env.instances.copy()
        inst[6]["material_index"] = 1  # named by index
        eid = int(inst[7]["entity_id"])
        inst[7]["material_index"] = 2  # named by entity id
        sc_inst = dataclasses.replace(sc_env, instances=inst)
        assert pt.apply_scene_delta(instance_materials=[(6, 1), (eid, abi.INVALID_INSTANCE_INDEX, 2)])
        assert np.array_equal(_render(pt, st.spp)[0], oracle_film(sc_env))  # baked at the last build
        assert pt.build_or_update_acceleration("sah", 4, 8)
        film_inst = _render(pt, st.spp)[0]
        assert np.array_equal(film_inst, oracle_film(sc_inst))

        m = sc_inst.materials.copy()
        m[0]["albedo"] = (0.2, 0.6, 0.9)
        lights = sc_inst.lights.copy()
        lights[0]["intensity"] *= 0.5
        sc_ml = dataclasses.replace(sc_inst, materials=m, lights=lights)
        assert pt.apply_scene_delta(materials=[(0, m[0])], lights=[(0, lights[0])])
        film_ml = _render(pt, st.spp)[0]
        assert np.array_equal(film_ml, oracle_film(sc_ml))

        # refused, and nothing applied: a bad index anywhere in the update, or nothing to change
        assert not pt.apply_scene_delta(materials=[(0, cornell.materials[0])], lights=[(len(lights), lights[0])])
        assert not pt.apply_scene_delta(instance_materials=[(6, len(m))])
        assert not pt.apply_scene_delta(instance_materials=[(10_000, 1)])
        assert not pt.apply_scene_delta(instance_materials=[(6, 1)])  # already 1: no change
        assert "did not apply" in pt.status().last_error
        assert not pt.apply_scene_delta()
        assert np.array_equal(_render(pt, st.spp)[0], film_ml)
    finally:
        pt.close()

"""Batched rendering: scenes of different sizes, images and spp through one scene pool and one
launch sequence give each scene exactly the film (and, for opaque scenes, the counters) that a
tracer holding only that scene renders, with or without HIP-graph replay and across pass splits."""
import dataclasses

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

COUNTERS = ("samples", "rays", "shadow_tests", "bvh_node_visits", "bvh_leaf_visits", "triangle_tests",
            "triangle_hits")


def _single(scene, w, h, spp, settings, frames=(0,), leaf=4):
    from vkpathtracer_amd.tracer import HipPathTracer

    st = dataclasses.replace(settings, width=w, height=h, spp=spp)
    with HipPathTracer(0) as pt:
        assert pt.configure(st) and pt.load_scene_snapshot(scene) and pt.build_or_update_acceleration("sah", leaf, 8)
        for f in frames:
            assert pt.render_tiles([(0, 0, w, h, s, 0, 0) for s in range(spp)], f)
        film, cnt = pt.film()
        return film, cnt, pt.read_counters()


def _jobs(cornell, small_procedural):
    from vkpathtracer_amd import scene as S

    sp2 = S.procedural_scene(seed=9, ico_subdivisions=3, bisect=False, sphere_count=20, light_count=3)
    return [(cornell, 64, 48, 3, 4), (small_procedural, 50, 40, 2, 2), (sp2, 33, 17, 5, 8), (cornell, 16, 16, 1, 4)]


@pytest.mark.parametrize("use_graph", [False, True])
def test_batch_films_and_counters_equal_single_scene_renders(cornell, small_procedural, use_graph):
    from vkpathtracer_amd import bvh, scene as S
    from vkpathtracer_amd.tracer import HipBatch

    settings = S.RenderSettings(width=1, height=1, spp=1, max_depth=8)
    jobs = _jobs(cornell, small_procedural)
    frames = (0, 1)
    with HipBatch(settings) as b:
        for sc, w, h, spp, leaf in jobs:
            b.add_scene(sc, w, h, spp, leaf_size=leaf)
        info = b.build()
        # the scenes' trees are built together on the device (one level-synchronous build, each scene
        # with its own leaf size): together they hold the host builder's nodes, leaves and depth
        host = [bvh.build_host(sc, "sah", leaf, 8) for sc, *_, leaf in jobs]
        assert info["device_built"] == 1
        assert info["node_count"] == sum(len(h["meta"]) for h in host)
        assert info["leaf_count"] == sum(int((h["meta"][:, 3] > 0).sum()) for h in host)
        assert info["max_depth"] == max(h["max_depth"] for h in host)
        for f in frames:
            b.render(f, use_graph=use_graph)
        total = {k: 0 for k in COUNTERS}
        for i, (sc, w, h, spp, leaf) in enumerate(jobs):
            film, cnt = b.film(i)
            rf, rc, rcnt = _single(sc, w, h, spp, settings, frames, leaf)
            assert np.array_equal(cnt, rc), i
            assert np.array_equal(film, rf), (i, float(np.abs(film - rf).max()))
            for k in COUNTERS:
                total[k] += rcnt[k]
        got = b.read_counters()
        for k in COUNTERS:
            assert got[k] == total[k], (k, got[k], total[k])


def test_batch_with_transmissive_scene_and_pass_splits(cornell, small_procedural):
    """A scene with transmissive materials and environment map in the pool, and a capacity small
    enough to split the frame into many passes: films still equal the single-scene renders."""
    from vkpathtracer_amd import scene as S
    from vkpathtracer_amd.tracer import HipBatch

    settings = S.RenderSettings(width=1, height=1, spp=1, max_depth=6)
    jobs = [(S.feature_scene(), 40, 30, 3, 4)] + _jobs(cornell, small_procedural)[:2]
    with HipBatch(settings) as b:
        for sc, w, h, spp, leaf in jobs:
            b.add_scene(sc, w, h, spp, leaf_size=leaf)
        b.build()
        b.set_tuning(wavefront_capacity=3000, refill_min=4)
        b.render(0)
        b.render(0, use_graph=True)
        for i, (sc, w, h, spp, leaf) in enumerate(jobs):
            film, cnt = b.film(i)
            rf, rc, _ = _single(sc, w, h, spp, settings, (0, 0), leaf)
            assert np.array_equal(cnt, rc), i
            assert np.array_equal(film, rf), (i, float(np.abs(film - rf).max()))


def test_batch_graph_replays_match_eager_frames(cornell, small_procedural):
    """Several graph replays with eager frames and synchronisations between them: every replay adds
    exactly what an eager render of the same frame adds."""
    from vkpathtracer_amd import scene as S
    from vkpathtracer_amd.tracer import HipBatch

    settings = S.RenderSettings(width=1, height=1, spp=1, max_depth=8)
    jobs = _jobs(cornell, small_procedural)

    def run(graph_frames):
        with HipBatch(settings) as b:
            for sc, w, h, spp, leaf in jobs:
                b.add_scene(sc, w, h, spp, leaf_size=leaf)
            b.build()
            films = []
            for f, g in graph_frames:
                b.reset_accumulation()
                b.render(f, use_graph=g)
                b.synchronize()
                films.append([b.film(i)[0] for i in range(len(jobs))])
            return films

    eager = run([(f, False) for f in (3, 4, 3, 5)])
    mixed = run([(3, True), (4, False), (3, True), (5, True)])
    for fa, fb in zip(eager, mixed):
        for a, c in zip(fa, fb):
            assert np.array_equal(a, c)


def test_autotuned_batch_renders_the_same_films(cornell, small_procedural):
    """The autotuner picks a leaf width and launch shapes by timing; whatever it picks, every film
    equals that of a batch with the same leaf width and default launch shapes."""
    from vkpathtracer_amd import scene as S
    from vkpathtracer_amd.autotune import tune_batch
    from vkpathtracer_amd.tracer import HipBatch

    settings = S.RenderSettings(width=1, height=1, spp=1, max_depth=6)
    jobs = [(sc, w, h, spp) for sc, w, h, spp, _ in _jobs(cornell, small_procedural)]
    res, tb = tune_batch(jobs, settings, leaf_sizes=(2, 4), reps=1)
    print(res.params, f"{res.default_ms:.2f} -> {res.ms:.2f} ms over {len(res.trials)} trials")
    assert res.params["leaf_size"] in (2, 4) and res.ms <= res.default_ms
    tb.reset_accumulation()
    tb.render(0)
    with HipBatch(settings) as b:
        for sc, w, h, spp in jobs:
            b.add_scene(sc, w, h, spp, leaf_size=res.params["leaf_size"])
        b.build()
        b.render(0)
        for i in range(len(jobs)):
            assert np.array_equal(tb.film(i)[0], b.film(i)[0]), i
    tb.close()


def test_graph_replay_invalidates_reused_forward_radiance(cornell):
    """A graph captured on frame 1, an eager render of frame 0, a replay of the graph for frame 2, then
    the adjoint of frame 0: the replay rewrote every path's radiance, so the adjoint must replay frame 0
    itself. Its gradients equal those of a batch
    that only runs the adjoint of frame 0 (up to the order of the fp32 atomic adds)."""
    import torch

    from vkpathtracer_amd import scene as S
    from vkpathtracer_amd.tracer import HipBatch

    settings = S.RenderSettings(width=1, height=1, spp=1, max_depth=8)
    w, h, spp = 40, 32, 2
    nm, nl = len(cornell.materials), len(cornell.lights)
    dev = torch.device("cuda", 0)
    fg = torch.from_numpy(np.random.default_rng(11).normal(size=(h * w, 3)).astype(np.float32)).to(dev)

    def grads(sequence):
        with HipBatch(settings) as b:
            b.add_scene(cornell, w, h, spp)
            b.build()
            for f, g in sequence:
                b.render(f, use_graph=g)
            b.synchronize()
            out = [torch.zeros((n, 3), device=dev) for n in (nm, nm, nl)]
            b.render_adjoint(fg, *out, frame_index=0)
            return [x.cpu().numpy().astype(np.float64) for x in out]

    ref = grads([])
    got = grads([(1, True), (0, False), (2, True)])
    for a, r in zip(got, ref):
        assert np.abs(a - r).max() <= 1e-5 * max(np.abs(r).max(), 1e-6)

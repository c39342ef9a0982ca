"""SDF primitives on the GPU (k_sdf_extend, the SDF shadow loop, the SDF shading branch) against the
oracle's intersect_scene: films, sample counts and every counter (the sdf_* ones included) identical for
a Cornell box holding all six shapes (scaled, rotated, glass, emissive, unknown), for the reference's
SDF documents, and for a batch mixing SDF and triangle-only scenes."""
import math
import pathlib

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

COUNTERS = ("samples", "rays", "shadow_tests", "bvh_node_visits", "triangle_tests", "sdf_tests", "sdf_steps",
            "sdf_hits", "sdf_misses")
GOLDEN = pathlib.Path(__file__).resolve().parent / "golden" / "scenes"


def _sdf_cornell():
    from vkpathtracer_amd import scene as S

    sc = S.cornell_box()
    mats = np.concatenate([sc.materials, S.make_materials(3)])
    glass, lamp, gloss = len(sc.materials), len(sc.materials) + 1, len(sc.materials) + 2
    mats[glass]["material_model"] = 5
    mats[glass]["transmission"] = 1.0
    mats[lamp]["emissive"] = (3.0, 2.5, 2.0)
    mats[gloss]["roughness"] = 0.2
    lo, hi = sc.vertices.min(0), sc.vertices.max(0)
    c = (lo + hi) / 2
    w = float((hi - lo).max())
    sdfs = [S.make_sdf("sphere", (c[0] - 0.25 * w, lo[1] + 0.15 * w, c[2]), scale=(1.5, 1, 1), rotation=(0, 0.7, 0),
                       radius=0.08 * w, material_index=glass),
            S.make_sdf("box", (c[0] + 0.2 * w, lo[1] + 0.1 * w, c[2] - 0.1 * w), scale=(0.3, 0.3, 0.3),
                       rotation=(0.2, 0.5, 0.1), material_index=gloss),
            S.make_sdf("rounded_box", (c[0], lo[1] + 0.35 * w, c[2] + 0.15 * w), scale=(0.25, 0.25, 0.25), param_a=0.03,
                       material_index=1),
            S.make_sdf("torus", (c[0] + 0.1 * w, c[1] + 0.2 * w, c[2]), scale=(0.1 * w, 0.1 * w, 0.1 * w),
                       rotation=(math.pi / 3, 0, 0), param_a=1.0, radius=0.3, material_index=gloss),
            S.make_sdf("capsule", (c[0] - 0.3 * w, c[1] + 0.25 * w, c[2] - 0.2 * w), scale=(0.05 * w,) * 3,
                       param_a=1.0, radius=0.5, material_index=lamp),
            S.make_sdf("plane", (0, 0, 0), param_b=-float(lo[1]) - 0.02 * w, material_index=0),
            S.make_sdf("unknown", tuple(c))]
    sc.materials = mats
    sc.sdf_primitives = np.array(sdfs, S.SDF_DTYPE)
    return sc


def _check(sc, st, tracer_factory):
    from oracle import oracle as O

    pt = tracer_factory(sc, st)
    assert pt.render_frame_samples(0, st.spp), pt.status().last_error
    film, cnt = pt.film()
    got = pt.read_counters()
    rf, rc, want = O.Oracle(sc).render(st)
    assert np.array_equal(cnt, rc)
    assert np.array_equal(film, rf), f"max |diff| {np.abs(film - rf).max()}"
    for k in COUNTERS:
        assert got[k] == want[k], (k, got[k], want[k])
    assert want["sdf_hits"] > 0


def test_all_shapes_in_cornell_bitexact(tracer_factory):
    from vkpathtracer_amd import scene as S

    _check(_sdf_cornell(), S.RenderSettings(width=48, height=40, spp=3, max_depth=6), tracer_factory)


@pytest.mark.parametrize("name", ["sdf_complexity.json", "glass_lab.json"])
def test_sdf_documents_bitexact(name, tracer_factory):
    from vkpathtracer_amd import scene as S
    from vkpathtracer_amd import scene_document as SD

    sc = SD.load_scene_document(GOLDEN / name)
    assert len(sc.sdf_primitives) > 0
    _check(sc, S.RenderSettings(width=40, height=32, spp=3, max_depth=6), tracer_factory)


@pytest.mark.parametrize("use_graph", [False, True])
def test_batch_mixing_sdf_and_triangle_scenes(use_graph):
    from oracle import oracle as O
    from vkpathtracer_amd import scene as S
    from vkpathtracer_amd.tracer import HipBatch

    jobs = [(_sdf_cornell(), 32, 24, 2), (S.cornell_box(), 24, 16, 1), (S.feature_scene(), 20, 20, 2)]
    with HipBatch(S.RenderSettings(width=1, height=1, spp=1, max_depth=6)) as b:
        for sc, w, h, spp in jobs:
            b.add_scene(sc, w, h, spp)
        b.build()
        b.render(0, use_graph=use_graph)
        for i, (sc, w, h, spp) in enumerate(jobs):
            film, cnt = b.film(i)
            rf, rc, _ = O.Oracle(sc).render(S.RenderSettings(width=w, height=h, spp=spp, max_depth=6))
            assert np.array_equal(cnt, rc) and np.array_equal(film, rf), i

"""The batched and 1M-triangle benchmark workloads at their full sizes, against the oracle.

* Batched forward: the 64 seeded scenes of scene.batch_scenes(seed=7, count=64) (4.8M triangles,
  images 128^2 to 1024^2, 4/16/64 spp, depth 8) in one batch with the autotuned leaf width and launch
  shapes, rendered eagerly and by HIP-graph replay. On 48 seeded pixels of every scene (its first and
  last pixel among them) the film sums and counts equal the fp32 oracle's accumulation of the same
  samples bit for bit, and the batch is scored against the fp64 oracle on the same samples.
* The 1M-triangle 1920x1080 frame at 64 spp (one interleaved wavefront pass): on 1,500 seeded pixels
  the film equals the fp32 oracle's bit for bit.
"""
import numpy as np
import pytest

from conftest import relrmse

pytestmark = pytest.mark.gpu

PIXELS_PER_SCENE = 48
C2_FP64_BOUND = 1.68e-3  # measured 1.60e-3 on these seeded samples, plus 5 %


def _subset(w, h, k, seed):
    px = np.random.default_rng(seed).choice(w * h, k, replace=False)
    return np.unique(np.concatenate([px, [0, w * h - 1]])).astype(np.uint32)


def _oracle_sums(orc, st, px, spp, dtype):
    """The film a sequence of render_tile calls accumulates at these pixels: samples 0..spp-1 added
    in order (non-finite samples dropped), in the oracle's real type."""
    acc = np.zeros((len(px), 3), dtype)
    cnt = np.zeros(len(px), np.uint32)
    for s in range(spp):
        rgb, valid, _ = orc.render_pixels(st, px, s)
        rgb = np.asarray(rgb, dtype)
        acc[valid] = acc[valid] + rgb[valid]
        cnt += valid
    return acc, cnt


def test_batched_64_scenes_bitexact_vs_oracle():
    import dataclasses

    from oracle import oracle as O
    from vkpathtracer_amd import autotune
    from vkpathtracer_amd import scene as S
    from vkpathtracer_amd.tracer import HipBatch

    jobs = S.batch_scenes(seed=7, count=64)
    tuned = autotune.tuned_params("batch64")
    leaf = tuned["leaf_size"]
    settings = S.RenderSettings(width=1, height=1, spp=1, max_depth=8)
    print(f"C2: {len(jobs)} scenes, {sum(sc.triangle_count for sc, *_ in jobs)} triangles, "
          f"{sum(w * h * spp for _, w, h, spp in jobs)} paths, leaf {leaf}", flush=True)
    with HipBatch(settings) as b:
        for sc, w, h, spp in jobs:
            b.add_scene(sc, w, h, spp, leaf_size=leaf)
        b.build()
        autotune.apply_params(b, tuned)
        films = {}
        for use_graph in (False, True, True):
            b.reset_accumulation()
            b.render(0, use_graph=use_graph)
            b.synchronize()
            cur = [b.film(i) for i in range(len(jobs))]
            if "eager" not in films:
                films["eager"] = cur
            else:  # every graph replay adds exactly what the eager frame added
                for (fa, ca), (fb, cb) in zip(films["eager"], cur):
                    assert np.array_equal(ca, cb) and np.array_equal(fa, fb)
        print("C2: eager frame and two graph replays rendered", flush=True)
    gpu64, ref64 = [], []
    for i, (sc, w, h, spp) in enumerate(jobs):
        st = dataclasses.replace(settings, width=w, height=h, spp=spp)
        px = _subset(w, h, PIXELS_PER_SCENE, 1000 + i)
        film, cnt = films["eager"][i]
        g = film.reshape(-1, 3)[px]
        gc = cnt.reshape(-1)[px]
        acc, rc = _oracle_sums(O.Oracle(sc, leaf_size=leaf), st, px, spp, np.float32)
        assert np.array_equal(gc, rc), i
        assert np.array_equal(g, acc), (i, float(np.abs(g - acc).max()))
        acc64, _ = _oracle_sums(O.Oracle(sc, leaf_size=leaf, variant="f64"), st, px, spp, np.float64)
        n = np.maximum(gc, 1)[:, None].astype(np.float64)
        gpu64.append(g.astype(np.float64) / n)
        ref64.append(acc64 / n)
    err = relrmse(np.concatenate(gpu64), np.concatenate(ref64))
    print(f"C2: films bit-identical to the fp32 oracle on {PIXELS_PER_SCENE}+ pixels of each of {len(jobs)} "
          f"scenes; relative RMSE vs the fp64 oracle {err:.4e}", flush=True)
    assert err <= C2_FP64_BOUND


def test_procedural_1080p_64spp_bitexact_on_seeded_pixels(tracer_factory):
    from oracle import oracle as O
    from vkpathtracer_amd import scene as S

    sc = S.procedural_scene(seed=1)
    st = S.RenderSettings(width=1920, height=1080, spp=64, max_depth=8)
    pt = tracer_factory(sc, st)
    assert pt.render_frame_samples(0, st.spp), pt.status().last_error
    film, cnt = pt.film()
    px = np.sort(np.random.default_rng(7).choice(st.width * st.height, 1500, replace=False)).astype(np.uint32)
    acc, rc = _oracle_sums(O.Oracle(sc), st, px, st.spp, np.float32)
    assert np.array_equal(cnt.reshape(-1)[px], rc)
    assert np.array_equal(film.reshape(-1, 3)[px], acc)
    print("1M-triangle 1080p x 64 spp: 1,500 seeded pixels bit-identical to the fp32 oracle", flush=True)

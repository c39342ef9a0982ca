"""Film resolve on the GPU (pt_hip_resolve_hdr / pt_hip_resolve_ldr and the batch variants) against the
oracle's restatement of FilmBuffer::resolve_hdr and ApplyFilmResolve: bit-identical images for every
tone map, output transform and clamp setting, on rendered films and on injected films holding
non-finite values, zero counts and values far outside [0, 1]."""
import itertools

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _tracer(scene, w, h, spp=4):
    from vkpathtracer_amd import scene as S
    from vkpathtracer_amd.tracer import HipPathTracer

    st = S.RenderSettings(width=w, height=h, spp=spp, max_depth=6)
    pt = HipPathTracer(0)
    assert pt.configure(st) and pt.load_scene_snapshot(scene) and pt.build_or_update_acceleration("sah", 4, 8)
    return pt


def _oracle_ldr(film, cnt, r):
    from oracle import oracle as O

    return O.resolve_ldr(film, cnt, r.exposure, r.white_balance_kelvin, r.tone_map, r.output_transform, r.gamma,
                         r.clamp_output)


SETTINGS = [dict(tone_map=tm, output_transform=ot, clamp_output=cl, exposure=ex, white_balance_kelvin=wb, gamma=g)
            for (tm, ot, cl), (ex, wb, g) in itertools.product(
                itertools.product(("linear", "reinhard", "filmic_approx", "aces_approx"), ("gamma", "linear"),
                                  (True, False)),
                [(1.0, 6500.0, 2.2), (0.37, 3100.0, 1.8)])]


def test_resolve_rendered_film_bitexact():
    from oracle import oracle as O
    from vkpathtracer_amd import scene as S

    pt = _tracer(S.feature_scene(), 64, 48)
    try:
        assert pt.render_frame_samples(0, 4)
        film, cnt = pt.film()
        np.testing.assert_array_equal(pt.resolve_hdr(), O.resolve_hdr(film, cnt))
        for kw in SETTINGS:
            r = S.FilmResolve(**kw)
            got = pt.resolve_ldr(r)
            want = _oracle_ldr(film, cnt, r)
            assert np.array_equal(got, want), (kw, int(np.abs(got.astype(int) - want).max()))
    finally:
        pt.close()


def test_resolve_injected_film_edge_values():
    """replace_film_history puts arbitrary sums and counts in the film: NaN, infinities, negatives,
    huge values, zero counts and uneven counts resolve exactly as the oracle resolves them."""
    from oracle import oracle as O
    from vkpathtracer_amd import scene as S

    w, h = 37, 23
    rng = np.random.default_rng(5)
    film = (rng.standard_normal((h, w, 3)) * rng.choice([0.01, 1.0, 30.0, 1e30], (h, w, 1))).astype(np.float32)
    film.reshape(-1, 3)[rng.choice(w * h, 40, replace=False)] = np.nan
    film.reshape(-1, 3)[rng.choice(w * h, 40, replace=False)] = np.inf
    film.reshape(-1, 3)[rng.choice(w * h, 40, replace=False)] = -np.inf
    cnt = rng.integers(0, 7, (h, w)).astype(np.uint32)
    pt = _tracer(S.cornell_box(), w, h)
    try:
        pt.replace_film_history(film, cnt)
        got_film, got_cnt = pt.film()
        assert np.array_equal(got_film, film, equal_nan=True) and np.array_equal(got_cnt, cnt)
        np.testing.assert_array_equal(pt.resolve_hdr(), O.resolve_hdr(film, cnt))
        for kw in SETTINGS:
            r = S.FilmResolve(**kw)
            assert np.array_equal(pt.resolve_ldr(r), _oracle_ldr(film, cnt, r)), kw
    finally:
        pt.close()


def test_resolve_into_device_memory():
    """A device tensor as the output receives the same image as host memory."""
    import torch

    from vkpathtracer_amd import scene as S

    pt = _tracer(S.cornell_box(), 40, 30)
    try:
        assert pt.render_frame_samples(0, 2)
        r = S.FilmResolve(tone_map="aces_approx", exposure=1.7)
        dev = torch.zeros((30, 40, 4), dtype=torch.uint8, device="cuda")
        pt.resolve_ldr(r, out=dev)
        assert np.array_equal(dev.cpu().numpy(), pt.resolve_ldr(r))
        devh = torch.zeros((30, 40, 3), dtype=torch.float32, device="cuda")
        pt.resolve_hdr(out=devh)
        assert np.array_equal(devh.cpu().numpy(), pt.resolve_hdr())
    finally:
        pt.close()


def test_batch_resolve_matches_tracer():
    from vkpathtracer_amd import scene as S
    from vkpathtracer_amd.tracer import HipBatch

    scenes = [(S.cornell_box(), 32, 24, 2), (S.feature_scene(), 40, 20, 3)]
    r = S.FilmResolve(tone_map="reinhard", white_balance_kelvin=4500.0)
    with HipBatch(S.RenderSettings(width=1, height=1, spp=1, max_depth=6)) as b:
        for sc, w, h, spp in scenes:
            b.add_scene(sc, w, h, spp)
        b.build()
        b.render(0)
        for i, (sc, w, h, spp) in enumerate(scenes):
            pt = _tracer(sc, w, h, spp)
            try:
                assert pt.render_frame_samples(0, spp)
                assert np.array_equal(b.resolve_hdr(i), pt.resolve_hdr())
                assert np.array_equal(b.resolve_ldr(i, r), pt.resolve_ldr(r))
            finally:
                pt.close()


def test_resolve_rejects_bad_settings():
    from vkpathtracer_amd import abi
    from vkpathtracer_amd import scene as S

    pt = _tracer(S.cornell_box(), 8, 8)
    try:
        with pytest.raises(ValueError):
            pt.resolve_ldr(S.FilmResolve(tone_map="hable"))
        bad = abi.make_resolve(None)
        bad.tone_map = 7
        import ctypes as C
        img = np.zeros((8, 8, 4), np.uint8)
        assert pt._lib.pt_hip_resolve_ldr(pt._h, C.byref(bad), img.ctypes.data) == 2
    finally:
        pt.close()

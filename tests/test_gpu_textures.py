"""Albedo textures on the GPU: textured renders bit-identical to the oracle, multi-view batches
equal to one tracer per camera, and texel / albedo / roughness gradients of the path-replay adjoint
against the oracle's (same paths, fp64 accumulation) for a single tracer and for a batch of views."""
import dataclasses

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

COUNTERS = ("samples", "rays", "shadow_tests", "bvh_node_visits", "bvh_leaf_visits", "triangle_tests",
            "triangle_hits")


def _scene(freq=10, size=64, seed=11):
    from vkpathtracer_amd import scene as S

    return S.textured_sphere_scene(seed=seed, frequency=freq, texture_size=size)


def test_textured_render_bitexact(tracer_factory):
    from oracle import oracle as O
    from vkpathtracer_amd import scene as S

    sc = _scene()
    st = S.RenderSettings(width=48, height=40, spp=3, max_depth=6)
    pt = tracer_factory(sc, st)
    assert pt.render_frame_samples(0, st.spp), pt.status().last_error
    film, cnt = pt.film()
    counters = pt.read_counters()
    rf, rc, rcnt = O.Oracle(sc).render(st)
    assert np.array_equal(cnt, rc)
    assert np.abs(rf).max() > 0
    assert np.array_equal(film, rf), f"max |diff| {np.abs(film - rf).max()}"
    for k in COUNTERS:
        assert counters[k] == rcnt[k], (k, counters[k], rcnt[k])


def test_update_texels_changes_the_render_as_a_reload_does(tracer_factory):
    """update_texels on a built tracer renders exactly what a tracer loaded with those texels renders."""
    from vkpathtracer_amd import scene as S

    sc = _scene(freq=8, size=32)
    st = S.RenderSettings(width=32, height=32, spp=2, max_depth=5)
    new_tex = S.perturbed_texture(sc.textures[0], 0.2, seed=4)
    pt = tracer_factory(sc, st)
    pt.update_texels(new_tex.reshape(-1, 3))
    assert pt.render_frame_samples(0, st.spp)
    a, _ = pt.film()
    pt2 = tracer_factory(dataclasses.replace(sc, textures=[new_tex]), st)
    assert pt2.render_frame_samples(0, st.spp)
    b, _ = pt2.film()
    assert np.array_equal(a, b)


def test_views_equal_single_tracers():
    """A batch holding the scene and three more cameras as views: every film equals a tracer's render of
    the scene with that camera, eager and from a graph."""
    from vkpathtracer_amd import scene as S
    from vkpathtracer_amd.tracer import HipBatch, HipPathTracer

    sc = _scene(freq=8, size=32)
    cams = S.ring_cameras(4)
    w, h, spp = 32, 24, 2
    settings = S.RenderSettings(width=1, height=1, spp=1, max_depth=6)
    for use_graph in (False, True):
        with HipBatch(settings) as b:
            b.add_scene(dataclasses.replace(sc, camera=cams[0]), w, h, spp)
            for c in cams[1:]:
                b.add_view(0, c, w, h, spp)
            b.build()
            b.render(0, use_graph=use_graph)
            for i, c in enumerate(cams):
                film, cnt = b.film(i)
                st = dataclasses.replace(settings, width=w, height=h, spp=spp)
                with HipPathTracer(0) as pt:
                    assert pt.configure(st) and pt.load_scene_snapshot(dataclasses.replace(sc, camera=c))
                    assert pt.build_or_update_acceleration()
                    assert pt.render_frame_samples(0, spp)
                    rf, rc = pt.film()
                assert np.array_equal(cnt, rc)
                assert np.array_equal(film, rf), (use_graph, i)


def _oracle_grads(sc, st, fg, variant="f32"):
    from oracle import oracle as O

    o = O.Oracle(sc, variant=variant)
    px = np.arange(st.width * st.height, dtype=np.uint32)
    tot = None
    for s in range(st.spp):
        g = O.adjoint_gradients(o, st, px, s, fg.reshape(-1, 3), threads=8)
        tot = g if tot is None else {k: tot[k] + g[k] for k in g}
    return tot


def _close(name, got, ref, tol=2e-5):
    scale = max(np.abs(ref).max(), 1e-6)
    err = np.abs(got - ref).max() / scale
    print(f"{name}: max|grad| {scale:.4g}, max abs err / scale {err:.2e}")
    assert err <= tol, name


@pytest.mark.parametrize("max_depth", [3, 6])
def test_texture_adjoint_matches_oracle(max_depth):
    import torch

    from vkpathtracer_amd import scene as S
    from vkpathtracer_amd.tracer import HipPathTracer

    sc = _scene()
    st = S.RenderSettings(width=40, height=32, spp=3, max_depth=max_depth)
    fg = np.random.default_rng(6).normal(size=(st.height, st.width, 3)).astype(np.float32)
    nt = sc.textures[0].shape[0] * sc.textures[0].shape[1]
    dev = torch.device("cuda", 0)
    with HipPathTracer(0) as pt:
        assert pt.configure(st) and pt.load_scene_snapshot(sc) and pt.build_or_update_acceleration()
        ga = torch.zeros((len(sc.materials), 3), device=dev)
        ge = torch.zeros_like(ga)
        gl = torch.zeros((len(sc.lights), 3), device=dev)
        gr = torch.zeros(len(sc.materials), device=dev)
        gt = torch.zeros((nt, 3), device=dev)
        pt.render_adjoint([(0, 0, st.width, st.height, s, 0, 0) for s in range(st.spp)],
                          torch.from_numpy(fg).to(dev), ga, ge, gl, grad_roughness=gr, grad_texels=gt)
        g = {k: v.cpu().numpy().astype(np.float64) for k, v in
             (("albedo", ga), ("emission", ge), ("light", gl), ("roughness", gr), ("texture", gt))}
    r = _oracle_grads(sc, st, fg)
    assert np.abs(r["texture"]).max() > 0
    for k in ("albedo", "emission", "light", "roughness", "texture"):
        _close(f"depth {max_depth} {k}", g[k], r[k], 2e-5 if k != "roughness" else 1e-4)


def test_batch_adjoint_equals_sum_of_view_adjoints():
    """The batch adjoint over a scene and its views adds, per parameter, what each camera's own tracer
    adds (film gradients laid out film after film)."""
    import torch

    from vkpathtracer_amd import scene as S
    from vkpathtracer_amd.tracer import HipBatch, HipPathTracer

    sc = _scene(freq=8, size=32)
    cams = S.ring_cameras(3)
    w, h, spp = 24, 20, 2
    settings = S.RenderSettings(width=1, height=1, spp=1, max_depth=5)
    nm, nl = len(sc.materials), len(sc.lights)
    nt = sc.textures[0].shape[0] * sc.textures[0].shape[1]
    dev = torch.device("cuda", 0)
    fg = np.random.default_rng(8).normal(size=(len(cams), h, w, 3)).astype(np.float32)

    def zeros():
        return [torch.zeros(s, device=dev) for s in ((nm, 3), (nm, 3), (nl, 3), (nm,), (nt, 3))]

    gb = zeros()
    with HipBatch(settings) as b:
        b.add_scene(dataclasses.replace(sc, camera=cams[0]), w, h, spp)
        for c in cams[1:]:
            b.add_view(0, c, w, h, spp)
        b.build()
        b.render(0)
        b.render_adjoint(torch.from_numpy(fg).to(dev), gb[0], gb[1], gb[2], 0, grad_roughness=gb[3],
                         grad_texels=gb[4])
    gs = zeros()
    st = dataclasses.replace(settings, width=w, height=h, spp=spp)
    for i, c in enumerate(cams):
        with HipPathTracer(0) as pt:
            assert pt.configure(st) and pt.load_scene_snapshot(dataclasses.replace(sc, camera=c))
            assert pt.build_or_update_acceleration()
            pt.render_adjoint([(0, 0, w, h, s, 0, 0) for s in range(spp)], torch.from_numpy(fg[i]).to(dev),
                              gs[0], gs[1], gs[2], grad_roughness=gs[3], grad_texels=gs[4])
    for name, a, c in zip(("albedo", "emission", "light", "roughness", "texture"), gb, gs):
        _close(name, a.cpu().numpy().astype(np.float64), c.cpu().numpy().astype(np.float64), 1e-5)


def test_multiview_autograd_texture_step():
    """MultiViewRenderer: an L2 loss against target views rendered from a perturbed texture; one
    gradient step on the texture lowers the loss."""
    import torch

    from vkpathtracer_amd import scene as S
    from vkpathtracer_amd.autograd import MultiViewRenderer

    sc = _scene(freq=10, size=32)
    cams = S.ring_cameras(4)
    settings = S.RenderSettings(width=1, height=1, spp=1, max_depth=4)
    r = MultiViewRenderer(sc, cams, 32, 32, 4, settings)
    a, e, l = r.initial_parameters()
    rough = torch.tensor(sc.materials["roughness"].astype(np.float32), device=r.device)
    target_tex = torch.from_numpy(S.perturbed_texture(sc.textures[0], 0.1, seed=2).reshape(-1, 3)).to(r.device)
    with torch.no_grad():
        target = r(a, e, l, rough, target_tex)
    tex = r.initial_texture().clone().requires_grad_(True)
    rr = rough.clone().requires_grad_(True)
    img = r(a, e, l, rr, tex)
    assert img.shape == (4, 32, 32, 3)
    loss = ((img - target) ** 2).sum()
    loss.backward()
    assert torch.isfinite(tex.grad).all() and tex.grad.abs().sum() > 0 and rr.grad.abs().sum() > 0
    with torch.no_grad():
        lr = 0.05 / float(tex.grad.abs().max())
        loss2 = ((r(a, e, l, rough, tex - lr * tex.grad) - target) ** 2).sum()
    print(f"loss {float(loss):.5f} -> {float(loss2):.5f}")
    assert float(loss2) < float(loss)
    r.close()

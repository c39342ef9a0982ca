"""GPU path-replay adjoint: against the oracle's adjoint of the same paths (accumulated in fp64), against
central finite differences of GPU renders, and through the torch.autograd.Function."""
import dataclasses

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _scene():
    from vkpathtracer_amd import scene as S

    # perturbed albedos: no channel ties in max(T) and no albedo on the [0, 1.5] clamp, where the
    # reference's estimator has kinks
    sc = S.perturbed_albedo_scene(S.cornell_box(), 0.1, seed=11)
    m = sc.materials.copy()
    m["albedo"] = np.clip(m["albedo"], 0.05, 0.95)
    return dataclasses.replace(sc, materials=m)


def _gpu_adjoint(sc, st, film_grad):
    import torch

    from vkpathtracer_amd.tracer import HipPathTracer

    with HipPathTracer(0) as pt:
        assert pt.configure(st) and pt.load_scene_snapshot(sc) and pt.build_or_update_acceleration()
        dev = torch.device("cuda", 0)
        fg = torch.from_numpy(np.ascontiguousarray(film_grad, np.float32)).to(dev)
        ga = torch.zeros((len(sc.materials), 3), device=dev)
        ge = torch.zeros((len(sc.materials), 3), device=dev)
        gl = torch.zeros((len(sc.lights), 3), device=dev)
        tiles = [(0, 0, st.width, st.height, s, 0, 0) for s in range(st.spp)]
        pt.render_adjoint(tiles, fg, ga, ge, gl)
        _, cnt = pt.film()
        assert cnt.sum() == 0, "the adjoint must not touch the film"
        return ga.cpu().numpy().astype(np.float64), ge.cpu().numpy().astype(np.float64), gl.cpu().numpy().astype(np.float64)


@pytest.mark.parametrize("max_depth", [3, 8])
def test_adjoint_matches_oracle_gradient(max_depth):
    """Same paths, same estimator: the GPU's fp32-atomic gradient equals the oracle's fp64-accumulated one."""
    from oracle import oracle as O
    from vkpathtracer_amd import scene as S

    sc = _scene()
    st = S.RenderSettings(width=48, height=40, spp=4, max_depth=max_depth)
    fg = np.random.default_rng(1).normal(size=(st.height, st.width, 3)).astype(np.float32)
    ga, ge, gl = _gpu_adjoint(sc, st, fg)
    o = O.Oracle(sc)
    px = np.arange(st.width * st.height, dtype=np.uint32)
    ra, re, rl = 0, 0, 0
    for s in range(st.spp):
        a, e, l, _ = O.adjoint_pixels(o, st, px, s, fg.reshape(-1, 3), threads=8)
        ra, re, rl = ra + a, re + e, rl + l
    for name, g, r in (("albedo", ga, ra), ("emission", ge, re), ("light", gl, rl)):
        scale = np.abs(r).max()
        err = np.abs(g - r).max() / scale
        print(f"depth {max_depth} {name}: max|grad| {scale:.4g}, max abs err / scale {err:.2e}")
        assert err < 2e-5, name


def test_adjoint_matches_finite_differences():
    """Below the roulette start depth no discrete decision depends on the parameters, so the adjoint
    equals the derivative of the rendered film: central differences with the effective fp32 step."""
    import torch  # noqa: F401

    from vkpathtracer_amd import scene as S
    from vkpathtracer_amd.tracer import HipPathTracer

    sc = _scene()
    st = S.RenderSettings(width=64, height=64, spp=8, max_depth=3)
    fg = np.random.default_rng(2).normal(size=(st.height, st.width, 3)).astype(np.float32)
    ga, ge, gl = _gpu_adjoint(sc, st, fg)

    pt = HipPathTracer(0)
    assert pt.configure(st) and pt.load_scene_snapshot(sc) and pt.build_or_update_acceleration()

    def f(mats, lights):
        assert pt.update_scene_delta(mats, lights)
        pt.reset_accumulation()
        assert pt.render_frame_samples(0, st.spp)
        film, _ = pt.film()
        return float((fg.astype(np.float64) * film).sum())

    worst = 0.0
    cases = [("albedo", i, c) for i, c in [(0, 0), (2, 1), (4, 2), (6, 0), (9, 1), (12, 2), (22, 0), (31, 1)]]
    cases += [("emissive", 34, c) for c in range(3)] + [("color", 0, c) for c in range(3)]
    for field, idx, c in cases:
        vals = []
        for sgn in (1.0, -1.0):
            mats, lights = sc.materials.copy(), sc.lights.copy()
            arr = lights if field == "color" else mats
            # relative step h = 1e-3 * max(1, |p|): the film is fp32, so a fixed step on a parameter of
            # magnitude 17 would be lost in rounding; everything here is linear or smooth in p
            h = 1e-3 * max(1.0, abs(float(arr[idx][field][c])))
            arr[idx][field][c] = np.float32(arr[idx][field][c] + sgn * h)
            vals.append((f(mats, lights), float(arr[idx][field][c])))
        fd = (vals[0][0] - vals[1][0]) / (vals[0][1] - vals[1][1])
        adj = {"albedo": ga, "emissive": ge, "color": gl}[field][idx, c]
        rel = abs(adj - fd) / max(abs(fd), 1e-6)
        print(f"{field}[{idx}][{c}] adjoint {adj: .6f} fd {fd: .6f} rel {rel:.2e}")
        if abs(fd) > 1e-2:
            worst = max(worst, rel)
    pt.close()
    assert worst < 1e-3


def test_autograd_function_inverse_rendering_step():
    import torch

    from vkpathtracer_amd import scene as S
    from vkpathtracer_amd.autograd import DifferentiableRenderer

    base = S.cornell_box()
    target_scene = S.perturbed_albedo_scene(base, 0.1, seed=3)
    st = S.RenderSettings(width=32, height=32, spp=4, max_depth=3)
    tr = DifferentiableRenderer(target_scene, st)
    with torch.no_grad():
        a, e, l = tr.initial_parameters()
        target = tr(a, e, l)
    tr.close()
    r = DifferentiableRenderer(_scene(), st)
    a, e, l = [x.clone().requires_grad_(True) for x in r.initial_parameters()]
    img = r(a, e, l)
    loss = ((img - target) ** 2).sum()
    loss.backward()
    assert torch.isfinite(a.grad).all() and a.grad.abs().sum() > 0 and e.grad.abs().sum() > 0
    with torch.no_grad():
        lr = 1e-3 / float(a.grad.abs().max())
        img2 = r(a - lr * a.grad, e, l)
        loss2 = ((img2 - target) ** 2).sum()
    print(f"loss {float(loss):.5f} -> {float(loss2):.5f}")
    assert float(loss2) < float(loss)
    r.close()


def test_autograd_backward_uses_its_own_forward_parameters():
    """Two forward calls with different parameters before the first one's backward: that backward
    equals the backward of a lone forward at the first parameters (the renderer uploads them again)."""
    import torch

    from vkpathtracer_amd import scene as S
    from vkpathtracer_amd.autograd import DifferentiableRenderer

    st = S.RenderSettings(width=24, height=24, spp=2, max_depth=4)
    r = DifferentiableRenderer(_scene(), st)
    a0, e0, l0 = r.initial_parameters()
    g = torch.randn((st.height, st.width, 3), device=a0.device)

    def grads_of(a, e, l, other=None):
        a, e, l = [x.clone().requires_grad_(True) for x in (a, e, l)]
        img = r(a, e, l)
        if other is not None:
            with torch.no_grad():
                r(*other)  # the renderer now holds other parameters
        (img * g).sum().backward()
        return [x.grad.cpu().numpy().astype(np.float64) for x in (a, e, l)]

    ref = grads_of(a0, e0, l0)
    got = grads_of(a0, e0, l0, other=(torch.clamp(a0 * 0.5 + 0.2, 0.05, 0.95), e0 * 2.0, l0 * 0.5))
    for x, y in zip(got, ref):
        assert np.abs(x - y).max() <= 1e-5 * max(np.abs(y).max(), 1e-6)
    r.close()


def test_adjoint_reuses_forward_radiance_only_when_valid():
    """A render_tiles of the same tiles leaves each path's radiance in place and the adjoint skips its
    forward replay, as does a second adjoint (the replay leaves that radiance alone); after
    update_scene_delta, or after an adjoint of another frame replayed that frame's forward in the same
    slots, it must replay. Gradients agree with a fresh tracer's (up to the order of the fp32 atomic
    adds)."""
    import torch

    from vkpathtracer_amd import scene as S
    from vkpathtracer_amd.tracer import HipPathTracer

    sc = _scene()
    st = S.RenderSettings(width=40, height=32, spp=3, max_depth=8)
    fg = np.random.default_rng(5).normal(size=(st.height, st.width, 3)).astype(np.float32)
    tiles = [(0, 0, st.width, st.height, s, 0, 0) for s in range(st.spp)]
    dev = torch.device("cuda", 0)
    m2 = sc.materials.copy()
    m2["albedo"] = np.clip(m2["albedo"] * 0.8 + 0.1, 0.05, 0.95)
    sc2 = dataclasses.replace(sc, materials=m2)

    def grads(pt, nm, nl, frame=0):
        ga = torch.zeros((nm, 3), device=dev)
        ge = torch.zeros((nm, 3), device=dev)
        gl = torch.zeros((nl, 3), device=dev)
        pt.render_adjoint(tiles, torch.from_numpy(fg).to(dev), ga, ge, gl, frame)
        return [x.cpu().numpy().astype(np.float64) for x in (ga, ge, gl)]

    def close(a, b):
        for x, y in zip(a, b):
            assert np.abs(x - y).max() <= 1e-5 * max(np.abs(y).max(), 1e-6)

    fresh1 = _gpu_adjoint(sc, st, fg)
    fresh2 = _gpu_adjoint(sc2, st, fg)
    with HipPathTracer(0) as pt:
        assert pt.configure(st) and pt.load_scene_snapshot(sc) and pt.build_or_update_acceleration()
        nm, nl = len(sc.materials), len(sc.lights)
        assert pt.render_tiles(tiles)
        close(grads(pt, nm, nl), fresh1)           # reuse
        close(grads(pt, nm, nl), fresh1)           # no forward in between: reuses again
        g1 = grads(pt, nm, nl, 1)                  # frame 1: replays its forward into the slots
        assert not all(np.array_equal(x, y) for x, y in zip(g1, fresh1))
        close(grads(pt, nm, nl), fresh1)           # frame 0 again: replays
        assert pt.render_tiles(tiles)
        assert pt.update_scene_delta(materials=m2)
        close(grads(pt, nm, nl), fresh2)           # scene changed after the render: replays
        assert pt.render_tiles(tiles)
        close(grads(pt, nm, nl), fresh2)           # reuse with the new scene
        assert pt.render_tiles(tiles[:2])
        close(grads(pt, nm, nl), fresh2)           # different tiles: replays


def _glossy_scene():
    """The box with glossy (model 3) walls and blocks: roughness in [0.2, 0.9], some metallic and
    clearcoat, one sheen block; the ceiling light widened so the glossy lobes reach it."""
    from vkpathtracer_amd import scene as S

    sc = _scene()
    m = sc.materials.copy()
    rng = np.random.default_rng(8)
    for i in range(len(m)):
        if m[i]["material_model"] == S.MODEL_EMISSIVE:
            continue
        m[i]["material_model"] = 6 if i % 11 == 3 else S.MODEL_GLOSSY
        m[i]["roughness"] = rng.uniform(0.2, 0.9)
        m[i]["metallic"] = 0.8 if i % 5 == 0 else 0.0
        m[i]["clearcoat"] = 0.3 if i % 7 == 1 else 0.0
        m[i]["sheen"] = 0.5 if i % 11 == 3 else 0.0
    v = sc.vertices.copy()
    tri = sc.indices.reshape(-1, 3)
    light_tris = [int(i["first_triangle"]) + k for i in sc.instances if i["material_index"] == len(m) - 1
                  for k in range(int(i["triangle_count"]))]
    lv = np.unique(tri[light_tris].reshape(-1))
    v[lv, 0] *= 5.0
    v[lv, 2] *= 5.0
    return dataclasses.replace(sc, materials=m, vertices=v)


@pytest.mark.parametrize("max_depth", [2, 6])
def test_roughness_gradient_matches_oracle(max_depth):
    """The GPU's fp32-atomic d/d roughness (NEE Blinn-Phong term + glossy sampling score) equals the
    oracle's fp64-accumulated one on the same paths."""
    import torch

    from oracle import oracle as O
    from vkpathtracer_amd import scene as S
    from vkpathtracer_amd.tracer import HipPathTracer

    sc = _glossy_scene()
    st = S.RenderSettings(width=40, height=32, spp=3, max_depth=max_depth)
    fg = np.random.default_rng(2).normal(size=(st.height, st.width, 3)).astype(np.float32)
    with HipPathTracer(0) as pt:
        assert pt.configure(st) and pt.load_scene_snapshot(sc) and pt.build_or_update_acceleration()
        dev = torch.device("cuda", 0)
        ga = torch.zeros((len(sc.materials), 3), device=dev)
        ge = torch.zeros_like(ga)
        gl = torch.zeros((len(sc.lights), 3), device=dev)
        gr = torch.zeros(len(sc.materials), device=dev)
        pt.render_adjoint([(0, 0, st.width, st.height, s, 0, 0) for s in range(st.spp)],
                          torch.from_numpy(fg).to(dev), ga, ge, gl, grad_roughness=gr)
        g = gr.cpu().numpy().astype(np.float64)
        galb = ga.cpu().numpy().astype(np.float64)
    o = O.Oracle(sc)
    px = np.arange(st.width * st.height, dtype=np.uint32)
    ref = 0
    ref_alb = 0
    for s in range(st.spp):
        out = O.adjoint_gradients(o, st, px, s, fg.reshape(-1, 3), threads=8)
        ref = ref + out["roughness"]
        ref_alb = ref_alb + out["albedo"]
    scale = np.abs(ref).max()
    print(f"depth {max_depth}: max |d/d roughness| {scale:.4g}, max err {np.abs(g - ref).max():.3g}")
    assert scale > 0
    assert np.abs(g - ref).max() <= 1e-4 * scale
    assert np.abs(galb - ref_alb).max() <= 1e-4 * np.abs(ref_alb).max()

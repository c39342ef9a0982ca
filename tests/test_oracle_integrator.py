"""The oracle's integrator pinned by mathematics it does not restate.

The reference ships no golden vectors for its CPU tracer and cannot be compiled here, so each piece of
ScalarCpuPathTracer that the oracle restates is checked against a closed form or a distribution fact:

* camera_rays (ScalarCpuPathTracer.Render.cpp:55-119): pinhole geometry (vertical and horizontal field
  of view, the centre ray, orientation, the regular image-plane grid) and the camera RNG key contract;
* trace at max_depth = 1 (ScalarCpuPathTracer.cpp:1158-1192): first-hit emission or the environment;
* trace at max_depth = 2 with point lights, a Lambert floor and a black environment (sample_direct_light
  and the MIS weight, :1022-1118, :1194-1214): every sample equals the next-event estimate in closed
  form, including the spot cone and the weight taken from lights[0] whichever light was sampled;
* Russian roulette (:1271-1280): throughput divided by rr at every depth, the kill rule, the clamps and
  the 0.001 cut-off, on a floor under a constant environment where the estimate is exact;
* light_pdf_for_direction (:1120-1138): the 0.9999 cone, the 1/N factor and the 1e-4 regulariser;
* sample_hemisphere and sample_phong_lobe (:978-1020): the moments of the cosine and cos^e densities;
* apply_material_effect 1-19 (:187-355): a second, independent transcription of the effects in float64;
* SampleSceneEnvironment (PathTracer.cpp:213-272): ramp maps whose bilinear lookup is a closed form;
* the portable transcendentals: the portable-math oracle against the glibc-math oracle on every
  benchmark configuration's seeded samples.
A planted slip in any of these (a sign, a dropped factor, a swapped axis, a wrong exponent) fails a test.
"""
import dataclasses
import math

import numpy as np
import pytest

from conftest import relrmse
from oracle import oracle as O


def _floor_scene(albedo=(0.6, 0.5, 0.4), lights=(), env=(0.0, 0.0, 0.0), size=200.0, roughness=1.0):
    """One large Lambert quad at y = 0 (model 0), seen from (0, 3, 4) looking at the origin."""
    from vkpathtracer_amd import scene as S

    verts = np.array([(-size, 0.0, -size), (-size, 0.0, size), (size, 0.0, size), (size, 0.0, -size)], np.float32)
    idx = np.array([0, 1, 2, 0, 2, 3], np.uint32)
    inst = np.zeros(1, S.INSTANCE_DTYPE)
    inst["triangle_count"] = 2
    inst["entity_id"] = 1
    inst["rotation"] = (0.0, 0.0, 0.0, 1.0)
    inst["scale"] = 1.0
    mats = S.make_materials(1)
    mats[0]["albedo"] = albedo
    mats[0]["roughness"] = roughness
    lt = np.array(list(lights), S.LIGHT_DTYPE) if lights else np.zeros(0, S.LIGHT_DTYPE)
    cam = S.Camera(position=(0.0, 3.0, 4.0), target=(0.0, 0.0, 0.0), up=(0.0, 1.0, 0.0), fov_deg=50.0)
    return S.Scene(verts, idx, inst, mats, lt, cam, environment_color=tuple(env), name="floor")


def _settings(**kw):
    from vkpathtracer_amd import scene as S

    base = dict(width=24, height=16, spp=1, max_depth=2, rr_start_depth=3)
    base.update(kw)
    return S.RenderSettings(**base)


def _samples(orc, st, pixels, count):
    """Radiance of samples 0..count-1 of every pixel: (count, N, 3) float64, and their primary rays."""
    out, rays = [], []
    for s in range(count):
        rgb, valid, _ = orc.render_pixels(st, pixels, s, threads=1)
        assert valid.all()
        out.append(np.asarray(rgb, np.float64))
        rays.append(orc.camera_rays(st, pixels, s)[0])
    return np.stack(out), np.stack(rays)


def _unit(v):
    v = np.asarray(v, np.float64)
    return v / np.linalg.norm(v, axis=-1, keepdims=True)


# ------------------------------------------------------------------ camera_rays
def test_camera_pinhole_geometry():
    """Vertical fov is the angle between the top and bottom image edges; the horizontal angle is
    2 atan(aspect tan(fov / 2)); the image centre looks along target - position; the top row leans
    towards the up vector and the left column away from forward x up; pixel centres land on a regular
    grid of the plane one unit in front of the camera."""
    from vkpathtracer_amd import scene as S

    for cam, (w, h) in (((0.3, 1.2, 3.5), (-0.2, 0.8, -1.0)), (32, 20)), (((2.0, 5.0, -1.0), (0.0, 0.0, 0.5)), (17, 29)):
        pos, tgt = cam
        sc = dataclasses.replace(_floor_scene(), camera=S.Camera(position=pos, target=tgt, up=(0.0, 1.0, 0.0),
                                                                 fov_deg=47.0))
        st = _settings(width=w, height=h)
        o = O.Oracle(sc)
        fov = math.radians(47.0)
        aspect = w / h
        fwd = _unit(np.subtract(tgt, pos))
        right = _unit(np.cross(fwd, (0.0, 1.0, 0.0)))
        up = np.cross(right, fwd)

        def d(x, y, u, v):
            return np.asarray(o.camera_dir(st, x, y, u, v), np.float64)

        def angle(a, b):
            return math.acos(min(1.0, float(np.dot(_unit(a), _unit(b)))))

        cx, cy = w // 2, h // 2
        ux, vy = (0.0, 0.0) if w % 2 == 0 else (0.5, 0.0)
        vy = 0.0 if h % 2 == 0 else 0.5
        assert angle(d(cx, cy, ux, vy), fwd) < 2e-6
        assert abs(angle(d(cx, 0, ux, 0.0), d(cx, h - 1, ux, 1.0)) - fov) < 1e-5
        assert abs(angle(d(0, cy, 0.0, vy), d(w - 1, cy, 1.0, vy)) - 2.0 * math.atan(aspect * math.tan(fov / 2))) < 1e-5
        assert np.dot(d(cx, 0, ux, 0.0), up) > 0.0 and np.dot(d(0, cy, 0.0, vy), right) < 0.0
        xs, ys = np.meshgrid(np.arange(w), np.arange(h))
        dirs = np.array([d(x, y, 0.5, 0.5) for x, y in zip(xs.ravel(), ys.ravel())])
        p = dirs / (dirs @ fwd)[:, None]
        gx, gy = p @ right, p @ up
        t = math.tan(fov / 2)
        assert np.allclose(gx, ((xs.ravel() + 0.5) / w * 2 - 1) * aspect * t, atol=2e-6)
        assert np.allclose(gy, (1 - (ys.ravel() + 0.5) / h * 2) * t, atol=2e-6)


def test_camera_rng_key_and_jitter():
    """camera_rays draws its pixel jitter from the key {pixel = y W + x, sample, frame, path id =
    uint32(((y << 32) | x) + sample + frame), settings.seed, dimension 0, depth 0} (Render.cpp:61-72,
    :246-251); the ray is the pinhole ray through (x + u, y + v) from the camera position."""
    L = O.lib("f32")
    sc = _floor_scene()
    st = _settings(width=13, height=7, seed=0x1234ABCD)
    o = O.Oracle(sc)
    px = np.arange(st.width * st.height, dtype=np.uint32)
    for sample, frame in ((0, 0), (5, 3)):
        rays, jit = o.camera_rays(st, px, sample, frame)
        for i in px:
            x, y = int(i) % st.width, int(i) // st.width
            path_id = (((y << 32) | x) + sample + frame) & 0xFFFFFFFF
            state = L.pto_rng_seed(st.seed, frame, sample, path_id, 0, 0, y * st.width + x)
            draws = np.zeros(2, np.float32)
            L.pto_rng_draws(state, 2, draws.ctypes.data, None)
            assert np.array_equal(jit[i], draws)
            assert np.array_equal(rays[i, :3], np.float32(sc.camera.position))
            assert np.array_equal(rays[i, 3:], o.camera_dir(st, x, y, float(draws[0]), float(draws[1])))


# ------------------------------------------------------------------ max_depth = 1
def test_depth_one_is_first_hit_emission_or_environment(cornell):
    """With one bounce the NEE gate depth < max_depth - 1 is closed: each sample is the emission of the
    material the camera ray hits (the last material for an index beyond the list) or the environment
    colour on a miss, exactly; one extension ray per sample and no shadow test."""
    env = (0.2, 0.3, 0.4)
    sc = dataclasses.replace(cornell, environment_color=env)
    st = _settings(width=32, height=32, max_depth=1)
    o = O.Oracle(sc)
    px = np.arange(st.width * st.height, dtype=np.uint32)
    emissive = sc.materials["emissive"]
    for s in range(3):
        rgb, valid, c = o.render_pixels(st, px, s, threads=1)
        rays, _ = o.camera_rays(st, px, s)
        t, _, mat, _, _ = o.intersect_scene(rays)
        miss = ~np.isfinite(t)
        assert miss.any() and (~miss).any() and (emissive[np.minimum(mat[~miss], len(emissive) - 1)] > 0).any()
        want = np.where(miss[:, None], np.float32(env), emissive[np.minimum(mat, len(emissive) - 1)])
        assert np.array_equal(np.asarray(rgb, np.float32), want.astype(np.float32))
        assert c["rays"] == len(px) and c["shadow_tests"] == 0 and c["samples"] == len(px)


# ------------------------------------------------------------------ max_depth = 2, next-event estimation
def _nee_closed_form(sc, rays, mis=True, albedo_fn=None):
    """Per ray: the depth-2 estimate of a Lambert floor under point lights with a black environment,
    for each light index (columns): N * albedo / pi * colour * intensity * spot / (d^2 + 1e-4) * cos
    * w, w = a^2 / (a^2 + b^2) with a = 1 / N and b = max(0, n . dir to lights[0]) / pi.
    albedo_fn(pos, normal, incoming) -> (N, 3) replaces the material albedo (material effects)."""
    o, d = rays[:, :3].astype(np.float64), rays[:, 3:].astype(np.float64)
    t = -o[:, 1] / d[:, 1]  # the floor y = 0
    pos = o + d * t[:, None]
    n = np.array([0.0, 1.0, 0.0])  # the face normal turned to the camera above
    lights = sc.lights
    nl = len(lights)
    alb = np.broadcast_to(np.asarray(sc.materials[0]["albedo"], np.float64), (len(rays), 3))
    to0 = _unit(lights[0]["position"].astype(np.float64) - pos)
    b = np.maximum(0.0, to0 @ n) / math.pi
    a = 1.0 / nl
    w = a * a / (a * a + b * b) if mis else np.ones(len(rays))
    out = np.zeros((len(rays), nl, 3))
    for k, l in enumerate(lights):
        to = l["position"].astype(np.float64) - pos
        dist2 = np.sum(to * to, axis=1)
        ldir = to / np.sqrt(dist2)[:, None]
        cos = ldir @ n
        spot = np.ones(len(rays))
        if l["spot_inner_cos"] > -0.999:
            cone = (-ldir) @ _unit(l["direction"])
            inner, outer = float(l["spot_inner_cos"]), float(l["spot_outer_cos"])
            x = np.clip((cone - outer) / (inner - outer), 0.0, 1.0)
            spot = np.where(cone <= outer, 0.0, np.where(cone >= inner, 1.0, x * x * (3 - 2 * x)))
        irr = l["color"].astype(np.float64)[None, :] * (float(l["intensity"]) * spot / (dist2 + 1e-4))[:, None]
        a_k = alb if albedo_fn is None else albedo_fn(pos, n, -ldir)
        val = nl * a_k / math.pi * irr * np.maximum(cos, 0.0)[:, None] * w[:, None]
        out[:, k] = np.where((cos > 0)[:, None], val, 0.0)
    return out


@pytest.mark.parametrize("mis", [True, False])
def test_depth_two_single_light_is_next_event_estimate(mis):
    from vkpathtracer_amd import scene as S

    sc = _floor_scene(lights=[S.make_light((0.7, 2.5, -0.4), (3.0, 2.0, 1.0), 4.0)])
    st = _settings(enable_mis=mis)
    o = O.Oracle(sc)
    px = np.arange(st.width * st.height, dtype=np.uint32)
    got, rays = _samples(o, st, px, 4)
    want = np.stack([_nee_closed_form(sc, r, mis)[:, 0] for r in rays])
    assert (want > 0).all()
    assert np.allclose(got, want, rtol=3e-5, atol=0)


def test_depth_two_spot_light_cone():
    """The spot attenuation: zero outside the outer cone, one inside the inner cone, smoothstep between
    (spot_attenuation, ScalarCpuPathTracer.cpp:168-185); the image holds all three regions."""
    from vkpathtracer_amd import scene as S

    l = S.make_light((0.0, 2.0, 0.5), (1.0, 1.0, 1.0), 6.0)
    l["direction"] = _unit((0.1, -1.0, -0.2)).astype(np.float32)
    l["spot_inner_cos"], l["spot_outer_cos"] = 0.97, 0.85
    sc = _floor_scene(lights=[l])
    st = _settings(width=32, height=24)
    o = O.Oracle(sc)
    px = np.arange(st.width * st.height, dtype=np.uint32)
    got, rays = _samples(o, st, px, 2)
    want = np.stack([_nee_closed_form(sc, r)[:, 0] for r in rays])
    lum = want.sum(-1)
    full = np.stack([_nee_closed_form(dataclasses.replace(sc, lights=np.array([S.make_light(
        (0.0, 2.0, 0.5), (1.0, 1.0, 1.0), 6.0)], S.LIGHT_DTYPE)), r)[:, 0] for r in rays]).sum(-1)
    assert (lum == 0).any() and np.isclose(lum, full, rtol=1e-12).any() and ((lum > 0) & (lum < full * 0.99)).any()
    # in the penumbra the smoothstep magnifies the fp32 rounding of the cone cosine (a few 1e-8): allow
    # the closed form's own slope there, d spot / d cone = 6 x (1 - x) / (inner - outer), times 5e-7
    slope = np.zeros_like(lum)
    for i, r in enumerate(rays):
        pos = r[:, :3].astype(np.float64) + r[:, 3:].astype(np.float64) * (-r[:, 1:2] / r[:, 4:5])
        cone = _unit(pos - l["position"].astype(np.float64)) @ _unit(l["direction"])
        x = np.clip((cone - 0.85) / (0.97 - 0.85), 0.0, 1.0)
        slope[i] = 6 * x * (1 - x) / (0.97 - 0.85)
    tol = 3e-5 * want + (full * slope * 5e-7)[..., None]
    assert (np.abs(got - want) <= tol + 1e-12).all()


def test_depth_two_two_lights_weight_uses_light_zero():
    """NEE picks one of two lights per sample (scaled by N = 2) and weights it with the bsdf pdf of the
    direction to lights[0] whatever light it picked (ScalarCpuPathTracer.cpp:1200-1209): every sample
    equals one of the two closed-form values, both occur about equally often, and the value of light 1
    differs from what its own direction's weight would give."""
    from vkpathtracer_amd import scene as S

    lights = [S.make_light((0.0, 2.0, 0.0), (1.0, 1.0, 1.0), 3.0), S.make_light((3.0, 0.6, -1.0), (1.0, 0.8, 0.6), 5.0)]
    sc = _floor_scene(lights=lights)
    st = _settings()
    o = O.Oracle(sc)
    px = np.arange(st.width * st.height, dtype=np.uint32)
    got, rays = _samples(o, st, px, 16)
    picks = np.zeros(2)
    for g, r in zip(got, rays):
        cand = _nee_closed_form(sc, r)  # (N, 2, 3)
        err = np.abs(g[:, None, :] - cand).max(-1) / np.abs(cand).max(-1)
        k = np.argmin(err, axis=1)
        assert (err[np.arange(len(k)), k] < 3e-5).all()
        picks += np.bincount(k, minlength=2)
        # light 1 weighted by its own direction would differ
        own = dataclasses.replace(sc, lights=sc.lights[::-1])
        alt = _nee_closed_form(own, r)[:, 0]
        assert (np.abs(alt - cand[:, 1]).max(-1) > 1e-3 * np.abs(cand[:, 1]).max(-1)).mean() > 0.9
    frac = picks[1] / picks.sum()
    assert abs(frac - 0.5) < 4 * math.sqrt(0.25 / picks.sum())


# ------------------------------------------------------------------ material effects
# apply_material_effect and procedural_wood_albedo (ScalarCpuPathTracer.cpp:187-355) transcribed a
# second time, directly from the reference and independently of the oracle's restatement, in float64.
# Only the position hash stays in float32 (x 12.9898f + y 78.233f + z 37.719f + seed 19.19f, sinf,
# x 43758.5453f, fract): its sine of large arguments amplifies any other rounding by 4e4.
def _hash01(x, y, z, seed):
    f = np.float32
    a = ((f(x) * f(12.9898) + f(y) * f(78.233)) + f(z) * f(37.719)) + f(seed) * f(19.19)
    v = np.sin(np.asarray(a, np.float64)).astype(f) * f(43758.5453)
    return np.asarray(v - np.floor(v), np.float64)


def _smoothstep(e0, e1, x):
    t = np.clip((x - e0) / max(1.0e-6, e1 - e0), 0.0, 1.0)
    return t * t * (3.0 - 2.0 * t)


def _wood(base, p, rough, coat, style):
    rough, coat = min(1.0, max(0.0, rough)), min(1.0, max(0.0, coat))
    polish = min(1.0, max(0.0, coat * (1.25 - rough * 0.75)))
    parquet, walnut, sandal = float(style == 17), float(style == 18), float(style == 19)
    x, y, z = p[:, 0], p[:, 1], p[:, 2]
    pw = 0.40 + 0.16 * parquet + 0.08 * polish - 0.18 * walnut
    pl = 1.15 + 0.85 * parquet + 0.55 * polish - 0.55 * walnut
    px = (x + sandal * y * 0.35) / max(0.12, pw)
    pz = (z + walnut * y * 0.70 + sandal * x * 0.22) / max(0.30, pl)
    cx, cz = np.floor(px), np.floor(pz)
    lx, lz = px - cx, pz - cz
    f = np.float32
    seed = _hash01(cx, cz, 0.0, 16.0)
    macro = _hash01(np.floor(f(x) * f(3.7)), np.floor(f(y) * f(3.7)), np.floor(f(z) * f(3.7)), 31.0 + style)
    micro = _hash01(np.floor(f(x) * f(34.0)), np.floor(f(y) * f(34.0)), np.floor(f(z) * f(34.0)), 47.0 + style)
    rotate = (parquet > 0.5) & (np.fmod(cx + cz, 2.0) >= 1.0)
    along, across = np.where(rotate, lz, lx), np.where(rotate, lx, lz)
    wobble = (macro - 0.5) * (0.06 + rough * 0.12 + walnut * 0.08 + sandal * 0.04)
    warped = along + wobble + np.sin(across * 6.0 + seed * 5.0) * 0.018
    freq = 12.0 + rough * 8.0 + walnut * 22.0 + sandal * 36.0
    wave = np.sin((warped * freq + seed * 6.2831853) +
                  np.sin(across * (6.0 + rough * 8.0) + seed * 4.0) * (0.25 + rough * 0.45 + walnut * 0.45))
    fine = np.sin(warped * (55.0 + rough * 55.0 + sandal * 110.0) + across * (9.0 + walnut * 25.0) + seed * 12.0)
    fw = 0.10 + rough * 0.18 + walnut * 0.18 + sandal * 0.22
    ring = _smoothstep(-0.35 + polish * 0.18, 0.82, wave) * (1.0 - fw) + (0.5 + 0.5 * fine) * fw
    seam = (1.0 - sandal) * (1.0 - _smoothstep(0.0, 0.028 + polish * 0.018, np.minimum(np.minimum(lx, 1 - lx),
                                                                                          np.minimum(lz, 1 - lz))))
    dark = base * (0.46 - rough * 0.10 - walnut * 0.18)
    amber = base * (1.14 + coat * 0.22 + sandal * 0.18) + np.array([0.08 + sandal * 0.18, 0.055 + sandal * 0.06, 0.025])
    wood = dark * (1.0 - ring)[:, None] + amber * ring[:, None]
    tint = 0.92 + (0.08 + walnut * 0.12 + sandal * 0.06) * seed + (macro - 0.5) * (0.08 + rough * 0.06)
    pores = _smoothstep(0.74 - walnut * 0.08, 1.0, micro) * (0.025 + rough * 0.055 + walnut * 0.055 + sandal * 0.018)
    wood = wood * tint[:, None] * (1.0 - pores)[:, None]
    wood = wood * (1.0 - seam * (0.10 + rough * 0.35 + parquet * 0.18) * (1.0 - coat * 0.45))[:, None]
    mix = 0.30 + rough * 0.34 - polish * 0.14 + walnut * 0.26 + sandal * 0.18
    return base * (1.0 - mix) + wood * mix


def _effect(mat, pos32, n, incoming):
    """apply_material_effect for each row of float32 positions (normal n, incoming direction rows)."""
    e = int(mat["material_effect"])
    p = pos32.astype(np.float64)
    c = np.broadcast_to(np.asarray(mat["albedo"], np.float64), p.shape).copy()
    h = _hash01(pos32[:, 0], pos32[:, 1], pos32[:, 2], float(e))
    x, y, z = p[:, 0], p[:, 1], p[:, 2]
    ndv = incoming @ -n  # dot(normal, -incoming)
    col = lambda *v: np.array(v, np.float64)  # noqa: E731
    if e == 1:
        rim = np.maximum(0.0, 1.0 - np.abs(ndv)) ** 2
        c = c * (0.65 + 0.25 * h)[:, None] + col(0.25, 0.22, 0.28) * (rim * (0.4 + float(mat["sheen"])))[:, None]
    elif e == 2:
        c = c * (0.45 + 0.55 * (0.5 + 0.5 * np.sin(x * 7.0 + z * 5.0 + h * 6.0)))[:, None]
    elif e == 3:
        c = c * np.where(h > 0.72, 0.18, 1.0)[:, None]
    elif e == 4:
        vein = 0.5 + 0.5 * np.sin((x + y * 0.4 + z * 0.7) * 9.0 + h * 3.0)
        c = c * (0.55 + 0.45 * vein)[:, None] + col(0.18, 0.20, 0.23) * (1.0 - vein)[:, None]
    elif e == 5:
        c = c * (0.55 + 0.35 * h)[:, None] + col(0.45, 0.16, 0.04) * (0.25 + 0.25 * h)[:, None]
    elif e == 6:
        c = c * (0.65 + 0.35 * h)[:, None]
    elif e == 7:
        c = c * 0.78 + col(1.0, 0.62, 0.42) * 0.16
    elif e == 8:
        c = c * 0.55 + np.stack([0.35 + 0.45 * h, 0.25 + 0.35 * (1.0 - h), np.full_like(h, 0.85)], 1) * 0.45
    elif e == 9:
        f = np.float32
        chk = np.fmod(np.floor(f(x) * f(4.0)) + np.floor(f(z) * f(4.0)), 2.0)
        c = c * np.where(chk < 0.5, 1.0, 0.28)[:, None]
    elif e == 10:
        c = c * 0.45 + np.stack([np.ones_like(h), 0.35 + 0.3 * h, np.full_like(h, 0.08)], 1) * 0.55
    elif e == 11:
        c = c * (0.55 + 0.45 * (0.5 + 0.5 * np.sin(y * 18.0 + h * 5.0)))[:, None]
    elif e == 12:
        c = c + col(0.55, 0.65, 0.95) * (np.maximum(0.0, ndv) ** 6)[:, None]
    elif e == 13:
        c = c * (0.65 + 0.35 * np.abs(np.sin(x * 10.0) * np.cos(z * 10.0)))[:, None]
    elif e == 14:
        c = col(0.15, 0.75, 1.0) * (0.2 + 0.8 * np.maximum(0.0, 1.0 - np.abs(ndv)) ** 0.8)[:, None]
    elif e == 15:
        bands = 0.5 + 0.5 * np.sin(y * 11.0 + x * 3.0 + h * 7.0)
        rim = np.maximum(0.0, 1.0 - np.abs(ndv)) ** 1.5
        c = c * (0.25 + 0.45 * bands)[:, None] + col(0.48, 0.56, 0.68) * (0.18 + 0.32 * rim)[:, None]
    elif 16 <= e <= 19:
        c = _wood(c, p, float(mat["roughness"]), float(mat["clearcoat"]), e)
    return np.clip(c, 0.0, 1.5)


@pytest.mark.parametrize("effect", list(range(1, 20)))
def test_material_effect_albedo(effect):
    """Each procedural effect, seen through the depth-2 next-event estimate on a Lambert floor: the
    oracle's samples equal the closed form with the effect's albedo from the second transcription
    (the hash's steps at discontinuities - floors, the crack threshold - may flip a few samples)."""
    from vkpathtracer_amd import scene as S

    sc = _floor_scene(albedo=(0.6, 0.5, 0.4), lights=[S.make_light((0.7, 2.5, -0.4), (3.0, 2.0, 1.0), 4.0)])
    sc.materials[0]["material_effect"] = effect
    sc.materials[0]["sheen"] = 0.3
    st = _settings(width=48, height=32)
    o = O.Oracle(sc)
    px = np.arange(st.width * st.height, dtype=np.uint32)
    got, rays = _samples(o, st, px, 2)
    mat = sc.materials[0]
    bad = 0
    for g, r in zip(got, rays):
        t, _, _, _, _ = o.intersect_scene(r)
        pos32 = r[:, :3] + r[:, 3:] * t[:, None]  # float32, as the hit position o + d t is formed
        want = _nee_closed_form(sc, r, albedo_fn=lambda pos, n, inc: _effect(mat, pos32, n, inc))[:, 0]
        plain = _nee_closed_form(sc, r)[:, 0]
        assert (np.abs(want - plain) > 1e-2 * plain).mean() > 0.2  # the effect changes the image
        bad += int((np.abs(g - want) > 2e-4 * np.abs(want) + 1e-9).any(axis=1).sum())
    print(f"effect {effect}: {bad} of {got.shape[0] * got.shape[1]} samples outside 2e-4")
    assert bad <= 0.01 * got.shape[0] * got.shape[1], bad


# ------------------------------------------------------------------ environment map
def test_environment_map_lookup():
    """SampleSceneEnvironment (PathTracer.cpp:213-272) on ramp maps: texel (x, y) = (x, y, 1) makes the
    bilinear lookup return (u w, v (h - 1), 1) times the scale, with u = 0.5 + atan2(d_z, d_x) / 2 pi
    (wrapped) and v = acos(d_y) / pi, wherever the four taps do not wrap around the seam."""
    w, h = 64, 33
    env = np.zeros((h, w, 3), np.float32)
    env[..., 0] = np.arange(w)[None, :]
    env[..., 1] = np.arange(h)[:, None]
    env[..., 2] = 1.0
    from vkpathtracer_amd import scene as S

    for scale, eff in (((0.0, 0.0, 0.0), (1.0, 1.0, 1.0)), ((2.0, 0.5, 3.0), (2.0, 0.5, 3.0))):
        base = _floor_scene(size=0.01)
        base.vertices[:, 1] -= 1000.0  # out of every camera's view: all rays miss
        for pos, tgt in (((0, 0, 0), (1, 0.3, 0.2)), ((0, 0, 0), (-0.2, 0.9, -0.5)), ((0, 0, 0), (0.1, -0.8, 0.6))):
            cam = S.Camera(position=pos, target=tgt, up=(0.0, 1.0, 0.0), fov_deg=80.0)
            sc = dataclasses.replace(base, camera=cam, environment_map=env, environment_map_scale=scale)
            st = _settings(width=40, height=30, max_depth=1)
            o = O.Oracle(sc)
            px = np.arange(st.width * st.height, dtype=np.uint32)
            rgb = np.asarray(o.render_pixels(st, px, 0, threads=1)[0], np.float64)
            d = _unit(o.camera_rays(st, px, 0)[0][:, 3:])
            u = 0.5 + np.arctan2(d[:, 2], d[:, 0]) / (2 * math.pi)
            u -= np.floor(u)
            v = np.arccos(np.clip(d[:, 1], -1, 1)) / math.pi
            xf, yf = u * w, v * (h - 1)
            ok = (np.floor(xf) < w - 1) & (np.floor(yf) < h - 1)
            assert ok.mean() > 0.9
            want = np.stack([xf, yf, np.ones_like(xf)], 1) * np.asarray(eff)
            assert np.allclose(rgb[ok], want[ok], rtol=1e-5, atol=2e-4)


# ------------------------------------------------------------------ Russian roulette
@pytest.mark.parametrize("albedo", [(0.5, 0.5, 0.5), (0.5, 0.25, 0.125), (0.05, 0.05, 0.05), (1.0, 1.0, 1.0),
                                    (0.0005, 0.0004, 0.0002)])
def test_roulette_divides_throughput_at_every_depth(albedo):
    """A Lambert floor under a constant environment E, no lights, two bounces, roulette from depth 3:
    the first bounce's ray always escapes, so each sample is exactly T E with T = albedo / rr,
    rr = clamp(max(albedo), 0.1, 0.99) — divided although depth 0 is below the roulette start — and 0
    when max(albedo) < 0.001 ends the path."""
    env = np.float32([0.25, 0.5, 1.0])
    sc = _floor_scene(albedo=albedo, env=env)
    st = _settings(max_depth=2, rr_start_depth=3)
    o = O.Oracle(sc)
    px = np.arange(st.width * st.height, dtype=np.uint32)
    a = np.float32(albedo)
    if a.max() < np.float32(0.001):
        want = np.zeros(3, np.float32)
    else:
        rr = np.float32(min(0.99, max(0.1, float(a.max()))))
        want = (a / rr) * env
    for s in range(3):
        rgb, valid, c = o.render_pixels(st, px, s, threads=1)
        assert valid.all()
        assert np.array_equal(np.asarray(rgb, np.float32), np.broadcast_to(want, (len(px), 3)))


def test_roulette_kill_rule():
    """With roulette from depth 0 a path survives its first bounce when u <= rr: grey albedo 0.5
    (rr = 0.5) leaves half the samples at E and the rest at 0."""
    env = np.float32([0.25, 0.5, 1.0])
    sc = _floor_scene(albedo=(0.5, 0.5, 0.5), env=env)
    st = _settings(max_depth=2, rr_start_depth=0, width=32, height=32)
    o = O.Oracle(sc)
    px = np.arange(st.width * st.height, dtype=np.uint32)
    alive = 0
    total = 0
    for s in range(4):
        rgb = np.asarray(o.render_pixels(st, px, s, threads=1)[0], np.float32)
        lit = (rgb == env).all(axis=1)
        assert (lit | (rgb == 0).all(axis=1)).all()
        alive += int(lit.sum())
        total += len(px)
    assert abs(alive / total - 0.5) < 4 * math.sqrt(0.25 / total)


# ------------------------------------------------------------------ light_pdf_for_direction
def test_light_pdf_cone_and_weights():
    """sum over lights whose centre direction has cos > 0.9999 with the query direction of
    (1 / N) / (d^2 + 1e-4); lights closer than 1e-4 are skipped."""
    from vkpathtracer_amd import scene as S

    lights = [S.make_light((0.0, 3.0, 0.0), (1, 1, 1), 1.0), S.make_light((0.0, 7.0, 0.0), (1, 1, 1), 1.0),
              S.make_light((4.0, 1.0, 0.0), (1, 1, 1), 1.0), S.make_light((0.0, 0.0, 0.0), (1, 1, 1), 1.0)]
    sc = _floor_scene(lights=lights)
    o = O.Oracle(sc)
    pos = np.zeros(3, np.float32)
    n = 4
    up = o.light_pdf(pos, (0.0, 1.0, 0.0))
    assert math.isclose(up, (1 / n) / (9 + 1e-4) + (1 / n) / (49 + 1e-4), rel_tol=1e-6)

    def tilted(c):  # a unit direction at cos c from +y
        s = math.sqrt(1 - c * c)
        return np.float32([s, c, 0.0])

    assert math.isclose(o.light_pdf(pos, tilted(0.99995)), up, rel_tol=1e-6)
    assert o.light_pdf(pos, tilted(0.99985)) == 0.0
    side = _unit((4.0, 1.0, 0.0)).astype(np.float32)
    assert math.isclose(o.light_pdf(pos, side), (1 / n) / (17 + 1e-4), rel_tol=1e-6)
    assert o.light_pdf(pos, (0.0, -1.0, 0.0)) == 0.0


# ------------------------------------------------------------------ BSDF samplers
@pytest.mark.parametrize("normal", [(0.0, 0.0, 1.0), (0.0, -0.99999, 0.004), (0.6, 0.0, 0.8), (-0.3, 0.9, -0.3)])
def test_cosine_hemisphere_moments(normal):
    """sample_hemisphere draws directions with density cos / pi about n: E[cos] = 2/3, E[cos^2] = 1/2,
    no mean tangential component, every direction a unit vector on n's side."""
    n = _unit(normal).astype(np.float32)
    m = 200_000
    d = O.hemisphere_samples(n, 0x9E3779B97F4A7C15, m).astype(np.float64)
    c = d @ n.astype(np.float64)
    assert np.allclose(np.linalg.norm(d, axis=1), 1.0, atol=1e-6) and (c >= -1e-7).all()
    sig = math.sqrt(1 / 2 - 4 / 9) / math.sqrt(m)
    assert abs(c.mean() - 2 / 3) < 5 * sig
    assert abs((c * c).mean() - 0.5) < 5 * math.sqrt(1 / 3 - 1 / 4) / math.sqrt(m)
    tang = d - c[:, None] * n.astype(np.float64)
    assert np.abs(tang.mean(axis=0)).max() < 5 * math.sqrt(1 / 6 / m)


@pytest.mark.parametrize("exponent", [1.0, 10.0, 120.0])
def test_phong_lobe_moments(exponent):
    """sample_phong_lobe about refl = n draws cos to refl with density proportional to cos^e:
    E[cos] = (e + 1)/(e + 2) and E[cos^2] = (e + 1)/(e + 3)."""
    n = np.float32([0.0, 0.0, 1.0])
    m = 200_000
    d, _ = O.phong_samples(exponent, n, n, 0.0, 0x51ED2701, m, variant="f32")
    c = d.astype(np.float64) @ n.astype(np.float64)
    e = exponent
    mean, mean2 = (e + 1) / (e + 2), (e + 1) / (e + 3)
    sig = math.sqrt(max(mean2 - mean * mean, 1e-12) / m)
    assert abs(c.mean() - mean) < 5 * sig
    assert (c > 0).all()


# ------------------------------------------------------------------ portable math vs glibc
def _subset_films(sc, st, pixels, spp):
    out = []
    for variant in ("f32", "f32_libm"):
        o = O.Oracle(sc, variant=variant)
        acc = np.zeros((len(pixels), 3))
        for s in range(spp):
            rgb, valid, _ = o.render_pixels(st, pixels, s)
            acc[valid] += np.asarray(rgb, np.float64)[valid]
        out.append(acc)
    return out


# configuration: (scene, width, height, spp, depth, pixel subset or None = full frame, measured relRMSE)
LIBM_CASES = {
    "cornell_256_16spp": ("cornell", 256, 256, 16, 8, None),
    "procedural_1m_1080p": ("procedural", 1920, 1080, 8, 8, 3000),
    "cornell_512_64spp": ("cornell", 512, 512, 8, 8, 4000),
    "textured_sphere_512_32spp": ("textured", 512, 512, 4, 6, 3000),
}


@pytest.mark.parametrize("name", list(LIBM_CASES))
def test_portable_math_matches_glibc_math(name):
    """The oracle (and the GPU, which equals it bit for bit) evaluates sin/cos/pow/atan2/acos with one
    portable method in double where the reference calls glibc. On every benchmark configuration's
    seeded samples the two builds agree within the path's 1e-4 relative-RMSE contract."""
    from vkpathtracer_amd import scene as S

    kind, w, h, spp, depth, k = LIBM_CASES[name]
    sc = {"cornell": S.cornell_box, "procedural": lambda: S.procedural_scene(seed=1),
          "textured": S.textured_sphere_scene}[kind]()
    st = S.RenderSettings(width=w, height=h, spp=spp, max_depth=depth)
    px = (np.arange(w * h, dtype=np.uint32) if k is None
          else np.sort(np.random.default_rng(7).choice(w * h, k, replace=False)).astype(np.uint32))
    port, libm = _subset_films(sc, st, px, spp)
    err = relrmse(port, libm)
    print(f"{name}: portable vs glibc math relRMSE {err:.3e} over {len(px)} pixels x {spp} spp")
    assert err <= 1e-4

"""The oracle's path-replay adjoint against central finite differences of the fp64 oracle.

Below the roulette start depth no discrete decision of the estimator depends on albedo, emission or
light colour, so the adjoint must equal the derivative of the rendered value; the step is the
effective one after the parameters are stored as fp32."""
import dataclasses

import numpy as np
import pytest

from oracle import oracle as O
from vkpathtracer_amd import scene as S


@pytest.fixture(scope="module")
def setup():
    sc = S.perturbed_albedo_scene(S.cornell_box(), 0.1, seed=7)
    m = sc.materials.copy()
    m["albedo"] = np.clip(m["albedo"], 0.05, 0.95)
    sc = dataclasses.replace(sc, materials=m)
    st = S.RenderSettings(width=24, height=24, spp=1, max_depth=3)
    px = np.arange(st.width * st.height, dtype=np.uint32)
    delta = np.random.default_rng(0).normal(size=(len(px), 3)).astype(np.float32)
    return sc, st, px, delta


def _f(scene, st, px, delta, spp):
    o = O.Oracle(scene, variant="f64")
    return sum(float((delta.astype(np.float64) * o.render_pixels(st, px, s, threads=4)[0]).sum()) for s in range(spp))


def test_adjoint_equals_central_differences(setup):
    sc, st, px, delta = setup
    spp = 4
    o = O.Oracle(sc, variant="f64")
    GA = GE = GL = 0
    for s in range(spp):
        a, e, l, _ = O.adjoint_pixels(o, st, px, s, delta, threads=4)
        GA, GE, GL = GA + a, GE + e, GL + l
    h = 1e-3
    for field, idx, c in [("albedo", 0, 0), ("albedo", 3, 1), ("albedo", 7, 0), ("albedo", 8, 1), ("albedo", 33, 2),
                          ("emissive", 34, 1), ("color", 0, 0), ("color", 0, 2)]:
        vals = []
        for sgn in (1.0, -1.0):
            s1 = dataclasses.replace(sc, materials=sc.materials.copy(), lights=sc.lights.copy())
            arr = s1.lights if field == "color" else s1.materials
            arr[idx][field][c] = np.float32(arr[idx][field][c] + sgn * h)
            vals.append((_f(s1, st, px, delta, spp), float(arr[idx][field][c])))
        fd = (vals[0][0] - vals[1][0]) / (vals[0][1] - vals[1][1])
        adj = {"albedo": GA, "emissive": GE, "color": GL}[field][idx, c]
        assert abs(adj - fd) <= 1e-5 * max(1.0, abs(fd)), (field, idx, c, adj, fd)


def _sc():
    sc = S.perturbed_albedo_scene(S.cornell_box(), 0.1, seed=7)
    m = sc.materials.copy()
    m["albedo"] = np.clip(m["albedo"], 0.05, 0.95)
    return dataclasses.replace(sc, materials=m)


def _adj_and_fd(sc, st, px, delta, s0, spp, h, idx, c):
    o = O.Oracle(sc, variant="f64")
    g = 0
    for s in range(s0, s0 + spp):
        g = g + O.adjoint_pixels(o, st, px, s, delta, threads=8)[0]
    vals = []
    for sgn in (1.0, -1.0):
        s1 = dataclasses.replace(sc, materials=sc.materials.copy())
        s1.materials[idx]["albedo"][c] = np.float32(s1.materials[idx]["albedo"][c] + sgn * h)
        o1 = O.Oracle(s1, variant="f64")
        f = sum(float((delta.astype(np.float64) * o1.render_pixels(st, px, s, threads=8)[0]).sum())
                for s in range(s0, s0 + spp))
        vals.append((f, float(s1.materials[idx]["albedo"][c])))
    return g[idx, c], (vals[0][0] - vals[1][0]) / (vals[0][1] - vals[1][1])


def test_adjoint_exact_through_constant_probability_roulette():
    """rr clamped to a constant: the roulette decisions no longer depend on the parameters, so the
    adjoint equals the derivative of the realisation at any depth."""
    st = S.RenderSettings(width=16, height=16, spp=1, max_depth=6, rr_min_survival=0.5, rr_max_survival=0.5)
    px = np.arange(256, dtype=np.uint32)
    delta = np.random.default_rng(5).normal(size=(256, 3)).astype(np.float32)
    for idx, c in [(0, 1), (4, 0), (7, 0)]:
        adj, fd = _adj_and_fd(_sc(), st, px, delta, 0, 32, 1e-3, idx, c)
        assert abs(adj - fd) <= 1e-5 * max(1.0, abs(fd)), (idx, c, adj, fd)


def test_adjoint_roulette_depth_consistent_in_expectation():
    """With rr = max(T) the roulette is unbiased (survival rr, weight 1/rr) and the adjoint differentiates
    its expectation; finite differences of a realisation also see survival flips, so they agree only in
    the mean. Sixteen independent blocks: the two means agree within four combined standard errors."""
    st = S.RenderSettings(width=16, height=16, spp=1, max_depth=6)
    px = np.arange(256, dtype=np.uint32)
    delta = np.ones((256, 3), np.float32)
    A, F = [], []
    for b in range(16):
        a, f = _adj_and_fd(_sc(), st, px, delta, 20000 + 128 * b, 128, 2e-2, 0, 1)
        A.append(a)
        F.append(f)
    A, F = np.array(A), np.array(F)
    se = np.sqrt(A.var() / len(A) + F.var() / len(F))
    assert abs(A.mean() - F.mean()) < 4 * se, (A.mean(), F.mean(), se)


def _glossy_box(seed=3):
    """The box with every surface glossy (model 3, roughness in [0.2, 0.6], so the Blinn-Phong NEE
    term is on), metals and clearcoats among them, and the emissive quad dimmed to black: at depth 2
    all radiance is depth-0 emission + NEE, a smooth function of roughness."""
    sc = _sc()
    m = sc.materials.copy()
    rng = np.random.default_rng(seed)
    for i in range(len(m)):
        if m[i]["material_model"] == S.MODEL_EMISSIVE:
            m[i]["emissive"] = (0.0, 0.0, 0.0)
            m[i]["material_model"] = S.MODEL_LAMBERT
            continue
        m[i]["material_model"] = S.MODEL_GLOSSY
        m[i]["roughness"] = rng.uniform(0.2, 0.6)
        m[i]["metallic"] = 0.8 if i % 5 == 0 else 0.0
        m[i]["clearcoat"] = 0.3 if i % 7 == 1 else 0.0
    return dataclasses.replace(sc, materials=m)


def test_roughness_nee_term_equals_central_differences():
    """max_depth 2, no emitters: the radiance is the depth-0 NEE term (the bounce direction reaches
    nothing that emits), deterministic in roughness; the adjoint's d/d roughness must equal the
    central difference of the fp64 render."""
    sc = _glossy_box()
    st = S.RenderSettings(width=24, height=24, spp=1, max_depth=2)
    px = np.arange(st.width * st.height, dtype=np.uint32)
    delta = np.random.default_rng(1).normal(size=(len(px), 3)).astype(np.float32)
    spp = 3
    o = O.Oracle(sc, variant="f64")
    G = sum(O.adjoint_gradients(o, st, px, s, delta, threads=4)["roughness"] for s in range(spp))
    assert np.abs(G).max() > 0
    h = 1e-4
    for idx in (0, 5, 8, 12, 21, 33):
        vals = []
        for sgn in (1.0, -1.0):
            m = sc.materials.copy()
            m[idx]["roughness"] = np.float32(m[idx]["roughness"] + sgn * h)
            s1 = dataclasses.replace(sc, materials=m)
            vals.append((_f(s1, st, px, delta, spp), float(m[idx]["roughness"])))
        fd = (vals[0][0] - vals[1][0]) / (vals[0][1] - vals[1][1])
        assert abs(G[idx] - fd) <= 2e-4 * max(1.0, abs(fd)), (idx, G[idx], fd)


def _normalize(v):
    v = np.asarray(v, np.float64)
    return (v / np.linalg.norm(v)).astype(np.float32)


# (exponent, reflection direction, normal, sheen probability): narrow and wide lobes, a reflection
# near grazing (many samples fall below the surface and are mirrored back), and a sheen mixture
LOBES = [(2.88, _normalize([0.3, 0.8, 0.1]), _normalize([0, 1, 0]), 0.0),
         (40.0, _normalize([0.9, 0.15, -0.2]), _normalize([0, 1, 0]), 0.0),
         (1.2, _normalize([0.99, 0.05, 0.0]), _normalize([0, 1, 0]), 0.0),
         (6.0, _normalize([0.2, 0.5, 0.8]), _normalize([0.1, 0.9, 0.2]), 0.3)]


@pytest.mark.parametrize("e,refl,n,ps", LOBES)
def test_phong_score_has_zero_mean_under_the_sampler(e, refl, n, ps):
    """E[d log p / de] = 0 when p is the density the sampler actually draws from: a density that
    missed the mirrored samples or the sheen mixture would fail this."""
    _, score = O.phong_samples(e, refl, n, ps, 12345, 400_000)
    se = score.std() / np.sqrt(len(score))
    assert abs(score.mean()) <= 4 * se, (score.mean(), se)


@pytest.mark.parametrize("e,refl,n,ps", LOBES)
def test_phong_score_gives_the_derivative_of_expectations(e, refl, n, ps):
    """E[score * f] = d E[f] / de for smooth f of the direction; the right side by a central
    difference in e over the same random numbers (the sampled direction is continuous in e)."""
    a = _normalize([0.4, 0.7, -0.6])
    f = lambda d: (d.astype(np.float64) @ a.astype(np.float64)) ** 2 + d[:, 1].astype(np.float64)
    n_s = 400_000
    d, score = O.phong_samples(e, refl, n, ps, 777, n_s)
    lhs = f(d) * score
    h = 1e-3 * e
    dp, _ = O.phong_samples(e + h, refl, n, ps, 777, n_s)
    dm, _ = O.phong_samples(e - h, refl, n, ps, 777, n_s)
    rhs = (f(dp) - f(dm)) / (2 * h)
    diff = lhs - rhs
    se = diff.std() / np.sqrt(n_s)
    assert abs(lhs.mean()) > 6 * se, "the test direction must carry signal"
    assert abs(diff.mean()) <= 4 * se, (lhs.mean(), rhs.mean(), se)

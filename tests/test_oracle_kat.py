"""Pins of the CPU oracle that do not depend on the oracle's own formulas:
published known answers, closed forms, textbook cases and brute force."""
import math

import numpy as np
import pytest

from oracle import oracle as O

MASK = (1 << 64) - 1


def test_splitmix64_known_answers():
    L = O.lib("f32")
    # SplitMix64 (Steele, Lea & Flood 2014) from state 0: the first two outputs of the generator
    # x += golden; z = mix(x); recomputed here from the published algorithm's definition.
    assert L.pto_splitmix64(0) == 0xE220A8397B1DCDAF
    # the reference iterates state = sm(state): the second draw from state 0 is sm(sm(0))
    assert L.pto_splitmix64(0xE220A8397B1DCDAF) == 0xA706DD2F4D197E6F


def _py_splitmix(x):
    x = (x + 0x9E3779B97F4A7C15) & MASK
    z = x
    z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & MASK
    z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & MASK
    return z ^ (z >> 31)


def test_splitmix64_matches_python_bignum():
    L = O.lib("f32")
    rng = np.random.default_rng(0)
    for x in [0, 1, MASK, 1 << 63] + [int(v) for v in rng.integers(0, 2**63, 200, dtype=np.uint64)]:
        assert L.pto_splitmix64(x) == _py_splitmix(x)


def test_next01_mapping_closed_form():
    """next01 = ((state >> 8) & 0xffffff) * (1/16777215) in fp32; 0xFFFFFF and 0xFFFFFE both round to 1.0f."""
    L = O.lib("f32")
    out = np.zeros(64, np.float32)
    state = 12345
    L.pto_rng_draws(state, 64, out.ctypes.data, None)
    s = state
    inv = np.float32(1.0) / np.float32(16777215.0)
    for i in range(64):
        s = _py_splitmix(s)
        assert out[i] == np.float32((s >> 8) & 0xFFFFFF) * inv
    assert np.float32(0xFFFFFF) * inv == np.float32(1.0)
    assert np.float32(0xFFFFFE) * inv == np.float32(1.0)
    assert np.float32(0xFFFFFD) * inv < np.float32(1.0)


def test_rng_seed_composition():
    """Rng(SampleKey): state = sm(sm(seed ^ sm(frame) ^ sm(sample) ^ sm(path) ^ sm(dim) ^ sm(depth)) + sm(pixel))."""
    L = O.lib("f32")
    sm = _py_splitmix
    for seed, frame, sample, path, dim, depth, pixel in [(0, 0, 0, 0, 0, 0, 0), (0xBADC0FFEE, 0, 3, 7, 0, 0, 99),
                                                         (MASK, 5, 1 << 31, (3 << 32) | 4, 2, 1, 1 << 40)]:
        mix = sm(seed ^ sm(frame) ^ sm(sample) ^ sm(path) ^ sm(dim) ^ sm(depth))
        assert L.pto_rng_seed(seed, frame, sample, path, dim, depth, pixel) == sm((mix + sm(pixel)) & MASK)


@pytest.mark.parametrize("fn,ref", [("pto_sinf", math.sin), ("pto_cosf", math.cos)])
def test_portable_sincos_within_half_ulp_plus(fn, ref):
    """Portable sin/cos over the tracer's argument range [0, 2*pi] (phi = 2*pi*u): the result is the
    correctly rounded value of the exact function (checked in double) except within 1 ulp."""
    L = O.lib("f32")
    f = getattr(L, fn)
    xs = np.concatenate([np.linspace(0.0, 2 * math.pi, 20001, dtype=np.float32),
                         np.float32(2 * math.pi) * np.random.default_rng(1).random(20000, np.float32),
                         np.array([0.0, math.pi / 2, math.pi, 1e-8, 6.2831855], np.float32)])
    got = np.array([f(float(x)) for x in xs], np.float32)
    exact = np.array([ref(float(x)) for x in xs], np.float64)
    cr = exact.astype(np.float32)
    ulp = np.spacing(np.abs(cr).astype(np.float32))
    err_ulps = np.abs(got.astype(np.float64) - exact) / np.maximum(ulp, np.float32(1e-45))
    assert err_ulps.max() <= 1.0
    assert np.mean(got == cr) > 0.999  # correctly rounded in essentially every case
    assert f(0.0) == 0.0 if fn == "pto_sinf" else f(0.0) == 1.0


def test_portable_powf_matches_exact():
    """pow over the tracer's uses: u^(1/(e+1)) with u in [0,1], e in [0, 5e6], and x^p with p in [4, 96]."""
    L = O.lib("f32")
    rng = np.random.default_rng(3)
    x = rng.random(20000).astype(np.float32)
    y = np.concatenate([(1.0 / (rng.uniform(0, 5e6, 10000) + 1.0)), rng.uniform(4, 96, 10000)]).astype(np.float32)
    got = np.array([L.pto_powf(float(a), float(b)) for a, b in zip(x, y)], np.float32)
    exact = np.array([float(a) ** float(b) for a, b in zip(x.astype(np.float64), y.astype(np.float64))])
    cr = exact.astype(np.float32)
    nz = cr != 0
    ulp = np.spacing(np.abs(cr[nz]))
    assert (np.abs(got[nz].astype(np.float64) - exact[nz]) / ulp).max() <= 1.0
    assert np.mean(got == cr) > 0.999
    assert L.pto_powf(0.0, 3.0) == 0.0 and L.pto_powf(1.0, 77.0) == 1.0 and L.pto_powf(0.3, 0.0) == 1.0
    assert L.pto_powf(0.5, 2.0) == 0.25 and L.pto_powf(0.25, 0.5) == 0.5


def test_portable_math_agrees_with_glibc_builds():
    """The libm-built oracle uses glibc sinf/cosf/powf like the reference: both agree to 1 ulp."""
    P, G = O.lib("f32"), O.lib("f32_libm")
    assert G.pto_uses_libm() == 1 and P.pto_uses_libm() == 0
    xs = np.float32(2 * math.pi) * np.random.default_rng(7).random(5000, np.float32)
    for fn in ("pto_sinf", "pto_cosf"):
        a = np.array([getattr(P, fn)(float(v)) for v in xs], np.float32)
        b = np.array([getattr(G, fn)(float(v)) for v in xs], np.float32)
        assert np.all(np.abs(a - b) <= np.spacing(np.abs(b)))


def test_portable_atan2_acos_exact_to_one_ulp():
    """atan2 over all quadrants (environment-map longitude) and acos over [-1, 1] (latitude)."""
    L = O.lib("f32")
    rng = np.random.default_rng(9)
    ys = rng.normal(size=20000).astype(np.float32)
    xs = rng.normal(size=20000).astype(np.float32)
    xs[:50] = 0.0
    got = np.array([L.pto_atan2f(float(y), float(x)) for y, x in zip(ys, xs)], np.float32)
    exact = np.arctan2(ys.astype(np.float64), xs.astype(np.float64))
    cr = exact.astype(np.float32)
    assert (np.abs(got - exact) / np.spacing(np.abs(cr))).max() <= 1.0
    assert np.mean(got == cr) > 0.999
    a = np.concatenate([np.linspace(-1, 1, 20001), [-1.0, 1.0, 0.0]]).astype(np.float32)
    g2 = np.array([L.pto_acosf(float(v)) for v in a], np.float32)
    e2 = np.arccos(a.astype(np.float64))
    c2 = e2.astype(np.float32)
    nz = c2 != 0
    assert (np.abs(g2[nz] - e2[nz]) / np.spacing(np.abs(c2[nz]))).max() <= 1.0
    assert L.pto_acosf(1.0) == 0.0 and L.pto_atan2f(0.0, 1.0) == 0.0
    assert L.pto_atan2f(1.0, 0.0) == np.float32(math.pi / 2) and L.pto_atan2f(0.0, -1.0) == np.float32(math.pi)

"""The oracle's film resolve (FilmBuffer::resolve_hdr, ApplyFilmResolve, WhiteBalanceScale,
CameraAdjustedFilmResolveSettings) against what the mathematics fixes: the published Tanner Helland
colour-temperature fit evaluated in double, the tone curves' fixed points, exact quantisation of
representable values, gamma against a double-precision power, the x86 conversion of unclamped values,
and the portable-math build against the libm build."""
import math

import numpy as np
import pytest

from oracle import oracle as O


def _helland(kelvin):
    """Tanner Helland's published blackbody fit (0-255 scale) divided by 255, in double, following the
    reference's branch structure (FilmBuffer.cpp:30-49: blue by the logarithm whenever t <= 66)."""
    t = min(max(kelvin, 1000.0), 40000.0) / 100.0
    clamp = lambda v: min(1.0, max(0.0, v))
    if t <= 66.0:
        r = 1.0
        g = clamp((99.4708025861 * math.log(max(1.0, t)) - 161.1195681661) / 255.0)
        b = 0.0 if t <= 19.0 else clamp((138.5177312231 * math.log(max(1.0, t - 10.0)) - 305.0447927307) / 255.0)
    else:
        r = clamp(329.698727446 * (t - 60.0) ** -0.1332047592 / 255.0)
        g = clamp(288.1221695283 * (t - 60.0) ** -0.0755148492 / 255.0)
        b = 1.0
    return np.array([r, g, b])


@pytest.mark.parametrize("kelvin", [800, 1000, 1500, 1900, 2500, 3200, 4100, 5000, 6500, 6700, 9000, 15000, 40000, 60000])
def test_color_temperature_matches_published_fit(kelvin):
    np.testing.assert_allclose(O.color_temperature_rgb(kelvin), _helland(kelvin), atol=2e-6)


def test_white_balance_scale():
    """6500 K is the identity; other temperatures are d65 / max(1e-3, rgb(K))."""
    np.testing.assert_array_equal(O.white_balance_scale(6500.0), [1, 1, 1])
    for k in (2000.0, 3500.0, 9000.0, 20000.0):
        want = _helland(6500.0) / np.maximum(1e-3, _helland(k))
        np.testing.assert_allclose(O.white_balance_scale(k), want, rtol=1e-5)


def test_resolve_hdr_is_the_mean():
    rng = np.random.default_rng(1)
    s = rng.uniform(0, 10, (5, 7, 3)).astype(np.float32)
    c = rng.integers(0, 9, (5, 7)).astype(np.uint32)
    want = s * (np.float32(1) / np.maximum(c, 1).astype(np.float32))[..., None]
    np.testing.assert_array_equal(O.resolve_hdr(s, c), want)


def test_quantisation_of_representable_values():
    """Linear output, linear tone map: k/255 quantises to k; alpha is 255."""
    k = np.arange(256, dtype=np.float32)
    s = np.repeat((k / np.float32(255))[:, None], 3, 1)[None]
    out = O.resolve_ldr(s, np.ones((1, 256), np.uint32), output_transform="linear")
    np.testing.assert_array_equal(out[0, :, 0], np.arange(256))
    assert (out[..., 3] == 255).all()


def test_tone_curve_fixed_points():
    """Reinhard(1) = 1/2, filmic(W = 11.2) = 1, ACES(0) = 0 and every curve is monotone."""
    one = np.ones((1, 1), np.uint32)
    assert O.resolve_ldr(np.full((1, 1, 3), 1, np.float32), one, tone_map="reinhard", output_transform="linear")[0, 0, 0] == 128
    assert O.resolve_ldr(np.full((1, 1, 3), 11.2, np.float32), one, tone_map="filmic_approx", output_transform="linear")[0, 0, 0] == 255
    assert O.resolve_ldr(np.zeros((1, 1, 3), np.float32), one, tone_map="aces_approx")[0, 0, 0] == 0
    x = np.linspace(0, 20, 4001, dtype=np.float32)
    s = np.repeat(x[:, None], 3, 1)[None]
    for tm in O.TONE_MAPS:
        v = O.resolve_ldr(s, np.ones((1, len(x)), np.uint32), tone_map=tm, output_transform="linear")[0, :, 0]
        assert (np.diff(v.astype(int)) >= 0).all(), tm


def test_gamma_against_double_power():
    """pow(x, 1/2.2) quantised: equal to the double-precision result wherever that is not within 1e-4
    of a rounding boundary."""
    x = np.random.default_rng(2).uniform(0, 1, 20000).astype(np.float32)
    s = np.repeat(x[:, None], 3, 1)[None]
    got = O.resolve_ldr(s, np.ones((1, len(x)), np.uint32))[0, :, 0].astype(int)
    inv_g = float(np.float32(1) / np.float32(2.2))
    v = np.power(x.astype(np.float64), inv_g) * 255 + 0.5
    safe = np.abs(v - np.round(v)) > 1e-4
    np.testing.assert_array_equal(got[safe], np.floor(v[safe]).astype(int))
    assert safe.mean() > 0.99


def test_exposure_white_balance_and_nonfinite():
    """Exposure and the white-balance scale multiply the mean; non-finite values resolve to black; an
    unclamped value above 1 wraps as the x86 conversion does (2.0 * 255 + 0.5 -> 510 -> 254)."""
    one = np.ones((1, 3), np.uint32)
    s = np.array([[[0.25, 0.25, 0.25], [np.nan, np.inf, -np.inf], [2.0, 2.0, 2.0]]], np.float32)
    out = O.resolve_ldr(s, one, exposure=2.0, output_transform="linear")
    np.testing.assert_array_equal(out[0, 0, :3], [128, 128, 128])
    np.testing.assert_array_equal(out[0, 1, :3], [0, 0, 0])
    out = O.resolve_ldr(s, one, output_transform="linear", clamp_output=False)
    np.testing.assert_array_equal(out[0, 2, :3], [254, 254, 254])
    wb = O.white_balance_scale(3000.0)
    out = O.resolve_ldr(np.full((1, 1, 3), 0.2, np.float32), np.ones((1, 1), np.uint32), white_balance_kelvin=3000.0,
                        output_transform="linear", clamp_output=False)
    want = np.floor(np.minimum(np.float32(0.2) * wb, 1e9) * 255 + 0.5).astype(int) % 256
    np.testing.assert_array_equal(out[0, 0, :3], want)


def test_portable_build_matches_libm_build():
    """The portable power (the GPU's) and libm's differ by at most one step in a small fraction of
    channels."""
    rng = np.random.default_rng(3)
    s = rng.gamma(1.0, 1.5, (64, 64, 3)).astype(np.float32)
    c = rng.integers(1, 5, (64, 64)).astype(np.uint32)
    for tm in O.TONE_MAPS:
        a = O.resolve_ldr(s, c, tone_map=tm, variant="f32").astype(int)
        b = O.resolve_ldr(s, c, tone_map=tm, variant="f32_libm").astype(int)
        d = np.abs(a - b)
        assert d.max() <= 1 and (d > 0).mean() < 1e-3, tm


def test_camera_adjusted_resolve():
    """f/2.8, 1/60 s, ISO 100 leave the exposure; f/1.4 quadruples it; +EV doubles per stop; an
    out-of-range white balance is ignored."""
    from vkpathtracer_amd import scene as S

    base = S.FilmResolve(exposure=1.5)
    assert S.camera_adjusted_resolve(base, 2.8, 1.0 / 60.0, 100.0).exposure == pytest.approx(1.5, rel=1e-6)
    assert S.camera_adjusted_resolve(base, 1.4, 1.0 / 60.0, 100.0).exposure == pytest.approx(6.0, rel=1e-6)
    assert S.camera_adjusted_resolve(base, 0.0, exposure_compensation=-2.0).exposure == 0.375
    assert S.camera_adjusted_resolve(base, 0.0, exposure_compensation=100.0).exposure == 1.5 * 2.0 ** 32
    assert S.camera_adjusted_resolve(base, white_balance_kelvin=500.0).white_balance_kelvin == 6500.0
    assert S.camera_adjusted_resolve(base, white_balance_kelvin=3000.0).white_balance_kelvin == 3000.0

"""The CPU baseline's AVX2 leaf intersector (8 triangles per instruction stream) gives exactly the
scalar oracle's films, counters and hits, with leaves of 4 and 8, NaN and axis-parallel rays included."""
import numpy as np
import pytest

from oracle import oracle as O
from vkpathtracer_amd import scene as S


def _has_avx2():
    try:
        return " avx2 " in open("/proc/cpuinfo").read().replace("\n", " ")
    except OSError:
        return False


pytestmark = pytest.mark.skipif(not _has_avx2(), reason="host CPU without AVX2")


@pytest.mark.parametrize("leaf", [4, 8])
def test_avx2_render_equals_scalar(small_procedural, leaf):
    st = S.RenderSettings(width=48, height=32, spp=2, max_depth=6)
    a = O.Oracle(small_procedural, leaf_size=leaf, variant="f32").render(st, threads=4)
    b = O.Oracle(small_procedural, leaf_size=leaf, variant="f32_avx2").render(st, threads=4)
    assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1])
    assert a[2] == b[2]


def test_avx2_trace_rays_equal_scalar(small_procedural):
    rays = S.random_rays(small_procedural, 4000, seed=3)
    rays[:50, 3:] = 0.0
    rays[:50, 4] = 1.0          # axis-parallel
    rays[50:60, 0] = np.nan     # NaN origins
    for any_hit in (False, True):
        a = O.Oracle(small_procedural, leaf_size=8, variant="f32").trace_rays(rays, any_hit=any_hit)
        b = O.Oracle(small_procedural, leaf_size=8, variant="f32_avx2").trace_rays(rays, any_hit=any_hit)
        assert np.array_equal(a[0], b[0], equal_nan=True) and np.array_equal(a[1], b[1])
        assert a[2] == b[2]

"""The oracle entry points behind the benchmark's CPU baseline and its report: the band schedule of the
reference's tiled CPU tracer renders exactly the film render_tile renders, whatever the bands or the
threads, and the |det| < 1e-4 rejection count follows Moller-Trumbore on hand cases."""
import numpy as np
import pytest

from oracle import oracle as O


@pytest.mark.parametrize("variant", ["f32", "f32_libm", "f32_avx2", "f32_avx2_libm"])
def test_band_schedule_equals_render_tile(variant):
    from vkpathtracer_amd import scene as S

    sc = S.feature_scene()
    st = S.RenderSettings(width=40, height=37, spp=3, max_depth=6)
    leaf = 8 if "avx2" in variant else 4
    o = O.Oracle(sc, leaf_size=leaf, variant=variant)
    ref, rcnt, rc = o.render(st, threads=1)
    film, cnt, c = o.render_bands(st, band_rows=16, threads=4)  # 3 bands, the last one short
    assert np.array_equal(cnt, rcnt) and np.array_equal(film, ref) and c == rc
    film2, cnt2, _ = o.render_bands(st, band_rows=4, band_first=1, band_stride=3, threads=3)
    rows = np.zeros(st.height, bool)
    for b in range(1, (st.height + 3) // 4, 3):
        rows[4 * b:4 * b + 4] = True
    assert np.array_equal(film2[rows], ref[rows]) and np.array_equal(cnt2[rows], rcnt[rows])
    assert not film2[~rows].any() and not cnt2[~rows].any()


def test_det_rejection_count():
    from vkpathtracer_amd import scene as S

    verts = np.array([(0.0, 0.0, 0.0), (1.0, 0.0, 0.0), (0.0, 0.0, 1.0)], np.float32)
    inst = np.zeros(1, S.INSTANCE_DTYPE)
    inst["triangle_count"] = 1
    inst["rotation"] = (0.0, 0.0, 0.0, 1.0)
    inst["scale"] = 1.0
    sc = S.Scene(verts, np.arange(3, dtype=np.uint32), inst, S.make_materials(1), np.zeros(0, S.LIGHT_DTYPE),
                 S.Camera())
    o = O.Oracle(sc)
    in_plane = np.array([[-1.0, 0.0, 0.25, 1.0, 0.0, 0.0]], np.float32)   # det = e1 . (d x e2) = 0
    crossing = np.array([[0.25, 1.0, 0.25, 0.0, -1.0, 0.0]], np.float32)  # det = 1
    grazing = np.array([[-1.0, 1e-5, 0.25, 1.0, -1e-5, 0.0]], np.float32)  # |det| = 1e-5 < 1e-4
    assert o.det_rejections(in_plane) == (1, 1)
    assert o.det_rejections(crossing) == (1, 0)
    assert o.det_rejections(grazing) == (1, 1)
    assert o.det_rejections(np.concatenate([in_plane, crossing, grazing, crossing])) == (4, 2)

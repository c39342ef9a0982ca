"""The row-band scheduler driving render_tiles as the reference's render coordinator does
(RenderCoordinator.cpp:1350-1391): every pass drained in same-GPU batches of up to 64 slices, with
priority feedback and foveated extra centre samples. The film and sample counts must equal the
oracle accumulating the same slices one at a time, in the scheduled order."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def test_scheduler_batches_render_like_sequential_tiles(cornell, tracer_factory):
    from oracle import oracle as O
    from vkpathtracer_amd import scene as S
    from vkpathtracer_amd.tile_scheduler import TilePriorityFeedback, TileScheduler, TileSchedulerConfig, drain_batches

    st = S.RenderSettings(width=64, height=40, spp=1, max_depth=6)
    pt = tracer_factory(cornell, st)
    sched = TileScheduler(TileSchedulerConfig(st.width, st.height, 8, 1, foveated_center_extra_samples=2,
                                              foveated_center_radius=0.11))
    order = []
    for sample in range(3):
        # feedback changes between passes: mixed dirty bands first, then variance and sample counts
        sched.set_feedback([TilePriorityFeedback(t, float((t * 7 + sample) % 5), sample + t % 2, t == sample)
                            for t in range(5)])
        sched.begin_sample(1, sample)
        for batch in drain_batches(sched, 4):
            assert pt.render_tiles(batch)
            order.extend(batch)
    assert len(order) == 3 * 7 and len({t[4] for t in order}) == 9
    film, cnt = pt.film()
    o = O.Oracle(cornell, builder="sah")
    ref = np.zeros_like(film)
    refc = np.zeros_like(cnt)
    for (x, y, w, h, s, _, _) in order:
        f, c, _ = o.render(st, sample_begin=s, sample_count=1, tile=(x, y, w, h))
        m = c > 0
        ref[m] = ref[m] + f[m]
        refc += c
    assert np.array_equal(cnt, refc)
    assert np.array_equal(film, ref)
    assert set(np.unique(cnt[16:24])) == {9} and set(np.unique(cnt[:16])) == {3}  # centre band: 3 slices a pass

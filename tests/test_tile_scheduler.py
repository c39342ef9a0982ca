"""Row-band tile scheduler: the reference's own expectations (tests/snapshot_bus_smoke.cpp:1277-1362,
tests/multi_gpu_accumulation_smoke.cpp:25-60, :143-175) and the batching rule of next_tile_batch."""
from vkpathtracer_amd.tile_scheduler import (TILE_BATCH_MAX, TilePriorityFeedback, TileScheduler,
                                             TileSchedulerConfig, drain_batches)


def _drain(s):
    out = []
    while (t := s.next_tile()) is not None:
        out.append(t)
    return out


def test_bands_clamp_gpu_and_sample_stamp():
    s = TileScheduler(TileSchedulerConfig(64, 35, 16, 2))
    s.begin_sample(9, 3)
    tiles = _drain(s)
    assert len(tiles) == 3
    assert tiles[0][1] == 0 and tiles[0][3] == 16
    assert tiles[2][1] == 32 and tiles[2][3] == 3
    assert tiles[1][6] == 1
    assert tiles[2][4] == 3
    assert all(t[0] == 0 and t[2] == 64 for t in tiles)


def test_priority_variance_samples_then_dirty():
    s = TileScheduler(TileSchedulerConfig(32, 32, 8, 1))
    s.set_feedback([TilePriorityFeedback(0, 0.10, 9, False), TilePriorityFeedback(1, 0.80, 9, False),
                    TilePriorityFeedback(2, 0.80, 2, False), TilePriorityFeedback(3, 0.30, 1, False)])
    s.begin_sample(10, 4)
    assert [t[5] for t in _drain(s)] == [2, 1, 3, 0]
    s.set_feedback([TilePriorityFeedback(0, 0.10, 9, False), TilePriorityFeedback(1, 0.80, 9, False),
                    TilePriorityFeedback(2, 0.80, 2, False), TilePriorityFeedback(3, 0.01, 99, True)])
    s.begin_sample(10, 5)
    order = [t[5] for t in _drain(s)]
    assert order[0] == 3 and order[1:] == [2, 1, 0]
    # all dirty: the dirty bit is moot and the variance order holds
    s.set_feedback([TilePriorityFeedback(i, v, 1, True) for i, v in enumerate((0.2, 0.9, 0.5, 0.1))])
    s.begin_sample(10, 6)
    assert [t[5] for t in _drain(s)] == [1, 2, 0, 3]


def test_foveated_extra_center_slices_and_stats():
    s = TileScheduler(TileSchedulerConfig(64, 40, 8, 2, foveated_center_extra_samples=2, foveated_center_radius=0.11))
    s.begin_sample(12, 7)
    tiles = _drain(s)
    assert [t[5] for t in tiles] == [0, 1, 2, 3, 4, 2, 2]
    assert [t[4] for t in tiles] == [21, 21, 21, 21, 21, 22, 23]
    st = s.stats()
    assert st.tile_count == 5 and st.scheduled_tile_count == 7 and st.next_tile == 7
    assert st.gpu_scheduled_tile_count == [5, 2]
    assert st.generation == 12 and st.sample_index == 7


def test_accumulation_layout_and_scaling_shares():
    s = TileScheduler(TileSchedulerConfig(8, 8, 2, 2))
    s.begin_sample(21, 0)
    tiles = _drain(s)
    assert len(tiles) == 4
    for t in tiles:  # every row of the frame belongs to band y // tile_height on GPU band % 2
        assert t[1] == t[5] * 2 and t[6] == t[5] % 2
    for g, share in ((2, [8, 8]), (4, [4, 4, 4, 4])):
        s = TileScheduler(TileSchedulerConfig(64, 64, 4, g))
        s.begin_sample(0, 0)
        assert s.stats().gpu_scheduled_tile_count == share


def test_configure_clamps_and_empty_frames():
    s = TileScheduler(TileSchedulerConfig(16, 10, 0, 0, foveated_center_radius=3.0))
    s.begin_sample(0, 0)
    assert len(_drain(s)) == 10 and s.stats().gpu_scheduled_tile_count == [10]
    assert s.is_foveated_center_tile(0) and s.is_foveated_center_tile(9)  # radius clamped to 0.5
    for cfg in (TileSchedulerConfig(0, 10), TileSchedulerConfig(10, 0)):
        e = TileScheduler(cfg)
        e.begin_sample(0, 0)
        assert e.next_tile() is None and e.next_tile_batch(4, []) == 0
    s.set_feedback([TilePriorityFeedback(99, 1.0, 0, True)])  # out of range: ignored


def test_batches_share_a_gpu_and_respect_the_cap():
    s = TileScheduler(TileSchedulerConfig(32, 64, 4, 1))
    s.begin_sample(0, 1)
    assert [len(b) for b in drain_batches(s, 5)] == [5, 5, 5, 1]
    s = TileScheduler(TileSchedulerConfig(32, 64, 4, 2))
    s.set_feedback([TilePriorityFeedback(i, 1.0 if i in (0, 2, 4) else 0.0, 0, False) for i in range(16)])
    s.begin_sample(0, 1)
    batches = list(drain_batches(s))
    assert [b[0][5] for b in batches][:2] == [0, 1] and len(batches[0]) == 3  # bands 0, 2, 4 on GPU 0
    for b in batches:
        assert len({t[6] for t in b}) == 1 and len(b) <= TILE_BATCH_MAX
    out = [("keep",)]
    s.begin_sample(0, 2)
    assert s.next_tile_batch(2, out) == 2 and out[0] == ("keep",) and len(out) == 3  # appends


def test_order_is_kept_while_feedback_stays_within_its_buckets():
    """TileScheduler.cpp:184-223: feedback that moves a band's variance within its logarithmic bucket
    (or its sample count within a doubling bucket) keeps the previous order, even where the exact
    comparator would now sort differently; leaving the bucket re-sorts."""
    from vkpathtracer_amd.tile_scheduler import (TilePriorityFeedback, TileScheduler, TileSchedulerConfig,
                                                 sample_count_bucket, variance_bucket)

    s = TileScheduler(TileSchedulerConfig(width=8, height=48, tile_height=16))
    s.set_feedback([TilePriorityFeedback(0, 1.00, 4), TilePriorityFeedback(1, 1.01, 4), TilePriorityFeedback(2, 0.1, 4)])
    s.begin_sample(1, 0)
    first = [s.next_tile()[5] for _ in range(3)]
    assert first == [1, 0, 2]
    assert variance_bucket(1.02) == variance_bucket(1.01) == variance_bucket(1.00)
    s.set_feedback([TilePriorityFeedback(0, 1.02, 4)])  # band 0 now has the larger variance, same bucket
    s.begin_sample(1, 1)
    assert [s.next_tile()[5] for _ in range(3)] == first
    # band 2's sample count leaves its bucket (4 -> 33): re-sorted with the exact comparator
    s.set_feedback([TilePriorityFeedback(2, 0.1, 33)])
    s.begin_sample(1, 2)
    assert [s.next_tile()[5] for _ in range(3)] == [0, 1, 2]
    # within buckets again (33 -> 40, variance 1.01 -> 1.03): the order stays although band 1 now leads
    assert sample_count_bucket(40) == sample_count_bucket(33) and variance_bucket(1.03) == variance_bucket(1.01)
    s.set_feedback([TilePriorityFeedback(2, 0.1, 40), TilePriorityFeedback(1, 1.03, 4)])
    s.begin_sample(1, 3)
    assert [s.next_tile()[5] for _ in range(3)] == [0, 1, 2]
    s.set_feedback([TilePriorityFeedback(1, 5.0, 4)])  # another bucket: re-sorted
    s.begin_sample(1, 4)
    assert [s.next_tile()[5] for _ in range(3)] == [1, 0, 2]

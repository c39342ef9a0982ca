"""Row-band tile scheduling for progressive rendering through `render_tiles`.

Restates the behaviour of the reference's `TileScheduler` (src/render/TileScheduler.cpp:68-321,
TileScheduler.h:11-68) so that a caller of this library can drive a tracer the way the reference's
render coordinator does (RenderCoordinator.cpp:1350-1391: drain up to 64 same-GPU tiles with
`next_tile_batch`, hand them to `render_tiles`):

* the frame is cut into row bands of `tile_height` rows at full width; band `tile_id` belongs to GPU
  `tile_id % gpu_count`;
* each `begin_sample(generation, sample_index)` orders the bands by the caller's feedback: dirty
  bands first (only when some but not all are dirty), then higher variance, then fewer samples,
  then lower id (TileScheduler.cpp:184-285);
* with `foveated_center_extra_samples = k`, bands whose centre lies within `foveated_center_radius`
  of the frame's vertical middle are scheduled k more times after the ordinary pass, and the sample
  index of a slice is `sample_index * (k + 1) + offset` so every slice renders a distinct sample
  (TileScheduler.cpp:104-129, :287-321).

The order is rebuilt only when a fingerprint of its inputs changes (TileScheduler.cpp:184-223): an
FNV-1a mix of the band count, the foveated extra samples, tile height, image height and GPU count, and
per band its dirty flag, its variance quantised to about 256 logarithmic buckets and its sample count
bucket (exact below 16, then one bucket per power of two). Feedback that changes only within a bucket
therefore keeps the previous order, as in the reference. Tiles are the 7-tuples `(x, y, width, height, sample_index, tile_id, gpu_id)`
that `HipPathTracer.render_tiles` takes (the fields of `PtRenderTile`).
"""
from __future__ import annotations

import dataclasses

TILE_BATCH_MAX = 64  # RenderCoordinator.cpp:1350
_U32_MAX = 0xFFFFFFFF
_FNV_OFFSET = 0xCBF29CE484222325  # FNV-1a 64 (TileScheduler.cpp:18-27)
_FNV_PRIME = 0x100000001B3
_U64 = (1 << 64) - 1


def _fnv_mix(h: int, v: int) -> int:
    """Mixes the 8 bytes of v, low byte first."""
    for i in range(8):
        h ^= (v >> (8 * i)) & 0xFF
        h = (h * _FNV_PRIME) & _U64
    return h


def variance_bucket(variance: float) -> int:
    """VarianceBucket (TileScheduler.cpp:35-46): about 256 bins over log10 of [1e-9, 1e9]."""
    import math

    if not (variance > 0.0):
        return 0
    scaled = (math.log10(min(max(variance, 1e-9), 1e9)) + 9.0) * (256.0 / 18.0)
    return int(min(max(scaled, 0.0), 255.0))


def sample_count_bucket(samples: int) -> int:
    """SampleCountBucket (TileScheduler.cpp:48-58): exact below 16, then one bucket per doubling."""
    if samples < 16:
        return samples
    bucket, v = 16, samples
    while v >= 32:
        v >>= 1
        bucket += 1
    return bucket


@dataclasses.dataclass
class TileSchedulerConfig:
    """TileSchedulerConfig (TileScheduler.h:11-18)."""
    width: int = 0
    height: int = 0
    tile_height: int = 16
    gpu_count: int = 1
    foveated_center_extra_samples: int = 0
    foveated_center_radius: float = 0.25


@dataclasses.dataclass
class TilePriorityFeedback:
    """TilePriorityFeedback (TileScheduler.h:29-34)."""
    tile_id: int = 0
    variance: float = 0.0
    sample_count: int = 0
    dirty: bool = False


@dataclasses.dataclass
class TileSchedulerStats:
    """TileSchedulerStats (TileScheduler.h:20-27)."""
    generation: int = 0
    sample_index: int = 0
    tile_count: int = 0
    next_tile: int = 0
    scheduled_tile_count: int = 0
    gpu_scheduled_tile_count: list = dataclasses.field(default_factory=list)


class TileScheduler:
    def __init__(self, config: TileSchedulerConfig | None = None):
        self._config = TileSchedulerConfig()
        self._generation = 0
        self._sample_index = 0
        self._next = 0
        self._tile_count = 0
        self._feedback: list[TilePriorityFeedback] = []
        self._order: list[tuple[int, int]] = []  # (tile_id, sample_offset)
        self._fingerprint: int | None = None  # of the inputs the current order was built from
        if config is not None:
            self.configure(config)

    def configure(self, config: TileSchedulerConfig) -> None:
        """TileScheduler::configure (TileScheduler.cpp:68-86): clamps, counts the bands and resets
        the feedback and the order."""
        c = dataclasses.replace(config)
        c.tile_height = max(1, int(c.tile_height))
        c.gpu_count = max(1, int(c.gpu_count))
        c.foveated_center_radius = min(max(float(c.foveated_center_radius), 0.0), 0.5)
        self._config = c
        self._tile_count = 0 if c.height == 0 else (c.height + c.tile_height - 1) // c.tile_height
        self._feedback = [TilePriorityFeedback(tile_id=i) for i in range(self._tile_count)]
        self._order = []
        self._fingerprint = None
        self._next = 0

    def set_feedback(self, feedback) -> None:
        """Replaces the feedback of each listed band; ids outside the frame are ignored."""
        for item in feedback:
            if 0 <= item.tile_id < len(self._feedback):
                self._feedback[item.tile_id] = dataclasses.replace(item)

    def begin_sample(self, generation: int, sample_index: int) -> None:
        self._generation = int(generation)
        self._sample_index = int(sample_index)
        self._next = 0
        self._rebuild_order()

    def next_tile(self):
        """The next scheduled slice as a render_tiles tuple, or None when the pass is drained."""
        c = self._config
        if c.width == 0 or c.height == 0 or self._next >= len(self._order):
            return None
        tile_id, offset = self._order[self._next]
        self._next += 1
        y = tile_id * c.tile_height
        stride = 1 if c.foveated_center_extra_samples == 0 else c.foveated_center_extra_samples + 1
        sample = min(self._sample_index * stride + offset, _U32_MAX)
        return (0, y, c.width, min(c.tile_height, c.height - y), sample, tile_id, tile_id % c.gpu_count)

    def next_tile_batch(self, max_tiles: int, out: list) -> int:
        """Appends up to max_tiles consecutive slices that share the first one's GPU to `out` (not
        cleared) and returns how many (TileScheduler.cpp:132-166)."""
        c = self._config
        if max_tiles == 0 or c.width == 0 or c.height == 0 or self._next >= len(self._order):
            return 0
        first = self.next_tile()
        out.append(first)
        drained = 1
        while drained < max_tiles and self._next < len(self._order):
            if self._order[self._next][0] % c.gpu_count != first[6]:
                break
            out.append(self.next_tile())
            drained += 1
        return drained

    def stats(self) -> TileSchedulerStats:
        per_gpu = [0] * max(1, self._config.gpu_count)
        for tile_id, _ in self._order:
            per_gpu[tile_id % self._config.gpu_count] += 1
        return TileSchedulerStats(self._generation, self._sample_index, self._tile_count, self._next,
                                  len(self._order), per_gpu)

    def is_foveated_center_tile(self, tile_id: int) -> bool:
        """A band whose centre row fraction is within the radius of 0.5 (TileScheduler.cpp:307-321)."""
        c = self._config
        if self._tile_count == 0 or c.height == 0:
            return False
        y = tile_id * c.tile_height
        if y >= c.height:
            return False
        h = min(c.tile_height, c.height - y)
        center = (float(y) + float(h) * 0.5) / float(c.height)
        return abs(center - 0.5) <= c.foveated_center_radius

    def _fingerprint_of_inputs(self) -> int:
        c = self._config
        h = _FNV_OFFSET
        for v in (self._tile_count, c.foveated_center_extra_samples, c.tile_height, c.height, c.gpu_count):
            h = _fnv_mix(h, int(v) & _U64)
        for t in range(self._tile_count):
            f = self._feedback[t]
            packed = ((1 if f.dirty else 0) << 56) | ((variance_bucket(f.variance) & 0xFF) << 48) | \
                     ((sample_count_bucket(int(f.sample_count)) & 0xFFFF) << 32) | t
            h = _fnv_mix(h, packed)
        return h

    def _rebuild_order(self) -> None:
        fingerprint = self._fingerprint_of_inputs()
        if self._order and fingerprint == self._fingerprint:
            return  # the bucketed sort keys are unchanged: keep the previous order
        self._fingerprint = fingerprint
        fb = self._feedback
        dirty = sum(1 for f in fb if f.dirty)
        mixed = 0 < dirty < self._tile_count
        # The comparator is a total order on distinct ids (finite variances): dirty first when the
        # bands are mixed, then variance descending, sample count ascending, id ascending.
        ranked = sorted(range(self._tile_count),
                        key=lambda t: ((0 if fb[t].dirty else 1) if mixed else 0, -fb[t].variance,
                                       fb[t].sample_count, t))
        order = [(t, 0) for t in ranked]
        extra = self._config.foveated_center_extra_samples
        for offset in range(1, extra + 1):
            order.extend((t, offset) for t in ranked if self.is_foveated_center_tile(t))
        self._order = order


def drain_batches(scheduler: TileScheduler, max_tiles: int = TILE_BATCH_MAX):
    """Yields the pass's batches as RenderCoordinator's batching loop forms them."""
    while True:
        batch: list = []
        if scheduler.next_tile_batch(max_tiles, batch) == 0:
            return
        yield batch

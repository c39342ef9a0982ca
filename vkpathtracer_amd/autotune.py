"""Launch-shape and BVH autotuner (no counterpart in the reference, whose CPU path has no launch shapes).

The launch shapes never change a result: films, counters and gradients are identical for every
choice (the GPU tests check this), only the time a frame takes. The leaf width changes the tree, so
the traversal counters change with it, and a film can change only where a ray meets two triangles at
exactly the same distance (which one the closest-hit test keeps depends on the visit order).

  leaf_size          BVH leaf width (binned SAH, triangles per leaf); needs a rebuild
  wavefront_capacity paths in flight per wavefront pass (queue and path-state sizes); a frame larger
                     than the capacity runs in several passes
  refill_min         idle lanes a persistent traversal wave gathers before it takes new rays
  traversal_blocks   persistent workgroups of the extend / shadow kernels (0 = all resident)
  shade_blocks       workgroups of the grid-stride shade kernel (0 = all resident)

`tune` is coordinate descent over a discrete space: each round tries every candidate of one
parameter with the others fixed and keeps the fastest, until a round changes nothing. Frames are
timed with the GPU synchronised on both sides; each candidate takes the best of `reps` frames after
one warm-up frame. `tune_tracer` tunes a HipPathTracer in place for a list of tiles; `tune_batch`
also tunes the leaf width of a batch of scenes (one build per leaf width) and returns a HipBatch
built and configured with the winner.
"""
from __future__ import annotations

import dataclasses
import time
from typing import Callable

LEAF_SIZES = (2, 4, 8)
REFILL = (4, 8, 16, 32)
BLOCK_FRACTIONS = (1.0, 0.75, 0.5)


@dataclasses.dataclass
class TuneResult:
    params: dict
    ms: float
    default_ms: float
    trials: list

    @property
    def speedup(self) -> float:
        return self.default_ms / self.ms if self.ms > 0 else 1.0


def tune(measure: Callable[[dict], float], space: dict, start: dict, max_rounds: int = 3,
         log: Callable[[str], None] | None = None) -> TuneResult:
    """Coordinate descent: `measure(params) -> ms` over `space` (parameter -> candidate list),
    starting from `start` (which must be a point of the space)."""
    cache: dict[tuple, float] = {}
    trials = []

    def run(p):
        key = tuple(sorted(p.items()))
        if key not in cache:
            cache[key] = measure(dict(p))
            trials.append({"params": dict(p), "ms": cache[key]})
            if log:
                log(f"autotune {p}: {cache[key]:.2f} ms")
        return cache[key]

    best = dict(start)
    best_ms = default_ms = run(best)
    for _ in range(max_rounds):
        changed = False
        for name, cands in space.items():
            for c in cands:
                if c == best[name]:
                    continue
                p = dict(best, **{name: c})
                ms = run(p)
                if ms < best_ms * 0.995:  # a win must beat timing noise
                    best, best_ms, changed = p, ms, True
        if not changed:
            break
    return TuneResult(best, best_ms, default_ms, trials)


def _timed(frame: Callable[[], None], sync: Callable[[], None], reps: int) -> float:
    frame()
    sync()
    best = float("inf")
    for _ in range(reps):
        sync()
        t0 = time.perf_counter()
        frame()
        sync()
        best = min(best, (time.perf_counter() - t0) * 1e3)
    return best


PATH_STATE_BYTES = 200  # device bytes per path of wavefront capacity (forward path state and queues)


def _capacities(paths: int, free_bytes: int | None = None) -> list:
    """The default (the whole frame in one pass, up to 128 Mi paths), the whole frame, twice the
    default, and 1/2 and 1/4 of the frame. At most 1 Gi paths (about 200 GB of path state, which an
    MI355X's 288 GB hold) and at most what 80 % of the device's free memory holds."""
    limit = 1 << 30
    if free_bytes is None:
        try:
            import torch

            free_bytes = torch.cuda.mem_get_info()[0]
        except Exception:  # no device: the static limit only
            free_bytes = None
    if free_bytes is not None:
        limit = min(limit, int(0.8 * free_bytes) // PATH_STATE_BYTES)
    default = max(1024, min(paths, 128 << 20))
    cands = [default, paths, min(paths, 256 << 20), paths // 2, paths // 4]
    out = []
    for c in cands:
        c = max(1024, min(int(c), limit))
        if c not in out:
            out.append(c)
    return out


def _runtime_space(target, paths: int) -> tuple[dict, dict]:
    """Candidates for the runtime parameters of a tracer or batch, and the default point."""
    target.set_tuning(0, 0, 0, 0)  # occupancy-derived grids
    d = target.tuning()
    trav, shade = d["traversal_blocks"], d["shade_blocks"]
    space = {
        "wavefront_capacity": _capacities(paths),
        "refill_min": list(REFILL),
        "traversal_blocks": sorted({max(1, int(trav * f)) for f in BLOCK_FRACTIONS}, reverse=True),
        "shade_blocks": sorted({max(1, int(shade * f)) for f in BLOCK_FRACTIONS}, reverse=True),
    }
    start = {"wavefront_capacity": space["wavefront_capacity"][0], "refill_min": 8 if 8 in REFILL else REFILL[0],
             "traversal_blocks": trav, "shade_blocks": shade}
    return space, start


def _apply(target, p: dict):
    target.set_tuning(p["wavefront_capacity"], p["refill_min"], p["traversal_blocks"], p["shade_blocks"])


def tune_tracer(pt, tiles, frame_index: int = 0, reps: int = 2, log=None) -> TuneResult:
    """Tune a built HipPathTracer for render_tiles(tiles); leaves the winner applied."""
    import torch

    paths = sum(int(t[2]) * int(t[3]) for t in tiles)
    space, start = _runtime_space(pt, paths)

    def frame():
        pt.reset_accumulation()
        if not pt.render_tiles(tiles, frame_index):
            raise RuntimeError(pt.status().last_error)

    def measure(p):
        _apply(pt, p)
        return _timed(frame, lambda: (pt.synchronize(), torch.cuda.synchronize()), reps)

    res = tune(measure, space, start, log=log)
    _apply(pt, res.params)
    return res


def tune_batch(jobs, settings, leaf_sizes=LEAF_SIZES, reps: int = 2, use_graph: bool = False, log=None):
    """Tune a batch of (scene, width, height, spp) jobs: the runtime parameters for each BVH leaf
    width, then the best leaf width. Returns (TuneResult with "leaf_size" in params, the HipBatch built
    with the winning leaf width and tuned). The default point is leaf 4 with default launch shapes."""
    import torch

    from .tracer import HipBatch

    paths = sum(w * h * spp for _, w, h, spp in jobs)
    best = None
    default_ms = None
    trials = []
    order = sorted(leaf_sizes, key=lambda L: L != 4)  # leaf 4 first: it holds the default point
    for leaf in order:
        b = HipBatch(settings)
        for sc, w, h, spp in jobs:
            b.add_scene(sc, w, h, spp, leaf_size=leaf)
        b.build()

        def frame(b=b):
            b.reset_accumulation()
            b.render(0, use_graph=use_graph)

        def measure(p, b=b, frame=frame):
            _apply(b, p)
            return _timed(frame, lambda: (b.synchronize(), torch.cuda.synchronize()), reps)

        space, start = _runtime_space(b, paths)
        res = tune(measure, space, start, log=(lambda m, leaf=leaf: log(f"leaf {leaf}: {m}")) if log else None)
        trials += [dict(t, leaf_size=leaf) for t in res.trials]
        if leaf == 4:
            default_ms = res.default_ms
        if best is None or res.ms < best[0].ms:
            if best is not None:
                best[1].close()
            _apply(b, res.params)
            best = (res, b, leaf)
        else:
            b.close()
    res, b, leaf = best
    return TuneResult(dict(res.params, leaf_size=leaf), res.ms,
                      default_ms if default_ms is not None else res.default_ms, trials), b


TABLE = __import__("pathlib").Path(__file__).resolve().parent / "tuning" / "mi355x.json"


def tuned_params(workload: str) -> dict | None:
    """The parameters the autotuner chose for a named workload on MI355X (the committed tuning table),
    or None when the table has no entry."""
    import json

    try:
        return dict(json.loads(TABLE.read_text())["workloads"][workload]["params"])
    except (OSError, KeyError, ValueError):
        return None


def fitting_capacity(capacity: int, free_bytes: int | None = None) -> int:
    """`capacity` paths, or fewer when their path state would not fit in 80 % of the device's free
    memory (at least 1 Mi paths): a tuned whole-frame pass on a device with less free memory renders in
    more, smaller passes instead of failing to allocate."""
    if free_bytes is None:
        try:
            import torch

            free_bytes = torch.cuda.mem_get_info()[0]
        except Exception:  # no device to ask
            return capacity
    return max(min(capacity, 1 << 20), min(capacity, int(0.8 * free_bytes) // PATH_STATE_BYTES))


def apply_params(target, params: dict, free_bytes: int | None = None) -> None:
    """Apply the launch-shape part of `params` to a HipPathTracer or HipBatch (the leaf width is a
    build parameter), with the wavefront capacity limited to what the device's free memory holds."""
    _apply(target, dict(params, wavefront_capacity=fitting_capacity(params["wavefront_capacity"], free_bytes)))


def main(argv=None):
    """Tune a named workload on this GPU and write the result as JSON:
    python -m vkpathtracer_amd.autotune batch64 --out FILE   (64 seeded scenes, batched forward)
    python -m vkpathtracer_amd.autotune procedural --out FILE (1M triangles, 1920x1080, 64 spp)"""
    import argparse
    import json
    import sys

    from . import scene as S

    ap = argparse.ArgumentParser(description=main.__doc__.split("\n")[0])
    ap.add_argument("workload", choices=["batch64", "procedural"])
    ap.add_argument("--out", required=True)
    ap.add_argument("--reps", type=int, default=2)
    args = ap.parse_args(argv)

    def log(m):
        print(f"[autotune] {m}", file=sys.stderr, flush=True)

    t0 = time.perf_counter()
    if args.workload == "batch64":
        jobs = S.batch_scenes(seed=7, count=64)
        settings = S.RenderSettings(width=1, height=1, spp=1, max_depth=8)
        res, b = tune_batch(jobs, settings, reps=args.reps, log=log)
        b.close()
        desc = "64 seeded scenes (batch_scenes(seed=7)), depth 8, eager batch frame"
    else:
        from .tracer import HipPathTracer

        sc = S.procedural_scene()
        st = S.RenderSettings(width=1920, height=1080, spp=64, max_depth=8)
        res = None
        for leaf in LEAF_SIZES:
            with HipPathTracer(0) as pt:
                assert pt.configure(st) and pt.load_scene_snapshot(sc) and pt.build_or_update_acceleration("sah", leaf)
                r = tune_tracer(pt, [(0, 0, st.width, st.height, s, 0, 0) for s in range(st.spp)], reps=args.reps,
                                log=lambda m, leaf=leaf: log(f"leaf {leaf}: {m}"))
            if leaf == 4:
                default_ms = r.default_ms
            if res is None or r.ms < res.ms:
                res = TuneResult(dict(r.params, leaf_size=leaf), r.ms, 0.0, r.trials)
        res.default_ms = default_ms
        desc = "procedural 1M triangles, 1920x1080, 64 spp, depth 8"
    out = {"workload": args.workload, "description": desc, "params": res.params, "ms": round(res.ms, 3),
           "default_ms": round(res.default_ms, 3), "speedup": round(res.speedup, 4), "trials": len(res.trials),
           "tuning_s": round(time.perf_counter() - t0, 1)}
    with open(args.out, "w") as f:
        json.dump(out, f, indent=1)
    log(json.dumps(out))


if __name__ == "__main__":
    main()

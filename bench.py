#!/usr/bin/env python3
"""Forward path-tracing throughput on one MI355X.

A step renders one full frame of the workload: reset_accumulation, then render_tiles with one
full-frame tile per sample (spp tiles) through the C ABI, exactly what ptbench's render loop
does per frame (BenchmarkRuntime.cpp:680-695). Inputs (scene, BVH) are resident in HBM before
the timed region; the BVH build is reported separately as the reference does.

Default workload: procedural 1M-triangle scene (displaced 655,360-triangle icosphere + 2048
instanced 168-triangle spheres + 16 quad lights), 1920x1080, 64 spp, depth 8, NEE + MIS, Russian
roulette from depth 3, binned-SAH BVH with leaves of at most 4 triangles, seed 0xBADC0FFEE.

Prints one JSON line. `value` is Mpaths/s (W*H*spp per second, the reference's samples_per_sec);
Mrays/s (extension + shadow rays) and the reference's paths_per_sec (= extension rays/s) are extra
fields. `roofline` is computed live from HIP events around every kernel launch on the tracer's
stream and the kernels' own work counters.
"""
from __future__ import annotations

import argparse
import dataclasses
import json
import os
import platform
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBPS = 8000.0  # MI355X HBM3E

# Algorithmic bytes per unit of work (the per-unit figures of the hot-path measurement contract,
# with reference SampleCounters semantics): 32 bytes per BVH node visit (one child box), 48 per
# triangle test (v0, e1, e2 padded to float4 records), an extension ray 32 in + 16 out, a shadow ray
# 32 in + 12 contribution + 4 result, a shade invocation 2 x 40 path state + 64 material, a film
# sample 16 (read-modify-write); raygen writes 36 bytes per path (ray 24, RNG 8, scene index 4; the
# first bounce takes throughput 1, radiance 0 and the identity path slot without reading them).
BYTES = {
    "node_visit": 32, "triangle": 48, "ext_ray": 32 + 16, "shadow_ray": 32 + 12 + 4, "shade": 2 * 40 + 64,
    "film_sample": 16, "raygen_path": 36,
}

# HBM bytes per launch measured with rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE on this workload
# (calibrated against a plain device copy: FETCH_SIZE x 2, WRITE_SIZE x 1), from
# profiles/wavefront_procedural_1080p_64spp.md. They describe the current kernels; re-measure after
# a kernel change.
# VALU issue of the same kernels from the same profile: VALU busy = SQ_ACTIVE_INST_VALU / (256 CUs x the
# launch's cycles on one XCD; rocprof's VALUBusy), lane utilisation = SQ_THREAD_CYCLES_VALU /
# (64 x SQ_ACTIVE_INST_VALU), L2 hit = TCC_HIT / (TCC_HIT + TCC_MISS).
MEASURED_VALU = {"extend": {"valu_busy": 0.86, "lane_utilisation": 0.47, "l2_hit": 0.802},
                 "shadow": {"valu_busy": 0.82, "lane_utilisation": 0.4, "l2_hit": 0.839},
                 "shade": {"valu_busy": 0.67, "lane_utilisation": 0.34, "l2_hit": 0.473}}
MEASURED_TRAFFIC = {"extend": 18.46e9, "shadow": 9.072e9, "shade": 6.118e9, "film": 1.677e9, "raygen": 4.778e9}
# The two tables above describe the kernels built from sources whose sha256 (over KERNEL_SOURCES, in
# order) is PROFILED_KERNEL_SHA256; when the sources differ they are dropped from the line.
KERNEL_SOURCES = ("vkpathtracer_amd/csrc/pt_kernels.hip", "vkpathtracer_amd/csrc/pt_device.h",
                  "vkpathtracer_amd/csrc/pt_math.h")
PROFILED_KERNEL_SHA256 = "06884899cbd6361fd3f4d1da49be652cc4e45fa8187c08ad1a5273a6f0c7cfea"


def kernel_source_sha256() -> str:
    import hashlib

    h = hashlib.sha256()
    for rel in KERNEL_SOURCES:
        with open(os.path.join(ROOT, rel), "rb") as f:
            h.update(f.read())
    return h.hexdigest()


def cpu_model() -> str:
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return platform.processor() or "unknown"


def workload(name: str, spp: int | None):
    from vkpathtracer_amd import scene as S

    if name == "procedural":
        sc = S.procedural_scene(seed=1)
        st = S.RenderSettings(width=1920, height=1080, spp=spp or 64, max_depth=8)
        desc = "procedural 1M-tri (displaced icosphere 655360 + 2048x168 spheres + 16 quad lights), 1920x1080"
    else:
        sc = S.cornell_box()
        st = S.RenderSettings(width=256, height=256, spp=spp or 16, max_depth=8)
        desc = "cornell box 34 tris + light quad, 256x256"
    return sc, st, desc


def kernel_bytes(ks: dict, counters: dict) -> dict:
    ext_rays = counters["rays"]
    sh_rays = ks["shadow_rays"]
    ext_visits = counters["bvh_node_visits"] - ks["shadow_node_visits"]
    ext_tris = counters["triangle_tests"] - ks["shadow_triangle_tests"]
    return {
        "extend": ext_rays * BYTES["ext_ray"] + ext_visits * BYTES["node_visit"] + ext_tris * BYTES["triangle"],
        "shadow": sh_rays * BYTES["shadow_ray"] + ks["shadow_node_visits"] * BYTES["node_visit"]
                  + ks["shadow_triangle_tests"] * BYTES["triangle"],
        "shade": ext_rays * BYTES["shade"],
        "raygen": ks["raygen"]["items"] * BYTES["raygen_path"],
        "film": ks["film"]["items"] * BYTES["film_sample"],
    }


def _work_per_ray(ks: dict, counters: dict) -> dict:
    """Traversal work per query of one step, extension and shadow apart (SampleCounters semantics:
    node visits count every box test, leaves every leaf entered, tests every triangle tested)."""
    ext = {"visits": counters["bvh_node_visits"] - ks["shadow_node_visits"],
           "leaves": counters["bvh_leaf_visits"] - ks["shadow_leaf_visits"],
           "tests": counters["triangle_tests"] - ks["shadow_triangle_tests"],
           "inner_fetches": ks["inner_node_fetches"] - ks["shadow_inner_node_fetches"]}
    sh = {"visits": ks["shadow_node_visits"], "leaves": ks["shadow_leaf_visits"],
          "tests": ks["shadow_triangle_tests"], "inner_fetches": ks["shadow_inner_node_fetches"]}
    out = {}
    for name, work, rays in (("extension", ext, counters["rays"]), ("shadow", sh, ks["shadow_rays"])):
        out[name] = {k: round(v / rays, 2) if rays else None for k, v in work.items()}
    return out


def _has_avx2() -> bool:
    try:
        with open("/proc/cpuinfo") as f:
            return " avx2 " in f.read().replace("\n", " ")
    except OSError:
        return False


def _cpu_rate(sc, st, variant: str, leaf: int, threads: int, target_seconds: float, band_rows: int = 16,
              samples: int = 4):
    """Paths/s of the oracle `variant` (BVH leaves <= leaf) with the reference's tiled CPU schedule:
    per sample, every band of band_rows whole rows is one job for one thread (pto_render_bands).
    Every band_stride-th band is rendered, the stride chosen from a one-band probe so that samples
    0..samples-1 take about target_seconds; returns (Mpaths/s, rows, bands, seconds)."""
    from oracle import oracle as O

    o = O.Oracle(sc, builder="sah", leaf_size=leaf, variant=variant)
    n_bands = (st.height + band_rows - 1) // band_rows
    mid = n_bands // 2
    o.render_bands(st, band_rows, mid, n_bands, 0, 1, threads=1)  # warm-up
    t0 = time.perf_counter()
    o.render_bands(st, band_rows, mid, n_bands, 0, 1, threads=1)
    per_band = max(time.perf_counter() - t0, 1e-6)  # one thread, one band, one sample
    bands_wanted = target_seconds * threads / (per_band * samples)
    stride = max(1, int(n_bands // max(bands_wanted, 1.0)))
    first = (stride // 2) if stride > 1 else 0
    bands = len(range(first, n_bands, stride))
    rows = sum(min(band_rows, st.height - b * band_rows) for b in range(first, n_bands, stride))
    t0 = time.perf_counter()
    o.render_bands(st, band_rows, first, stride, 0, samples, threads=threads)
    dt = time.perf_counter() - t0
    return rows * st.width * samples / dt / 1e6, rows, bands, dt


def cpu_baseline(sc, st, target_seconds: float, log) -> dict:
    """The CPU restatement (oracle, fp32, glibc sinf / cosf / powf as the reference is built) on the host
    cores with TiledCpuPathTracer's schedule (16-row bands, one band per job per sample), three ways, each
    on a bounded set of bands of the same frame, samples 0-3: scalar = one thread, leaves of 4;
    multicore = every core, leaves of 4; avx2 = every core, leaves of 8 intersected 8 triangles at a time
    with AVX2 (the reference's AVX2 configuration, CpuBvhAccelerator.cpp:252-256). `value` is the AVX2 rate
    (multicore when the host lacks AVX2)."""
    cores = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else (os.cpu_count() or 1)
    modes = {}
    plan = [("scalar", "f32_libm", 4, 1), ("multicore", "f32_libm", 4, cores)]
    if _has_avx2():
        plan.append(("avx2", "f32_avx2_libm", 8, cores))
    for name, variant, leaf, threads in plan:
        rate, rows, bands, dt = _cpu_rate(sc, st, variant, leaf, threads, target_seconds)
        modes[name] = {"mpaths_per_s": round(rate, 4), "threads": threads, "leaf": leaf, "oracle_build": variant,
                       "sample": f"{bands} bands of 16 rows ({rows} x {st.width} pixels) x 4 samples, {dt:.1f} s"}
        log(f"cpu {name}: {rate:.3f} Mpaths/s ({threads} threads, {bands} bands)")
    top = "avx2" if "avx2" in modes else "multicore"
    m = modes[top]
    return {"value": m["mpaths_per_s"], "unit": "Mpaths/s", "cores": m["threads"], "kind": "port",
            "mode": top,
            "sample": f"{m['sample']} of the same frame; oracle fp32 restatement with glibc math "
                      f"({'AVX2 leaf-8 intersector, ' if top == 'avx2' else ''}16-row band jobs as TiledCpuPathTracer, "
                      f"OpenMP over {m['threads']} threads), {cpu_model()}",
            "modes": modes}


def parity_report(pt, sc, st, leaf: int, log, pixels: int = 1500, rays: int = 250_000) -> dict:
    """The benchmark frame against the CPU oracle (fp32, the arithmetic the kernels restate), outside
    the timed region: the film's per-pixel mean radiance on `pixels` seeded pixels against the oracle's
    over the same samples (relative RMSE and the fraction of pixels bit-identical); closest-hit primitive
    ids and t of `rays` seeded random rays (mismatch counts); and the fraction of their triangle tests
    that the |det| < 1e-4 cut-off rejects (BvhTriangleIntersector.cpp:48)."""
    from oracle import oracle as O
    from vkpathtracer_amd import scene as S

    o = O.Oracle(sc, builder="sah", leaf_size=leaf)
    film, cnt = pt.film()
    px = np.sort(np.random.default_rng(7).choice(st.width * st.height, pixels, replace=False)).astype(np.uint32)
    acc = np.zeros((len(px), 3), np.float32)
    n = np.zeros(len(px), np.uint32)
    for s_ in range(st.spp):
        rgb, valid, _ = o.render_pixels(st, px, s_)
        rgb = np.asarray(rgb, np.float32)
        acc[valid] = acc[valid] + rgb[valid]
        n += valid
    g = film.reshape(-1, 3)[px]
    gm = g.astype(np.float64) / np.maximum(cnt.reshape(-1)[px], 1)[:, None]
    om = acc.astype(np.float64) / np.maximum(n, 1)[:, None]
    den = float(np.sqrt(np.mean(om * om))) or 1.0
    relrmse = float(np.sqrt(np.mean((gm - om) ** 2))) / den
    same = float(np.mean(np.all(g == acc, axis=1) & (cnt.reshape(-1)[px] == n)))
    r = S.random_rays(sc, rays, seed=11)
    t, p, _ = pt.trace_rays(r)
    rt, rp, _ = o.trace_rays(r)
    tests, rejects = o.det_rejections(r)
    out = {"relrmse": relrmse, "pixels": len(px), "spp": st.spp, "bitexact_pixel_fraction": round(same, 6),
           "rays": rays, "primid_mismatch": int(np.count_nonzero(p != rp)),
           "t_mismatch": int(np.count_nonzero((t != rt) & ~(np.isnan(t) & np.isnan(rt)))),
           "det_reject_fraction": rejects / tests if tests else None, "triangle_tests": tests,
           "oracle": "fp32 restatement (oracle/pto.c, portable math)"}
    log(f"parity: relRMSE {relrmse:.3e} over {len(px)} pixels x {st.spp} spp, {same:.4f} of pixels bit-identical, "
        f"{out['primid_mismatch']} primitive-id mismatches in {rays} rays, |det| rejections {rejects}/{tests}")
    return out


def adjoint_ratio(log, reps: int = 3) -> dict:
    """Adjoint/forward time on the inverse-rendering box: 512x512, 64 spp, depth 8, the 35 albedos, the
    emission and the light colour as parameters. forward = reset + render_tiles over the 64 sample
    tiles. backward = render_adjoint of the same tiles right after that forward, as in an autograd
    step: it reuses each path's forward radiance and replays the paths once, recomputing every vertex.
    standalone = render_adjoint of a frame that no forward rendered (frame 1: forward replay + adjoint
    replay; an adjoint keeps the forward radiance it reused, so a repeated call would reuse it again).
    ratio = (forward + backward) / forward: the time of a forward and its adjoint over the forward's
    alone; standalone_ratio = standalone / forward; backward_over_forward = backward / forward."""
    import torch

    from vkpathtracer_amd import scene as S
    from vkpathtracer_amd.tracer import HipPathTracer

    sc = S.cornell_box()
    st = S.RenderSettings(width=512, height=512, spp=64, max_depth=8)
    tiles = [(0, 0, st.width, st.height, s, 0, 0) for s in range(st.spp)]
    dev = torch.device("cuda", 0)
    with HipPathTracer(0) as pt:
        assert pt.configure(st) and pt.load_scene_snapshot(sc) and pt.build_or_update_acceleration()
        g = torch.randn((st.height, st.width, 3), device=dev)
        ga = torch.zeros((len(sc.materials), 3), device=dev)
        ge = torch.zeros_like(ga)
        gl = torch.zeros((len(sc.lights), 3), device=dev)

        def fwd():
            pt.reset_accumulation()
            assert pt.render_tiles(tiles)
            pt.synchronize()

        def adj(frame=0):
            pt.render_adjoint(tiles, g, ga, ge, gl, frame)

        def timed(fn):
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            fn()
            torch.cuda.synchronize()
            return (time.perf_counter() - t0) * 1e3

        fwd()
        adj()
        f_ms, b_ms, s_ms = [], [], []
        for _ in range(reps):
            f_ms.append(timed(fwd))
            b_ms.append(timed(adj))   # right after the forward: reuses its radiance
            s_ms.append(timed(lambda: adj(1)))  # nothing to reuse: replays forward, then adjoint
    f, b, sa = min(f_ms), min(b_ms), min(s_ms)
    log(f"forward {f:.1f} ms, backward {b:.1f} ms, standalone adjoint {sa:.1f} ms")
    return {"workload": "cornell box 512x512, 64 spp, depth 8, d/d(35 albedos, 35 emissions, light colour)",
            "forward_ms": round(f, 2), "backward_ms": round(b, 2), "standalone_adjoint_ms": round(sa, 2),
            "ratio": round((f + b) / f, 3), "standalone_ratio": round(sa / f, 3),
            "backward_over_forward": round(b / f, 3)}


def textured_multiview(log, reps: int = 2, views: int = 16, size: int = 512, spp: int = 32) -> dict:
    """Batched differentiable rendering: the 200,000-triangle displaced geodesic sphere with a 1024^2 rgb
    fp32 value-noise albedo texture (8 glossy materials with their own roughness, a ground quad, 6 quad
    lights), `views` cameras on a ring in one batch (the scene plus views-1 views), size^2 x spp each,
    depth 6. forward = reset + pt_batch_render of every view's frame; backward = pt_batch_render_adjoint
    of the same frame right after it (d/d albedo, emission, light colour, roughness and every texel,
    fp32 atomics; the forward radiance of the single pass is reused); standalone = the adjoint of a frame
    no forward rendered (frame 1). ratio = (forward + backward) / forward; standalone_ratio = standalone /
    forward."""
    import torch

    from vkpathtracer_amd import scene as S
    from vkpathtracer_amd.tracer import HipBatch

    sc = S.textured_sphere_scene()
    cams = S.ring_cameras(views)
    settings = S.RenderSettings(width=1, height=1, spp=1, max_depth=6)
    paths = views * size * size * spp
    dev = torch.device("cuda", 0)

    def timed(fn):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        fn()
        torch.cuda.synchronize()
        return (time.perf_counter() - t0) * 1e3

    with HipBatch(settings) as b:
        b.add_scene(dataclasses.replace(sc, camera=cams[0]), size, size, spp)
        for c in cams[1:]:
            b.add_view(0, c, size, size, spp)
        info = b.build()
        nm, nl = len(sc.materials), len(sc.lights)
        nt = sc.textures[0].shape[0] * sc.textures[0].shape[1]
        g = torch.randn((views, size, size, 3), device=dev)
        grads = [torch.zeros(shape, device=dev) for shape in ((nm, 3), (nm, 3), (nl, 3), (nm,), (nt, 3))]

        def fwd():
            b.reset_accumulation()
            b.render(0)
            b.synchronize()

        def adj(frame=0):
            b.render_adjoint(g, grads[0], grads[1], grads[2], frame, grad_roughness=grads[3], grad_texels=grads[4])

        fwd()
        adj()
        f_ms, b_ms, s_ms = [], [], []
        for _ in range(reps):
            f_ms.append(timed(fwd))
            b_ms.append(timed(adj))
            s_ms.append(timed(lambda: adj(1)))
        counters = b.read_counters()
    f, bk, sa = min(f_ms), min(b_ms), min(s_ms)
    log(f"textured multi-view: forward {f:.1f} ms, backward {bk:.1f} ms, standalone adjoint {sa:.1f} ms")
    return {"workload": f"{sc.triangle_count}-triangle displaced textured sphere, 1024^2 rgb fp32 texture, {views} "
                        f"views x {size}^2 x {spp} spp, depth 6, d/d(albedo, emission, light, roughness, "
                        f"{nt} texels)",
            "paths": paths, "bvh_build_ms": round(info["build_ms"], 1),
            "forward_ms": round(f, 2), "forward_mpaths_per_s": round(paths / f / 1e3, 2),
            "backward_ms": round(bk, 2), "standalone_adjoint_ms": round(sa, 2),
            "ratio": round((f + bk) / f, 3), "standalone_ratio": round(sa / f, 3),
            "backward_over_forward": round(bk / f, 3), "rays_per_frame": counters["rays"]}


# The render times the reference publishes for this path, all on its Cornell scene
# assets/scenes/cornell_native.json (48 triangles) with ptbench's defaults: max depth 6, NEE + MIS,
# seed 0xBADC0FFEE. Samples per pixel = spp x rays per pixel per dispatch. Each entry: published
# render time (ms) of a frame, and where it is quoted.
REFERENCE_PUBLISHED = [
    {"width": 96, "height": 96, "samples": 1, "published": {
        "cpu_scalar_ms": (5.4613, "experiments/pathtracer_optimization_graph/NOTES.md:165"),
        "vulkan_compute_arc_b580_ms": (4.5266, "experiments/pathtracer_optimization_graph/NOTES.md:166-167")}},
    {"width": 64, "height": 64, "samples": 2, "published": {
        "cpu_scalar_ms": (6.0917, "experiments/d3d12_dxr_optimization_graph/NOTES.md:36"),
        "d3d12_compute_ms": (1.6897, "experiments/d3d12_dxr_optimization_graph/NOTES.md:36")}},
    {"width": 64, "height": 64, "samples": 16, "published": {
        "d3d12_compute_rpp8_best_bvh_ms": (4.11, "experiments/d3d12_dxr_algorithm_graph/NOTES.md:22,53-55")}},
    {"width": 128, "height": 128, "samples": 16, "published": {
        "d3d12_compute_rpp8_ms": (5.04725, "experiments/d3d12_dxr_algorithm_graph/NOTES.md:95")}},
]


def reference_configs(log, reps: int = 30) -> dict:
    """The reference's own published configurations: its Cornell scene (the committed copy of
    cornell_native.json, converted by the scene-document ingest) at each published resolution and sample
    count, ptbench's protocol (BVH built before timing; a frame = reset_accumulation + the frame's samples,
    here one render_tiles call), the median of `reps` frames, beside the published render time of the same
    frame on the reference's CPU and GPU backends."""
    import torch

    from vkpathtracer_amd import scene as S
    from vkpathtracer_amd import scene_document as SD
    from vkpathtracer_amd.tracer import HipPathTracer

    sc = SD.load_scene_document(os.path.join(ROOT, "tests", "golden", "scenes", "cornell_native.json"))
    rows = []
    for cfg in REFERENCE_PUBLISHED:
        st = S.RenderSettings(width=cfg["width"], height=cfg["height"], spp=cfg["samples"], max_depth=6)
        tiles = [(0, 0, st.width, st.height, s, 0, 0) for s in range(st.spp)]
        with HipPathTracer(0) as pt:
            assert pt.configure(st) and pt.load_scene_snapshot(sc) and pt.build_or_update_acceleration()

            def frame():
                pt.reset_accumulation()
                assert pt.render_tiles(tiles)

            for _ in range(3):
                frame()
            pt.synchronize()
            times = []
            for _ in range(reps):
                torch.cuda.synchronize()
                t0 = time.perf_counter()
                frame()
                pt.synchronize()
                times.append((time.perf_counter() - t0) * 1e3)
        ms = float(np.median(times))
        paths = st.width * st.height * st.spp
        best = min(v[0] for v in cfg["published"].values())
        rows.append({"resolution": f"{st.width}x{st.height}", "samples_per_pixel": st.spp, "ms": round(ms, 4),
                     "mpaths_per_s": round(paths / ms / 1e3, 2),
                     "published_ms": {k: v[0] for k, v in cfg["published"].items()},
                     "published_source": {k: v[1] for k, v in cfg["published"].items()},
                     "speedup_vs_best_published": round(best / ms, 2)})
        log(f"cornell_native {st.width}x{st.height}x{st.spp}: {ms:.3f} ms (best published {best} ms)")
    return {"scene": "cornell_native.json (48 triangles), max depth 6, NEE + MIS, seed 0xbadc0ffee, SAH leaf 4",
            "configs": rows}


def batched_comparison(log, count: int = 64, reps: int = 2) -> dict:
    """Many scenes in one batch against the same scenes rendered one tracer at a time.

    Workload: `count` seeded scenes (displaced icosphere + instanced spheres + 8 quad lights) with
    triangle counts log-uniform in [1k, 500k], square images from {128, 256, 512, 1024}, spp from
    {4, 16, 64}, depth 8, NEE + MIS. batched = one pt_batch_render of every scene's full frame (eager
    launches, then HIP-graph replay); per_scene = one tracer per scene, each rendering its frame with
    render_tiles, timed frame by frame and summed (scene upload and BVH builds are outside all timings).
    The batch uses the leaf width and launch shapes the autotuner chose for this workload (the
    committed tuning table); the per-scene tracers use the defaults, which the tuner found within 1 %
    of their best on the single-scene workload."""
    import torch

    from vkpathtracer_amd import autotune
    from vkpathtracer_amd import scene as S
    from vkpathtracer_amd.tracer import HipBatch, HipPathTracer

    jobs = S.batch_scenes(seed=7, count=count)
    tuned = autotune.tuned_params("batch64") if count == 64 else None
    settings = S.RenderSettings(width=1, height=1, spp=1, max_depth=8)
    paths = sum(w * h * spp for _, w, h, spp in jobs)
    tris = sum(sc.triangle_count for sc, _, _, _ in jobs)
    log(f"batch: {count} scenes, {tris} triangles, {paths} paths per frame")

    def timed(fn):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        fn()
        torch.cuda.synchronize()
        return (time.perf_counter() - t0) * 1e3

    out = {"workload": f"{count} seeded scenes, 1k-500k triangles (log-uniform), images 128^2..1024^2, spp 4/16/64, "
                       "depth 8", "scenes": count, "triangles": tris, "paths_per_frame": paths}
    leaf = tuned["leaf_size"] if tuned else 4
    out["tuning"] = {"source": "vkpathtracer_amd/tuning/mi355x.json (autotune.py)" if tuned else "defaults",
                     "params": tuned}
    with HipBatch(settings) as b:
        for sc, w, h, spp in jobs:
            b.add_scene(sc, w, h, spp, leaf_size=leaf)
        info = b.build()
        out["bvh_build_ms"] = round(info["build_ms"], 1)
        if tuned:
            autotune.apply_params(b, tuned)
        for use_graph, key in ((False, "batched"), (True, "batched_graph")):
            def frame():
                b.reset_accumulation()
                b.render(0, use_graph=use_graph)
                b.synchronize()
            frame()
            ms = min(timed(frame) for _ in range(reps))
            out[key + "_ms"] = round(ms, 2)
            out[key + "_mpaths_per_s"] = round(paths / ms / 1e3, 2)
    per_scene = 0.0
    for sc, w, h, spp in jobs:
        st = dataclasses.replace(settings, width=w, height=h, spp=spp)
        tiles = [(0, 0, w, h, s, 0, 0) for s in range(spp)]
        with HipPathTracer(0) as pt:
            assert pt.configure(st) and pt.load_scene_snapshot(sc) and pt.build_or_update_acceleration("sah", 4, 8)

            def frame():
                pt.reset_accumulation()
                assert pt.render_tiles(tiles)
                pt.synchronize()
            frame()
            per_scene += min(timed(frame) for _ in range(reps))
    out["per_scene_ms"] = round(per_scene, 2)
    out["per_scene_mpaths_per_s"] = round(paths / per_scene / 1e3, 2)
    out["speedup_batched_vs_per_scene"] = round(per_scene / min(out["batched_ms"], out["batched_graph_ms"]), 3)
    log(f"batched {out['batched_ms']} ms, graph {out['batched_graph_ms']} ms, per-scene {out['per_scene_ms']} ms")
    return out


def main():
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", choices=["procedural", "cornell"], default="procedural")
    ap.add_argument("--spp", type=int, default=None)
    ap.add_argument("--capacity", type=int, default=None, help="paths in flight per wavefront pass")
    ap.add_argument("--tuning", choices=["table", "default"], default="table",
                    help="leaf width and launch shapes: the autotuner's table for the workload, or the defaults")
    ap.add_argument("--cpu-seconds", type=float, default=8.0, help="target CPU time per baseline mode")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-parity", action="store_true", help="skip the oracle comparison of the benchmark frame")
    ap.add_argument("--no-profile", action="store_true", help="time without per-kernel events")
    ap.add_argument("--no-adjoint", action="store_true", help="skip the adjoint/forward time ratio")
    ap.add_argument("--no-batched", action="store_true", help="skip the many-scene batch comparison")
    ap.add_argument("--no-textured", action="store_true", help="skip the textured multi-view adjoint")
    ap.add_argument("--no-reference-configs", action="store_true",
                    help="skip the reference's published Cornell configurations")
    ap.add_argument("--copy-probe-mib", type=int, default=0,
                    help="run a plain device copy of this size first (calibrates HBM byte counters)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: this build renders on a single MI355X (no multi-GPU path); "
              "run with --gpus 1", file=sys.stderr)
        sys.exit(2)

    import torch

    from vkpathtracer_amd.tracer import HipPathTracer

    def log(msg):
        print(f"[bench] {msg}", file=sys.stderr, flush=True)

    if args.copy_probe_mib:
        n = args.copy_probe_mib << 18  # float32 elements
        a = torch.ones(n, device="cuda:0")
        b = torch.empty_like(a)
        for _ in range(3):
            b.copy_(a)
        torch.cuda.synchronize()
        log(f"copy probe: 3 x {args.copy_probe_mib} MiB device-to-device copies")
        del a, b

    from vkpathtracer_amd import autotune

    sc, st, desc = workload(args.workload, args.spp)
    log(f"scene {sc.name}: {sc.triangle_count} triangles")
    tuned = autotune.tuned_params(args.workload) if args.tuning == "table" else None
    leaf = tuned["leaf_size"] if tuned else 4
    pt = HipPathTracer(0)
    assert pt.configure(st)
    assert pt.load_scene_snapshot(sc)
    t0 = time.perf_counter()
    bst = pt.build_or_update_acceleration("sah", leaf, 8)
    build_wall_ms = (time.perf_counter() - t0) * 1e3
    assert bst, bst.message
    if tuned:
        autotune.apply_params(pt, tuned)
    if args.capacity:
        pt.set_wavefront_capacity(args.capacity)
    tiles = [(0, 0, st.width, st.height, s, 0, 0) for s in range(st.spp)]

    def step():
        pt.reset_accumulation()
        if not pt.render_tiles(tiles, 0):
            raise RuntimeError(pt.status().last_error)

    log(f"bvh built in {build_wall_ms:.0f} ms: {pt.build_info}")
    for _ in range(args.warmup):
        step()
    log("warmup done")
    pt.synchronize()
    torch.cuda.synchronize()
    pt.set_profiling(not args.no_profile)
    pt.reset_kernel_stats()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        step()
    pt.synchronize()
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    counters = pt.read_counters()  # counters of the last step (reset_accumulation clears them)
    ks = pt.kernel_stats()
    pt.set_profiling(False)

    paths = st.width * st.height * st.spp
    ms_per_step = dt * 1e3 / args.steps
    mpaths = paths * args.steps / dt / 1e6
    ext_rays = counters["rays"]
    shadow_rays = ks["shadow_rays"]  # device counters are per step (reset_accumulation clears them)
    mrays = (ext_rays + shadow_rays) * args.steps / dt / 1e6

    roofline = None
    per_kernel = {}
    if not args.no_profile:
        # device work counters cover the last step; event times and launch counts cover all K steps
        step_ks = dict(ks)
        for k in ("inner_node_fetches", "shadow_node_visits", "shadow_triangle_tests", "shadow_inner_node_fetches",
                  "shadow_leaf_visits", "shadow_rays"):
            step_ks[k] = ks[k] * args.steps
        tot_counters = {k: v * args.steps for k, v in counters.items()}
        kb = kernel_bytes(step_ks, tot_counters)
        for name in ("raygen", "extend", "shade", "shadow", "film"):
            k = ks[name]
            if k["launches"]:
                per_kernel[name] = {"launches": k["launches"], "avg_ms": round(k["total_ms"] / k["launches"], 4),
                                    "total_ms": round(k["total_ms"], 2),
                                    "gbps": round(kb[name] / (k["total_ms"] * 1e6), 1) if k["total_ms"] else None}
        dom = max(per_kernel, key=lambda n: per_kernel[n]["total_ms"])
        k = ks[dom]
        achieved = kb[dom] / (k["total_ms"] * 1e-3) / 1e9
        avg_ms = k["total_ms"] / k["launches"]
        sha = kernel_source_sha256()
        same_shape = args.workload == "procedural" and args.spp is None and not args.capacity
        profiled = same_shape and sha == PROFILED_KERNEL_SHA256
        traffic = MEASURED_TRAFFIC.get(dom) if profiled else None
        roofline = {"bound": "hbm", "kernel": dom, "achieved": round(achieved, 1), "peak": HBM_PEAK_GBPS,
                    "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBPS, 4),
                    "traffic": int(traffic) if traffic else None,
                    "traffic_source": "profiles/wavefront_procedural_1080p_64spp.md (rocprofv3 FETCH_SIZE, WRITE_SIZE)"
                    if traffic else None,
                    "measured_hbm_gbps": round(traffic / (avg_ms * 1e6), 1) if traffic else None,
                    "bytes_per_launch": int(kb[dom] / k["launches"]), "bytes_per_unit": BYTES,
                    "avg_launch_ms": round(avg_ms, 4),
                    "valu": {n: MEASURED_VALU[n] for n in ("extend", "shadow", "shade")} if profiled else None,
                    "valu_source": "profiles/wavefront_procedural_1080p_64spp.md (rocprofv3 SQ_* and TCC_* counters)"
                    if profiled else None,
                    "kernel_source_sha256": sha,
                    "profiled_kernel_sha256": PROFILED_KERNEL_SHA256,
                    "profile_note": None if profiled else
                    ("kernel sources differ from the profiled ones: traffic and valu omitted" if same_shape
                     else "not the profiled workload shape: traffic and valu omitted")}

    log(f"timed {args.steps} steps: {dt:.3f} s")
    bvh_info = {k: pt.build_info[k] for k in ("node_count", "leaf_count", "max_depth", "device_bytes")}
    parity = None if args.no_parity else parity_report(pt, sc, st, leaf, log)
    pt.close()
    adjoint = None if args.no_adjoint else adjoint_ratio(log)
    batched = None if args.no_batched else batched_comparison(log)
    textured = None if args.no_textured else textured_multiview(log)
    refcfg = None if args.no_reference_configs else reference_configs(log)
    cpu = None
    if not args.no_cpu_baseline:
        cpu = cpu_baseline(sc, st, args.cpu_seconds, log)

    out = {
        "metric": "Mpaths/s & Mrays/s fwd at 1 GPU vs CPU AVX2; adjoint/forward time ratio",
        "value": round(mpaths, 3),
        "unit": "Mpaths/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded procedural scene generated in-process)",
        "config": {"workload": desc, "width": st.width, "height": st.height, "spp": st.spp,
                   "max_depth": st.max_depth, "triangles": sc.triangle_count,
                   "bvh": f"binned SAH, leaf<={leaf}, 8 buckets",
                   "nee": True, "mis": True, "rr_start_depth": 3, "seed": hex(st.seed)},
        "mrays_per_s": round(mrays, 3),
        "ref_paths_per_sec": round(ext_rays * args.steps / dt, 1),
        "rays_per_step": {"extension": ext_rays, "shadow": shadow_rays},
        "work_per_ray": _work_per_ray(ks, counters),
        "bvh_build_ms": round(build_wall_ms, 1),
        "bvh": bvh_info,
        "tuning": {"source": "vkpathtracer_amd/tuning/mi355x.json (autotune.py)" if tuned else "defaults",
                   "params": tuned},
        "roofline": roofline,
        "kernels": per_kernel,
        "parity": parity,
        "cpu_baseline": cpu,
        "adjoint": adjoint,
        "batched": batched,
        "textured_multiview": textured,
        "reference_configs": refcfg,
    }
    if cpu:
        out["speedup_vs_cpu"] = round(mpaths / cpu["value"], 1)
    print(json.dumps(out))


if __name__ == "__main__":
    main()

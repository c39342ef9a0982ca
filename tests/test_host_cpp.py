"""The C++23 host side (vkpathtracer_amd/host): snapshot files written by Python read back by the C++
reader with every count and array intact, and — on the GPU — pt_hip_bench's film and counters equal
to the oracle's, with one launch per frame or one render_tile per sample."""
import json
import pathlib
import subprocess

import numpy as np
import pytest

ROOT = pathlib.Path(__file__).resolve().parent.parent
BENCH = ROOT / "vkpathtracer_amd" / "bin" / "pt_hip_bench"


def _scenes():
    from vkpathtracer_amd import scene as S

    return {"cornell": S.cornell_box(), "feature": S.feature_scene(),
            "procedural": S.procedural_scene(seed=3, ico_subdivisions=3, bisect=True, sphere_count=16, light_count=2),
            "textured": S.textured_sphere_scene(frequency=6, texture_size=16)}


@pytest.mark.parametrize("name", ["cornell", "feature", "procedural", "textured"])
def test_snapshot_file_round_trip(tmp_path, name):
    sc = _scenes()[name]
    path = tmp_path / f"{name}.ptsnap"
    sc.save_snapshot(path)
    out = subprocess.run([str(BENCH), "--check-snapshot", str(path)], check=True, capture_output=True, text=True)
    got = json.loads(out.stdout)
    texels, descs = sc.texture_arrays()
    env = sc.environment_map
    want = {"vertices": len(sc.vertices), "indices": len(sc.indices), "local_vertices": len(sc.local_vertices),
            "local_indices": len(sc.local_indices), "instances": len(sc.instances), "materials": len(sc.materials),
            "lights": len(sc.lights), "environment_map": [0, 0] if env is None else [env.shape[1], env.shape[0]],
            "texcoords": 0 if sc.texcoords is None else len(sc.texcoords), "texels": len(texels),
            "textures": len(descs), "triangles": sc.triangle_count}
    for k, v in want.items():
        assert got[k] == v, (k, got[k], v)
    vsum = float(np.sum(sc.vertices, dtype=np.float64))
    assert abs(got["vertex_sum"] - vsum) <= 1e-6 * max(1.0, abs(vsum))
    assert abs(got["fov_deg"] - sc.camera.fov_deg) < 1e-6


def test_bench_cli_rejects_bad_arguments():
    r = subprocess.run([str(BENCH), "--width"], capture_output=True, text=True)
    assert r.returncode == 2 and "usage" in r.stderr
    r = subprocess.run([str(BENCH), "--check-snapshot", "/nonexistent.ptsnap"], capture_output=True, text=True)
    assert r.returncode == 1


def _read_film(d):
    raw = (d / "film.bin").read_bytes()
    w, h = np.frombuffer(raw[:8], np.uint32)
    rgb = np.frombuffer(raw[8:8 + 12 * w * h], np.float32).reshape(h, w, 3)
    cnt = np.frombuffer(raw[8 + 12 * w * h:], np.uint32).reshape(h, w)
    return rgb, cnt


@pytest.mark.gpu
@pytest.mark.parametrize("per_sample", [False, True])
def test_pt_hip_bench_film_equals_oracle(tmp_path, per_sample):
    from oracle import oracle as O
    from vkpathtracer_amd import scene as S

    sc = S.feature_scene()
    path = tmp_path / "feature.ptsnap"
    sc.save_snapshot(path)
    st = S.RenderSettings(width=48, height=32, spp=3, max_depth=8)
    out = tmp_path / "out"
    cmd = [str(BENCH), "--scene", str(path), "--width", "48", "--height", "32", "--spp", "3", "--max-depth", "8",
           "--seed", str(st.seed), "--warmup", "0", "--out", str(out)] + (["--per-sample-launches"] if per_sample else [])
    r = subprocess.run(cmd, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    res = json.loads((out / "results.json").read_text())
    rgb, cnt = _read_film(out)
    rf, rc, rcnt = O.Oracle(sc).render(st)
    assert np.array_equal(cnt, rc)
    assert np.array_equal(rgb, rf), float(np.abs(rgb - rf).max())
    for k in ("samples", "rays", "shadow_tests", "bvh_node_visits", "triangle_tests"):
        assert res["counters"][k] == rcnt[k], k
    assert res["throughput"]["samples_per_sec"] > 0 and res["renderer_path"] == "hip-wavefront"
    pfm = (out / "beauty.pfm").read_bytes()
    assert pfm.startswith(b"PF\n48 32\n-1.0\n")


def _artifact_json(dir_):
    r = subprocess.run([str(BENCH), "--validate-artifacts", str(dir_)], capture_output=True, text=True, timeout=60)
    return r.returncode, json.loads(r.stdout)


def test_validate_artifacts_contract(tmp_path):
    """--validate-artifacts restates ValidateBenchmarkArtifactsOnDisk: every required file of the contract
    is reported missing in an empty directory; a complete set passes; an empty file, a results.json path
    that resolves to nothing and a trace without events are each reported."""
    rc, v = _artifact_json(tmp_path)
    assert rc == 1 and not v["ok"]
    assert set(v["missing_files"]) == {"results.json", "results.csv", "metadata.json", "scene_snapshot.json",
                                       "shader_manifest.json", "asset_manifest.json", "beauty.png", "beauty.exr",
                                       "logs.jsonl", "profiler_trace.json"}
    for name in v["missing_files"]:
        (tmp_path / name).write_text("x")
    (tmp_path / "results.json").write_text(json.dumps({k: f"{tmp_path}/{f}" for k, f in (
        ("beauty_png", "beauty.png"), ("beauty_exr", "beauty.exr"), ("logs_jsonl", "logs.jsonl"),
        ("profiler_trace_json", "profiler_trace.json"))}))
    (tmp_path / "metadata.json").write_text('{"benchmark_capabilities": {}}')
    (tmp_path / "profiler_trace.json").write_text('{"events": []}')
    rc, v = _artifact_json(tmp_path)
    assert rc == 0 and v["ok"] and not v["invalid_files"], v
    (tmp_path / "beauty.png").write_text("")
    (tmp_path / "profiler_trace.json").write_text("{}")
    res = json.loads((tmp_path / "results.json").read_text())
    res["beauty_exr"] = "/nonexistent/beauty.exr"
    (tmp_path / "results.json").write_text(json.dumps(res))
    rc, v = _artifact_json(tmp_path)
    assert rc == 1 and v["empty_files"] == ["beauty.png"]
    assert "profiler_trace.json: missing events array" in v["invalid_files"]
    assert "results.json: beauty_exr does not resolve to an artifact file" in v["invalid_files"]


def _read_png(path):
    """RGBA8 pixels of a PNG, decoded with zlib and the PNG specification's chunk rules (CRC checked)."""
    import struct
    import zlib

    raw = pathlib.Path(path).read_bytes()
    assert raw[:8] == b"\x89PNG\r\n\x1a\n"
    pos, idat, w, h = 8, b"", 0, 0
    while pos < len(raw):
        n, = struct.unpack(">I", raw[pos:pos + 4])
        typ, data = raw[pos + 4:pos + 8], raw[pos + 8:pos + 8 + n]
        assert struct.unpack(">I", raw[pos + 8 + n:pos + 12 + n])[0] == zlib.crc32(typ + data)
        if typ == b"IHDR":
            w, h, depth, ctype = struct.unpack(">IIBB", data[:10])
            assert depth == 8 and ctype == 6
        elif typ == b"IDAT":
            idat += data
        pos += 12 + n
    rows = np.frombuffer(zlib.decompress(idat), np.uint8).reshape(h, 4 * w + 1)
    assert (rows[:, 0] == 0).all()
    return rows[:, 1:].reshape(h, w, 4)


def _read_exr(path):
    """RGB float pixels of an uncompressed scanline OpenEXR file, read by the OpenEXR layout rules."""
    import struct

    raw = pathlib.Path(path).read_bytes()
    assert raw[:4] == b"\x76\x2f\x31\x01" and raw[4] == 2
    pos, attrs = 8, {}
    while raw[pos] != 0:
        name_end = raw.index(b"\0", pos)
        type_end = raw.index(b"\0", name_end + 1)
        size, = struct.unpack("<i", raw[type_end + 1:type_end + 5])
        attrs[raw[pos:name_end].decode()] = (raw[name_end + 1:type_end].decode(), raw[type_end + 5:type_end + 5 + size])
        pos = type_end + 5 + size
    pos += 1
    assert attrs["compression"][1] == b"\0"
    x0, y0, x1, y1 = struct.unpack("<4i", attrs["dataWindow"][1])
    w, h = x1 - x0 + 1, y1 - y0 + 1
    chl, names, p = attrs["channels"][1], [], 0
    while chl[p] != 0:
        e = chl.index(b"\0", p)
        names.append(chl[p:e].decode())
        assert struct.unpack("<i", chl[e + 1:e + 5])[0] == 2  # FLOAT
        p = e + 17
    offsets = struct.unpack(f"<{h}Q", raw[pos:pos + 8 * h])
    img = np.zeros((h, w, len(names)), np.float32)
    for off in offsets:
        y, size = struct.unpack("<ii", raw[off:off + 8])
        data = np.frombuffer(raw[off + 8:off + 8 + size], "<f4").reshape(len(names), w)
        img[y - y0] = data.T
    return {n: img[..., i] for i, n in enumerate(names)}


@pytest.mark.gpu
def test_pt_hip_bench_artifact_set(tmp_path):
    """A run writes ptbench's whole artifact set and passes its own contract check; beauty.png holds the
    oracle's ApplyFilmResolve of the written film, beauty.exr its resolve_hdr, image_hash is FNV-1a of
    the PNG bytes, and the trace carries the per-kernel GPU events."""
    from oracle import oracle as O
    from vkpathtracer_amd import scene as S

    sc = S.cornell_box()
    path = tmp_path / "cornell.ptsnap"
    sc.save_snapshot(path)
    out = tmp_path / "out"
    cmd = [str(BENCH), "--scene", str(path), "--width", "40", "--height", "30", "--spp", "2", "--warmup", "0",
           "--tone-map", "aces_approx", "--exposure", "1.5", "--white-balance", "4000", "--out", str(out)]
    r = subprocess.run(cmd, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    rc, v = _artifact_json(out)
    assert rc == 0 and v["ok"], v
    rgb, cnt = _read_film(out)
    png = _read_png(out / "beauty.png")
    want = O.resolve_ldr(rgb, cnt, exposure=1.5, white_balance_kelvin=4000.0, tone_map="aces_approx")
    assert np.array_equal(png, want)
    exr = _read_exr(out / "beauty.exr")
    hdr = O.resolve_hdr(rgb, cnt)
    for i, ch in enumerate("RGB"):
        assert np.array_equal(exr[ch], hdr[..., i]), ch
    res = json.loads((out / "results.json").read_text())
    data = (out / "beauty.png").read_bytes()
    h = 1469598103934665603
    for b in data:
        h = ((h ^ b) * 1099511628211) & 0xFFFFFFFFFFFFFFFF
    assert res["image_hash"] == f"{h:016x}"
    assert res["resolve"]["tone_map"] == "aces_approx"
    rows = (out / "results.csv").read_text().splitlines()
    assert rows[0].startswith("run_id,scene,backend") and rows[1].endswith(res["image_hash"])
    ev = json.loads((out / "profiler_trace.json").read_text())["events"]
    gpu = {e["name"]: e for e in ev if e["kind"] == "gpu"}
    assert set(gpu) == {"raygen", "extend", "shade", "shadow", "film"} and gpu["extend"]["launches"] > 0


@pytest.mark.gpu
@pytest.mark.parametrize("builder", ["midpoint", "sah"])
def test_hip_ray_accelerator_equals_oracle(tmp_path, builder):
    """The C++ HipRayAccelerator (IRayAccelerator, PathTracer.h:1001-1036) driven by pt_hip_bench
    --ray-queries: build(snapshot, deterministic) reports RayAcceleratorBuildInfo, one batched intersect
    fills every RayQueryHit as the oracle's CpuBvhAccelerator restatement does, and single-ray
    intersect(Ray, RayQueryHit&, RayQueryStats*) calls give the same hits and add the same stats."""
    from oracle import oracle as O
    from vkpathtracer_amd import abi, scene as S

    sc = S.procedural_scene(seed=5, ico_subdivisions=3, bisect=True, sphere_count=24, light_count=2)
    snap = tmp_path / "scene.ptsnap"
    sc.save_snapshot(snap)
    rays = S.random_rays(sc, 20_000, seed=4)
    rays[::5, 3:] *= -1.0
    rpath = tmp_path / "rays.f32"
    rpath.write_bytes(np.ascontiguousarray(rays, np.float32).tobytes())
    out = tmp_path / "out"
    r = subprocess.run([str(BENCH), "--scene", str(snap), "--ray-queries", str(rpath), "--single", "2000",
                        "--builder", builder, "--leaf", "4", "--out", str(out)], capture_output=True, text=True,
                       timeout=120)
    assert r.returncode == 0, r.stderr
    res = json.loads((out / "ray_queries.json").read_text())
    hits = np.frombuffer((out / "hits.bin").read_bytes(), abi.RAY_QUERY_HIT_DTYPE)
    orc = O.Oracle(sc, builder=builder, leaf_size=4)
    t, p, c = orc.trace_rays(rays)
    geo = orc.export()
    hit = p >= 0
    assert np.array_equal(hits["hit"].astype(bool), hit)
    assert np.array_equal(hits["t"][hit], t[hit])
    assert np.array_equal(hits["primitive_index"][hit], p[hit].astype(np.uint32))
    assert np.array_equal(hits["position"][hit], rays[hit, :3] + rays[hit, 3:] * t[hit, None])
    assert np.array_equal(hits["normal"][hit], geo["geo"][p[hit], 9:12].astype(np.float32))
    assert np.array_equal(hits["material_index"][hit], geo["material"][p[hit]])
    assert res["build_info"] == {"built": True, "primitive_count": orc.prim_count, "node_count": orc.node_count,
                                 "leaf_count": int((geo["meta"][:, 3] > 0).sum()), "deterministic": True}
    for k in ("triangle_tests", "triangle_hits", "bvh_node_visits", "bvh_leaf_visits"):
        assert res["batch_stats"][k] == c[k], k
    assert res["single"] == 2000 and res["single_mismatched"] == 0
    assert res["single_hits"] == int(hit[:2000].sum())
    _, _, c1 = orc.trace_rays(rays[:2000])
    for k in ("triangle_tests", "triangle_hits", "bvh_node_visits", "bvh_leaf_visits"):
        assert res["single_stats"][k] == c1[k], k

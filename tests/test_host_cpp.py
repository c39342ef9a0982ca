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

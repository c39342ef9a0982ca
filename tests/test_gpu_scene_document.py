"""The reference's scene documents (tests/golden/scenes) converted by vkpathtracer_amd.scene_document
and rendered through libvkpt_hip.so: film sums, sample counts and traversal counters identical to
the oracle's render of the same converted scene, one tracer per document and all documents in one
batch (eager and replayed from a HIP graph), plus a converted snapshot file rendered by the C++ host."""
import pathlib
import subprocess

import numpy as np
import pytest

from conftest import relrmse

pytestmark = pytest.mark.gpu

GOLDEN = pathlib.Path(__file__).resolve().parent / "golden" / "scenes"
NAMES = sorted(p.name for p in GOLDEN.glob("*.json"))
ROOT = pathlib.Path(__file__).resolve().parent.parent
REL_TOL = 1e-4
ALWAYS = ("samples", "rays", "shadow_tests", "bvh_node_visits", "triangle_tests")
OPAQUE = ALWAYS + ("bvh_leaf_visits", "triangle_hits")


def _scene(name):
    from vkpathtracer_amd import scene_document as SD

    return SD.load_scene_document(GOLDEN / name)


def _transmissive(sc) -> bool:
    m = sc.materials
    return bool(((m["material_model"] == 5) | (m["transmission"] > 0.05)).any())


@pytest.mark.parametrize("name", NAMES)
def test_document_render_bitexact(name, tracer_factory):
    from oracle import oracle as O
    from vkpathtracer_amd import scene as S

    sc = _scene(name)
    st = S.RenderSettings(width=48, height=36, spp=4, max_depth=6)
    print(f"document {name}: {sc.triangle_count} triangles", flush=True)
    pt = tracer_factory(sc, st)
    assert pt.render_frame_samples(0, st.spp), pt.status().last_error
    film, cnt = pt.film()
    counters = pt.read_counters()
    rf, rc, rcnt = O.Oracle(sc).render(st)
    assert np.array_equal(cnt, rc)
    assert relrmse(film, rf) <= REL_TOL
    assert np.array_equal(film, rf), f"max |diff| {np.abs(film - rf).max()}"
    for k in (ALWAYS if _transmissive(sc) else OPAQUE):
        assert counters[k] == rcnt[k], (k, counters[k], rcnt[k])


@pytest.mark.parametrize("use_graph", [False, True])
def test_documents_in_one_batch(use_graph):
    """Every fixture at its own image size and sample count in one scene pool: each film equals the
    oracle's render of that document."""
    from oracle import oracle as O
    from vkpathtracer_amd import scene as S
    from vkpathtracer_amd.tracer import HipBatch

    settings = S.RenderSettings(width=1, height=1, spp=1, max_depth=6)
    jobs = [(name, 24 + 8 * (i % 3), 20 + 4 * (i % 2), 1 + i % 3) for i, name in enumerate(NAMES)]
    scenes = [_scene(n) for n, *_ in jobs]
    with HipBatch(settings) as b:
        for sc, (name, w, h, spp) in zip(scenes, jobs):
            b.add_scene(sc, w, h, spp)
        b.build()
        b.render(0, use_graph=use_graph)
        for i, (sc, (name, w, h, spp)) in enumerate(zip(scenes, jobs)):
            film, cnt = b.film(i)
            st = S.RenderSettings(width=w, height=h, spp=spp, max_depth=6)
            rf, rc, _ = O.Oracle(sc).render(st)
            assert np.array_equal(cnt, rc), name
            assert np.array_equal(film, rf), (name, float(np.abs(film - rf).max()))


@pytest.mark.parametrize("name", ["cornell_native", "sdf_complexity"])
def test_host_renders_converted_snapshot(tmp_path, name):
    """python -m vkpathtracer_amd.scene_document writes a snapshot that pt_hip_bench --scene renders:
    its film equals the oracle's render of the converted document (SDF primitives travel in a
    version-2 snapshot)."""
    import json

    from oracle import oracle as O
    from vkpathtracer_amd import scene as S

    exe = ROOT / "vkpathtracer_amd" / "bin" / "pt_hip_bench"
    assert exe.exists(), "pt_hip_bench is not built"
    snap = tmp_path / f"{name}.ptsnap"
    out = tmp_path / "out"
    subprocess.run(["python", "-m", "vkpathtracer_amd.scene_document", str(GOLDEN / f"{name}.json"),
                    "--out", str(snap)], check=True, cwd=ROOT, timeout=120)
    st = S.RenderSettings(width=40, height=30, spp=3, max_depth=6)
    r = subprocess.run([str(exe), "--scene", str(snap), "--width", "40", "--height", "30", "--spp", "3",
                        "--max-depth", "6", "--seed", str(st.seed), "--warmup", "0", "--out", str(out)],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    raw = (out / "film.bin").read_bytes()
    w, h = np.frombuffer(raw[:8], np.uint32)
    rgb = np.frombuffer(raw[8:8 + 12 * w * h], np.float32).reshape(h, w, 3)
    cnt = np.frombuffer(raw[8 + 12 * w * h:], np.uint32).reshape(h, w)
    rf, rc, rcnt = O.Oracle(_scene(f"{name}.json")).render(st)
    assert np.array_equal(cnt, rc)
    assert np.array_equal(rgb, rf), float(np.abs(rgb - rf).max())
    res = json.loads((out / "results.json").read_text())
    for k in (ALWAYS if _transmissive(_scene(f"{name}.json")) else OPAQUE):
        assert res["counters"][k] == rcnt[k], k

"""bench.py's host-side contract without a GPU: --gpus above 1 exits with a message, the workloads and
the byte model are defined, and the CPU baseline runs on a tiny workload."""
import subprocess
import sys
import pathlib

ROOT = pathlib.Path(__file__).resolve().parent.parent


def test_more_than_one_gpu_is_refused():
    r = subprocess.run([sys.executable, str(ROOT / "bench.py"), "--gpus", "2"], capture_output=True, text=True)
    assert r.returncode == 2 and "one GPU" in r.stderr


def test_workloads_and_byte_model():
    import bench

    sc, st, desc = bench.workload("cornell", None)
    assert (st.width, st.height, st.spp, st.max_depth) == (256, 256, 16, 8) and "cornell" in desc
    ks = {"shadow_rays": 10, "shadow_node_visits": 100, "shadow_triangle_tests": 20,
          "raygen": {"items": 7}, "film": {"items": 7}}
    counters = {"rays": 5, "bvh_node_visits": 300, "triangle_tests": 50}
    kb = bench.kernel_bytes(ks, counters)
    B = bench.BYTES
    assert kb["extend"] == 5 * B["ext_ray"] + 200 * B["node_visit"] + 30 * B["triangle"]
    assert kb["shadow"] == 10 * B["shadow_ray"] + 100 * B["node_visit"] + 20 * B["triangle"]


def test_cpu_baseline_modes_run():
    import bench
    from vkpathtracer_amd import scene as S

    st = S.RenderSettings(width=32, height=32, spp=4, max_depth=4)
    out = bench.cpu_baseline(S.cornell_box(), st, 0.05, lambda m: None)
    assert out["kind"] == "port" and out["value"] > 0 and {"scalar", "multicore"} <= set(out["modes"])
    assert out["modes"]["scalar"]["threads"] == 1


def test_profiled_constants_are_stamped_with_the_kernel_sources():
    """The traffic and VALU figures pasted from the committed rocprofv3 profile are reported only while
    the kernel sources hash to the profiled ones."""
    import bench

    sha = bench.kernel_source_sha256()
    assert len(sha) == 64 and sha == bench.kernel_source_sha256()
    assert bench.PROFILED_KERNEL_SHA256 is None or len(bench.PROFILED_KERNEL_SHA256) == 64

import pathlib
import sys

import pytest

ROOT = pathlib.Path(__file__).resolve().parent.parent
if str(ROOT) not in sys.path:
    sys.path.insert(0, str(ROOT))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through libvkpt_hip.so on cuda:0)")
    config.addinivalue_line("markers", "slow: CPU test taking more than a few seconds")


def relrmse(a, b):
    import numpy as np

    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    den = np.sqrt(np.mean(b * b))
    return float(np.sqrt(np.mean((a - b) ** 2)) / (den if den > 0 else 1.0))


@pytest.fixture(scope="session")
def cornell():
    from vkpathtracer_amd import scene as S

    return S.cornell_box()


@pytest.fixture(scope="session")
def small_procedural():
    """A 20k-triangle version of the procedural workload (4 subdivisions, 64 spheres)."""
    from vkpathtracer_amd import scene as S

    return S.procedural_scene(seed=3, ico_subdivisions=4, bisect=True, sphere_count=64, light_count=4)


@pytest.fixture(scope="session")
def tracer_factory():
    from vkpathtracer_amd.tracer import HipPathTracer

    made = []

    def make(scene, settings, builder="sah", leaf_size=4, bucket_count=8):
        pt = HipPathTracer(0)
        made.append(pt)
        assert pt.configure(settings)
        assert pt.load_scene_snapshot(scene)
        st = pt.build_or_update_acceleration(builder, leaf_size, bucket_count)
        assert st, st.message
        return pt

    yield make
    for pt in made:
        pt.close()

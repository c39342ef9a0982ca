"""The C-ABI library loads without a GPU, exports every symbol include/vkpt_hip.h declares, and its
record sizes agree with the Python mirrors and with the oracle's records."""
import ctypes as C
import pathlib
import re

import numpy as np

from oracle import oracle as O
from vkpathtracer_amd import abi, scene as S

HEADER = pathlib.Path(__file__).resolve().parent.parent / "include" / "vkpt_hip.h"


def declared_functions():
    text = HEADER.read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(pt_(?:hip|bvh|batch)_\w+)\s*\(", text)))


def test_every_declared_symbol_is_exported_and_bound():
    lib = abi.load_library()
    names = declared_functions()
    assert len(names) >= 20
    for n in names:
        assert hasattr(lib, n), n
        assert n in abi.SIGNATURES, f"{n} has no ctypes binding"
    assert set(abi.SIGNATURES) == set(names)


def test_record_sizes_match_everywhere():
    lib = abi.load_library()
    L = O.lib("f32")
    py = [S.MATERIAL_DTYPE.itemsize, S.LIGHT_DTYPE.itemsize, S.INSTANCE_DTYPE.itemsize, C.sizeof(abi.PtCamera),
          C.sizeof(abi.PtSceneView), C.sizeof(abi.PtRenderSettings), C.sizeof(abi.PtSampleCounters)]
    for i, size in enumerate(py):
        assert lib.pt_hip_sizeof(i) == size, i
        assert L.pto_sizeof(i) == size, i
    assert lib.pt_hip_sizeof(7) == C.sizeof(abi.PtRenderTile) == 28
    assert lib.pt_hip_sizeof(8) == C.sizeof(abi.PtTuning) == 16
    assert lib.pt_hip_sizeof(10) == C.sizeof(abi.PtFilmResolve) == 24
    assert lib.pt_hip_sizeof(11) == C.sizeof(abi.PtInstanceTransformUpdate) == 64
    assert C.sizeof(abi.PtInstanceTransformResult) == 24
    assert lib.pt_hip_sizeof(13) == abi.RAY_QUERY_HIT_DTYPE.itemsize == 40
    assert lib.pt_hip_sizeof(18) == C.sizeof(abi.PtBuildInfo) == 48
    assert lib.pt_hip_sizeof(14) == C.sizeof(abi.PtSceneDeltaUpdate) == 56
    assert lib.pt_hip_sizeof(15) == abi.MATERIAL_UPDATE_DTYPE.itemsize == 88
    assert lib.pt_hip_sizeof(16) == abi.LIGHT_UPDATE_DTYPE.itemsize == 68
    assert lib.pt_hip_sizeof(17) == abi.INSTANCE_MATERIAL_UPDATE_DTYPE.itemsize == 16
    assert C.sizeof(O.PtoScene) == C.sizeof(abi.PtSceneView)
    assert C.sizeof(O.PtoSettings) == C.sizeof(abi.PtRenderSettings) == 56


def test_version_string():
    assert b"gfx950" in abi.load_library().pt_hip_version()


def test_create_without_gpu_fails_cleanly():
    """With no visible device, pt_hip_create reports unsupported instead of crashing."""
    import torch

    if torch.cuda.is_available():
        return
    lib = abi.load_library()
    h = C.c_void_p()
    rc = lib.pt_hip_create(0, C.byref(h))
    assert rc in (1, 2) and not h.value

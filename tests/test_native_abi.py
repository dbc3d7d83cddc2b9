"""The C-ABI libraries load without a GPU and export every function that
include/tracer_hip.h and include/tracer/scenes.h declare."""
import ctypes as C
import re
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parents[1]


def _declared(header):
    text = (ROOT / "include" / header).read_text()
    decl = r"^\s*(?:tr_status|void|int32_t|const char\*|const tr_scene_desc\*|trs_scene\*)\s*(trs?_[a-z0-9_]+)\s*\("
    return sorted(set(re.findall(decl, text, flags=re.M)))


def test_hip_library_exports_header():
    from tracer_amd import native
    lib = native.hip_lib()
    names = [n for n in _declared("tracer_hip.h") if n.startswith("tr_")]
    assert len(names) >= 18
    for n in names:
        assert hasattr(lib, n), n


def test_scene_library_exports_header():
    from tracer_amd import native
    lib = native.scenes_lib()
    for n in _declared("tracer/scenes.h"):
        assert hasattr(lib, n), n


def test_abi_struct_sizes():
    from tracer_amd import abi
    assert C.sizeof(abi.TrMaterial) == 132
    assert C.sizeof(abi.TrMedium) == 152
    assert C.sizeof(abi.TrCamera) == 52
    assert C.sizeof(abi.TrRenderStats) == 160
    assert C.sizeof(abi.TrRenderParams) == 96
    assert C.sizeof(abi.TrBdptParams) == 64
    assert C.sizeof(abi.TrGuideConfig) == 48


def test_version_string():
    from tracer_amd import native
    assert "gfx950" in native.version()

"""Build the native libraries in-tree.

    python -m tracer_amd.build            # libtracer_hip.so (+ facade, oracle)

libtracer_hip.so: HIP kernels for gfx950 + the C ABI (include/tracer_hip.h).
libtracer.so: the C++ facade with the reference's factory API and the
synthetic scene generators (include/tracer/).
"""
from __future__ import annotations

import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor
from pathlib import Path

ROOT = Path(__file__).resolve().parent.parent
CSRC = ROOT / "tracer_amd" / "csrc"
FACADE = ROOT / "tracer_amd" / "facade"
LIB = ROOT / "tracer_amd" / "lib"
OBJ = ROOT / "tracer_amd" / "_build"
ARCH = os.environ.get("TRACER_OFFLOAD_ARCH", "gfx950")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

COMMON = ["-O3", "-std=c++17", "-fPIC", "-Wall", "-Wno-unused-function", "-Wno-unused-variable",
          f"-I{ROOT / 'include'}", f"-I{CSRC}"]
# Packed-FP32 instructions are disabled for all device code (the host pass
# ignores the feature with a warning).
HIP_FLAGS = [f"--offload-arch={ARCH}", "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
             "-munsafe-fp-atomics"]

HIP_SOURCES = {
    # the traversal TU must round exactly like the CPU restatement
    "k_trace.hip": ["-ffp-contract=off"],
    "k_render.hip": [],
    "k_guide.hip": [],
    "k_bdpt.hip": [],
    "capi.cpp": ["-x", "hip"],
    "guide_capi.cpp": ["-x", "hip"],
}
CXX_SOURCES = {"bvh_build.cpp": [], "envmap.cpp": ["-ffp-contract=off"]}
FACADE_SOURCES = ["facade.cpp", "scenes.cpp", "obj_load.cpp", "image_load.cpp", "image_write.cpp"]


def _run(cmd: list[str]) -> None:
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(" ".join(cmd) + "\n" + r.stdout + r.stderr)
        raise RuntimeError(f"compile failed: {cmd[-1] if cmd else ''}")
    noise = "is not a recognized feature for this target"
    msg = "\n".join(l for l in (r.stdout + r.stderr).splitlines() if noise not in l)
    if msg.strip():
        sys.stderr.write(msg + "\n")


def build_hip(jobs: int = 8) -> Path:
    OBJ.mkdir(parents=True, exist_ok=True)
    LIB.mkdir(parents=True, exist_ok=True)
    cmds, objs = [], []
    for src, extra in HIP_SOURCES.items():
        obj = OBJ / (src + ".o")
        objs.append(obj)
        cmds.append([HIPCC, *COMMON, *HIP_FLAGS, *extra, "-c", str(CSRC / src), "-o", str(obj)])
    for src, extra in CXX_SOURCES.items():
        obj = OBJ / (src + ".o")
        objs.append(obj)
        cmds.append(["g++", *COMMON, "-march=x86-64-v3", "-pthread", *extra, "-c", str(CSRC / src), "-o", str(obj)])
    with ThreadPoolExecutor(max_workers=jobs) as ex:
        list(ex.map(_run, cmds))
    out = LIB / "libtracer_hip.so"
    _run([HIPCC, *HIP_FLAGS, "-shared", "-fPIC", "-o", str(out), *map(str, objs), "-lpthread"])
    return out


def build_facade() -> Path | None:
    srcs = [FACADE / s for s in FACADE_SOURCES if (FACADE / s).exists()]
    if not srcs:
        return None
    out = LIB / "libtracer.so"
    _run(["g++", *COMMON, "-march=x86-64-v3", "-pthread", "-shared", *map(str, srcs), "-o", str(out),
          f"-L{LIB}", "-ltracer_hip", "-lz", "-Wl,-rpath,$ORIGIN"])
    return out


def build_harness() -> Path | None:
    """The compiled drop-in harness of the facade tests (tests/cpp/facade_harness.cpp)."""
    src = ROOT / "tests" / "cpp" / "facade_harness.cpp"
    if not src.exists():
        return None
    out_dir = ROOT / "tests" / "_build"
    out_dir.mkdir(parents=True, exist_ok=True)
    out = out_dir / "facade_harness"
    _run(["g++", "-O2", "-std=c++17", f"-I{ROOT / 'include'}", str(src), "-o", str(out), f"-L{LIB}", "-ltracer",
          "-ltracer_hip", "-Wl,-rpath,$ORIGIN/../../tracer_amd/lib"])
    return out


def build_oracle() -> None:
    _run(["make", "-s", "-C", str(ROOT / "oracle"), "-j4"])
    if Path("/root/reference/dep/pcg_random.hpp").exists():
        _run(["make", "-s", "-C", str(ROOT / "oracle"), "ref"])


def build_all() -> None:
    build_hip()
    build_facade()
    build_harness()
    build_oracle()


if __name__ == "__main__":
    build_all()
    print("built:", *sorted(p.name for p in LIB.glob("*.so")))

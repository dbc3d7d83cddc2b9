"""AddressSanitizer (with leak detection) and UndefinedBehaviorSanitizer over
the host code: the oracle (oracle/refcpu/ref_pt.cpp) and the library's host
side (facade scene builders, OBJ / MTL ingest, PNG and Radiance HDR readers
and writers, post-processing, the 4-wide BVH builder, the environment-map
distribution).  tests/cpp/sanitize_harness.cpp drives them over the five
benchmark configurations at reduced detail, an OBJ room with textures, glass
and fog, a density grid, degenerate geometry (zero-area and duplicate
triangles, an emitter of zero area, zero-length and reversed normals), every
material kind at extreme parameters and grazing directions, films of every
filter radius with NaN samples, and truncated or corrupted image and OBJ
files.  The device code is not instrumented (no GPU sanitizer here); the
harness links the uninstrumented libtracer_hip.so only for the symbols the
facade names and never calls into it."""
import os
import subprocess
from pathlib import Path

ROOT = Path(__file__).resolve().parents[1]
EXE = ROOT / "tests" / "_build" / "sanitize_harness"
SOURCES = [ROOT / "tests" / "cpp" / "sanitize_harness.cpp", ROOT / "oracle" / "refcpu" / "ref_pt.cpp",
           *(ROOT / "tracer_amd" / "facade" / f for f in ("facade.cpp", "scenes.cpp", "obj_load.cpp", "image_load.cpp",
                                                          "image_write.cpp")),
           ROOT / "tracer_amd" / "csrc" / "bvh_build.cpp", ROOT / "tracer_amd" / "csrc" / "envmap.cpp"]
HEADERS = [*(ROOT / "include").rglob("*.h*"), *(ROOT / "oracle" / "refcpu").glob("*.hpp"),
           *(ROOT / "tracer_amd" / "facade").glob("*.h"), ROOT / "tracer_amd" / "csrc" / "bvh_build.h",
           ROOT / "tracer_amd" / "csrc" / "envmap.h"]


def _build():
    newest = max(p.stat().st_mtime for p in SOURCES + HEADERS)
    if EXE.exists() and EXE.stat().st_mtime > newest:
        return
    EXE.parent.mkdir(parents=True, exist_ok=True)
    lib = ROOT / "tracer_amd" / "lib"
    cmd = ["g++", "-O1", "-g", "-std=c++17", "-fno-omit-frame-pointer", "-fsanitize=address,undefined",
           "-fno-sanitize-recover=undefined", "-pthread", f"-I{ROOT / 'include'}", f"-I{ROOT / 'tracer_amd' / 'csrc'}",
           *map(str, SOURCES), "-o", str(EXE), f"-L{lib}", "-ltracer_hip", "-lz", f"-Wl,-rpath,{lib}"]
    r = subprocess.run(cmd, capture_output=True, text=True, timeout=900)
    assert r.returncode == 0, r.stderr[-4000:]


def test_host_code_under_address_and_undefined_behaviour_sanitizers(tmp_path):
    _build()
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=0:halt_on_error=1",
               UBSAN_OPTIONS="print_stacktrace=1:halt_on_error=1")
    r = subprocess.run([str(EXE), str(tmp_path)], capture_output=True, text=True, timeout=600, env=env)
    print(r.stdout[-500:])
    assert r.returncode == 0, (r.stdout + r.stderr)[-6000:]
    assert "runtime error" not in r.stderr and "AddressSanitizer" not in r.stderr
    assert r.stdout.startswith("ok:")

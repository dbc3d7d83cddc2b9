"""Regenerate the committed fixtures under tests/golden/.

pcg32_kat.jsonl: first outputs of pcg32 for seeds 0..4 (default stream) and
three explicit (state, stream) pairs, printed by oracle/_ref/pcg_kat, which
oracle/Makefile (target `ref`) compiles against the reference's vendored
third-party header dep/pcg_random.hpp (the generator instantiated at
src/core/sampler.hpp:32).  Needs the reference tree, so it runs in the build
container only.
"""
import subprocess
from pathlib import Path

ROOT = Path(__file__).resolve().parents[2]

if __name__ == "__main__":
    subprocess.run(["make", "-s", "-C", str(ROOT / "oracle"), "ref"], check=True)
    out = subprocess.run([str(ROOT / "oracle" / "_ref" / "pcg_kat")], check=True, capture_output=True, text=True).stdout
    (Path(__file__).parent / "pcg32_kat.jsonl").write_text(out)
    print("wrote pcg32_kat.jsonl")

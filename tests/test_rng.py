"""PCG32 pins.  The reference draws from pcg32 (dep/pcg_random.hpp:1862,
used at src/core/sampler.hpp:32); tests/golden/pcg32_kat.jsonl holds its
outputs, produced from the vendored header by tests/golden/make_fixtures.py.
The oracle's restatement must reproduce them, and the product's per-sample
stream (pcg_for_sample in tracer_amd/csrc/tr_device.h) must equal the oracle
PCG32 seeded the same way."""
import ctypes as C
import json
from pathlib import Path

import numpy as np
import pytest

GOLD = Path(__file__).parent / "golden" / "pcg32_kat.jsonl"


def _oracle_pcg(state, stream, default, n):
    from oracle import refcpu
    L = refcpu.lib()
    L.ref_pcg32.argtypes = [C.c_uint64, C.c_uint64, C.c_int, C.c_int64, C.POINTER(C.c_uint32)]
    L.ref_pcg32.restype = None
    out = np.zeros(n, np.uint32)
    L.ref_pcg32(state, stream, default, n, out.ctypes.data_as(C.POINTER(C.c_uint32)))
    return out


def test_oracle_pcg32_matches_reference_kat():
    rows = [json.loads(l) for l in GOLD.read_text().splitlines()]
    assert len(rows) == 8
    for r in rows:
        default = r["stream"] is None
        got = _oracle_pcg(r["seed"], 0 if default else r["stream"], int(default), len(r["out"]))
        assert got.tolist() == r["out"], r


def test_kat_harness_reproduces_fixture():
    exe = Path(__file__).resolve().parents[1] / "oracle" / "_ref" / "pcg_kat"
    if not exe.exists():
        pytest.skip("oracle/_ref/pcg_kat is only built where the reference tree exists")
    import subprocess
    assert subprocess.run([str(exe)], capture_output=True, text=True, check=True).stdout == GOLD.read_text()


def _mix64(z):
    M = (1 << 64) - 1
    z = (z + 0x9E3779B97F4A7C15) & M
    z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & M
    z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & M
    return z ^ (z >> 31)


@pytest.mark.parametrize("seed,pixel,sample", [(0, 0, 0), (1, 12345, 7), (2**40 + 3, 1048575, 63)])
def test_product_stream_is_pcg32(seed, pixel, sample):
    from tracer_amd import native
    L = native.hip_lib()
    L.tr_rng_sequence.argtypes = [C.c_uint64, C.c_uint32, C.c_uint32, C.c_int64, C.POINTER(C.c_uint32)]
    L.tr_rng_sequence.restype = None
    out = np.zeros(16, np.uint32)
    L.tr_rng_sequence(seed, pixel, sample, 16, out.ctypes.data_as(C.POINTER(C.c_uint32)))
    M = (1 << 64) - 1
    state = _mix64(seed ^ _mix64((pixel * 0x9E3779B97F4A7C15 + sample) & M))
    ref = _oracle_pcg(state, pixel, 0, 16)
    assert out.tolist() == ref.tolist()

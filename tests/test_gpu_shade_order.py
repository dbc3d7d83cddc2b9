"""The plain shade kernels order each window of their queue by the shading
class of the hit (rays that left the scene, Lambert, Disney, other) before
shading it.  A path's result does not depend on which thread shades it, so the
film must come out bit for bit the same with the ordering forced on every
launch, at its default threshold (windows of at least two rounds) and switched
off.  The threshold is read once per process (TRACER_HIP_SHADE_SORT), so each
setting renders in a child process of its own, one after the other.

The cases run queues large enough for the default threshold to sort (512 x 512
at 16 spp: 4M paths in flight) through the plain kernel, an image environment
(the extended kernel), the light cache and three light samples per vertex."""
import hashlib
import os
import subprocess
import sys
from pathlib import Path

import pytest

pytestmark = pytest.mark.gpu
ROOT = Path(__file__).resolve().parent.parent

CHILD = r"""
import hashlib, sys
import numpy as np
from tracer_amd import native
ctx = native.Context(0)
out = []
def run(name, scene, **kw):
    print(name, file=sys.stderr, flush=True)
    ctx.load_scene(scene)
    sums, _, _ = ctx.render(scene.render_params(**kw), want_image=False)
    out.append(hashlib.sha1(np.ascontiguousarray(sums).tobytes()).hexdigest())
s2 = native.config_scene(2)
run("c2", s2, width=512, height=512, spp=16, seed=3)
run("c2 light cache", s2, width=512, height=512, spp=16, seed=4, guide=2, guide_train_passes=2,
    samples_per_pass=4, guide_light_grid=8)
run("c2 three light samples", s2, width=512, height=512, spp=16, seed=5, direct_light_sample_num=3)
H, W = 64, 128
v = (np.arange(H) + 0.5) / H
s2.env_image = np.repeat((0.2 + 0.6 * (1 - v))[:, None, None] * np.array([0.6, 0.75, 1.0]), W, 1).astype(np.float32)
run("c2 image environment", s2, width=512, height=512, spp=16, seed=6)
s3 = native.config_scene(3, detail=0.25)
run("c3", s3, width=512, height=512, spp=16, seed=7, max_depth=12)
print(" ".join(out))
"""


def _hashes(setting: str) -> list[str]:
    env = dict(os.environ, TRACER_HIP_SHADE_SORT=setting)
    print(f"shade order setting {setting}", flush=True)
    r = subprocess.run([sys.executable, "-c", CHILD], env=env, capture_output=True, text=True, timeout=280,
                       cwd=str(ROOT))
    assert r.returncode == 0, r.stderr[-2000:]
    return r.stdout.split()


def test_shade_order_leaves_film_unchanged():
    off = _hashes("5")
    assert len(off) == 5
    assert _hashes("2") == off
    assert _hashes("1") == off

"""Benchmark of the wavefront path tracer on one MI355X.

    python bench.py [--steps K] [--warmup W] [--gpus 1] [--config 2] [--no-cpu]

One step renders the configuration's full frame (configuration 2: 64
displaced level-5 icospheres, 1,310,754 triangles, Disney materials, 16 quad
lights + constant environment, 1024x1024, 64 spp, max depth 16) with the
scene and BVH already resident in HBM.  The timed region is K steps between
device synchronisations; value = camera samples per second over it.

Also reported:
  roofline   k_extend (closest-hit traversal, the dominant kernel): algorithmic
             bytes per extension ray = 32 (ray) + 16 (hit) + 128 * nodes
             visited + 48 * triangles tested, with visit counts from an
             instrumented replay of the same frame (profile = 2); divided by
             the k_extend time measured with HIP events on the library's
             stream inside the timed region.  traffic (HBM bytes per launch)
             comes from the committed rocprofv3 --pmc summary when present.
  cpu_baseline  the CPU restatement of the reference renderer (all lights per
             vertex, binary SAH BVH, std::thread over 16x16 tiles) on the
             same scene at a reduced film, timed on this host.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time
from pathlib import Path

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

METRIC = "equal-time relMSE vs 64k-spp ref (plain vs guided); Mrays/s; samples/s"
HBM_PEAK_GBS = 8000.0
NODE_BYTES, TRI_BYTES, RAY_BYTES, HIT_BYTES = 128, 48, 32, 16


def cpu_model() -> str:
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(scene, params, threads: int, target_s: float = 15.0) -> dict:
    """The oracle's reference-mode renderer on a reduced film of the same scene."""
    from oracle import refcpu
    t0 = time.time()
    ref = refcpu.RefScene(scene, 3)
    build_s = time.time() - t0
    W, H = 256, 256
    probe = scene.render_params(width=W, height=H, spp=1, seed=99)
    _, ms, _ = ref.render(probe, quirks=True, nee=0, threads=threads)
    spp = max(1, min(64, int(target_s * 1000.0 / max(ms, 1.0))))
    p = scene.render_params(width=W, height=H, spp=spp, seed=100)
    _, ms, rays = ref.render(p, quirks=True, nee=0, threads=threads)
    samples = W * H * spp
    return {
        "value": samples / (ms / 1e3),
        "unit": "samples/s",
        "cores": threads,
        "kind": "port",
        "sample": (f"reference-mode CPU restatement (all-lights NEE + MIS BSDF ray, binary SAH BVH leaf 3, "
                   f"16x16 tiles) on the same scene, {W}x{H} film at {spp} spp ({samples} samples, "
                   f"{ms / 1e3:.1f} s, {(rays[0] + rays[1]) / (ms / 1e3) / 1e6:.1f} Mrays/s); BVH build "
                   f"{build_s:.1f} s not timed; CPU: {cpu_model()}"),
    }


def read_traffic(path: Path):
    try:
        d = json.loads(path.read_text())
        return d.get("bytes_per_launch")
    except (OSError, ValueError):
        return None


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--config", type=int, default=2)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-threads", type=int, default=0)
    ap.add_argument("--copy-probe", action="store_true",
                    help="run one 1 GiB device-to-device copy (calibrates rocprofv3 FETCH_SIZE / WRITE_SIZE)")
    args = ap.parse_args()
    if args.copy_probe:
        import torch
        n = 1 << 28
        a = torch.ones(n, dtype=torch.float32, device="cuda")
        b = torch.empty_like(a)
        torch.cuda.synchronize()
        b.copy_(a)
        torch.cuda.synchronize()
        print(json.dumps({"copy_bytes_read": n * 4, "copy_bytes_written": n * 4}))
        return 0
    if args.gpus != 1:
        print("bench.py measures one GPU: this build has no multi-GPU path (rendering shards by sample index; "
              "see DESIGN.md). Run with --gpus 1.", file=sys.stderr)
        return 2

    from tracer_amd import native

    scene = native.config_scene(args.config)
    ctx = native.Context(0)
    t0 = time.time()
    ctx.load_scene(scene)
    bvh_s = time.time() - t0
    info = ctx.bvh_info()
    base = scene.render_params(seed=0)
    W, H, spp = base.width, base.height, base.spp

    for i in range(args.warmup):
        ctx.render(scene.render_params(seed=1000 + i), want_image=False)

    # timed region: K full frames, profile = 1 (HIP events around every launch)
    totals = dict(ext=0, shadow=0, ms_extend=0.0, ms_shade=0.0, ms_shadow=0.0, ms_other=0.0, ms_dev=0.0, launches=0)
    t0 = time.perf_counter()
    for k in range(args.steps):
        _, _, st = ctx.render(scene.render_params(seed=k, profile=1), want_image=False)
        totals["ext"] += st.extension_rays
        totals["shadow"] += st.shadow_rays
        totals["ms_extend"] += st.ms_extend
        totals["ms_shade"] += st.ms_shade
        totals["ms_shadow"] += st.ms_shadow
        totals["ms_other"] += st.ms_other
        totals["ms_dev"] += st.ms_total
        totals["launches"] += st.extend_launches
    elapsed = time.perf_counter() - t0

    samples = args.steps * W * H * spp
    value = samples / elapsed
    mrays = (totals["ext"] + totals["shadow"]) / elapsed / 1e6

    # instrumented replay of one frame: node visits and triangle tests per extension ray
    _, _, cst = ctx.render(scene.render_params(seed=0, profile=2), want_image=False)
    nodes_per_ray = cst.node_visits / max(1, cst.extension_rays)
    tris_per_ray = cst.tri_tests / max(1, cst.extension_rays)
    bytes_per_ray = RAY_BYTES + HIT_BYTES + NODE_BYTES * nodes_per_ray + TRI_BYTES * tris_per_ray
    launches = max(1, totals["launches"])
    bytes_per_launch = bytes_per_ray * totals["ext"] / launches
    ms_per_launch = totals["ms_extend"] / launches
    achieved = bytes_per_launch / (ms_per_launch / 1e3) / 1e9 if ms_per_launch > 0 else 0.0
    traffic = read_traffic(ROOT / "profiles" / f"k_extend_c{args.config}_traffic.json")

    out = {
        "metric": METRIC,
        "value": value,
        "unit": "samples/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": elapsed * 1e3 / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {
            "workload": (f"configuration {args.config}: {scene.num_triangles} triangles, {W}x{H} film, {spp} spp, "
                         f"max depth {base.max_depth}, plain wavefront path tracer (one light by power + MIS)"),
            "film": f"{W}x{H}", "spp": spp, "max_depth": base.max_depth, "triangles": scene.num_triangles,
        },
        "mrays_per_s": mrays,
        "samples_per_s": value,
        "rays": {"extension": totals["ext"], "shadow": totals["shadow"]},
        "kernel_ms": {k: round(totals[k], 3) for k in ("ms_extend", "ms_shade", "ms_shadow", "ms_other")},
        "device_ms_per_step": totals["ms_dev"] / args.steps,
        "bvh": {"build_s": round(bvh_s, 3), **info, "nodes_per_ray": nodes_per_ray, "tris_per_ray": tris_per_ray},
        "roofline": {
            "kernel": "k_extend",
            "bound": "hbm",
            "achieved": achieved,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS,
            "traffic": traffic,
            "bytes_per_ray": bytes_per_ray,
            "launches": launches,
            "ms_per_launch": ms_per_launch,
        },
    }
    if not args.no_cpu:
        threads = args.cpu_threads or min(16, os.cpu_count() or 1)
        out["cpu_baseline"] = cpu_baseline(scene, base, threads)
    ctx.close()
    print(json.dumps(out))
    return 0


if __name__ == "__main__":
    sys.exit(main())

"""Benchmark of the wavefront path tracer on one MI355X.

    python bench.py [--steps K] [--warmup W] [--gpus 1] [--no-cpu] [--no-relmse]

Throughput (the reported value).  One step renders configuration 2's full
frame (64 displaced level-5 icospheres, 1,310,754 triangles, Disney
materials, 16 quad lights + constant environment, 1024x1024, 64 spp, max depth
16) with the scene and BVH resident in HBM; value = camera samples per second
over K steps between device synchronisations.

Also reported on the same JSON line:
  roofline       k_extend (closest-hit traversal, the dominant kernel):
                 algorithmic bytes per extension ray = 32 (ray) + 16 (hit) +
                 128 * nodes visited + 48 * triangles tested, visit counts from
                 an instrumented replay of the same frame (which runs every
                 bounce through k_extend), times the rays of the k_extend
                 launches (the tail kernel traces the last bounces' rays
                 itself); divided by the k_extend time measured with HIP
                 events on the library's stream inside the timed region.
                 traffic: HBM bytes per launch
                 from the committed rocprofv3 FETCH_SIZE / WRITE_SIZE summary.
  equal_time     configuration 4 (the only light shines through a 0.5 x 0.5
                 aperture) at its 1024 x 1024 film, 256 spp: relMSE against a
                 16384-spp plain reference of the guided renderer (learned
                 light sampling, training included) and of the plain renderer
                 given the same device time; the guide network alone and with
                 the learned light sampling as ablations.
  guide_network  GFLOP/s of guide inference and training against the bf16
                 MFMA peak.
  cpu_baseline   the CPU restatement of the reference renderer (all lights
                 per vertex, binary SAH BVH, std::thread over 16x16 tiles) on
                 configuration 2 at a reduced film, timed on this host, with
                 the GPU's time-to-equal-relMSE speedup on the same film.
  other_configs  configuration 3 (fog) full frame, configuration 5 (18.9M
                 triangles) full frame at 256 spp, plain and guided (network
                 width 128 with learned light sampling; relMSE on a 480x270
                 crop), the bidirectional path tracer on configuration 1,
                 configuration 2 under a 2048 x 1024
                 image environment instead of its constant one, configuration
                 2 with a 1024 x 1024 8-bit base-colour texture, configuration
                 3 with a 64^3 density grid in place of its fog, and OBJ
                 ingest: a 328K-triangle OBJ/MTL pair read and flattened by
                 the facade (MB/s, triangles/s) and rendered.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

METRIC = "equal-time relMSE vs 64k-spp ref (plain vs guided); Mrays/s; samples/s"
HBM_PEAK_GBS = 8000.0
BF16_PEAK_TFLOPS = 2500.0
NODE_BYTES, TRI_BYTES, RAY_BYTES, HIT_BYTES = 128, 48, 32, 16


def cpu_model() -> str:
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_topology() -> dict:
    """Logical CPUs this process may run on, physical cores among them, SMT."""
    allowed = sorted(os.sched_getaffinity(0))
    cores = set()
    for cpu in allowed:
        try:
            base = f"/sys/devices/system/cpu/cpu{cpu}/topology/"
            pkg = open(base + "physical_package_id").read().strip()
            core = open(base + "core_id").read().strip()
            cores.add((pkg, core))
        except OSError:
            cores.add(("?", str(cpu)))
    try:
        smt = open("/sys/devices/system/cpu/smt/active").read().strip() == "1"
    except OSError:
        smt = len(cores) < len(allowed)
    return {"model": cpu_model(), "logical_cpus_allowed": len(allowed), "physical_cores_allowed": len(cores),
            "smt_active": smt, "os_cpu_count": os.cpu_count()}


def cpu_baseline(scene, threads: int, target_s: float = 8.0, runs: int = 3) -> dict:
    """The oracle's CPU renderer (fixed mode: the reference's algorithm with its
    estimator defects corrected) on a reduced film of the same scene, timed
    `runs` times (the median is reported), and the time-to-equal-relMSE ratio
    of the plain GPU path on that film."""
    from oracle import refcpu
    from tracer_amd import native
    t0 = time.time()
    ref = refcpu.RefScene(scene, 3)
    build_s = time.time() - t0
    W, H = 256, 256
    probe = scene.render_params(width=W, height=H, spp=1, seed=99)
    _, ms, _ = ref.render(probe, quirks=False, nee=0, threads=threads)
    spp = max(1, min(64, int(target_s * 1000.0 / max(ms, 1.0))))
    timings = []
    for r in range(runs):
        p = scene.render_params(width=W, height=H, spp=spp, seed=100 + r)
        sums_r, ms_r, rays_r = ref.render(p, quirks=False, nee=0, threads=threads)
        timings.append(ms_r)
        if r == 0:
            cpu_sums, rays = sums_r, rays_r
    ms = float(np.median(timings))
    samples = W * H * spp
    # GPU on the same film: a 16384-spp reference, then the plain path at 64 spp
    ctx = native.Context(0)
    ctx.load_scene(scene)
    R = None
    for c in range(4):
        R, _, _ = ctx.render(scene.render_params(width=W, height=H, spp=4096, seed=7, sample_offset=4096 * c), film=R,
                             want_image=False)
    Rimg = R[..., :3] / R[..., 3:]

    def relmse(sums):
        img = sums[..., :3] / sums[..., 3:]
        return float(np.mean((img - Rimg) ** 2 / (Rimg ** 2 + 1e-2)))

    gpu_spp = 64
    g_err, g_ms = [], []
    for k in range(4):
        sums, _, st = ctx.render(scene.render_params(width=W, height=H, spp=gpu_spp, seed=300 + k), want_image=False)
        g_err.append(relmse(sums))
        g_ms.append(st.ms_total)
    ctx.close()
    cpu_err = relmse(cpu_sums)
    # relMSE falls as 1/time for both: the ratio of (relMSE x time) is the time-to-equal-relMSE speedup
    speedup = (cpu_err * ms) / (float(np.mean(g_err)) * float(np.mean(g_ms)))
    return {
        "value": samples / (ms / 1e3),
        "unit": "samples/s",
        "cores": threads,
        "kind": "port",
        "runs_ms": timings,
        "host": cpu_topology(),
        "sample": (f"CPU restatement of the reference renderer in fixed mode (all-lights NEE + MIS BSDF ray, binary SAH "
                   f"BVH leaf 3, 16x16 tiles) on configuration 2, {W}x{H} film at {spp} spp ({samples} samples, "
                   f"median of {runs} runs {ms / 1e3:.1f} s, {(rays[0] + rays[1]) / (ms / 1e3) / 1e6:.1f} Mrays/s) on "
                   f"{threads} threads; BVH build {build_s:.1f} s not timed; CPU: {cpu_model()}"),
        "equal_relmse": {
            "film": f"{W}x{H}", "reference": "plain GPU path, 16384 spp, FP64 film",
            "cpu_spp": spp, "cpu_ms": ms, "cpu_relmse": cpu_err,
            "gpu_spp": gpu_spp, "gpu_ms": float(np.mean(g_ms)), "gpu_relmse": float(np.mean(g_err)),
            "gpu_speedup_at_equal_relmse": speedup,
        },
    }


def config5_guided(ctx, scene, hidden: int = 128, light_grid: int = 20, train_spp: int = 8, train_passes: int = 4,
                   seeds: int = 12) -> dict:
    """Configuration 5 at its full film and 256 spp with the guide network
    widened to `hidden` (and learned light sampling on a light_grid^3 grid),
    against the plain renderer: device time of the full frame, and relMSE at
    equal spp and equal time on a 480x270 crop of the scene (the same camera
    at a quarter of the resolution) against a 16384-spp plain reference.  The
    plain estimator's error is heavy-tailed on this scene (a rare bright path
    in one 256-spp image can raise its relMSE twentyfold), so the comparison
    runs `seeds` independent renders and reports the mean and the median."""
    P = scene.render_params(seed=0)
    W, H, spp = P.width, P.height, P.spp

    def guided(w, h, seed, n_spp, profile=0):
        ctx.guide_init(hidden=hidden, seed=11 + seed, batch=1 << 16, steps_per_pass=8)
        ts = train_spp * train_passes
        s1, _, a = ctx.render(scene.render_params(width=w, height=h, spp=ts, seed=seed, guide=1,
                                                  guide_train_passes=train_passes, samples_per_pass=train_spp,
                                                  guide_light_grid=light_grid, profile=profile), want_image=False)
        s2, _, b = ctx.render(scene.render_params(width=w, height=h, spp=n_spp - ts, seed=seed, sample_offset=ts,
                                                  guide=1, guide_light_grid=light_grid, profile=profile),
                              film=s1, want_image=False)
        return s2, a.ms_total + b.ms_total, a, b

    _, g_ms, a, b = guided(W, H, 900, spp, profile=1)
    _, _, pst = ctx.render(scene.render_params(seed=901), want_image=False)
    full = {"film": f"{W}x{H}", "spp": spp, "plain_ms": pst.ms_total, "guided_ms": g_ms,
            "guided_samples_per_s": W * H * spp / (g_ms / 1e3), "plain_samples_per_s": W * H * spp / (pst.ms_total / 1e3),
            "guide_inference_ms": a.ms_guide + b.ms_guide, "guide_training_ms": a.ms_guide_train + b.ms_guide_train}
    w, h = W // 4, H // 4
    R = None
    for c in range(32):
        R, _, _ = ctx.render(scene.render_params(width=w, height=h, spp=512, seed=1, sample_offset=512 * c), film=R,
                             want_image=False)
    Rimg = R[..., :3] / R[..., 3:]

    def relmse(sums):
        img = sums[..., :3] / sums[..., 3:]
        return float(np.mean((img - Rimg) ** 2 / (Rimg ** 2 + 1e-2)))

    g_err, g_t, p_err, p_t = [], [], [], []
    for k in range(seeds):
        sg, ms, _, _ = guided(w, h, 500 + k, spp)
        g_err.append(relmse(sg))
        g_t.append(ms)
        sp, _, st = ctx.render(scene.render_params(width=w, height=h, spp=spp, seed=600 + k), want_image=False)
        p_err.append(relmse(sp))
        p_t.append(st.ms_total)
    pe, pt_, ge, gt = map(lambda v: float(np.mean(v)), (p_err, p_t, g_err, g_t))
    pm, gm = float(np.median(p_err)), float(np.median(g_err))
    return {"full_frame": full,
            "crop": {"film": f"{w}x{h}", "reference": "plain, 16384 spp", "spp": spp, "seeds": seeds,
                     "plain_relmse": pe, "plain_relmse_median": pm, "plain_relmse_max": float(np.max(p_err)),
                     "plain_ms": pt_, "guided_relmse": ge, "guided_relmse_median": gm,
                     "guided_relmse_max": float(np.max(g_err)), "guided_ms": gt, "equal_spp_ratio": pe / ge,
                     "equal_time_ratio_plain_over_guided": (pe * pt_) / (ge * gt),
                     "equal_time_ratio_median": (pm * pt_) / (gm * gt)},
            "guide": f"network width {hidden}, {train_passes} training passes of {train_spp} spp, learned light "
                     f"sampling on a {light_grid}^3 grid"}


def bdpt_config1(ctx) -> dict:
    """Bidirectional path tracing (tr_render_bdpt) of configuration 1 at
    1024x1024 with the reference's default vertex counts (10 camera and 10
    light vertices): samples, subpath rays and connection rays per second."""
    from tracer_amd import native
    scene = native.config_scene(1)
    ctx.load_scene(scene)
    p = scene.bdpt_params(10, 10, width=1024, height=1024, spp=8, seed=1, profile=1)
    ctx.render_bdpt(scene.bdpt_params(10, 10, width=1024, height=1024, spp=1, seed=2), want_image=False)
    t0 = time.perf_counter()
    _, _, st = ctx.render_bdpt(p, want_image=False)
    el = time.perf_counter() - t0
    return {"workload": "configuration 1, 1024x1024, 8 spp, max_camera_vertex_count 10, max_light_vertex_count 10",
            "samples_per_s": 1024 * 1024 * 8 / el, "subpath_mrays_per_s": st.extension_rays / el / 1e6,
            "connection_mrays_per_s": st.shadow_rays / el / 1e6, "s_per_frame": el,
            "device_ms": {"subpath_traversal": st.ms_extend, "subpath_vertices": st.ms_shade,
                          "connections": st.ms_tail, "visibility": st.ms_shadow, "start_gather_film": st.ms_other}}


def other_configs() -> dict:
    """Full frames of configuration 3 (fog, delta tracking) and configuration 5
    (18.9M triangles, 256 spp), configuration 5 guided, and the bidirectional
    path tracer."""
    from tracer_amd import native
    out = {}
    ctx = native.Context(0)
    for cfg, spp in ((3, None), (5, None)):
        scene = native.config_scene(cfg)
        t0 = time.time()
        ctx.load_scene(scene)
        bvh_s = time.time() - t0
        p = scene.render_params(seed=0) if spp is None else scene.render_params(seed=0, spp=spp)
        ctx.render(scene.render_params(seed=9, spp=1, width=p.width, height=p.height), want_image=False)
        t0 = time.perf_counter()
        _, _, st = ctx.render(p, want_image=False)
        el = time.perf_counter() - t0
        out[f"config{cfg}"] = {
            "workload": (f"{scene.num_triangles} triangles, {p.width}x{p.height}, {p.spp} spp, max depth {p.max_depth}"),
            "samples_per_s": p.width * p.height * p.spp / el, "mrays_per_s": (st.extension_rays + st.shadow_rays) / el / 1e6,
            "s_per_frame": el, "bvh_build_s": bvh_s,
        }
        print(f"[bench] configuration {cfg}: {out[f'config{cfg}']['samples_per_s'] / 1e6:.1f} M samples/s",
              file=sys.stderr, flush=True)
        if cfg == 5:
            out["config5_guided"] = config5_guided(ctx, scene)
            print(f"[bench] configuration 5 guided: {json.dumps(out['config5_guided']['crop'])}", file=sys.stderr,
                  flush=True)
    out["bdpt_config1"] = bdpt_config1(ctx)
    print(f"[bench] BDPT configuration 1: {out['bdpt_config1']['samples_per_s'] / 1e6:.1f} M samples/s",
          file=sys.stderr, flush=True)

    def frame(scene, p):
        ctx.load_scene(scene)
        ctx.render(scene.render_params(seed=9, spp=1, width=p.width, height=p.height), want_image=False)
        t0 = time.perf_counter()
        _, _, st = ctx.render(p, want_image=False)
        el = time.perf_counter() - t0
        return {"samples_per_s": p.width * p.height * p.spp / el,
                "mrays_per_s": (st.extension_rays + st.shadow_rays) / el / 1e6, "s_per_frame": el}

    # configuration 2 under an image environment: a sky gradient with a sun disc
    scene = native.config_scene(2)
    H, W = 1024, 2048
    v = (np.arange(H) + 0.5) / H
    u = (np.arange(W) + 0.5) / W
    U, V = np.meshgrid(u, v)
    sky = (0.2 + 0.6 * np.clip(1 - V, 0, 1))[..., None] * np.array([0.6, 0.75, 1.0])
    sky = sky + 200.0 * (((U - 0.3) ** 2 + (V - 0.25) ** 2) < 2e-4)[..., None]
    scene.env_image = sky.astype(np.float32)
    p = scene.render_params(seed=0)
    out["config2_env_image"] = {"workload": f"configuration 2 with a {W}x{H} image environment, {p.width}x{p.height}, "
                                            f"{p.spp} spp, max depth {p.max_depth}", **frame(scene, p)}
    print(f"[bench] configuration 2, image environment: {out['config2_env_image']['samples_per_s'] / 1e6:.1f} M samples/s",
          file=sys.stderr, flush=True)

    # configuration 2 with the Disney base colour from a 1024x1024 8-bit texture
    # through spherical texture coordinates of the mesh vertices
    scene = native.config_scene(2)
    P = scene.positions.reshape(-1, 3).astype(np.float64)
    D = P - P.mean(0)
    R = np.linalg.norm(D, axis=1) + 1e-12
    scene.uvs = np.stack([0.5 + np.arctan2(D[:, 2], D[:, 0]) / (2 * np.pi) * 8,
                          np.arccos(np.clip(D[:, 1] / R, -1, 1)) / np.pi * 4], 1).astype(np.float32)
    tex = np.random.default_rng(1).integers(0, 256, (1024, 1024, 3), dtype=np.uint8)
    scene.textures = [tex]
    for m in scene.materials:
        m.base_color_texture = 1
    p = scene.render_params(seed=0)
    out["config2_textured"] = {"workload": f"configuration 2 with every material's base colour from a 1024x1024 8-bit "
                                           f"texture, {p.width}x{p.height}, {p.spp} spp, max depth {p.max_depth}",
                               **frame(scene, p)}
    print(f"[bench] configuration 2, textured: {out['config2_textured']['samples_per_s'] / 1e6:.1f} M samples/s",
          file=sys.stderr, flush=True)

    # configuration 3 with its homogeneous fog replaced by a 64^3 density grid
    # spanning the room (a heterogeneous medium: delta and ratio tracking)
    from tracer_amd import abi
    scene = native.config_scene(3)
    P = scene.positions.reshape(-1, 3)
    lo, hi = P.min(0), P.max(0)
    n = 64
    z, y, x = np.meshgrid(*(np.arange(n) / (n - 1),) * 3, indexing="ij")
    grid = 0.02 + 0.4 * np.exp(-((x - 0.5) ** 2 + (y - 0.4) ** 2 + (z - 0.5) ** 2) / 0.05)
    fog = scene.media[0]
    l2w = np.zeros((3, 4))
    l2w[np.arange(3), np.arange(3)] = hi - lo
    l2w[:, 3] = lo
    scene.media = [abi.make_grid_medium(grid, albedo=tuple(fog.sigma_s[k] / max(fog.sigma_s[k] + fog.sigma_a[k], 1e-12)
                                                         for k in range(3)),
                                        g=fog.g, local_to_world=l2w, max_scattering_count=fog.max_scattering_count)]
    p = scene.render_params(seed=0)
    out["config3_grid_medium"] = {"workload": f"configuration 3 with a {n}^3 density grid in place of its fog, "
                                              f"{p.width}x{p.height}, {p.spp} spp, max depth {p.max_depth}",
                                  **frame(scene, p)}
    print(f"[bench] configuration 3, grid medium: {out['config3_grid_medium']['samples_per_s'] / 1e6:.1f} M samples/s",
          file=sys.stderr, flush=True)

    # OBJ ingest: configuration 2's spheres at a quarter of the size, written as OBJ + MTL
    import tempfile
    src = native.config_scene(2, detail=0.25)
    with tempfile.TemporaryDirectory() as tmp:
        P = src.positions.reshape(-1, 3)
        N = src.normals.reshape(-1, 3)
        I = src.indices.reshape(-1, 3) + 1
        lines = ["mtllib m.mtl", "usemtl a"]
        body = "\n".join(lines) + "\n"
        body += "".join(f"v {x:.7g} {y:.7g} {z:.7g}\n" for x, y, z in P)
        body += "".join(f"vn {x:.7g} {y:.7g} {z:.7g}\n" for x, y, z in N)
        body += "".join(f"f {a}//{a} {b}//{b} {c}//{c}\n" for a, b, c in I)
        path = Path(tmp) / "m.obj"
        path.write_text(body)
        (Path(tmp) / "m.mtl").write_text("newmtl a\nKd 0.6 0.5 0.4\nKs 0.2 0.2 0.2\nNs 30\n")
        cam = abi.TrCamera.from_buffer_copy(src.camera)   # a copy the loader may keep
        t0 = time.perf_counter()
        obj = native.obj_scene(path, cam, outside_fog=False)
        parse_s = time.perf_counter() - t0
        size_mb = path.stat().st_size / 1e6
    obj.env = (0.5, 0.5, 0.5)
    p = obj.render_params(width=1024, height=1024, spp=16, max_depth=16, seed=0)
    out["obj_ingest"] = {
        "workload": f"{obj.num_triangles} triangles in a {size_mb:.1f} MB OBJ (v//vn corners, one Phong material), "
                    f"rendered at {p.width}x{p.height}, {p.spp} spp under a constant environment",
        "ingest_s": parse_s, "ingest_mb_per_s": size_mb / parse_s, "ingest_triangles_per_s": obj.num_triangles / parse_s,
        "ingest_covers": "load_model_from_file, run()-style wiring, flattening to tr_scene_desc, copy to Python",
        **frame(obj, p)}
    print(f"[bench] OBJ ingest: {size_mb / parse_s:.1f} MB/s, {obj.num_triangles / parse_s / 1e6:.2f} M triangles/s; "
          f"render {out['obj_ingest']['samples_per_s'] / 1e6:.1f} M samples/s", file=sys.stderr, flush=True)
    ctx.close()
    return out


def limiter_text(info, traffic, ms_per_launch):
    """What the counter pass says bounds the traversal kernel."""
    if not info or not traffic or ms_per_launch <= 0:
        return None
    frac = traffic / (ms_per_launch / 1e3) / 1e9 / HBM_PEAK_GBS
    return (f"VALU issue under divergence and memory latency: the bytes that leave L2 use {frac:.2f} of the HBM "
            f"bandwidth, while the VALU is {info.get('valu_busy', 0):.0%} busy with "
            f"{info.get('lanes_active_per_valu_instruction', 0):.0f} of 64 lanes active per instruction")


def read_traffic(path: Path):
    try:
        return json.loads(path.read_text())
    except (OSError, ValueError):
        return None


def guide_flops(H: int) -> tuple[float, float]:
    """Algorithmic flops of one network evaluation and of one training sample
    (81 used inputs, 40 used outputs): forward; forward + input gradients of
    layers 4..1 + weight gradients."""
    fwd = 2.0 * (81 * H + 3 * H * H + H * 40)
    bwd_in = 2.0 * (3 * H * H + H * 40)
    return fwd, fwd + bwd_in + fwd


def throughput(ctx, scene, steps: int, warmup: int) -> dict:
    base = scene.render_params(seed=0)
    W, H, spp = base.width, base.height, base.spp
    film = ctx.host_array((H, W, 4))   # FP64 film sums of each frame, copied to page-locked host memory
    for i in range(warmup):
        ctx.render(scene.render_params(seed=1000 + i), want_image=False, out=film)
    tot = dict(ext=0, shadow=0, tail_ext=0, ms_extend=0.0, ms_shade=0.0, ms_shadow=0.0, ms_tail=0.0, ms_other=0.0, ms_dev=0.0,
               launches=0)
    t0 = time.perf_counter()
    for k in range(steps):
        _, _, st = ctx.render(scene.render_params(seed=k, profile=1), want_image=False, out=film)
        tot["ext"] += st.extension_rays
        tot["shadow"] += st.shadow_rays
        tot["tail_ext"] += st.tail_extension_rays
        tot["ms_tail"] += st.ms_tail
        tot["ms_extend"] += st.ms_extend
        tot["ms_shade"] += st.ms_shade
        tot["ms_shadow"] += st.ms_shadow
        tot["ms_other"] += st.ms_other
        tot["ms_dev"] += st.ms_total
        tot["launches"] += st.extend_launches
    elapsed = time.perf_counter() - t0
    return {"W": W, "H": H, "spp": spp, "max_depth": base.max_depth, "elapsed": elapsed, **tot}


def equal_time(res: int, ref_spp: int, spp: int, seeds: int, ablation_seeds: int, hidden: int, light_grid: int,
               light_train_passes: int, light_train_pp: int, net_train_pp: int, net_steps: int) -> dict:
    """Configuration 4 at res x res (its stated 1024 x 1024 by default): the
    guided renderer against the plain one at the same device time, both
    against a ref_spp plain reference.

    Guided (headline): learned light sampling (guide = 2): light_train_passes
    training passes of light_train_pp spp each (the first samples lights by
    power, the later ones the distribution learned so far), then the rest of
    the spp sampling the learned distribution; training included in the time.
    Plain: a render of the same device time (its spp scaled by the time ratio
    measured here), independent seeds.
    Ablations (relMSE x time against the plain render's, fewer seeds): the
    guide network alone (guide = 1) and the network with the learned light
    sampling."""
    from tracer_amd import native
    scene = native.config_scene(4)
    ctx = native.Context(0)
    ctx.load_scene(scene)
    t0 = time.perf_counter()
    chunk = min(1024, ref_spp)
    ref = None
    for c in range(ref_spp // chunk):
        p = scene.render_params(width=res, height=res, spp=chunk, seed=1, sample_offset=c * chunk)
        ref, _, _ = ctx.render(p, film=ref, want_image=False)
        print(f"[bench] reference chunk {c + 1}/{ref_spp // chunk} ({time.perf_counter() - t0:.1f} s)", file=sys.stderr,
              flush=True)
    ref_s = time.perf_counter() - t0
    R = ref[..., :3] / ref[..., 3:]

    def relmse(sums):
        img = sums[..., :3] / sums[..., 3:]
        return float(np.mean((img - R) ** 2 / (R ** 2 + 1e-2)))

    def light_render(k):
        ts = light_train_passes * light_train_pp
        s1, _, a = ctx.render(scene.render_params(width=res, height=res, spp=ts, seed=200 + k, guide=2,
                                                  guide_train_passes=light_train_passes,
                                                  samples_per_pass=light_train_pp, guide_light_grid=light_grid),
                              want_image=False)
        s2, _, b = ctx.render(scene.render_params(width=res, height=res, spp=spp - ts, seed=200 + k, sample_offset=ts,
                                                  guide=2, guide_light_grid=light_grid), film=s1, want_image=False)
        return relmse(s2), a.ms_total + b.ms_total

    def net_render(k, grid):
        ctx.guide_init(hidden=hidden, seed=11 + k, batch=1 << 16, steps_per_pass=net_steps)
        s1, _, a = ctx.render(scene.render_params(width=res, height=res, spp=net_train_pp, seed=200 + k, guide=1,
                                                  guide_train_passes=1, samples_per_pass=net_train_pp,
                                                  guide_light_grid=grid, profile=1), want_image=False)
        s2, _, b = ctx.render(scene.render_params(width=res, height=res, spp=spp - net_train_pp, seed=200 + k,
                                                  sample_offset=net_train_pp, guide=1, guide_light_grid=grid,
                                                  profile=1), film=s1, want_image=False)
        return relmse(s2), a.ms_total + b.ms_total, a, b

    ctx.render(scene.render_params(width=res, height=res, spp=1, seed=5), want_image=False)   # warm-up
    p_ms, p_err_same = [], []
    for k in range(seeds):
        sums, _, st = ctx.render(scene.render_params(width=res, height=res, spp=spp, seed=300 + k), want_image=False)
        p_ms.append(st.ms_total)
        p_err_same.append(relmse(sums))
    g = [light_render(k) for k in range(seeds)]
    g_err, g_ms = [x[0] for x in g], [x[1] for x in g]
    spp_eq = max(1, int(round(spp * np.mean(g_ms) / np.mean(p_ms))))
    p_err, p_eq_ms = [], []
    for k in range(seeds):
        sums, _, st = ctx.render(scene.render_params(width=res, height=res, spp=spp_eq, seed=400 + k), want_image=False)
        p_err.append(relmse(sums))
        p_eq_ms.append(st.ms_total)
    print(f"[bench] equal time: plain {np.mean(p_err):.5f} ({spp_eq} spp), learned light sampling {np.mean(g_err):.5f}",
          file=sys.stderr, flush=True)
    # ablations: relMSE x device time against the plain render at spp
    plain_work = float(np.mean(p_err_same)) * float(np.mean(p_ms))
    ablations = {}
    g_inf, g_inf_ms, g_smp, g_train_ms = 0, 0.0, 0, 0.0
    for name, grid in (("network", 0), ("network_and_light", light_grid)):
        if ablation_seeds <= 0:
            break
        r = [net_render(k, grid) for k in range(ablation_seeds)]
        e, ms = [x[0] for x in r], [x[1] for x in r]
        ablations[name] = {"relmse": float(np.mean(e)), "ms": float(np.mean(ms)),
                           "ratio_plain_over_guided_by_time_scaling": plain_work / (float(np.mean(e)) * float(np.mean(ms)))}
        if grid == 0:
            for _, _, a, b in r:
                g_inf += a.guide_inferences + b.guide_inferences
                g_inf_ms += a.ms_guide + b.ms_guide
                g_smp += a.guide_train_samples
                g_train_ms += a.ms_guide_train
    ctx.close()
    fwd, trn = guide_flops(hidden)
    inf_tflops = g_inf * fwd / (g_inf_ms / 1e3) / 1e12 if g_inf_ms > 0 else 0.0
    trn_tflops = g_smp * trn / (g_train_ms / 1e3) / 1e12 if g_train_ms > 0 else 0.0
    ref_floor = float(np.mean(p_err_same)) * spp / ref_spp
    return {
        "equal_time": {
            "workload": (f"configuration 4 ({scene.num_triangles} triangles, the only light behind a 0.5x0.5 aperture) at "
                         f"{res}x{res}, {spp} spp, max depth 16; reference: plain GPU path, {ref_spp} spp, FP64 film "
                         f"({ref_s:.0f} s; its own relMSE is about {ref_floor:.1e}, the plain 256-spp error scaled by "
                         f"spp); {seeds} seeds each"),
            "guided": {"mode": "learned light sampling (guide = 2)", "spp": spp, "ms": float(np.mean(g_ms)),
                       "relmse": float(np.mean(g_err)), "relmse_median": float(np.median(g_err)),
                       "training": f"{light_train_passes} passes of {light_train_pp} spp ({light_grid}^3 cells), "
                                   f"the first by power, later ones sampling what was learned so far; time included"},
            "plain_equal_time": {"spp": spp_eq, "ms": float(np.mean(p_eq_ms)), "relmse": float(np.mean(p_err)),
                                 "relmse_median": float(np.median(p_err))},
            "plain_equal_spp": {"spp": spp, "ms": float(np.mean(p_ms)), "relmse": float(np.mean(p_err_same)),
                                "relmse_median": float(np.median(p_err_same))},
            "ratio_plain_over_guided": float(np.mean(p_err) / np.mean(g_err)),
            "ratio_median": float(np.median(p_err) / np.median(g_err)),
            "ablations": {**ablations, "seeds": ablation_seeds,
                          "network_training": f"one {net_train_pp}-spp pass, {net_steps} Adam steps of 65536 samples"},
        },
        "guide_network": {
            "hidden": hidden, "flops_per_inference": fwd, "flops_per_training_sample": trn,
            "inference_tflops": inf_tflops, "training_tflops": trn_tflops, "peak_tflops_bf16": BF16_PEAK_TFLOPS,
            "inference_frac_of_peak": inf_tflops / BF16_PEAK_TFLOPS,
            "training_frac_of_peak": trn_tflops / BF16_PEAK_TFLOPS,
            "inferences": int(g_inf), "training_samples": int(g_smp),
        },
    }


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--config", type=int, default=2)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-threads", type=int, default=0)
    ap.add_argument("--no-relmse", action="store_true")
    ap.add_argument("--no-other", action="store_true", help="skip the configuration 3 and 5 frames")
    ap.add_argument("--relmse-res", type=int, default=1024)
    ap.add_argument("--ref-spp", type=int, default=16384)
    ap.add_argument("--relmse-spp", type=int, default=256)
    ap.add_argument("--relmse-seeds", type=int, default=6)
    ap.add_argument("--ablation-seeds", type=int, default=2)
    ap.add_argument("--light-grid", type=int, default=20, help="cells per axis of the learned light sampling")
    ap.add_argument("--light-train-passes", type=int, default=8)
    ap.add_argument("--light-train-pass-spp", type=int, default=8)
    ap.add_argument("--train-pass-spp", type=int, default=32, help="spp of the guide network's training pass")
    ap.add_argument("--train-steps", type=int, default=16, help="Adam steps after the network's training pass")
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
        print("bench.py measures one GPU: there is no multi-GPU path (rendering would shard by sample index "
              "across GPUs, each rendering its own samples of every pixel into its own FP64 film). "
              "Run with --gpus 1.", file=sys.stderr)
        return 2

    from tracer_amd import native

    scene = native.config_scene(args.config)
    ctx = native.Context(0)
    t0 = time.time()
    ctx.load_scene(scene)
    bvh_s = time.time() - t0
    info = ctx.bvh_info()
    print(f"[bench] configuration {args.config} loaded ({scene.num_triangles} triangles, BVH {bvh_s:.2f} s)",
          file=sys.stderr, flush=True)
    tp = throughput(ctx, scene, args.steps, args.warmup)
    samples = args.steps * tp["W"] * tp["H"] * tp["spp"]
    value = samples / tp["elapsed"]
    mrays = (tp["ext"] + tp["shadow"]) / tp["elapsed"] / 1e6
    print(f"[bench] throughput {value / 1e6:.1f} M samples/s, {mrays:.0f} Mrays/s", file=sys.stderr, flush=True)

    # instrumented replay of one frame: node visits and triangle tests per extension ray
    _, _, cst = ctx.render(scene.render_params(seed=0, profile=2), want_image=False)
    nodes_per_ray = cst.node_visits / max(1, cst.extension_rays)
    tris_per_ray = cst.tri_tests / max(1, cst.extension_rays)
    bytes_per_ray = RAY_BYTES + HIT_BYTES + NODE_BYTES * nodes_per_ray + TRI_BYTES * tris_per_ray
    launches = max(1, tp["launches"])
    # rays of the k_extend launches: the tail kernel traces the last bounces' rays itself
    bytes_per_launch = bytes_per_ray * (tp["ext"] - tp["tail_ext"]) / launches
    ms_per_launch = tp["ms_extend"] / launches
    achieved = bytes_per_launch / (ms_per_launch / 1e3) / 1e9 if ms_per_launch > 0 else 0.0
    traffic_info = read_traffic(ROOT / "profiles" / f"k_extend_c{args.config}_traffic.json")
    traffic = traffic_info.get("bytes_per_launch") if traffic_info else None
    ctx.close()

    out = {
        "metric": METRIC,
        "value": value,
        "unit": "samples/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": tp["elapsed"] * 1e3 / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {
            "workload": (f"configuration {args.config}: {scene.num_triangles} triangles, {tp['W']}x{tp['H']} film, "
                         f"{tp['spp']} spp, max depth {tp['max_depth']}, plain wavefront path tracer (one light by "
                         f"power + MIS)"),
            "film": f"{tp['W']}x{tp['H']}", "spp": tp["spp"], "max_depth": tp["max_depth"],
            "triangles": scene.num_triangles,
        },
        "mrays_per_s": mrays,
        "samples_per_s": value,
        "rays": {"extension": tp["ext"], "shadow": tp["shadow"]},
        "kernel_ms": {k: round(tp[k], 3) for k in ("ms_extend", "ms_shade", "ms_shadow", "ms_tail", "ms_other")},
        "tail_extension_rays": tp["tail_ext"],
        "device_ms_per_step": tp["ms_dev"] / args.steps,
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
            # what the counters show (profiles/k_extend_c2_kernel_stats.md): the bytes that miss L2, over
            # this run's launch time, against the HBM peak; the L2 hit rate; lanes active per VALU instruction
            "counter_achieved_gbs": (traffic / (ms_per_launch / 1e3) / 1e9) if traffic and ms_per_launch > 0 else None,
            "counter_frac": (traffic / (ms_per_launch / 1e3) / 1e9 / HBM_PEAK_GBS) if traffic and ms_per_launch > 0
            else None,
            "l2_hit_rate": traffic_info.get("l2_hit_rate") if traffic_info else None,
            "lanes_active_per_valu_instruction": traffic_info.get("lanes_active_per_valu_instruction")
            if traffic_info else None,
            "valu_busy": traffic_info.get("valu_busy") if traffic_info else None,
            "limiter": limiter_text(traffic_info, traffic, ms_per_launch),
        },
    }
    if not args.no_other:
        out["other_configs"] = other_configs()
    if not args.no_relmse:
        print("[bench] equal-time relMSE on configuration 4 ...", file=sys.stderr, flush=True)
        out.update(equal_time(args.relmse_res, args.ref_spp, args.relmse_spp, args.relmse_seeds, args.ablation_seeds, 64,
                              args.light_grid, args.light_train_passes, args.light_train_pass_spp, args.train_pass_spp,
                              args.train_steps))
        print(f"[bench] equal-time ratio {out['equal_time']['ratio_plain_over_guided']:.2f}", file=sys.stderr,
              flush=True)
    if not args.no_cpu:
        # every logical CPU the job may use, up to the 16 that make up one GPU's share of the host
        threads = args.cpu_threads or min(16, len(os.sched_getaffinity(0)))
        print(f"[bench] CPU baseline on {threads} threads ...", file=sys.stderr, flush=True)
        out["cpu_baseline"] = cpu_baseline(scene, threads)
        out["gpu_over_cpu_samples_per_s"] = value / out["cpu_baseline"]["value"]
        print(f"[bench] GPU/CPU at equal relMSE: {out['cpu_baseline']['equal_relmse']['gpu_speedup_at_equal_relmse']:.0f}x",
              file=sys.stderr, flush=True)
    print(json.dumps(out))
    return 0


if __name__ == "__main__":
    sys.exit(main())

"""ctypes binding of the CPU oracle (oracle/_build/librefcpu.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, as the checker or the timed CPU baseline.  The
product (tracer_amd/) never imports this module.
"""
from __future__ import annotations

import ctypes as C
import subprocess
from pathlib import Path

import numpy as np

from tracer_amd.abi import TrMaterial, TrRenderParams, TrSceneDesc, dptr, fptr, iptr

HERE = Path(__file__).resolve().parent
LIB = HERE / "_build" / "librefcpu.so"
_lib = None


def lib() -> C.CDLL:
    global _lib
    if _lib is None:
        if not LIB.exists():
            subprocess.run(["make", "-s", "-C", str(HERE)], check=True)
        L = C.CDLL(str(LIB))
        P, I64, I32, VP = C.POINTER, C.c_int64, C.c_int32, C.c_void_p
        L.ref_scene_create.argtypes = [P(TrSceneDesc), I32]
        L.ref_scene_create.restype = VP
        L.ref_scene_destroy.argtypes = [VP]
        L.ref_scene_destroy.restype = None
        L.ref_bvh_node_count.argtypes = [VP]
        L.ref_bvh_node_count.restype = I64
        L.ref_bvh_export.argtypes = [VP, VP, P(C.c_int32)]
        L.ref_bvh_export.restype = None
        L.ref_light_pmf.argtypes = [VP, P(C.c_float)]
        L.ref_light_pmf.restype = None
        L.ref_light_pick.argtypes = [VP, P(C.c_float), I64, P(C.c_int32), P(C.c_float)]
        L.ref_light_pick.restype = None
        L.ref_film_splat.argtypes = [C.c_int, C.c_int, C.c_int, C.c_float, C.c_float, P(C.c_float), P(C.c_float), I64,
                                     P(C.c_double)]
        L.ref_film_splat.restype = None
        for name in ("ref_trace_closest_mt", "ref_trace_closest_wt_brute", "ref_trace_closest_wt_bvh"):
            f = getattr(L, name)
            f.argtypes = [VP, P(C.c_float), P(C.c_float), I64, I32]
            f.restype = None
        L.ref_trace_any_wt.argtypes = [VP, P(C.c_float), P(C.c_int32), I64, I32]
        L.ref_trace_any_wt.restype = None
        L.ref_mt_single.argtypes = [P(C.c_float), P(C.c_float), P(C.c_float)]
        L.ref_wt_single.argtypes = [P(C.c_float), P(C.c_float), P(C.c_float)]
        L.ref_bsdf_eval.argtypes = [P(TrMaterial), P(C.c_int32), P(C.c_float), P(C.c_float), P(C.c_float), I64, I32, I32,
                                    P(C.c_double), P(C.c_double)]
        L.ref_bsdf_eval.restype = None
        L.ref_bsdf_sample.argtypes = [P(TrMaterial), P(C.c_int32), P(C.c_float), P(C.c_float), P(C.c_float), I64, I32,
                                      I32, P(C.c_double), P(C.c_double), P(C.c_double)]
        L.ref_bsdf_sample.restype = None
        L.ref_hg_eval.argtypes = [C.c_double, C.c_double]
        L.ref_hg_eval.restype = C.c_double
        L.ref_hg_sample.argtypes = [C.c_float, P(C.c_float), P(C.c_float), I64, P(C.c_float), P(C.c_float)]
        L.ref_hg_sample.restype = None
        L.ref_camera_rays.argtypes = [VP, P(C.c_float), P(C.c_float), I64, P(C.c_float)]
        L.ref_camera_rays.restype = None
        L.ref_shading_frames.argtypes = [VP, P(C.c_int32), P(C.c_float), I64, P(C.c_float)]
        L.ref_shading_frames.restype = None
        L.ref_render.argtypes = [VP, P(TrRenderParams), I32, I32, I32, P(C.c_double), P(C.c_uint64)]
        L.ref_render.restype = C.c_double
        L.ref_env_create.argtypes = [C.c_int, C.c_int, P(C.c_float), P(C.c_float)]
        L.ref_env_create.restype = VP
        L.ref_env_destroy.argtypes = [VP]
        L.ref_env_destroy.restype = None
        L.ref_env_sample.argtypes = [VP, P(C.c_float), I64, P(C.c_float), P(C.c_float), P(C.c_float)]
        L.ref_env_sample.restype = None
        L.ref_env_eval.argtypes = [VP, P(C.c_float), I64, I32, P(C.c_float), P(C.c_float)]
        L.ref_env_eval.restype = None
        L.ref_env_radiance_integral.argtypes = [VP, P(C.c_float)]
        L.ref_env_radiance_integral.restype = None
        L.ref_tex_eval8.argtypes = [C.c_int, C.c_int, P(C.c_uint8), P(C.c_float), I64, I32, P(C.c_float)]
        L.ref_tex_eval8.restype = None
        L.ref_grid_density.argtypes = [C.c_void_p, P(C.c_float), I64, P(C.c_float)]
        L.ref_grid_density.restype = None
        L.ref_grid_albedo_g.argtypes = [C.c_void_p, P(C.c_float), I64, P(C.c_float), P(C.c_float)]
        L.ref_grid_albedo_g.restype = None
        L.ref_grid_stats.argtypes = [C.c_void_p, P(C.c_float), P(C.c_float), I64, C.c_uint64, I32, P(C.c_double)]
        L.ref_grid_stats.restype = None
        _lib = L
    return _lib


def tex_eval8(rgb8: np.ndarray, uv: np.ndarray, quirks: bool) -> np.ndarray:
    """Lookups of an 8-bit texture (h, w, 3) at uv (n, 2)."""
    t = np.ascontiguousarray(rgb8, np.uint8)
    uv = np.ascontiguousarray(uv, np.float32)
    out = np.zeros((uv.shape[0], 3), np.float32)
    lib().ref_tex_eval8(t.shape[1], t.shape[0], t.ctypes.data_as(C.POINTER(C.c_uint8)), fptr(uv), uv.shape[0],
                        1 if quirks else 0, fptr(out))
    return out


def grid_density(medium, pts: np.ndarray) -> np.ndarray:
    """Densities of a heterogeneous medium (a TrMedium with a grid) at world
    points (n, 3)."""
    pts = np.ascontiguousarray(pts, np.float32)
    out = np.zeros(pts.shape[0], np.float32)
    lib().ref_grid_density(C.addressof(medium), fptr(pts), pts.shape[0], fptr(out))
    return out


def grid_albedo_g(medium, pts: np.ndarray):
    """Albedo (n, 3) and asymmetry (n,) of a heterogeneous medium at world
    points (n, 3): its grids, or its constants."""
    pts = np.ascontiguousarray(pts, np.float32)
    a = np.zeros((pts.shape[0], 3), np.float32)
    g = np.zeros(pts.shape[0], np.float32)
    lib().ref_grid_albedo_g(C.addressof(medium), fptr(pts), pts.shape[0], fptr(a), fptr(g))
    return a, g


def grid_stats(medium, a, b, trials: int, seed: int = 1, indirect: bool = False):
    """(mean ratio-tracking transmittance, scatter fraction of delta
    tracking, mean scatter weight) over `trials` segments a -> b."""
    a = np.ascontiguousarray(a, np.float32)
    b = np.ascontiguousarray(b, np.float32)
    out = np.zeros(3, np.float64)
    lib().ref_grid_stats(C.addressof(medium), fptr(a), fptr(b), trials, seed, 1 if indirect else 0,
                         out.ctypes.data_as(C.POINTER(C.c_double)))
    return tuple(out)


class RefEnv:
    """The oracle's image environment light (IBL over an HDR texture)."""

    def __init__(self, rgb: np.ndarray, world_to_light=None):
        self.rgb = np.ascontiguousarray(rgb, np.float32)
        h, w, _ = self.rgb.shape
        rot = np.eye(3, dtype=np.float32) if world_to_light is None else np.asarray(world_to_light, np.float32)
        self.rot = np.ascontiguousarray(rot.reshape(9))
        self.h = lib().ref_env_create(w, h, fptr(self.rgb), fptr(self.rot))

    def __del__(self):
        if getattr(self, "h", None):
            lib().ref_env_destroy(self.h)
            self.h = None

    def sample(self, u: np.ndarray):
        """u: (n, 2) -> (wi (n, 3), pdf (n,), L (n, 3)); pdf 0 = no sample."""
        u = np.ascontiguousarray(u, np.float32)
        n = u.shape[0]
        wi, pdf, L = np.zeros((n, 3), np.float32), np.zeros(n, np.float32), np.zeros((n, 3), np.float32)
        lib().ref_env_sample(self.h, fptr(u), n, fptr(wi), fptr(pdf), fptr(L))
        return wi, pdf, L

    def eval(self, dirs: np.ndarray, quirks: bool = False):
        d = np.ascontiguousarray(dirs, np.float32)
        n = d.shape[0]
        L, pdf = np.zeros((n, 3), np.float32), np.zeros(n, np.float32)
        lib().ref_env_eval(self.h, fptr(d), n, 1 if quirks else 0, fptr(L), fptr(pdf))
        return L, pdf

    def radiance_integral(self) -> np.ndarray:
        out = np.zeros(3, np.float32)
        lib().ref_env_radiance_integral(self.h, fptr(out))
        return out


class RefScene:
    """The oracle's copy of a scene (reference BVH with max_leaf prims per leaf)."""

    def __init__(self, scene, max_leaf: int = 3):
        self.scene = scene
        d = scene.desc()
        self.h = lib().ref_scene_create(C.byref(d), max_leaf)

    def __del__(self):
        if getattr(self, "h", None):
            lib().ref_scene_destroy(self.h)
            self.h = None

    def _rays(self, fn, rays, threads):
        rays = np.ascontiguousarray(rays, np.float32)
        hits = np.zeros((rays.shape[0], 4), np.float32)
        fn(self.h, fptr(rays), fptr(hits), rays.shape[0], threads)
        return hits

    def closest_mt(self, rays, threads=8):
        return self._rays(lib().ref_trace_closest_mt, rays, threads)

    def closest_brute(self, rays, threads=8):
        return self._rays(lib().ref_trace_closest_wt_brute, rays, threads)

    def closest_wt(self, rays, threads=8):
        return self._rays(lib().ref_trace_closest_wt_bvh, rays, threads)

    def any_wt(self, rays, threads=8):
        rays = np.ascontiguousarray(rays, np.float32)
        occ = np.zeros(rays.shape[0], np.int32)
        lib().ref_trace_any_wt(self.h, fptr(rays), iptr(occ), rays.shape[0], threads)
        return occ

    def bvh(self):
        n = lib().ref_bvh_node_count(self.h)
        nodes = np.zeros(n, dtype=np.dtype([("lo", "<f4", 3), ("hi", "<f4", 3), ("offset", "<i4"), ("count", "<u2"),
                                            ("axis", "u1"), ("pad", "u1")]))
        prims = np.zeros(self.scene.num_triangles, np.int32)
        lib().ref_bvh_export(self.h, nodes.ctypes.data_as(C.c_void_p), iptr(prims))
        return nodes, prims

    def light_pmf(self, n_area: int):
        out = np.zeros(n_area + 2, np.float32)
        lib().ref_light_pmf(self.h, fptr(out))
        return out

    def light_pick(self, u):
        """Distribution1D::sample_discrete over the light powers: (index, pmf)."""
        u = np.ascontiguousarray(u, np.float32).reshape(-1)
        idx = np.zeros(u.size, np.int32)
        pmf = np.zeros(u.size, np.float32)
        lib().ref_light_pick(self.h, fptr(u), u.size, iptr(idx), fptr(pmf))
        return idx, pmf

    def camera_rays(self, film_uv, lens_uv):
        film_uv = np.ascontiguousarray(film_uv, np.float32)
        lens_uv = np.ascontiguousarray(lens_uv, np.float32)
        rays = np.zeros((film_uv.shape[0], 8), np.float32)
        lib().ref_camera_rays(self.h, fptr(film_uv), fptr(lens_uv), film_uv.shape[0], fptr(rays))
        return rays

    def shading_frames(self, prims, bary):
        """(n, 3, 3) shading frames (rows x, y, z) at hits of prims with barycentrics bary (n, 2)."""
        prims = np.ascontiguousarray(prims, np.int32)
        bary = np.ascontiguousarray(bary, np.float32)
        out = np.zeros((prims.shape[0], 9), np.float32)
        lib().ref_shading_frames(self.h, prims.ctypes.data_as(C.POINTER(C.c_int32)), fptr(bary), prims.shape[0],
                                 fptr(out))
        return out.reshape(-1, 3, 3)

    def render(self, params: TrRenderParams, quirks: bool, nee: int, threads: int = 0, sums=None):
        if sums is None:
            sums = np.zeros((params.height, params.width, 4), np.float64)
        rays = (C.c_uint64 * 2)()
        ms = lib().ref_render(self.h, C.byref(params), int(quirks), int(nee), threads, dptr(sums), rays)
        return sums, ms, (rays[0], rays[1])


def film_splat(width: int, height: int, filter_type: int, radius: float, alpha: float, pos, L) -> np.ndarray:
    """Tile::add_sample of samples at film positions pos (n, 2, pixel units)
    with radiance L (n, 3), in order, over a width x height film -> (H, W, 4)."""
    pos = np.ascontiguousarray(pos, np.float32).reshape(-1, 2)
    L = np.ascontiguousarray(L, np.float32).reshape(-1, 3)
    sums = np.zeros((height, width, 4), np.float64)
    lib().ref_film_splat(width, height, filter_type, radius, alpha, fptr(pos), fptr(L), pos.shape[0], dptr(sums))
    return sums


def bsdf_eval(mats, ids, wi, wo, ng=None, precision=1, fixed=1):
    n = len(ids)
    arr = (TrMaterial * len(mats))(*mats)
    ids = np.ascontiguousarray(ids, np.int32)
    wi = np.ascontiguousarray(wi, np.float32)
    wo = np.ascontiguousarray(wo, np.float32)
    ngc = np.ascontiguousarray(ng, np.float32) if ng is not None else None
    f = np.zeros((n, 3), np.float64)
    pdf = np.zeros(n, np.float64)
    lib().ref_bsdf_eval(arr, iptr(ids), fptr(wi), fptr(wo), fptr(ngc), n, precision, fixed, dptr(f), dptr(pdf))
    return f, pdf


def bsdf_sample(mats, ids, wo, u, ng=None, precision=1, fixed=1):
    n = len(ids)
    arr = (TrMaterial * len(mats))(*mats)
    ids = np.ascontiguousarray(ids, np.int32)
    wo = np.ascontiguousarray(wo, np.float32)
    u = np.ascontiguousarray(u, np.float32)
    ngc = np.ascontiguousarray(ng, np.float32) if ng is not None else None
    wi = np.zeros((n, 3), np.float64)
    f = np.zeros((n, 3), np.float64)
    pdf = np.zeros(n, np.float64)
    lib().ref_bsdf_sample(arr, iptr(ids), fptr(wo), fptr(ngc), fptr(u), n, precision, fixed, dptr(wi), dptr(f),
                          dptr(pdf))
    return wi, f, pdf


def mt_single(ray, tri):
    out = np.zeros(3, np.float32)
    ok = lib().ref_mt_single(fptr(np.asarray(ray, np.float32)), fptr(np.asarray(tri, np.float32)), fptr(out))
    return bool(ok), out


def wt_single(ray, tri):
    out = np.zeros(3, np.float32)
    ok = lib().ref_wt_single(fptr(np.asarray(ray, np.float32)), fptr(np.asarray(tri, np.float32)), fptr(out))
    return bool(ok), out

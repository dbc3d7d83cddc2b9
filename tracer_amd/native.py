"""Python bindings over the C ABI (include/tracer_hip.h, include/tracer/scenes.h).

The GPU path always runs through libtracer_hip.so; there is no fallback.  If
the library is missing, loading raises NativeLibraryMissing.
"""
from __future__ import annotations

import ctypes as C
import weakref
from dataclasses import dataclass, field
from pathlib import Path

import numpy as np

from . import abi
from .abi import TrMaterial, TrMedium, TrRenderParams, TrRenderStats, TrSceneDesc, dptr, fptr, iptr

LIB_DIR = Path(__file__).resolve().parent / "lib"


class NativeLibraryMissing(RuntimeError):
    pass


class TracerError(RuntimeError):
    def __init__(self, status: int, msg: str):
        super().__init__(f"status {status}: {msg}")
        self.status = status


_hip = None
_scenes = None


def hip_lib() -> C.CDLL:
    global _hip
    if _hip is None:
        p = LIB_DIR / "libtracer_hip.so"
        if not p.exists():
            raise NativeLibraryMissing(f"{p} is not built (run python -m tracer_amd.build)")
        lib = C.CDLL(str(p))
        P, I64, I32, VP = C.POINTER, C.c_int64, C.c_int32, C.c_void_p
        lib.tr_create.argtypes = [P(VP), I32]
        lib.tr_destroy.argtypes = [VP]
        lib.tr_destroy.restype = None
        lib.tr_host_alloc.argtypes = [VP, C.c_int64, C.POINTER(C.c_void_p)]
        lib.tr_host_alloc.restype = C.c_int32
        lib.tr_host_free.argtypes = [VP, VP]
        lib.tr_host_free.restype = C.c_int32
        lib.tr_last_error.argtypes = [VP]
        lib.tr_last_error.restype = C.c_char_p
        lib.tr_version.restype = C.c_char_p
        lib.tr_scene_upload.argtypes = [VP, P(TrSceneDesc)]
        lib.tr_bvh_build.argtypes = [VP, P(abi.TrBvhOpts)]
        lib.tr_bvh_info.argtypes = [VP, P(I64), P(I64), P(I64)]
        lib.tr_render.argtypes = [VP, P(TrRenderParams), P(C.c_double), I32, P(C.c_float), P(TrRenderStats)]
        lib.tr_render_bdpt.argtypes = [VP, P(abi.TrBdptParams), P(C.c_double), I32, P(C.c_float), P(TrRenderStats)]
        lib.tr_trace_closest.argtypes = [VP, P(C.c_float), P(C.c_float), I64]
        lib.tr_trace_any.argtypes = [VP, P(C.c_float), P(C.c_int32), I64]
        lib.tr_trace_closest_device.argtypes = [VP, VP, VP, I64]
        lib.tr_trace_any_device.argtypes = [VP, VP, VP, I64]
        lib.tr_trace_closest_counted.argtypes = [VP, P(C.c_float), P(C.c_float), P(C.c_uint32), P(C.c_uint32), I64]
        lib.tr_stream.argtypes = [VP, P(VP)]
        lib.tr_synchronize.argtypes = [VP]
        lib.tr_bsdf_eval_batch.argtypes = [VP, P(TrMaterial), I32, P(C.c_int32), P(C.c_float), P(C.c_float),
                                           P(C.c_float), I64, P(C.c_float), P(C.c_float)]
        lib.tr_bsdf_sample_batch.argtypes = [VP, P(TrMaterial), I32, P(C.c_int32), P(C.c_float), P(C.c_float),
                                             P(C.c_float), I64, P(C.c_float), P(C.c_float), P(C.c_float)]
        lib.tr_env_sample_batch.argtypes = [VP, P(C.c_float), I64, P(C.c_float), P(C.c_float), P(C.c_float)]
        lib.tr_film_to_ldr.argtypes = [VP, C.c_float, C.c_float, P(C.c_uint8), I64]
        lib.tr_env_eval_batch.argtypes = [VP, P(C.c_float), I64, P(C.c_float), P(C.c_float)]
        lib.tr_medium_density_batch.argtypes = [VP, C.c_int32, P(C.c_float), I64, P(C.c_float)]
        lib.tr_medium_albedo_g_batch.argtypes = [VP, C.c_int32, P(C.c_float), I64, P(C.c_float), P(C.c_float)]
        lib.tr_texture_eval_batch.argtypes = [VP, C.c_int32, P(C.c_float), I64, P(C.c_float)]
        F = P(C.c_float)
        lib.tr_camera_ray_batch.argtypes = [VP, F, F, I64, F]
        lib.tr_shading_frame_batch.argtypes = [VP, P(I32), F, I64, F]
        lib.tr_set_camera.argtypes = [VP, P(abi.TrCamera)]
        lib.tr_light_pmf.argtypes = [VP, F, I64, P(I64)]
        lib.tr_light_pick_batch.argtypes = [VP, F, I64, P(C.c_int32), F]
        lib.tr_hg_sample_batch.argtypes = [VP, C.c_float, F, F, I64, F, F]
        lib.tr_film_accumulate_batch.argtypes = [VP, P(TrRenderParams), F, P(C.c_double)]
        lib.tr_guide_default_config.argtypes = [P(abi.TrGuideConfig)]
        lib.tr_guide_default_config.restype = None
        lib.tr_guide_init.argtypes = [VP, P(abi.TrGuideConfig)]
        lib.tr_guide_num_params.argtypes = [VP, P(I64)]
        lib.tr_guide_get_params.argtypes = [VP, F, I64]
        lib.tr_guide_set_params.argtypes = [VP, F, I64]
        lib.tr_guide_set_box.argtypes = [VP, F, F]
        lib.tr_guide_forward.argtypes = [VP, F, I64, F]
        lib.tr_guide_train_batch.argtypes = [VP, F, I64, I32, F, F]
        lib.tr_guide_mixture_eval.argtypes = [VP, F, F, F, I64, F, F, F]
        lib.tr_guide_last_samples.argtypes = [VP, F, I64, P(I64)]
        lib.tr_guide_mfma_selftest.argtypes = [VP, F, F, F]
        _hip = lib
    return _hip


def scenes_lib() -> C.CDLL:
    global _scenes
    if _scenes is None:
        hip_lib()  # libtracer.so links against it
        p = LIB_DIR / "libtracer.so"
        if not p.exists():
            raise NativeLibraryMissing(f"{p} is not built (run python -m tracer_amd.build)")
        lib = C.CDLL(str(p))
        lib.trs_create.argtypes = [C.c_int32, C.c_uint64, C.c_double, C.POINTER(C.c_char_p)]
        lib.trs_create.restype = C.c_void_p
        lib.trs_load_obj.argtypes = [C.c_char_p, C.POINTER(abi.TrCamera), C.c_int32, C.POINTER(C.c_char_p)]
        lib.trs_load_obj.restype = C.c_void_p
        lib.trs_load_obj_grid.argtypes = [C.c_char_p, C.POINTER(abi.TrCamera), C.POINTER(C.c_float), C.c_int32,
                                          C.c_int32, C.c_int32, C.POINTER(C.c_float), C.POINTER(C.c_float),
                                          C.c_float, C.POINTER(C.c_float), C.POINTER(C.c_float), C.c_int32,
                                          C.POINTER(C.c_char_p)]
        lib.trs_load_obj_grid.restype = C.c_void_p
        lib.trs_read_hdr.argtypes = [C.c_char_p, C.POINTER(C.c_int32), C.POINTER(C.c_int32), C.POINTER(C.c_float),
                                     C.c_int64, C.POINTER(C.c_char_p)]
        lib.trs_read_hdr.restype = C.c_int32
        lib.trs_read_png.argtypes = [C.c_char_p, C.POINTER(C.c_int32), C.POINTER(C.c_int32), C.POINTER(C.c_uint8),
                                     C.c_int64, C.POINTER(C.c_char_p)]
        lib.trs_read_png.restype = C.c_int32
        lib.trs_write_hdr.argtypes = [C.c_char_p, C.c_int32, C.c_int32, C.POINTER(C.c_float), C.POINTER(C.c_char_p)]
        lib.trs_write_hdr.restype = C.c_int32
        lib.trs_write_png.argtypes = [C.c_char_p, C.c_int32, C.c_int32, C.POINTER(C.c_uint8), C.POINTER(C.c_char_p)]
        lib.trs_write_png.restype = C.c_int32
        lib.trs_postprocess.argtypes = [C.POINTER(C.c_float), C.c_int32, C.c_int32, C.c_float, C.c_float]
        lib.trs_postprocess.restype = C.c_int32
        lib.trs_desc.argtypes = [C.c_void_p]
        lib.trs_desc.restype = C.POINTER(TrSceneDesc)
        lib.trs_render_defaults.argtypes = [C.c_void_p, C.POINTER(TrRenderParams)]
        lib.trs_render_defaults.restype = None
        lib.trs_max_leaf.argtypes = [C.c_void_p]
        lib.trs_destroy.argtypes = [C.c_void_p]
        lib.trs_destroy.restype = None
        _scenes = lib
    return _scenes


def loaded_libraries() -> list[str]:
    return [str(LIB_DIR / n) for n, l in (("libtracer_hip.so", _hip), ("libtracer.so", _scenes)) if l is not None]


@dataclass
class Scene:
    """A flattened scene held in numpy arrays; desc() gives the C view."""
    positions: np.ndarray
    indices: np.ndarray
    material_ids: np.ndarray
    materials: list
    normals: np.ndarray | None = None
    uvs: np.ndarray | None = None
    emission: np.ndarray | None = None
    medium_inside: np.ndarray | None = None
    medium_outside: np.ndarray | None = None
    media: list = field(default_factory=list)
    env: tuple | None = None
    env_image: np.ndarray | None = None   # (h, w, 3) linear RGB: an image environment (needs env set)
    textures: list = field(default_factory=list)   # (h, w, 3) uint8 or float32 arrays; materials name 1 + index
    env_world_to_light: np.ndarray | None = None   # 3x3 rotation
    sun: tuple | None = None          # (direction, irradiance)
    camera: abi.TrCamera | None = None
    defaults: TrRenderParams | None = None
    max_leaf: int = 3

    def __post_init__(self):
        self._keep = []

    @property
    def num_triangles(self) -> int:
        return int(self.indices.size // 3)

    def desc(self) -> TrSceneDesc:
        d = TrSceneDesc()
        self.positions = np.ascontiguousarray(self.positions, np.float32).reshape(-1)
        self.indices = np.ascontiguousarray(self.indices, np.int32).reshape(-1)
        self.material_ids = np.ascontiguousarray(self.material_ids, np.int32).reshape(-1)
        d.num_vertices = self.positions.size // 3
        d.positions = fptr(self.positions)
        if self.normals is not None:
            self.normals = np.ascontiguousarray(self.normals, np.float32).reshape(-1)
        d.normals = fptr(self.normals)
        if self.uvs is not None:
            self.uvs = np.ascontiguousarray(self.uvs, np.float32).reshape(-1)
        d.uvs = fptr(self.uvs)
        d.num_triangles = self.indices.size // 3
        d.indices = iptr(self.indices)
        d.material_ids = iptr(self.material_ids)
        if self.emission is not None:
            self.emission = np.ascontiguousarray(self.emission, np.float32).reshape(-1)
        d.emission = fptr(self.emission)
        for k in ("medium_inside", "medium_outside"):
            v = getattr(self, k)
            if v is not None:
                setattr(self, k, np.ascontiguousarray(v, np.int32).reshape(-1))
        d.medium_inside = iptr(self.medium_inside)
        d.medium_outside = iptr(self.medium_outside)
        mats = (TrMaterial * len(self.materials))(*self.materials)
        d.num_materials = len(self.materials)
        d.materials = C.cast(mats, C.POINTER(TrMaterial))
        media = (TrMedium * max(1, len(self.media)))(*self.media) if self.media else None
        d.num_media = len(self.media)
        d.media = C.cast(media, C.POINTER(TrMedium)) if media is not None else C.POINTER(TrMedium)()
        if self.env is not None:
            d.has_env = 1
            d.env_radiance = abi.f3(*self.env)
            if self.env_image is not None:
                self.env_image = np.ascontiguousarray(self.env_image, np.float32)
                d.env_height, d.env_width = self.env_image.shape[:2]
                d.env_pixels = fptr(self.env_image)
                rot = np.eye(3) if self.env_world_to_light is None else np.asarray(self.env_world_to_light)
                d.env_world_to_light = (C.c_float * 9)(*map(float, np.asarray(rot, np.float32).reshape(9)))
        if self.sun is not None:
            d.has_sun = 1
            d.sun_direction = abi.f3(*self.sun[0])
            d.sun_irradiance = abi.f3(*self.sun[1])
        d.camera = self.camera if self.camera is not None else abi.TrCamera()
        texs = None
        if self.textures:
            self.textures = [np.ascontiguousarray(t, np.uint8 if t.dtype == np.uint8 else np.float32)
                             for t in self.textures]
            texs = (abi.TrTexture * len(self.textures))()
            for i, t in enumerate(self.textures):
                texs[i].height, texs[i].width = t.shape[:2]
                if t.dtype == np.uint8:
                    texs[i].rgb8 = t.ctypes.data_as(C.POINTER(C.c_uint8))
                else:
                    texs[i].rgbf = fptr(t)
            d.num_textures = len(self.textures)
            d.textures = C.cast(texs, C.POINTER(abi.TrTexture))
        self._keep = [mats, media, texs]
        return d

    def render_params(self, **kw) -> TrRenderParams:
        p = TrRenderParams()
        if self.defaults is not None:
            C.memmove(C.byref(p), C.byref(self.defaults), C.sizeof(p))
        else:
            p.width, p.height, p.spp, p.min_depth, p.max_depth = 64, 64, 1, 3, 8
            p.direct_light_sample_num, p.filter_type, p.filter_radius = 1, abi.TR_FILTER_BOX, 0.5
            p.guide_fraction, p.guide_lr = 0.5, 1e-3
        for k, v in kw.items():
            setattr(p, k, v)
        return p


    def bdpt_params(self, max_camera_vertex_count: int = 10, max_light_vertex_count: int = 10,
                    **kw) -> abi.TrBdptParams:
        """BDPTRendererParams (src/factory/renderer.hpp:39-45) with the scene's film,
        spp, filter and seed."""
        d = self.render_params()
        p = abi.TrBdptParams()
        p.width, p.height, p.spp, p.seed = d.width, d.height, d.spp, d.seed
        p.filter_type, p.filter_radius, p.filter_alpha = d.filter_type, d.filter_radius, d.filter_alpha
        p.max_camera_vertex_count, p.max_light_vertex_count = max_camera_vertex_count, max_light_vertex_count
        for k, v in kw.items():
            setattr(p, k, v)
        return p


def config_scene(config: int, seed: int | None = None, detail: float = 1.0) -> Scene:
    """Synthetic configuration 1..5 built through the C++ factory API."""
    lib = scenes_lib()
    err = C.c_char_p()
    h = lib.trs_create(config, config - 1 if seed is None else seed, detail, C.byref(err))
    if not h:
        raise TracerError(abi.TR_ERR_INVALID, (err.value or b"").decode())
    return _scene_from_handle(lib, h)


def obj_scene(path, camera: abi.TrCamera, outside_fog: bool = True) -> Scene:
    """A Wavefront OBJ scene read by the C++ facade (load_model_from_file,
    create_texture_from_file) and wired as the reference's run() wires one."""
    lib = scenes_lib()
    err = C.c_char_p()
    h = lib.trs_load_obj(str(path).encode(), C.byref(camera), 1 if outside_fog else 0, C.byref(err))
    if not h:
        raise TracerError(abi.TR_ERR_INVALID, (err.value or b"").decode())
    return _scene_from_handle(lib, h)


def obj_grid_scene(path, camera: abi.TrCamera, density, local_to_world, albedo=(0.9, 0.9, 0.9), g=0.0,
                   max_scattering_count: int = 2 ** 31 - 1) -> Scene:
    """obj_scene with a heterogeneous medium outside every surface, built by
    the C++ facade through create_image_texture3d and
    create_heterogeneous_medium; density is indexed [z, y, x] and
    local_to_world (3x4 or 4x4) maps the unit grid cube into the world.
    albedo is an RGB triple or a grey grid, g a number or a grid, both grids
    of the density's shape."""
    lib = scenes_lib()
    err = C.c_char_p()
    grid = np.ascontiguousarray(density, np.float32)
    if grid.ndim != 3:
        raise ValueError("density must be a 3-D array indexed [z, y, x]")
    l2w = np.ascontiguousarray(np.asarray(local_to_world, np.float64)[:3, :4], np.float32)
    alb_grid = g_grid = None
    if np.ndim(albedo) == 3:
        alb_grid = np.ascontiguousarray(albedo, np.float32)
        alb = np.ones(3, np.float32)
    else:
        alb = np.ascontiguousarray(albedo, np.float32)
    if np.ndim(g) == 3:
        g_grid = np.ascontiguousarray(g, np.float32)
        g = 0.0
    for a in (alb_grid, g_grid):
        if a is not None and a.shape != grid.shape:
            raise ValueError("albedo and g grids must have the density's shape")
    nz, ny, nx = grid.shape
    h = lib.trs_load_obj_grid(str(path).encode(), C.byref(camera), fptr(grid), nx, ny, nz, fptr(l2w), fptr(alb),
                              float(g), fptr(alb_grid), fptr(g_grid), int(max_scattering_count), C.byref(err))
    if not h:
        raise TracerError(abi.TR_ERR_INVALID, (err.value or b"").decode())
    return _scene_from_handle(lib, h)


def read_hdr(path) -> np.ndarray:
    """A Radiance .hdr image as (height, width, 3) linear RGB, read by the
    C++ facade (the loader behind create_texture2d_from_file)."""
    lib = scenes_lib()
    err = C.c_char_p()
    w, h = C.c_int32(), C.c_int32()
    p = str(path).encode()
    if lib.trs_read_hdr(p, C.byref(w), C.byref(h), None, 0, C.byref(err)) != 0:
        raise TracerError(abi.TR_ERR_INVALID, (err.value or b"").decode())
    out = np.zeros((h.value, w.value, 3), np.float32)
    if lib.trs_read_hdr(p, C.byref(w), C.byref(h), fptr(out), out.size, C.byref(err)) != 0:
        raise TracerError(abi.TR_ERR_INVALID, (err.value or b"").decode())
    return out


def read_png(path) -> np.ndarray:
    """An 8-bit PNG as (height, width, 3) uint8, rows bottom to top."""
    lib = scenes_lib()
    err = C.c_char_p()
    w, h = C.c_int32(), C.c_int32()
    p = str(path).encode()
    if lib.trs_read_png(p, C.byref(w), C.byref(h), None, 0, C.byref(err)) != 0:
        raise TracerError(abi.TR_ERR_INVALID, (err.value or b"").decode())
    out = np.zeros((h.value, w.value, 3), np.uint8)
    if lib.trs_read_png(p, C.byref(w), C.byref(h), out.ctypes.data_as(C.POINTER(C.c_uint8)), out.size,
                        C.byref(err)) != 0:
        raise TracerError(abi.TR_ERR_INVALID, (err.value or b"").decode())
    return out


def write_hdr(path, rgb: np.ndarray) -> None:
    """write_image_to_hdr of a (height, width, 3) float image."""
    a = np.ascontiguousarray(rgb, np.float32)
    err = C.c_char_p()
    if scenes_lib().trs_write_hdr(str(path).encode(), a.shape[1], a.shape[0], fptr(a), C.byref(err)) != 0:
        raise TracerError(abi.TR_ERR_INVALID, (err.value or b"").decode())


def write_png(path, rgb: np.ndarray) -> None:
    """write_image_to_png of a (height, width, 3) uint8 image."""
    a = np.ascontiguousarray(rgb, np.uint8)
    err = C.c_char_p()
    if scenes_lib().trs_write_png(str(path).encode(), a.shape[1], a.shape[0], a.ctypes.data_as(C.POINTER(C.c_uint8)),
                                  C.byref(err)) != 0:
        raise TracerError(abi.TR_ERR_INVALID, (err.value or b"").decode())


def postprocess(rgb: np.ndarray, aces_exposure: float = -1.0, gamma: float = 0.0) -> np.ndarray:
    """create_aces_tone_mapper(aces_exposure) (when >= 0) then
    create_gamma_corrector(gamma) (when > 0) on a copy of a float image."""
    a = np.array(rgb, np.float32, order="C", copy=True)
    if scenes_lib().trs_postprocess(fptr(a), a.shape[1], a.shape[0], aces_exposure, gamma) != 0:
        raise TracerError(abi.TR_ERR_INVALID, "trs_postprocess: bad arguments")
    return a


def _scene_from_handle(lib, h) -> Scene:
    try:
        d = lib.trs_desc(h).contents
        nv, nt = d.num_vertices, d.num_triangles

        def arr(ptr, n, dt):
            if not ptr:
                return None
            return np.ctypeslib.as_array(ptr, shape=(n,)).astype(dt, copy=True)

        mats = [TrMaterial.from_buffer_copy(d.materials[i]) for i in range(d.num_materials)]
        media = [TrMedium.from_buffer_copy(d.media[i]) for i in range(d.num_media)]
        for m in media:   # copy each density, albedo and g grid out of the handle's memory
            keep = []
            if m.density:
                grid = np.ctypeslib.as_array(m.density, shape=(m.nz * m.ny * m.nx,)).reshape(m.nz, m.ny, m.nx).copy()
                m.density = grid.ctypes.data_as(C.POINTER(C.c_float))
                keep.append(grid)
            if m.albedo_grid:
                a = np.ctypeslib.as_array(m.albedo_grid, shape=(3 * m.anz * m.any * m.anx,)).copy()
                m.albedo_grid = a.ctypes.data_as(C.POINTER(C.c_float))
                keep.append(a)
            if m.g_grid:
                g = np.ctypeslib.as_array(m.g_grid, shape=(m.gnz * m.gny * m.gnx,)).copy()
                m.g_grid = g.ctypes.data_as(C.POINTER(C.c_float))
                keep.append(g)
            m._grid = keep
        defaults = TrRenderParams()
        lib.trs_render_defaults(h, C.byref(defaults))
        textures = []
        for i in range(d.num_textures):
            t = d.textures[i]
            n = t.height * t.width * 3
            if t.rgb8:
                textures.append(np.ctypeslib.as_array(t.rgb8, shape=(n,)).reshape(t.height, t.width, 3).copy())
            else:
                textures.append(np.ctypeslib.as_array(t.rgbf, shape=(n,)).reshape(t.height, t.width, 3).copy())
        s = Scene(
            positions=arr(d.positions, 3 * nv, np.float32), indices=arr(d.indices, 3 * nt, np.int32),
            material_ids=arr(d.material_ids, nt, np.int32), materials=mats,
            normals=arr(d.normals, 3 * nv, np.float32), uvs=arr(d.uvs, 2 * nv, np.float32),
            emission=arr(d.emission, 3 * nt, np.float32), medium_inside=arr(d.medium_inside, nt, np.int32),
            medium_outside=arr(d.medium_outside, nt, np.int32), media=media,
            env=tuple(d.env_radiance) if d.has_env else None,
            sun=(tuple(d.sun_direction), tuple(d.sun_irradiance)) if d.has_sun else None,
            camera=abi.TrCamera.from_buffer_copy(d.camera), defaults=defaults, max_leaf=lib.trs_max_leaf(h),
            textures=textures)
        return s
    finally:
        lib.trs_destroy(h)


class Context:
    """One device context (tr_ctx): a stream, an uploaded scene and its BVH."""

    def __init__(self, device: int = 0):
        self.lib = hip_lib()
        self.h = C.c_void_p()
        st = self.lib.tr_create(C.byref(self.h), device)
        if st != abi.TR_OK:
            raise TracerError(st, f"tr_create failed on device {device}")

    def close(self):
        if self.h:
            self.lib.tr_destroy(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _check(self, st: int):
        if st != abi.TR_OK:
            raise TracerError(st, self.lib.tr_last_error(self.h).decode())

    def load_scene(self, scene: Scene, max_leaf: int | None = None, threads: int = 0):
        d = scene.desc()
        self._check(self.lib.tr_scene_upload(self.h, C.byref(d)))
        o = abi.TrBvhOpts(max_leaf if max_leaf is not None else 4, threads)
        self._check(self.lib.tr_bvh_build(self.h, C.byref(o)))

    def bvh_info(self) -> dict:
        n, nb, tb = C.c_int64(), C.c_int64(), C.c_int64()
        self._check(self.lib.tr_bvh_info(self.h, C.byref(n), C.byref(nb), C.byref(tb)))
        return {"nodes": n.value, "node_bytes": nb.value, "tri_bytes": tb.value}

    def host_array(self, shape, dtype=np.float64) -> np.ndarray:
        """A zero-filled numpy array in page-locked host memory (tr_host_alloc),
        freed when the array is garbage-collected."""
        dt = np.dtype(dtype)
        nbytes = int(np.prod(shape)) * dt.itemsize
        ptr = C.c_void_p()
        self._check(self.lib.tr_host_alloc(self.h, nbytes, C.byref(ptr)))
        buf = (C.c_char * nbytes).from_address(ptr.value)
        arr = np.frombuffer(buf, dtype=dt).reshape(shape)
        arr[...] = 0
        lib, p = self.lib, ptr.value
        # freed without the context, which may be closed before the array goes
        weakref.finalize(buf, lambda: lib.tr_host_free(None, C.c_void_p(p)))
        return arr

    def render(self, params: TrRenderParams, film: np.ndarray | None = None, want_image: bool = True,
               out: np.ndarray | None = None):
        """film: FP64 sums to accumulate onto (updated in place).  out: an
        (H, W, 4) float64 array to receive the sums of a fresh render, reused
        across calls instead of a new array per call."""
        W, H = params.width, params.height
        self._film_shape = (H, W)
        acc = film is not None
        if acc:
            sums = film
        elif out is not None:
            if out.shape != (H, W, 4) or out.dtype != np.float64 or not out.flags.c_contiguous:
                raise ValueError("out must be a C-contiguous (height, width, 4) float64 array")
            sums = out
        else:
            sums = np.zeros((H, W, 4), np.float64)
        img = np.zeros((H, W, 3), np.float32) if want_image else None
        st = TrRenderStats()
        self._check(self.lib.tr_render(self.h, C.byref(params), dptr(sums), 1 if acc else 0,
                                       fptr(img) if img is not None else fptr(None), C.byref(st)))
        return sums, img, st

    def render_bdpt(self, params: abi.TrBdptParams, film: np.ndarray | None = None, want_image: bool = True):
        """Bidirectional path tracing (tr_render_bdpt); film: FP64 sums to
        accumulate onto (updated in place)."""
        W, H = params.width, params.height
        self._film_shape = (H, W)
        acc = film is not None
        sums = film if acc else np.zeros((H, W, 4), np.float64)
        img = np.zeros((H, W, 3), np.float32) if want_image else None
        st = TrRenderStats()
        self._check(self.lib.tr_render_bdpt(self.h, C.byref(params), dptr(sums), 1 if acc else 0,
                                            fptr(img) if img is not None else fptr(None), C.byref(st)))
        return sums, img, st

    def trace_closest(self, rays: np.ndarray) -> np.ndarray:
        rays = np.ascontiguousarray(rays, np.float32)
        n = rays.shape[0]
        hits = np.zeros((n, 4), np.float32)
        self._check(self.lib.tr_trace_closest(self.h, fptr(rays), fptr(hits), n))
        return hits

    def trace_closest_counted(self, rays: np.ndarray):
        rays = np.ascontiguousarray(rays, np.float32)
        n = rays.shape[0]
        hits = np.zeros((n, 4), np.float32)
        nv = np.zeros(n, np.uint32)
        nt = np.zeros(n, np.uint32)
        self._check(self.lib.tr_trace_closest_counted(self.h, fptr(rays), fptr(hits),
                                                      nv.ctypes.data_as(C.POINTER(C.c_uint32)),
                                                      nt.ctypes.data_as(C.POINTER(C.c_uint32)), n))
        return hits, nv, nt

    def trace_any(self, rays: np.ndarray) -> np.ndarray:
        rays = np.ascontiguousarray(rays, np.float32)
        n = rays.shape[0]
        occ = np.zeros(n, np.int32)
        self._check(self.lib.tr_trace_any(self.h, fptr(rays), iptr(occ), n))
        return occ

    def bsdf_eval(self, mats: list, ids: np.ndarray, wi: np.ndarray, wo: np.ndarray, ng: np.ndarray | None = None):
        n = len(ids)
        arr = (TrMaterial * len(mats))(*mats)
        ids = np.ascontiguousarray(ids, np.int32)
        wi = np.ascontiguousarray(wi, np.float32)
        wo = np.ascontiguousarray(wo, np.float32)
        ngc = np.ascontiguousarray(ng, np.float32) if ng is not None else None
        f = np.zeros((n, 3), np.float32)
        pdf = np.zeros(n, np.float32)
        self._check(self.lib.tr_bsdf_eval_batch(self.h, arr, len(mats), iptr(ids), fptr(wi), fptr(wo), fptr(ngc), n,
                                                fptr(f), fptr(pdf)))
        return f, pdf

    def film_to_ldr(self, aces_exposure: float = -1.0, inv_gamma: float = 1.0 / 2.2) -> np.ndarray:
        """The last render's film as (height, width, 3) uint8 (tr_film_to_ldr)."""
        if not getattr(self, "_film_shape", None):
            raise TracerError(abi.TR_ERR_STATE, "film_to_ldr before render")
        h, w = self._film_shape
        out = np.zeros((h, w, 3), np.uint8)
        self._check(self.lib.tr_film_to_ldr(self.h, aces_exposure, inv_gamma, out.ctypes.data_as(C.POINTER(C.c_uint8)),
                                            out.size))
        return out

    def env_sample(self, u: np.ndarray):
        """The loaded scene's environment light: u (n, 2) -> (wi, pdf, L)."""
        u = np.ascontiguousarray(u, np.float32)
        n = u.shape[0]
        wi, pdf, L = np.zeros((n, 3), np.float32), np.zeros(n, np.float32), np.zeros((n, 3), np.float32)
        self._check(self.lib.tr_env_sample_batch(self.h, fptr(u), n, fptr(wi), fptr(pdf), fptr(L)))
        return wi, pdf, L

    def medium_density(self, medium: int, pos: np.ndarray) -> np.ndarray:
        """Density of grid medium `medium` of the loaded scene at world positions (n, 3)."""
        x = np.ascontiguousarray(pos, np.float32).reshape(-1, 3)
        out = np.zeros(x.shape[0], np.float32)
        self._check(self.lib.tr_medium_density_batch(self.h, int(medium), fptr(x), x.shape[0], fptr(out)))
        return out

    def medium_albedo_g(self, medium: int, pos: np.ndarray):
        """Albedo (n, 3) and asymmetry (n,) of grid medium `medium` at world positions (n, 3)."""
        x = np.ascontiguousarray(pos, np.float32).reshape(-1, 3)
        a = np.zeros((x.shape[0], 3), np.float32)
        g = np.zeros(x.shape[0], np.float32)
        self._check(self.lib.tr_medium_albedo_g_batch(self.h, int(medium), fptr(x), x.shape[0], fptr(a), fptr(g)))
        return a, g

    def texture_eval(self, texture: int, uv: np.ndarray) -> np.ndarray:
        """Texture `texture` (1-based) of the loaded scene at uv (n, 2) -> (n, 3)."""
        t = np.ascontiguousarray(uv, np.float32).reshape(-1, 2)
        out = np.zeros((t.shape[0], 3), np.float32)
        self._check(self.lib.tr_texture_eval_batch(self.h, int(texture), fptr(t), t.shape[0], fptr(out)))
        return out

    def env_eval(self, dirs: np.ndarray):
        """Radiance and light pdf of the environment at unit directions (n, 3)."""
        d = np.ascontiguousarray(dirs, np.float32)
        n = d.shape[0]
        L, pdf = np.zeros((n, 3), np.float32), np.zeros(n, np.float32)
        self._check(self.lib.tr_env_eval_batch(self.h, fptr(d), n, fptr(L), fptr(pdf)))
        return L, pdf

    def bsdf_sample(self, mats: list, ids: np.ndarray, wo: np.ndarray, u: np.ndarray, ng: np.ndarray | None = None):
        n = len(ids)
        arr = (TrMaterial * len(mats))(*mats)
        ids = np.ascontiguousarray(ids, np.int32)
        wo = np.ascontiguousarray(wo, np.float32)
        u = np.ascontiguousarray(u, np.float32)
        ngc = np.ascontiguousarray(ng, np.float32) if ng is not None else None
        wi = np.zeros((n, 3), np.float32)
        f = np.zeros((n, 3), np.float32)
        pdf = np.zeros(n, np.float32)
        self._check(self.lib.tr_bsdf_sample_batch(self.h, arr, len(mats), iptr(ids), fptr(wo), fptr(ngc), fptr(u), n,
                                                  fptr(wi), fptr(f), fptr(pdf)))
        return wi, f, pdf


    # ---- test batches through the renderer's device code ----------------
    def set_camera(self, camera: abi.TrCamera):
        """Scene::set_camera for the uploaded scene (tr_set_camera)."""
        self._check(self.lib.tr_set_camera(self.h, C.byref(camera)))

    def camera_rays(self, film_uv: np.ndarray, lens_uv: np.ndarray) -> np.ndarray:
        """Camera rays of the loaded scene: film_uv, lens_uv (n, 2) -> (n, 8)."""
        f = np.ascontiguousarray(film_uv, np.float32).reshape(-1, 2)
        l = np.ascontiguousarray(lens_uv, np.float32).reshape(-1, 2)
        out = np.zeros((f.shape[0], 8), np.float32)
        self._check(self.lib.tr_camera_ray_batch(self.h, fptr(f), fptr(l), f.shape[0], fptr(out)))
        return out

    def shading_frames(self, prims: np.ndarray, bary: np.ndarray) -> np.ndarray:
        """Device shading frames at hits of prims (scene order) with barycentrics bary (n, 2) -> (n, 3, 3), rows x, y, z."""
        t = np.ascontiguousarray(prims, np.int32).reshape(-1)
        b = np.ascontiguousarray(bary, np.float32).reshape(-1, 2)
        out = np.zeros((t.shape[0], 9), np.float32)
        self._check(self.lib.tr_shading_frame_batch(self.h, t.ctypes.data_as(C.POINTER(C.c_int32)), fptr(b), t.shape[0],
                                                    fptr(out)))
        return out.reshape(-1, 3, 3)

    def light_pmf(self) -> np.ndarray:
        """Light-selection pmf: emissive triangles in scene order, environment, sun."""
        n = C.c_int64()
        self._check(self.lib.tr_light_pmf(self.h, fptr(None), 0, C.byref(n)))
        out = np.zeros(n.value, np.float32)
        self._check(self.lib.tr_light_pmf(self.h, fptr(out), out.size, C.byref(n)))
        return out

    def light_pick(self, u: np.ndarray):
        """Light index (numbered as light_pmf) and pmf picked by uniforms u (n,)."""
        uu = np.ascontiguousarray(u, np.float32).reshape(-1)
        idx = np.zeros(uu.size, np.int32)
        pmf = np.zeros(uu.size, np.float32)
        self._check(self.lib.tr_light_pick_batch(self.h, fptr(uu), uu.size, iptr(idx), fptr(pmf)))
        return idx, pmf

    def hg_sample(self, g: float, wo: np.ndarray, u: np.ndarray):
        """Henyey-Greenstein samples around wo (n, 3) with uniforms u (n, 2) -> (wi, pdf)."""
        w = np.ascontiguousarray(wo, np.float32).reshape(-1, 3)
        uu = np.ascontiguousarray(u, np.float32).reshape(-1, 2)
        wi = np.zeros_like(w)
        pdf = np.zeros(w.shape[0], np.float32)
        self._check(self.lib.tr_hg_sample_batch(self.h, float(g), fptr(w), fptr(uu), w.shape[0], fptr(wi), fptr(pdf)))
        return wi, pdf

    def film_accumulate(self, params: TrRenderParams, radiance: np.ndarray) -> np.ndarray:
        """Film sums (H, W, 4) of given sample radiance (spp, H, W, 3) placed where tr_render places them."""
        L = np.ascontiguousarray(radiance, np.float32)
        if L.shape != (params.spp, params.height, params.width, 3):
            raise ValueError("radiance must be (spp, height, width, 3)")
        sums = np.zeros((params.height, params.width, 4), np.float64)
        self._check(self.lib.tr_film_accumulate_batch(self.h, C.byref(params), fptr(L), dptr(sums)))
        return sums

    # ---- learned path guiding --------------------------------------------
    def guide_config(self, **kw) -> abi.TrGuideConfig:
        cfg = abi.TrGuideConfig()
        self.lib.tr_guide_default_config(C.byref(cfg))
        for k, v in kw.items():
            setattr(cfg, k, v)
        return cfg

    def guide_init(self, **kw):
        cfg = self.guide_config(**kw)
        self._check(self.lib.tr_guide_init(self.h, C.byref(cfg)))

    def guide_num_params(self) -> int:
        n = C.c_int64()
        self._check(self.lib.tr_guide_num_params(self.h, C.byref(n)))
        return n.value

    def guide_params(self) -> np.ndarray:
        p = np.zeros(self.guide_num_params(), np.float32)
        self._check(self.lib.tr_guide_get_params(self.h, fptr(p), p.size))
        return p

    def guide_set_params(self, p: np.ndarray):
        p = np.ascontiguousarray(p, np.float32)
        self._check(self.lib.tr_guide_set_params(self.h, fptr(p), p.size))

    def guide_set_box(self, lo, hi):
        lo = np.ascontiguousarray(lo, np.float32)
        hi = np.ascontiguousarray(hi, np.float32)
        self._check(self.lib.tr_guide_set_box(self.h, fptr(lo), fptr(hi)))

    def guide_forward(self, points: np.ndarray) -> np.ndarray:
        """points: (n, 9) = (x, n, wo) -> (n, 48) head outputs."""
        pts = np.ascontiguousarray(points, np.float32)
        out = np.zeros((pts.shape[0], 48), np.float32)
        self._check(self.lib.tr_guide_forward(self.h, fptr(pts), pts.shape[0], fptr(out)))
        return out

    def guide_train_batch(self, records: np.ndarray, update: bool = False, normalize: bool = False):
        """records: (n, 16) = (x, n, wo, wi, weight, 3 unused) -> (loss, gradient)."""
        rec = np.ascontiguousarray(records, np.float32)
        grad = np.zeros(self.guide_num_params(), np.float32)
        loss = C.c_float()
        flags = (1 if update else 0) | (2 if normalize else 0)
        self._check(self.lib.tr_guide_train_batch(self.h, fptr(rec), rec.shape[0], flags, C.byref(loss), fptr(grad)))
        return loss.value, grad

    def guide_mixture(self, raw: np.ndarray, dirs: np.ndarray, u: np.ndarray):
        """Device mixture head: (samples, pdf at samples, pdf at dirs)."""
        raw = np.ascontiguousarray(raw, np.float32)
        dirs = np.ascontiguousarray(dirs, np.float32)
        u = np.ascontiguousarray(u, np.float32)
        n = raw.shape[0]
        smp = np.zeros((n, 3), np.float32)
        ps = np.zeros(n, np.float32)
        pd = np.zeros(n, np.float32)
        self._check(self.lib.tr_guide_mixture_eval(self.h, fptr(raw), fptr(dirs), fptr(u), n, fptr(smp), fptr(ps),
                                                   fptr(pd)))
        return smp, ps, pd

    def guide_last_samples(self, cap: int = 1 << 20) -> np.ndarray:
        """Weighted training samples of the latest training pass, (n, 16)."""
        n = C.c_int64()
        self._check(self.lib.tr_guide_last_samples(self.h, fptr(None), 0, C.byref(n)))
        k = min(cap, n.value)
        out = np.zeros((k, 16), np.float32)
        if k:
            self._check(self.lib.tr_guide_last_samples(self.h, fptr(out), k, C.byref(n)))
        return out

    def mfma_selftest(self, A: np.ndarray, B: np.ndarray) -> np.ndarray:
        A = np.ascontiguousarray(A, np.float32)
        B = np.ascontiguousarray(B, np.float32)
        D = np.zeros((16, 16), np.float32)
        self._check(self.lib.tr_guide_mfma_selftest(self.h, fptr(A), fptr(B), fptr(D)))
        return D


def gpu_available() -> bool:
    try:
        import torch  # noqa: F401  (device discovery only)
        return torch.cuda.device_count() > 0
    except Exception:
        return False


def version() -> str:
    return hip_lib().tr_version().decode()


__all__ = ["Context", "Scene", "config_scene", "hip_lib", "scenes_lib", "NativeLibraryMissing", "TracerError",
           "gpu_available", "version", "loaded_libraries"]


"""ctypes mirror of the plain-data types in include/tracer_hip.h.

Pure layout declarations (no library is loaded here); the product bindings in
tracer_amd.native and the test harness both use them.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

TR_OK, TR_ERR_INVALID, TR_ERR_HIP, TR_ERR_OOM, TR_ERR_STATE = range(5)
TR_MAT_LAMBERT, TR_MAT_DISNEY, TR_MAT_PHONG, TR_MAT_GLASS = 0, 1, 2, 3
TR_FILTER_BOX, TR_FILTER_GAUSSIAN = 0, 1
TR_TR_ANALYTIC, TR_TR_RATIO = 0, 1

f3 = C.c_float * 3


class TrMaterial(C.Structure):
    _fields_ = [
        ("type", C.c_int32),
        ("base_color", f3),
        ("metallic", C.c_float), ("roughness", C.c_float), ("transmission", C.c_float),
        ("transmission_roughness", C.c_float), ("ior", C.c_float),
        ("specular_scale", f3),
        ("specular_tint", C.c_float), ("anisotropic", C.c_float), ("sheen", C.c_float),
        ("sheen_tint", C.c_float), ("clearcoat", C.c_float), ("clearcoat_gloss", C.c_float),
        ("shininess", C.c_float),
        ("base_color_texture", C.c_int32),
        ("param_texture", C.c_int32 * 12),
        ("normal_texture", C.c_int32),
    ]


class TrTexture(C.Structure):
    _fields_ = [("width", C.c_int32), ("height", C.c_int32), ("rgb8", C.POINTER(C.c_uint8)),
                ("rgbf", C.POINTER(C.c_float))]


class TrMedium(C.Structure):
    _fields_ = [("sigma_a", f3), ("sigma_s", f3), ("g", C.c_float), ("max_scattering_count", C.c_int32),
                ("density", C.POINTER(C.c_float)), ("nx", C.c_int32), ("ny", C.c_int32), ("nz", C.c_int32),
                ("white_for_indirect", C.c_int32), ("local_to_world", C.c_float * 12),
                ("albedo_grid", C.POINTER(C.c_float)), ("anx", C.c_int32), ("any", C.c_int32), ("anz", C.c_int32),
                ("g_grid", C.POINTER(C.c_float)), ("gnx", C.c_int32), ("gny", C.c_int32), ("gnz", C.c_int32)]


def make_grid_medium(density, albedo=(1.0, 1.0, 1.0), g=0.0, local_to_world=None, max_scattering_count=2 ** 31 - 1,
                     white_for_indirect=False) -> TrMedium:
    """create_heterogeneous_medium (src/medium/heterogeneous.cpp:101-108) over
    a float density grid indexed [z, y, x]; local_to_world (3x4 or 4x4) maps
    the unit grid cube into the world (identity by default).  albedo is an RGB
    triple or an RGB grid indexed [z, y, x, c]; g a number or a grid indexed
    [z, y, x] (their own sizes, the same unit cube).  The returned structure
    keeps the grids alive."""
    import numpy as np
    grid = np.ascontiguousarray(density, np.float32)
    if grid.ndim != 3:
        raise ValueError("density must be a 3-D array indexed [z, y, x]")
    m = TrMedium()
    keep = [grid]
    if np.ndim(albedo) == 4:
        a = np.ascontiguousarray(albedo, np.float32)
        m.albedo_grid = a.ctypes.data_as(C.POINTER(C.c_float))
        m.anz, m.any, m.anx = a.shape[:3]
        m.sigma_s = f3(1.0, 1.0, 1.0)
        keep.append(a)
    else:
        m.sigma_s = f3(*albedo)
    if np.ndim(g) == 3:
        gg = np.ascontiguousarray(g, np.float32)
        m.g_grid = gg.ctypes.data_as(C.POINTER(C.c_float))
        m.gnz, m.gny, m.gnx = gg.shape
        m.g = 0.0
        keep.append(gg)
    else:
        m.g = g
    m.max_scattering_count = max_scattering_count
    m.density = grid.ctypes.data_as(C.POINTER(C.c_float))
    m.nz, m.ny, m.nx = grid.shape
    m.white_for_indirect = 1 if white_for_indirect else 0
    M = np.eye(4) if local_to_world is None else np.asarray(local_to_world, np.float64)
    m.local_to_world[:] = [float(x) for x in M[:3, :4].reshape(-1)]
    m._grid = keep
    return m


class TrCamera(C.Structure):
    _fields_ = [("pos", f3), ("target", f3), ("up", f3), ("fov_y", C.c_float), ("lens_radius", C.c_float),
                ("focal_distance", C.c_float), ("aspect", C.c_float)]


class TrSceneDesc(C.Structure):
    _fields_ = [
        ("num_vertices", C.c_int64),
        ("positions", C.POINTER(C.c_float)),
        ("normals", C.POINTER(C.c_float)),
        ("uvs", C.POINTER(C.c_float)),
        ("num_triangles", C.c_int64),
        ("indices", C.POINTER(C.c_int32)),
        ("material_ids", C.POINTER(C.c_int32)),
        ("emission", C.POINTER(C.c_float)),
        ("medium_inside", C.POINTER(C.c_int32)),
        ("medium_outside", C.POINTER(C.c_int32)),
        ("num_materials", C.c_int32),
        ("num_media", C.c_int32),
        ("materials", C.POINTER(TrMaterial)),
        ("media", C.POINTER(TrMedium)),
        ("has_env", C.c_int32),
        ("env_radiance", f3),
        ("has_sun", C.c_int32),
        ("sun_direction", f3),
        ("sun_irradiance", f3),
        ("camera", TrCamera),
        ("env_width", C.c_int32), ("env_height", C.c_int32),
        ("env_pixels", C.POINTER(C.c_float)),
        ("env_world_to_light", C.c_float * 9),
        ("num_textures", C.c_int32),
        ("textures", C.POINTER(TrTexture)),
    ]


class TrBvhOpts(C.Structure):
    _fields_ = [("max_leaf_primitives", C.c_int32), ("num_threads", C.c_int32)]


class TrRenderParams(C.Structure):
    _fields_ = [
        ("width", C.c_int32), ("height", C.c_int32), ("spp", C.c_int32),
        ("min_depth", C.c_int32), ("max_depth", C.c_int32), ("direct_light_sample_num", C.c_int32),
        ("filter_type", C.c_int32), ("filter_radius", C.c_float), ("filter_alpha", C.c_float),
        ("shadow_transmittance", C.c_int32),
        ("seed", C.c_uint64),
        ("sample_offset", C.c_int32), ("samples_per_pass", C.c_int32),
        ("guide", C.c_int32), ("guide_train_passes", C.c_int32),
        ("guide_fraction", C.c_float), ("guide_lr", C.c_float),
        ("profile", C.c_int32), ("guide_max_depth", C.c_int32),
        ("guide_train_fraction", C.c_float), ("tail_paths", C.c_int32),
        ("guide_light_grid", C.c_int32),
    ]


class TrBdptParams(C.Structure):
    """tr_bdpt_params (include/tracer_hip.h): BDPTRendererParams plus the film."""
    _fields_ = [
        ("width", C.c_int32), ("height", C.c_int32), ("spp", C.c_int32),
        ("max_camera_vertex_count", C.c_int32), ("max_light_vertex_count", C.c_int32),
        ("filter_type", C.c_int32), ("filter_radius", C.c_float), ("filter_alpha", C.c_float),
        ("seed", C.c_uint64),
        ("sample_offset", C.c_int32), ("samples_per_pass", C.c_int32),
        ("only_s", C.c_int32), ("only_t", C.c_int32), ("profile", C.c_int32), ("reserved", C.c_int32),
    ]


class TrRenderStats(C.Structure):
    _fields_ = [
        ("camera_paths", C.c_uint64), ("extension_rays", C.c_uint64), ("shadow_rays", C.c_uint64),
        ("nonfinite_dropped", C.c_uint64), ("guide_fallbacks", C.c_uint64), ("guide_train_samples", C.c_uint64),
        ("ms_total", C.c_double), ("ms_extend", C.c_double), ("ms_shade", C.c_double), ("ms_shadow", C.c_double),
        ("ms_other", C.c_double), ("ms_guide", C.c_double),
        ("extend_launches", C.c_uint64), ("node_visits", C.c_uint64), ("tri_tests", C.c_uint64),
        ("guide_vertices", C.c_uint64), ("guide_inferences", C.c_uint64), ("ms_guide_train", C.c_double),
        ("tail_extension_rays", C.c_uint64), ("ms_tail", C.c_double),
    ]

    def as_dict(self) -> dict:
        return {k: getattr(self, k) for k, _ in self._fields_}


class TrGuideConfig(C.Structure):
    """tr_guide_config (include/tracer_hip.h)."""
    _fields_ = [
        ("hidden", C.c_int32), ("batch", C.c_int32), ("steps_per_pass", C.c_int32), ("record_capacity", C.c_int32),
        ("lr", C.c_float), ("beta1", C.c_float), ("beta2", C.c_float), ("eps", C.c_float),
        ("weight_clamp", C.c_float), ("reserved0", C.c_int32),
        ("seed", C.c_uint64),
    ]


def fptr(a: np.ndarray | None):
    if a is None:
        return C.POINTER(C.c_float)()
    assert a.dtype == np.float32 and a.flags.c_contiguous
    return a.ctypes.data_as(C.POINTER(C.c_float))


def iptr(a: np.ndarray | None):
    if a is None:
        return C.POINTER(C.c_int32)()
    assert a.dtype == np.int32 and a.flags.c_contiguous
    return a.ctypes.data_as(C.POINTER(C.c_int32))


def dptr(a: np.ndarray | None):
    if a is None:
        return C.POINTER(C.c_double)()
    assert a.dtype == np.float64 and a.flags.c_contiguous
    return a.ctypes.data_as(C.POINTER(C.c_double))


# order of tr_material::param_texture (TR_PARAM_*)
DISNEY_PARAMS = ["metallic", "roughness", "transmission", "transmission_roughness", "ior", "specular_scale",
                 "specular_tint", "anisotropic", "sheen", "sheen_tint", "clearcoat", "clearcoat_gloss"]

MATERIAL_FIELDS = ["type", "base_color", "metallic", "roughness", "transmission", "transmission_roughness", "ior",
                   "specular_scale", "specular_tint", "anisotropic", "sheen", "sheen_tint", "clearcoat",
                   "clearcoat_gloss", "shininess", "base_color_texture"]


def make_material(**kw) -> TrMaterial:
    m = TrMaterial()
    m.type = kw.get("type", TR_MAT_DISNEY)
    for k in ("base_color", "specular_scale"):
        v = kw.get(k, (1.0, 1.0, 1.0) if k == "specular_scale" else (0.5, 0.5, 0.5))
        if np.isscalar(v):
            v = (v, v, v)
        setattr(m, k, f3(*map(float, v)))
    defaults = dict(metallic=0.0, roughness=0.5, transmission=0.0, transmission_roughness=0.5, ior=1.5,
                    specular_tint=0.0, anisotropic=0.0, sheen=0.0, sheen_tint=0.5, clearcoat=0.0, clearcoat_gloss=1.0,
                    shininess=1.0)
    for k, v in defaults.items():
        setattr(m, k, float(kw.get(k, v)))
    m.base_color_texture = int(kw.get("base_color_texture", 0))
    for k, t in (kw.get("param_textures") or {}).items():   # {"roughness": 1-based texture id, ...}
        m.param_texture[DISNEY_PARAMS.index(k)] = int(t)
    m.normal_texture = int(kw.get("normal_texture", 0))
    return m


def material_array(mats: list[TrMaterial]):
    arr = (TrMaterial * len(mats))(*mats)
    return arr

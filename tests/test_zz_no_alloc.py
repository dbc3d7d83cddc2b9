"""tr_trace_closest_device_noalloc (the allocation-free variant of
tr_trace_closest_device) gives the same hits as the allocating entry point,
bit for bit, on two calls in a row, the second with fewer rays."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _rays(scene, n, seed):
    rng = np.random.default_rng(seed)
    P = scene.positions.reshape(-1, 3)
    lo, hi = P.min(0), P.max(0)
    rays = np.zeros((n, 8), np.float32)
    rays[:, :3] = rng.uniform(lo, hi, (n, 3))
    d = rng.normal(size=(n, 3))
    rays[:, 4:7] = d / np.linalg.norm(d, axis=1, keepdims=True)
    rays[:, 7] = 3e38
    return rays


class _DeviceBuffer:
    """Caller-owned device memory from the HIP runtime the library itself uses."""

    def __init__(self, hip, nbytes):
        self.hip = hip
        self.ptr = C.c_void_p()
        assert hip.hipMalloc(C.byref(self.ptr), C.c_size_t(nbytes)) == 0
        self.nbytes = nbytes

    def upload(self, a):
        assert a.nbytes == self.nbytes
        assert self.hip.hipMemcpy(self.ptr, a.ctypes.data_as(C.c_void_p), C.c_size_t(a.nbytes), 1) == 0

    def download(self, a):
        assert self.hip.hipMemcpy(a.ctypes.data_as(C.c_void_p), self.ptr, C.c_size_t(a.nbytes), 2) == 0
        return a

    def free(self):
        if self.ptr:
            self.hip.hipFree(self.ptr)
            self.ptr = C.c_void_p()


def test_noalloc_closest_equals_allocating_entry_point():
    from tracer_amd import native
    lib = native.hip_lib()
    hip = C.CDLL("/opt/rocm/lib/libamdhip64.so.7")   # the runtime libtracer_hip.so links
    hip.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
    hip.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    hip.hipFree.argtypes = [C.c_void_p]
    VP, I64 = C.c_void_p, C.c_int64
    lib.tr_trace_closest_device_noalloc.argtypes = [VP, VP, VP, I64]
    scene = native.config_scene(2, detail=1 / 4)
    ctx = native.Context(0)
    bufs = []
    try:
        ctx.load_scene(scene)
        for n, seed in ((1 << 18, 1), (3001, 2)):
            print(f"noalloc: {n} rays, seed {seed}", flush=True)
            rays = _rays(scene, n, seed)
            d_rays, d_a, d_b = (_DeviceBuffer(hip, n * 32), _DeviceBuffer(hip, n * 16), _DeviceBuffer(hip, n * 16))
            bufs += [d_rays, d_a, d_b]
            d_rays.upload(rays)
            nan = np.full((n, 4), np.nan, np.float32)
            d_a.upload(nan)
            d_b.upload(nan)
            ctx._check(lib.tr_trace_closest_device(ctx.h, d_rays.ptr, d_a.ptr, n))
            ctx._check(lib.tr_synchronize(ctx.h))
            ctx._check(lib.tr_trace_closest_device_noalloc(ctx.h, d_rays.ptr, d_b.ptr, n))
            ctx._check(lib.tr_synchronize(ctx.h))
            ha = d_a.download(np.zeros((n, 4), np.float32))
            hb = d_b.download(np.zeros((n, 4), np.float32))
            assert np.array_equal(ha.view(np.int32), hb.view(np.int32))
            assert (ha[:, 1].view(np.int32) >= 0).mean() > 0.3
        # bad arguments are rejected without touching the GPU
        assert lib.tr_trace_closest_device_noalloc(ctx.h, VP(0), VP(0), 5) != 0
        assert lib.tr_trace_closest_device_noalloc(ctx.h, VP(0), VP(0), -1) != 0
    finally:
        for b in bufs:
            b.free()
        ctx.close()

"""ahip_trace_rays_noalloc against ahip_trace_rays, through the C ABI, on the GPU.

The no-allocation entry point takes caller-owned device arrays and keeps the traversal stack in a
fixed device array; its closest hits must equal ahip_trace_rays' bit for bit (same kernel, same
tree).  Two calls in a row, the second with fewer rays and an output array pre-filled with junk,
so a stale or unwritten record from the first call would show; a triangle scene whose BVH uses
the global stack part and a hair scene (curve BVH) are both covered.
"""
import numpy as np
import pytest

from anacapa_amd import scenes
from anacapa_amd import _abi as A
from rays import random_rays, camera_rays

pytestmark = pytest.mark.gpu


class _DeviceArray:
    """Caller-owned device memory for the test, through the HIP runtime the library itself links
    (torch bundles a runtime of its own, which does not see the GPU once that one holds it)."""

    _hip = None

    @classmethod
    def hip(cls):
        if cls._hip is None:
            import ctypes as C
            h = C.CDLL("libamdhip64.so.7")
            h.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
            h.hipFree.argtypes = [C.c_void_p]
            h.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
            h.hipDeviceSynchronize.argtypes = []
            cls._hip = h
        return cls._hip

    def __init__(self, host: np.ndarray):
        import ctypes as C
        self.nbytes = host.nbytes
        self.p = C.c_void_p()
        assert self.hip().hipMalloc(C.byref(self.p), max(self.nbytes, 4)) == 0
        self.put(host)

    def put(self, host: np.ndarray):
        host = np.ascontiguousarray(host)
        assert host.nbytes == self.nbytes
        assert self.hip().hipMemcpy(self.p, host.ctypes.data, self.nbytes, 1) == 0   # host to device

    def get(self, like: np.ndarray) -> np.ndarray:
        out = np.empty_like(like)
        assert self.hip().hipMemcpy(out.ctypes.data, self.p, self.nbytes, 2) == 0    # device to host
        return out

    def free(self):
        if self.p:
            self.hip().hipFree(self.p)
            self.p = None


def _noalloc(gpu, rays):
    rays = np.ascontiguousarray(rays, np.float32)
    junk = np.full(len(rays), 0xFF, np.uint8).repeat(A.HIT_DTYPE.itemsize).view(A.HIT_DTYPE)
    d_rays, d_hits = _DeviceArray(rays), _DeviceArray(junk)   # every record must be written
    try:
        assert _DeviceArray.hip().hipDeviceSynchronize() == 0
        st = gpu.lib.ahip_trace_rays_noalloc(gpu._ctx, d_rays.p, len(rays), d_hits.p)
        assert st == A.AHIP_OK, gpu.last_error()
        return d_hits.get(junk)
    finally:
        d_rays.free()
        d_hits.free()


def _same(a, b):
    for k in ("mesh_id", "prim_id", "tri_index"):
        assert np.array_equal(a[k], b[k]), k
    for k in ("t", "u", "v"):
        assert np.array_equal(a[k].view(np.uint32), b[k].view(np.uint32)), k


@pytest.mark.parametrize("name", ["ggx", "hair"])
def test_noalloc_trace_equals_trace_rays(gpu, name):
    if name == "ggx":
        # 32x32 spheres: a tree deep enough that traversal uses the global part of the stack
        sc = scenes.ggx_spheres(width=128, height=72, spp=1, grid=32, freq=3)
    else:
        sc = scenes.small_hair()
    gpu.prepare(sc)
    info = gpu.bvh_info()
    print(f"[noalloc] {name}: {info}", flush=True)
    first = np.concatenate([camera_rays(gpu.camera, 12000, seed=3), random_rays(sc, 8000, seed=4)])
    second = np.concatenate([camera_rays(gpu.camera, 2000, seed=5), random_rays(sc, 1000, seed=6)])
    for k, rays in enumerate((first, second)):
        got = _noalloc(gpu, rays)
        ref = gpu.trace_rays(rays)
        hit = ref["mesh_id"] != 0xFFFFFFFF
        print(f"[noalloc] {name} call {k}: {len(rays)} rays, {int(hit.sum())} hits", flush=True)
        assert hit.sum() > len(rays) // 10
        if name == "hair":
            assert ((ref["tri_index"] & 0x80000000) != 0).sum() > 0   # curve hits present
        _same(got, ref)


def test_noalloc_argument_errors(gpu):
    sc = scenes.triangle_soup(n=200)
    gpu.prepare(sc)
    buf = _DeviceArray(np.zeros(64, np.float32))
    lib, ctx = gpu.lib, gpu._ctx
    try:
        assert lib.ahip_trace_rays_noalloc(ctx, None, 4, buf.p) == A.AHIP_E_INVALID_ARG
        assert lib.ahip_trace_rays_noalloc(ctx, buf.p, 4, None) == A.AHIP_E_INVALID_ARG
        assert lib.ahip_trace_rays_noalloc(ctx, buf.p, 0, buf.p) == A.AHIP_OK
        st = lib.ahip_trace_rays_noalloc(ctx, buf.p, (1 << 30) + 1, buf.p)
        assert st == A.AHIP_E_INVALID_ARG
        assert b"AHIP_NOALLOC_MAX_RAYS" in lib.ahip_last_error(ctx)
    finally:
        buf.free()

"""ctypes binding of oracle/_ref/libadapter_harness.so: the reference-side adapter
(anacapa_amd/adapter/HipPathIntegrator) driven by the reference's own render loop (RenderSession's
tile loop on its ThreadPool, HaltonSampler clones, TileBuffers), next to the reference
PathIntegrator on the same SceneView.  Test infrastructure only."""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

from anacapa_amd import _abi as A

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HARNESS_SO = os.path.join(ROOT, "oracle", "_ref", "libadapter_harness.so")


def available() -> bool:
    return os.path.exists(HARNESS_SO)


class AdapterHarness:
    def __init__(self, scene, cam: A.Camera, max_depth: int, integrator: int = 0, clamp: float = 10.0):
        L = C.CDLL(HARNESS_SO)
        L.ah_create.restype = C.c_void_p
        L.ah_create.argtypes = [C.POINTER(A.SceneDesc), C.POINTER(A.Camera), C.c_uint32, C.c_uint32, C.c_float]
        L.ah_destroy.argtypes = [C.c_void_p]
        L.ah_status.argtypes = [C.c_void_p, C.POINTER(C.c_int), C.POINTER(C.c_int), C.c_char_p, C.c_size_t,
                                C.c_char_p, C.c_size_t]
        L.ah_render.argtypes = [C.c_void_p, C.c_int, C.c_uint32, C.c_uint32, C.c_uint64, C.c_uint32, C.c_uint32,
                                C.c_float, C.c_void_p, C.c_void_p, C.c_void_p]
        L.ah_film_merges.restype = C.c_int
        self.lib, self.cam, self.max_depth, self.clamp = L, cam, max_depth, clamp
        self.desc, self._keep = scene.desc()
        self.h = L.ah_create(C.byref(self.desc), C.byref(cam), max_depth, integrator, clamp)
        if not self.h:
            raise RuntimeError("the reference cannot represent this scene")

    def __del__(self):
        if getattr(self, "h", None):
            self.lib.ah_destroy(self.h)
            self.h = None

    def status(self):
        v, p = C.c_int(), C.c_int()
        name, err = C.create_string_buffer(256), C.create_string_buffer(512)
        self.lib.ah_status(self.h, C.byref(v), C.byref(p), name, 256, err, 512)
        return {"valid": bool(v.value), "prepared": bool(p.value), "name": name.value.decode(),
                "error": err.value.decode()}

    def render(self, which: str, spp: int, threads: int = 8, tile: int = 64, seed: int = 0):
        """The frame through the reference tile loop with the adapter ("adapter"), the reference
        PathIntegrator ("reference") or the reference BDPTIntegrator ("reference_bdpt", with the
        harness's clamp): (accum [h, w, 5], albedo sums [h, w, 3], normal sums [h, w, 3])."""
        h, w = self.cam.height, self.cam.width
        out = np.zeros((h, w, 5), np.float32)
        alb, nrm = np.zeros((h, w, 3), np.float32), np.zeros((h, w, 3), np.float32)
        which_i = {"adapter": 0, "reference": 1, "reference_bdpt": 2}[which]
        self.lib.ah_render(self.h, which_i, spp, self.max_depth, seed, threads, tile, self.clamp,
                           A.ptr(out), A.ptr(alb), A.ptr(nrm))
        return out, alb, nrm

    def film_merges(self) -> int:
        return self.lib.ah_film_merges()

"""Python mirror of the reference's GPU backend interface, over the C ABI.

``HipPathIntegrator`` follows CudaPathIntegrator (src/gpu/cuda/CudaPathIntegrator.h:21-52)
and IIntegrator (include/anacapa/integrator/IIntegrator.h:172-198): construct, ``is_valid``,
``prepare(scene)``, ``clear_accum``, ``render_frame`` (adds into a persistent device buffer),
``render_tile`` and the acceleration-structure batch calls (IAccelerationStructure::traceBatch /
occludedBatch).  Errors raise ``BackendError`` carrying the ABI status and message; the
renderer-side fallback to the CPU integrator is exercised in the tests through
``is_valid`` / ``render_frame`` failures exactly as RenderSession does (RenderSession.cpp:441-464,
546-547).
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _abi as A


class BackendError(RuntimeError):
    def __init__(self, status: int, message: str):
        super().__init__(f"ahip status {status}: {message}")
        self.status = status


class HipPathIntegrator:
    def __init__(self, device: int = 0, lib: C.CDLL | None = None):
        self.lib = lib or A.load_library()
        self._ctx = C.c_void_p()
        self.create_status = self.lib.ahip_create(device, C.byref(self._ctx))
        self._keep = None
        self.scene = None
        self.width = self.height = 0

    # -------------------------------------------------------------- lifetime
    def close(self):
        if self._ctx:
            self.lib.ahip_destroy(self._ctx)
            self._ctx = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def is_valid(self) -> bool:
        return bool(self._ctx) and bool(self.lib.ahip_is_valid(self._ctx))

    def name(self) -> str:
        return self.lib.ahip_device_name(self._ctx).decode()

    def backend_type(self) -> str:
        return self.lib.ahip_backend_type().decode()

    def last_error(self) -> str:
        return self.lib.ahip_last_error(self._ctx).decode()

    def _check(self, status: int):
        if status != A.AHIP_OK:
            raise BackendError(status, self.last_error())

    # -------------------------------------------------------------- scene / settings
    def prepare(self, scene, camera: A.Camera | None = None, max_depth=None, min_depth=None, seed=None,
                flags: int = 0, curve_depth: int = 0, integrator: int = 0, firefly_clamp: float = 10.0):
        desc, keep = scene.desc()
        self._check(self.lib.ahip_prepare(self._ctx, C.byref(desc)))
        self._keep = keep
        self.scene = scene
        if camera is None:
            camera = self.pinhole(scene.cam_from, scene.cam_at, scene.cam_up, scene.vfov, scene.width, scene.height)
        self.set_camera(camera)
        self.set_params(scene.max_depth if max_depth is None else max_depth,
                        scene.min_depth if min_depth is None else min_depth,
                        scene.seed if seed is None else seed, flags, curve_depth, integrator, firefly_clamp)

    def pinhole(self, frm, at, up, vfov, w, h) -> A.Camera:
        cam = A.Camera()
        f, a, u = (np.ascontiguousarray(v, np.float32) for v in (frm, at, up))
        self._check(self.lib.ahip_camera_pinhole(f.ctypes.data_as(A.F32P), a.ctypes.data_as(A.F32P),
                                                 u.ctypes.data_as(A.F32P), vfov, w, h, C.byref(cam)))
        return cam

    def set_camera(self, cam: A.Camera):
        self._check(self.lib.ahip_set_camera(self._ctx, C.byref(cam)))
        self.camera = cam
        self.width, self.height = cam.width, cam.height

    def set_params(self, max_depth=8, min_depth=2, seed=0, flags=0, curve_depth=0, integrator=0, firefly_clamp=10.0):
        """integrator: A.AHIP_INTEGRATOR_PATH or A.AHIP_INTEGRATOR_BDPT; firefly_clamp: BDPT's largest
        luminance of one s, t >= 2 contribution (0 = off; the reference's RenderSettings default 10)."""
        p = A.RenderParams(max_depth, min_depth, seed, flags, curve_depth, integrator, firefly_clamp)
        self._check(self.lib.ahip_set_params(self._ctx, C.byref(p)))

    def set_launch(self, max_paths=0, tile=0):
        self._check(self.lib.ahip_set_launch(self._ctx, max_paths, tile))

    def bvh_info(self):
        n, t, d = C.c_uint32(), C.c_uint32(), C.c_uint32()
        self._check(self.lib.ahip_bvh_info(self._ctx, C.byref(n), C.byref(t), C.byref(d)))
        return n.value, t.value, d.value

    def bvh_dump(self):
        """Reference-layout binary BVH (8 words per node) and the reference prim index list."""
        nodes_flat, ntris, _ = self.bvh_info()
        # the binary tree has 2*interior+1 nodes when the root is interior, else one leaf
        nbin = 2 * nodes_flat + 1 if nodes_flat else (1 if ntris else 0)
        nodes = np.zeros((nbin, 8), np.uint32)
        prims = np.zeros(2 * ntris, np.uint32)
        self._check(self.lib.ahip_bvh_dump(self._ctx, A.ptr(nodes), A.ptr(prims)))
        return nodes, prims

    # -------------------------------------------------------------- rendering
    def clear_accum(self):
        self._check(self.lib.ahip_clear_accum(self._ctx))

    def render_frame(self, sample_start: int, sample_count: int, width=None, height=None):
        self._check(self.lib.ahip_render_frame(self._ctx, width or self.width, height or self.height,
                                               sample_start, sample_count))

    def render_adaptive(self, total_spp: int, tile: int = 64, base_spp: int = 0):
        """RenderSession's adaptive two-pass render on the device (base pass, then the remaining
        samples by tile variance) into the accumulation buffer; returns (extra spp per tile,
        tile scores), both [tilesY, tilesX]."""
        ty, tx = (-(-self.height // tile), -(-self.width // tile)) if tile > 0 else (0, 0)
        extra = np.zeros((ty, tx), np.uint32)
        var = np.zeros((ty, tx), np.float32)
        self._check(self.lib.ahip_render_adaptive(self._ctx, self.width, self.height, max(tile, 0), total_spp,
                                                  base_spp, A.ptr(extra) if extra.size else None,
                                                  A.ptr(var) if var.size else None))
        return extra, var

    def resolve(self, exposure=0.0, srgb8=True, rgb=True, variance=False, aovs=False):
        """Device resolve of the accumulation: dict of the requested outputs (srgb8 [h,w,3] u8,
        rgb [h,w,3], variance [h,w], albedo / normal [h,w,3])."""
        h, w = self.height, self.width
        out = {}
        if srgb8:
            out["srgb8"] = np.zeros((h, w, 3), np.uint8)
        if rgb:
            out["rgb"] = np.zeros((h, w, 3), np.float32)
        if variance:
            out["variance"] = np.zeros((h, w), np.float32)
        if aovs:
            out["albedo"] = np.zeros((h, w, 3), np.float32)
            out["normal"] = np.zeros((h, w, 3), np.float32)
        self._check(self.lib.ahip_resolve(self._ctx, exposure, A.ptr(out.get("srgb8")), A.ptr(out.get("rgb")),
                                          A.ptr(out.get("variance")), A.ptr(out.get("albedo")),
                                          A.ptr(out.get("normal"))))
        return out

    def download_accum(self) -> np.ndarray:
        out = np.zeros((self.height, self.width, 5), np.float32)
        self._check(self.lib.ahip_download_accum(self._ctx, A.ptr(out)))
        return out

    def render_tile(self, x0, y0, w, h, sample_start, sample_count) -> np.ndarray:
        w = min(w, self.width - x0)
        h = min(h, self.height - y0)
        out = np.zeros((h, w, 5), np.float32)
        self._check(self.lib.ahip_render_tile(self._ctx, x0, y0, w, h, sample_start, sample_count, A.ptr(out)))
        return out

    def render_tile_aovs(self, x0, y0, w, h, sample_start, sample_count):
        """render_tile plus the tile's denoising AOVs as PathIntegrator::renderTile forms them:
        (accum [h, w, 5], albedo [h, w, 3] mean first-hit reflectance, normal [h, w, 3] mean
        first-hit normal scaled to unit length)."""
        w = min(w, self.width - x0)
        h = min(h, self.height - y0)
        out = np.zeros((h, w, 5), np.float32)
        alb = np.zeros((h, w, 3), np.float32)
        nrm = np.zeros((h, w, 3), np.float32)
        self._check(self.lib.ahip_render_tile_aovs(self._ctx, x0, y0, w, h, sample_start, sample_count,
                                                   A.ptr(out), A.ptr(alb), A.ptr(nrm)))
        return out, alb, nrm

    def set_aovs(self, enable: bool = True):
        """Accumulate first-hit reflectance and normal sums in render_frame (cleared by clear_accum)."""
        self._check(self.lib.ahip_set_aovs(self._ctx, 1 if enable else 0))

    def download_aovs(self) -> np.ndarray:
        """[h, w, 6]: sums of first-hit reflectance rgb and normal xyz over the accumulated samples."""
        out = np.zeros((self.height, self.width, 6), np.float32)
        self._check(self.lib.ahip_download_aovs(self._ctx, A.ptr(out)))
        return out

    def render_batch(self, cameras, sample_starts, sample_counts, max_depth=0):
        """Render many views of the prepared scene in one batched call (shared passes, one captured
        HIP graph replayed per pass); returns one [h, w, 5] accumulation array per view."""
        outs, views = [], []
        for cam, s0, ns in zip(cameras, sample_starts, sample_counts):
            out = np.zeros((cam.height, cam.width, 5), np.float32)
            outs.append(out)
            views.append(A.View(cam, s0, ns, max_depth, 0, out.ctypes.data_as(A.F32P)))
        arr = (A.View * len(views))(*views)
        self._check(self.lib.ahip_render_batch(self._ctx, arr, len(views)))
        return outs

    # -------------------------------------------------------------- differentiation
    def set_eval_params(self, tri_albedo=None, texture=0, texels=None):
        """Evaluation parameters of the detached replay estimator (None = the sampling values)."""
        a = None if tri_albedo is None else np.ascontiguousarray(tri_albedo, np.float32).reshape(-1)
        t = None if texels is None else np.ascontiguousarray(texels, np.float32).reshape(-1)
        self._check(self.lib.ahip_set_eval_params(self._ctx, A.ptr(a), texture, A.ptr(t)))

    def replay(self, sample_start, sample_count) -> np.ndarray:
        """Image of the detached replay estimator over the current camera's film, [h, w, 5]."""
        out = np.zeros((self.height, self.width, 5), np.float32)
        self._check(self.lib.ahip_replay(self._ctx, sample_start, sample_count, A.ptr(out)))
        return out

    def adjoint(self, sample_start, sample_count, dl_dimage, texture_shape=None):
        """Gradients (d_albedo [tris, 3], d_texture [h, w] or None) of sum(dl_dimage * I), I the
        per-pixel mean radiance of the replay estimator, at the current evaluation parameters."""
        d = np.ascontiguousarray(dl_dimage, np.float32)
        assert d.shape == (self.height, self.width, 3)
        da = np.zeros((self.scene.num_triangles, 3), np.float32)
        dt = np.zeros(texture_shape, np.float32) if texture_shape is not None else None
        self._check(self.lib.ahip_adjoint(self._ctx, self.width, self.height, sample_start, sample_count,
                                          A.ptr(d), A.ptr(da), A.ptr(dt)))
        return da, dt

    # -------------------------------------------------------------- accel / sampler boundary
    def trace_rays(self, rays: np.ndarray) -> np.ndarray:
        rays = np.ascontiguousarray(rays, np.float32).reshape(-1, 8)
        hits = np.zeros(len(rays), A.HIT_DTYPE)
        self._check(self.lib.ahip_trace_rays(self._ctx, A.ptr(rays), len(rays), A.ptr(hits)))
        return hits

    def occluded_rays(self, rays: np.ndarray) -> np.ndarray:
        rays = np.ascontiguousarray(rays, np.float32).reshape(-1, 8)
        out = np.zeros(len(rays), np.uint8)
        self._check(self.lib.ahip_occluded_rays(self._ctx, A.ptr(rays), len(rays), A.ptr(out)))
        return out

    def halton(self, seed, px, py, idx, dim) -> np.ndarray:
        px, py, idx, dim = (np.ascontiguousarray(a, np.uint32) for a in (px, py, idx, dim))
        out = np.zeros(len(px), np.float32)
        self._check(self.lib.ahip_halton_eval(self._ctx, seed, A.ptr(px), A.ptr(py), A.ptr(idx), A.ptr(dim),
                                              len(px), A.ptr(out)))
        return out

    def material_probe(self, mat: A.Material, rows: np.ndarray, flags: int = 0) -> np.ndarray:
        """IMaterial::evaluate / sample of one material on the device (ahip_material_probe)."""
        rows = np.ascontiguousarray(rows, np.float32).reshape(-1, 19)
        out = np.zeros((len(rows), 16), np.float32)
        self._check(self.lib.ahip_material_probe(self._ctx, C.byref(mat), flags, len(rows), A.ptr(rows),
                                                 A.ptr(out)))
        return out

    # -------------------------------------------------------------- measurement / testing
    def stats(self) -> dict:
        s = A.Stats()
        self._check(self.lib.ahip_get_stats(self._ctx, C.byref(s)))
        return {k: getattr(s, k) for k, _ in A.Stats._fields_}

    def bounce_stats(self):
        """Per bounce since reset_stats: rays [n, 3] (continuation, shadow, MIS) and ms [n]."""
        n = C.c_uint32()
        self._check(self.lib.ahip_get_bounce_stats(self._ctx, 0, None, None, C.byref(n)))
        rays = np.zeros((n.value, 3), np.uint64)
        ms = np.zeros(n.value, np.float64)
        self._check(self.lib.ahip_get_bounce_stats(self._ctx, n.value, A.ptr(rays), A.ptr(ms), C.byref(n)))
        return rays, ms

    def reset_stats(self):
        self._check(self.lib.ahip_reset_stats(self._ctx))

    def count_traffic(self, sample_start, sample_count):
        self._check(self.lib.ahip_count_traffic(self._ctx, self.width, self.height, sample_start, sample_count))
        return self.stats()

    def fail_next_launch(self):
        self._check(self.lib.ahip_fail_next_launch(self._ctx))


def resolve(accum: np.ndarray) -> np.ndarray:
    """Film pixel value = pre-weighted sum / weight (PixelAccumulator::resolve, Film.h:46-55)."""
    w = accum[..., 3:4]
    return np.where(w > 0, accum[..., :3] / np.maximum(w, 1e-30), 0.0)


def write_png(path: str, srgb8: np.ndarray, lib: C.CDLL | None = None):
    """8-bit RGB PNG of a resolved film ([h, w, 3] uint8), renamed into place when complete."""
    lib = lib or A.load_library()
    a = np.ascontiguousarray(srgb8, np.uint8)
    st = lib.ahip_write_png(str(path).encode(), a.shape[1], a.shape[0], A.ptr(a))
    if st != A.AHIP_OK:
        raise BackendError(st, f"cannot write {path}")


def write_exr(path: str, layers: dict, lib: C.CDLL | None = None):
    """Float OpenEXR of RGB layers {prefix: [h, w, 3]} ('' = beauty, 'albedo.', 'normals.')."""
    lib = lib or A.load_library()
    names = list(layers)
    arrs = [np.ascontiguousarray(layers[k], np.float32) for k in names]
    h, w = arrs[0].shape[:2]
    assert all(a.shape == (h, w, 3) for a in arrs)
    pre = (C.c_char_p * len(names))(*[k.encode() for k in names])
    ptrs = (C.c_void_p * len(arrs))(*[a.ctypes.data for a in arrs])
    st = lib.ahip_write_exr(str(path).encode(), w, h, len(names), pre, ptrs)
    if st != A.AHIP_OK:
        raise BackendError(st, f"cannot write {path}")

"""USD-lite scene ingest over the C ABI: one text USD layer (.usda) as a ``Scene``.

The reading is native (``ahip_usd_load`` / ``ahip_usd_load_text`` in libanacapa_hip.so,
anacapa_amd/csrc/usd_lite.cpp); this module only copies the flat arrays the library returns into
the same ``Scene`` the procedural builders produce, so the loaded scene goes through
``HipPathIntegrator.prepare`` like any other.  It restates the reference's
``loadUSD(path, filmWidth, filmHeight, cameraOverridePath)`` (src/scene/usd/USDLoader.h:26-29)
for a single layer without composition; see the header comment of usd_lite.cpp for what is
read and what is refused.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _abi as A
from .scenes import Scene


class UsdLoadError(RuntimeError):
    def __init__(self, status: int, message: str):
        super().__init__(f"ahip status {status}: {message}")
        self.status = status


def _copy_material(m: A.Material) -> A.Material:
    out = A.Material()
    C.memmove(C.byref(out), C.byref(m), C.sizeof(A.Material))
    return out


def _copy_light(l: A.Light) -> A.Light:
    out = A.Light()
    C.memmove(C.byref(out), C.byref(l), C.sizeof(A.Light))
    return out


def load_usda(path: str | None = None, *, text: str | None = None, width: int, height: int,
              spp: int = 1, max_depth: int = 8, min_depth: int = 2, seed: int = 0,
              camera_path: str | None = None, lib: C.CDLL | None = None) -> Scene:
    """Load a .usda file (``path``) or layer text (``text``) for a film of width x height.

    Raises ``UsdLoadError`` with the library's status (AHIP_E_UNSUPPORTED for layers that need
    what the device path lacks, AHIP_E_INVALID_ARG for malformed text).  A layer without a camera
    keeps the ``Scene`` default camera and sets ``scene.has_camera = False``.
    """
    if (path is None) == (text is None):
        raise ValueError("give exactly one of path= or text=")
    lib = lib or A.load_library()
    handle = C.c_void_p()
    err = C.create_string_buffer(1024)
    cam = camera_path.encode() if camera_path else None
    if path is not None:
        st = lib.ahip_usd_load(str(path).encode(), width, height, cam, C.byref(handle), err, len(err))
    else:
        raw = text.encode()
        st = lib.ahip_usd_load_text(raw, len(raw), width, height, cam, C.byref(handle), err, len(err))
    if st != A.AHIP_OK:
        raise UsdLoadError(st, err.value.decode(errors="replace"))
    try:
        d = lib.ahip_usd_scene_desc(handle).contents
        nm, nv, ni = d.num_meshes, d.num_vertices, d.num_indices

        def take(p, n, dt):
            if n == 0:
                return np.zeros(0, dt)
            return np.ctypeslib.as_array(p, shape=(n,)).astype(dt, copy=True)

        frm, at, up = (C.c_float * 3)(), (C.c_float * 3)(), (C.c_float * 3)()
        vfov = C.c_float()
        F = C.POINTER(C.c_float)
        cst = lib.ahip_usd_camera(handle, C.cast(frm, F), C.cast(at, F), C.cast(up, F), C.byref(vfov), None)
        has_camera = cst == A.AHIP_OK
        cam_kw = dict(cam_from=tuple(frm), cam_at=tuple(at), cam_up=tuple(up), vfov=float(vfov.value)) \
            if has_camera else {}
        sc = Scene(
            name=str(path) if path is not None else "usda",
            positions=take(d.positions, nv * 3, np.float32).reshape(-1, 3),
            normals=take(d.normals, nv * 3, np.float32).reshape(-1, 3),
            uvs=take(d.uvs, nv * 2, np.float32).reshape(-1, 2),
            indices=take(d.indices, ni, np.uint32),
            vertex_offset=take(d.mesh_vertex_offset, nm + 1, np.uint32),
            index_offset=take(d.mesh_index_offset, nm + 1, np.uint32),
            normal_count=take(d.mesh_normal_count, nm, np.uint32),
            uv_count=take(d.mesh_uv_count, nm, np.uint32),
            mesh_material=take(d.mesh_material, nm, np.int32),
            materials=[_copy_material(d.materials[i]) for i in range(d.num_materials)],
            lights=[_copy_light(d.lights[i]) for i in range(d.num_lights)],
            env_light=d.env_light,
            env_radiance=tuple(d.env_radiance),
            width=width, height=height, spp=spp, max_depth=max_depth, min_depth=min_depth, seed=seed,
            **cam_kw,
        )
        sc.mesh_names = [lib.ahip_usd_mesh_name(handle, i).decode() for i in range(nm)]
        sc.has_camera = has_camera
        return sc
    finally:
        lib.ahip_usd_free(handle)

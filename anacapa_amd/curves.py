"""Hair curves in interchange bases as device strands, over the C ABI (ahip_curves_to_bezier).

The conversion is native (anacapa_amd/csrc/curve_basis.cpp): it restates the curve handling of
the reference's Alembic reader (src/scene/alembic/AlembicLoader.cpp:58-306) for the data one
ICurves sample holds, without the Alembic SDK.  ``to_bezier`` returns the strand arrays and
``attach`` puts them on a ``Scene``.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _abi as A

BASES = {"linear": A.AHIP_CURVE_LINEAR, "bezier": A.AHIP_CURVE_BEZIER, "bspline": A.AHIP_CURVE_BSPLINE,
         "catmullrom": A.AHIP_CURVE_CATMULL_ROM}


class CurveError(RuntimeError):
    def __init__(self, status: int, message: str):
        super().__init__(f"ahip status {status}: {message}")
        self.status = status


def to_bezier(basis, counts, points, widths=None, colors=None, world=None, default_width=0.005,
              width_scale=1.0, material_index=0, lib: C.CDLL | None = None) -> dict:
    """Endpoint-sharing cubic Bezier strands from one curve object.

    basis: AHIP_CURVE_* or a key of BASES; counts: CVs per curve; points: [sum(counts), 3];
    widths: per-vertex, per-curve, a single value or None; colors: [curves, 3] or None;
    world: 4x4 acting on column vectors (Mat4f) or None.  Returns the strand arrays of
    ahip_scene_desc plus ``num_skipped``.
    """
    lib = lib or A.load_library()
    b = BASES[basis] if isinstance(basis, str) else int(basis)
    counts = np.ascontiguousarray(counts, np.uint32)
    pts = np.ascontiguousarray(points, np.float32).reshape(-1, 3)
    w = None if widths is None else np.ascontiguousarray(np.atleast_1d(widths), np.float32)
    col = None if colors is None else np.ascontiguousarray(colors, np.float32).reshape(-1, 3)
    m = None if world is None else np.ascontiguousarray(world, np.float32).reshape(4, 4)
    opts = A.CurveOptions(default_width, width_scale, material_index, 0)
    h = C.c_void_p()
    err = C.create_string_buffer(512)
    st = lib.ahip_curves_to_bezier(b, len(counts), A.ptr(counts), A.ptr(pts), A.ptr(w),
                                   0 if w is None else len(w), A.ptr(col), A.ptr(m), C.byref(opts),
                                   C.byref(h), err, len(err))
    if st != A.AHIP_OK:
        raise CurveError(st, err.value.decode(errors="replace"))
    try:
        s = lib.ahip_curve_set_strands(h).contents
        ns, nc = s.num_strands, s.num_cvs

        def take(p, n, dt):
            return np.zeros(0, dt) if n == 0 else np.ctypeslib.as_array(p, shape=(n,)).astype(dt, copy=True)

        return {
            "strand_cv_offset": take(s.strand_cv_offset, ns + 1, np.uint32),
            "curve_cvs": take(s.curve_cvs, 3 * nc, np.float32).reshape(-1, 3),
            "strand_width_offset": take(s.strand_width_offset, ns + 1, np.uint32),
            "strand_widths": take(s.strand_widths, nc, np.float32),
            "strand_mesh_index": take(s.strand_mesh_index, ns, np.uint32),
            "strand_color": take(s.strand_color, 3 * ns, np.float32).reshape(-1, 3),
            "num_skipped": int(s.num_skipped),
        }
    finally:
        lib.ahip_curve_set_free(h)


def attach(scene, strands: dict):
    """Replace the scene's strands with converted ones (per-CV widths)."""
    for k in ("strand_cv_offset", "curve_cvs", "strand_width_offset", "strand_widths", "strand_mesh_index",
              "strand_color"):
        setattr(scene, k, strands[k])
    scene.strand_width = None
    return scene

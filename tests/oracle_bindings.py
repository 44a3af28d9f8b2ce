"""ctypes bindings of the CPU oracle (oracle/liboracle.so) and of the reference build
(oracle/_ref/libanacapa_ref.so).  Test infrastructure only."""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

from anacapa_amd import _abi as A

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_SO = os.path.join(ROOT, "oracle", "liboracle.so")
ORACLE_RELEASE_SO = os.path.join(ROOT, "oracle", "liboracle_release.so")   # same source, Release flags
REF_SO = os.path.join(ROOT, "oracle", "_ref", "libanacapa_ref.so")

VP = C.c_void_p
P = A.ptr


def _load(path):
    if not os.path.exists(path):
        return None
    return C.CDLL(path)


PROBE_IN, PROBE_OUT = 19, 16   # material probe layout (oracle.cpp / ref_driver.cpp / ahip_material_probe)


def probe_inputs(n_, t, color, wo, wi, u, uc, curve):
    """Pack the shared material-probe input rows: normal, tangent, colour, wo, wi, u(2), uc, curve."""
    m = len(wo)
    out = np.zeros((m, PROBE_IN), np.float32)
    out[:, 0:3], out[:, 3:6], out[:, 6:9] = n_, t, color
    out[:, 9:12], out[:, 12:15], out[:, 15:17] = wo, wi, u
    out[:, 17], out[:, 18] = uc, curve
    return out


class _ShadingProbes:
    """Shading probes both CPU libraries export under their own prefix (orc_ / ref_)."""

    prefix = ""

    def _bind_probes(self):
        L, p = self.lib, self.prefix
        getattr(L, p + "hair_consts").argtypes = [C.POINTER(A.Material), VP]
        getattr(L, p + "mh_fn").argtypes = [C.c_int, C.c_uint32, VP, VP, VP, VP, VP, VP]
        getattr(L, p + "mh_ap").argtypes = [C.c_uint32, VP, VP]
        getattr(L, p + "hair_lobes").argtypes = [C.POINTER(A.Material), C.c_uint32, VP, VP]
        getattr(L, p + "hair_geom").argtypes = [C.c_uint32, VP, VP]
        getattr(L, p + "dome_create").restype = VP
        getattr(L, p + "dome_create").argtypes = [C.c_uint32, C.c_uint32, VP, C.c_float, VP]
        getattr(L, p + "dome_destroy").argtypes = [VP]

    def _fn(self, name):
        return getattr(self.lib, self.prefix + name)

    def hair_consts(self, mat):
        out = np.zeros(7, np.float32)
        self._fn("hair_consts")(C.byref(mat), P(out))
        return out

    def mh_fn(self, which, a, b=None, c=None, d=None, e=None):
        a = np.ascontiguousarray(a, np.float32)
        args = [np.ascontiguousarray(a if x is None else x, np.float32) for x in (b, c, d, e)]
        out = np.zeros(len(a), np.float32)
        self._fn("mh_fn")(which, len(a), P(a), *[P(x) for x in args], P(out))
        return out

    def mh_ap(self, inp):
        inp = np.ascontiguousarray(inp, np.float32)
        out = np.zeros((len(inp), 9), np.float32)
        self._fn("mh_ap")(len(inp), P(inp), P(out))
        return out

    def hair_lobes(self, mat, inp):
        inp = np.ascontiguousarray(inp, np.float32)
        out = np.zeros((len(inp), 4), np.float32)
        self._fn("hair_lobes")(C.byref(mat), len(inp), P(inp), P(out))
        return out

    def hair_geom(self, inp):
        inp = np.ascontiguousarray(inp, np.float32)
        out = np.zeros((len(inp), 6), np.float32)
        self._fn("hair_geom")(len(inp), P(inp), P(out))
        return out

    def dome(self, rgb, intensity=1.0, rot=None):
        return _Dome(self, rgb, intensity, rot)


class _Dome:
    def __init__(self, owner, rgb, intensity, rot):
        self.o = owner
        self.rgb = np.ascontiguousarray(rgb, np.float32)
        h, w = self.rgb.shape[:2]
        self.w, self.h = w, h
        r = np.ascontiguousarray(np.eye(3) if rot is None else rot, np.float32).reshape(9)
        self.p = owner._fn("dome_create")(w, h, P(self.rgb), float(intensity), P(r))

    def __del__(self):
        if getattr(self, "p", None):
            self.o._fn("dome_destroy")(self.p)
            self.p = None

    def sample(self, u):
        u = np.ascontiguousarray(u, np.float32).reshape(-1, 2)
        out = np.zeros((len(u), 7), np.float32)
        self.o._fn(self.o.dome_sample_name)(self.p, len(u), P(u), P(out))
        return out

    def eval(self, dirs):
        dirs = np.ascontiguousarray(dirs, np.float32).reshape(-1, 3)
        out = np.zeros((len(dirs), 4), np.float32)
        self.o._fn(self.o.dome_eval_name)(self.p, len(dirs), P(dirs), P(out))
        return out


class Oracle(_ShadingProbes):
    """This repository's C++ restatement of the reference CPU path (oracle/oracle.cpp)."""

    def __init__(self, release: bool = False):
        """release: the build at the reference's Release flags (timing only; tests use the default)."""
        path = ORACLE_RELEASE_SO if release else ORACLE_SO
        self.lib = _load(path)
        if self.lib is None:
            raise RuntimeError(f"{path} not built (make -C oracle)")
        L = self.lib
        L.orc_scene_create.restype = VP
        L.orc_scene_create.argtypes = [C.POINTER(A.SceneDesc)]
        L.orc_scene_destroy.argtypes = [VP]
        L.orc_radical_inverse.restype = C.c_float
        L.orc_radical_inverse.argtypes = [C.c_uint32, C.c_uint64]
        L.orc_width_at.restype = C.c_float
        L.orc_width_at.argtypes = [VP, C.c_uint32, C.c_float]
        for n in ("orc_pcg_u32",):
            getattr(L, n).argtypes = [C.c_uint64, C.c_uint64, C.c_uint32, VP]
        L.orc_halton.argtypes = [C.c_uint64, C.c_uint32, VP, VP, VP, VP, VP]
        L.orc_halton_seq.argtypes = [C.c_uint64, C.c_uint32, C.c_uint32, C.c_uint32, C.c_uint32, VP]
        L.orc_make_pinhole.argtypes = [VP, VP, VP, C.c_float, C.c_uint32, C.c_uint32, C.POINTER(A.Camera)]
        L.orc_generate_rays.argtypes = [C.POINTER(A.Camera), C.c_uint32, VP, VP, VP, VP, VP]
        L.orc_set_render.argtypes = [VP, C.POINTER(A.Camera), C.c_uint32, C.c_uint32, C.c_uint64, C.c_uint32]
        L.orc_bvh_sizes.argtypes = [VP, VP, VP]
        L.orc_bvh_dump.argtypes = [VP, VP, VP]
        L.orc_trace.argtypes = [VP, VP, C.c_uint32, VP, VP, VP]
        L.orc_occluded.argtypes = [VP, VP, C.c_uint32, VP]
        L.orc_render_tile.argtypes = [VP] + [C.c_uint32] * 6 + [VP]
        L.orc_render_tile_bdpt.argtypes = [VP] + [C.c_uint32] * 6 + [C.c_float, VP, VP, VP, VP]
        L.orc_render_tile_aovs.argtypes = [VP] + [C.c_uint32] * 6 + [VP, VP, VP]
        L.orc_render_samples.argtypes = [VP, C.c_uint32, VP, VP, VP, VP, VP]
        L.orc_render_frame.argtypes = [VP, C.c_uint32, C.c_uint32, C.c_uint32, C.c_uint32, VP, VP]
        L.orc_material_probe.argtypes = [C.POINTER(A.Material), C.c_uint32, VP, VP]
        L.orc_material_eval.argtypes = [C.POINTER(A.Material), C.c_uint32, VP, VP]
        L.orc_ggx_helper.argtypes = [C.c_int, C.c_uint32, VP, VP, VP, VP]
        L.orc_distribution1d.argtypes = [VP, C.c_uint32, C.c_uint32, VP, VP, VP, VP, VP, VP]
        L.orc_dome_sample.argtypes = [VP, C.c_uint32, VP, VP]
        L.orc_dome_eval.argtypes = [VP, C.c_uint32, VP, VP, VP]
        L.orc_set_eval.argtypes = [VP, VP, C.c_uint32, VP]
        L.orc_replay_tile.argtypes = [VP] + [C.c_uint32] * 6 + [VP]
        L.orc_adjoint.argtypes = [VP, C.c_uint32, C.c_uint32, VP, VP, VP]
        L.orc_probe.argtypes = [C.POINTER(A.Material), C.c_uint32, C.c_uint32, VP, VP]
        L.orc_eval_ggx.argtypes = [C.c_uint32, VP, VP]
        L.orc_dome_sample_raw.argtypes = [VP, C.c_uint32, VP, VP]
        L.orc_dome_eval_raw.argtypes = [VP, C.c_uint32, VP, VP]
        self._bind_probes()

    prefix = "orc_"
    dome_sample_name, dome_eval_name = "dome_sample_raw", "dome_eval_raw"

    def probe(self, mat, inp, flags=0):
        inp = np.ascontiguousarray(inp, np.float32)
        out = np.zeros((len(inp), PROBE_OUT), np.float32)
        self.lib.orc_probe(C.byref(mat), flags, len(inp), P(inp), P(out))
        return out

    def eval_ggx(self, inp):
        inp = np.ascontiguousarray(inp, np.float32)
        out = np.zeros((len(inp), 5), np.float32)
        self.lib.orc_eval_ggx(len(inp), P(inp), P(out))
        return out

    # ---- stateless
    def pcg(self, seed, seq, n):
        out = np.zeros(n, np.uint32)
        self.lib.orc_pcg_u32(seed, seq, n, P(out))
        return out

    def halton(self, seed, px, py, idx, dim):
        px, py, idx, dim = (np.ascontiguousarray(a, np.uint32) for a in (px, py, idx, dim))
        out = np.zeros(len(px), np.float32)
        self.lib.orc_halton(seed, len(px), P(px), P(py), P(idx), P(dim), P(out))
        return out

    def halton_seq(self, seed, px, py, idx, nd):
        out = np.zeros(nd, np.float32)
        self.lib.orc_halton_seq(seed, px, py, idx, nd, P(out))
        return out

    def pinhole(self, frm, at, up, vfov, w, h):
        cam = A.Camera()
        f, a, u = (np.array(v, np.float32) for v in (frm, at, up))
        self.lib.orc_make_pinhole(P(f), P(a), P(u), vfov, w, h, C.byref(cam))
        return cam

    def generate_rays(self, cam, px, py, jx, jy):
        px, py = (np.ascontiguousarray(a, np.uint32) for a in (px, py))
        jx, jy = (np.ascontiguousarray(a, np.float32) for a in (jx, jy))
        out = np.zeros((len(px), 6), np.float32)
        self.lib.orc_generate_rays(C.byref(cam), len(px), P(px), P(py), P(jx), P(jy), P(out))
        return out

    def ggx_helper(self, which, a, b=None, c=None):
        a = np.ascontiguousarray(a, np.float32)
        b = np.ascontiguousarray(a if b is None else b, np.float32)
        c = np.ascontiguousarray(a if c is None else c, np.float32)
        out = np.zeros(len(a) * (3 if which == 6 else 1), np.float32)
        self.lib.orc_ggx_helper(which, len(a), P(a), P(b), P(c), P(out))
        return out

    def distribution1d(self, values, us):
        values = np.ascontiguousarray(values, np.float32)
        us = np.ascontiguousarray(us, np.float32)
        n = len(us)
        idx, pdf, ur = np.zeros(n, np.uint32), np.zeros(n, np.float32), np.zeros(n, np.float32)
        cdf, integ = np.zeros(len(values) + 1, np.float32), np.zeros(1, np.float32)
        self.lib.orc_distribution1d(P(values), len(values), n, P(us), P(idx), P(pdf), P(ur), P(cdf), P(integ))
        return idx, pdf, ur, cdf, float(integ[0])

    def material_probe(self, mat, wo, u):
        """wo [n,3], u [n,3] (u.x, u.y, uComponent) -> dict of wi, f, pdf, valid, eval_pdf."""
        q = np.ascontiguousarray(np.concatenate([wo, u], 1), np.float32)
        out = np.zeros((len(q), 9), np.float32)
        self.lib.orc_material_probe(C.byref(mat), len(q), P(q), P(out))
        return dict(wi=out[:, :3], f=out[:, 3:6], pdf=out[:, 6], valid=out[:, 7] > 0, eval_pdf=out[:, 8])

    def material_eval(self, mat, wo, wi):
        q = np.ascontiguousarray(np.concatenate([wo, wi], 1), np.float32)
        out = np.zeros((len(q), 4), np.float32)
        self.lib.orc_material_eval(C.byref(mat), len(q), P(q), P(out))
        return out[:, :3], out[:, 3]

    # ---- scenes
    def width_at(self, widths, v):
        w = np.ascontiguousarray(widths, np.float32)
        return self.lib.orc_width_at(P(w) if len(w) else None, len(w), v)

    def scene(self, sc):
        return OracleScene(self, sc)


class OracleScene:
    def __init__(self, orc: Oracle, sc):
        self.orc, self.sc = orc, sc
        self.desc, self._keep = sc.desc()
        self.h = orc.lib.orc_scene_create(C.byref(self.desc))
        if not self.h:
            raise RuntimeError("oracle rejected the scene")
        self.set_render()

    def __del__(self):
        if getattr(self, "h", None):
            self.orc.lib.orc_scene_destroy(self.h)
            self.h = None

    def set_render(self, cam=None, max_depth=None, min_depth=2, seed=None, flags=0):
        sc = self.sc
        if cam is None:
            cam = self.orc.pinhole(sc.cam_from, sc.cam_at, sc.cam_up, sc.vfov, sc.width, sc.height)
        self.cam = cam
        self.orc.lib.orc_set_render(self.h, C.byref(cam), sc.max_depth if max_depth is None else max_depth,
                                    min_depth, sc.seed if seed is None else seed, flags)

    def bvh(self):
        nn, npr = np.zeros(1, np.uint32), np.zeros(1, np.uint32)
        self.orc.lib.orc_bvh_sizes(self.h, P(nn), P(npr))
        nodes = np.zeros((int(nn[0]), 8), np.uint32)
        prims = np.zeros(int(npr[0]), np.uint32)
        self.orc.lib.orc_bvh_dump(self.h, P(nodes), P(prims))
        return nodes, prims

    def trace(self, rays, counts=False):
        rays = np.ascontiguousarray(rays, np.float32)
        hits = np.zeros(len(rays), A.HIT_DTYPE)
        nc = np.zeros(len(rays), np.uint32) if counts else None
        tc = np.zeros(len(rays), np.uint32) if counts else None
        self.orc.lib.orc_trace(self.h, P(rays), len(rays), P(hits), P(nc), P(tc))
        return (hits, nc, tc) if counts else hits

    def occluded(self, rays):
        rays = np.ascontiguousarray(rays, np.float32)
        out = np.zeros(len(rays), np.uint8)
        self.orc.lib.orc_occluded(self.h, P(rays), len(rays), P(out))
        return out

    def render_tile(self, x0, y0, w, h, s0, ns):
        out = np.zeros((h, w, 5), np.float32)
        self.orc.lib.orc_render_tile(self.h, x0, y0, w, h, s0, ns, P(out))
        return out

    def render_tile_aovs(self, x0, y0, w, h, s0, ns):
        out = np.zeros((h, w, 5), np.float32)
        alb = np.zeros((h, w, 3), np.float32)
        nrm = np.zeros((h, w, 3), np.float32)
        self.orc.lib.orc_render_tile_aovs(self.h, x0, y0, w, h, s0, ns, P(out), P(alb), P(nrm))
        return out, alb, nrm

    def render_tile_bdpt(self, x0, y0, w, h, s0, ns, clamp=10.0, border=False):
        """BDPTIntegrator::renderTile: (accum [h, w, 5], albedo [h, w, 3], normal [h, w, 3]) and with
        border=True also [h, w, ns] per-sample flags: bit 0 where a connection's visibility hangs on
        the last bits of a hit distance (Scene::connVisible), bit 1 where a subpath sampled a smooth
        StandardSurface, whose throughput hangs on the last bit of a half-vector (Scene::smooth)."""
        out = np.zeros((h, w, 5), np.float32)
        alb = np.zeros((h, w, 3), np.float32)
        nrm = np.zeros((h, w, 3), np.float32)
        bd = np.zeros((h, w, ns), np.uint8)
        st = self.orc.lib.orc_render_tile_bdpt(self.h, x0, y0, w, h, s0, ns, clamp, P(out), P(alb), P(nrm), P(bd))
        if st:
            raise ValueError("the oracle's BDPT takes Lambert / Emissive / untextured StandardSurface triangles "
                             "lit by area lights, the dome and directional lights with a scene sphere")
        return (out, alb, nrm, bd) if border else (out, alb, nrm)

    def set_eval(self, tri_albedo=None, texture=0, texels=None):
        """Evaluation parameters of the detached replay estimator (None = sampling values)."""
        self._eval_keep = [None if a is None else np.ascontiguousarray(a, np.float32).reshape(-1)
                           for a in (tri_albedo, texels)]
        a, t = self._eval_keep
        self.orc.lib.orc_set_eval(self.h, P(a), texture, P(t))

    def replay_tile(self, x0, y0, w, h, s0, ns):
        out = np.zeros((h, w, 5), np.float32)
        self.orc.lib.orc_replay_tile(self.h, x0, y0, w, h, s0, ns, P(out))
        return out

    def adjoint(self, s0, ns, dldi, texels_shape=None):
        """Gradients (d_albedo [T,3], d_texture [h,w] or None) of sum(dldi * image)."""
        dldi = np.ascontiguousarray(dldi, np.float32)
        da = np.zeros((self.sc.num_triangles, 3), np.float32)
        dt = np.zeros(texels_shape, np.float32) if texels_shape is not None else None
        self.orc.lib.orc_adjoint(self.h, s0, ns, P(dldi), P(da), P(dt))
        return da, dt

    def render_samples(self, px, py, idx):
        px, py, idx = (np.ascontiguousarray(a, np.uint32) for a in (px, py, idx))
        out = np.zeros((len(px), 3), np.float32)
        rc = np.zeros(len(px), np.uint32)
        self.orc.lib.orc_render_samples(self.h, len(px), P(px), P(py), P(idx), P(out), P(rc))
        return out, rc

    def render_frame(self, spp, row0=0, row1=None, threads=0):
        W, H = self.cam.width, self.cam.height
        out = np.zeros((H, W, 5), np.float32)
        sec = np.zeros(1, np.float64)
        self.orc.lib.orc_render_frame(self.h, row0, H if row1 is None else row1, spp, threads, P(out), P(sec))
        return out, float(sec[0])

    def dome_sample(self, u):
        u = np.ascontiguousarray(u, np.float32)
        out = np.zeros((len(u), 5), np.float32)
        assert self.orc.lib.orc_dome_sample(self.h, len(u), P(u), P(out)) == 0
        return out[:, :3], out[:, 3], out[:, 4]

    def dome_eval(self, dirs):
        dirs = np.ascontiguousarray(dirs, np.float32)
        rgb = np.zeros((len(dirs), 3), np.float32)
        pdf = np.zeros(len(dirs), np.float32)
        assert self.orc.lib.orc_dome_eval(self.h, len(dirs), P(dirs), P(rgb), P(pdf)) == 0
        return rgb, pdf


class Reference(_ShadingProbes):
    """The reference's own BVHBackend / PathIntegrator / Halton, built by oracle/Makefile."""

    def __init__(self):
        self.lib = _load(REF_SO)
        if self.lib is None:
            raise RuntimeError("oracle/_ref/libanacapa_ref.so not built")
        L = self.lib
        L.ref_pcg_u32.argtypes = [C.c_uint64, C.c_uint64, C.c_uint32, VP]
        L.ref_pcg_float.argtypes = [C.c_uint64, C.c_uint32, VP]
        L.ref_halton_seq.argtypes = [C.c_uint64, C.c_uint32, C.c_uint32, C.c_uint32, C.c_uint32, VP]
        L.ref_halton_seq2d.argtypes = [C.c_uint64, C.c_uint32, C.c_uint32, C.c_uint32, C.c_uint32, VP]
        L.ref_make_pinhole.argtypes = [VP, VP, VP, C.c_float, C.c_uint32, C.c_uint32, C.POINTER(A.Camera)]
        L.ref_generate_rays.argtypes = [C.POINTER(A.Camera), C.c_uint32, VP, VP, VP, VP, VP]
        L.ref_scene_create.restype = VP
        L.ref_scene_create.argtypes = [C.POINTER(A.SceneDesc)]
        L.ref_scene_destroy.argtypes = [VP]
        L.ref_bvh_sizes.argtypes = [VP, VP, VP]
        L.ref_bvh_dump.argtypes = [VP, VP, VP]
        L.ref_trace.argtypes = [VP, VP, C.c_uint32, VP]
        L.ref_occluded.argtypes = [VP, VP, C.c_uint32, VP]
        L.ref_render_tile.argtypes = [VP, C.POINTER(A.Camera)] + [C.c_uint32] * 7 + [C.c_uint64, VP]
        L.ref_render_tile_aovs.argtypes = [VP, C.POINTER(A.Camera)] + [C.c_uint32] * 7 + [C.c_uint64, VP, VP, VP]
        L.ref_render_tile_bdpt.argtypes = ([VP, C.POINTER(A.Camera)] + [C.c_uint32] * 7 +
                                           [C.c_uint64, C.c_float, VP, VP, VP])
        L.ref_render_frame_bdpt.argtypes = ([VP, C.POINTER(A.Camera), C.c_uint32, C.c_uint32, C.c_uint64, C.c_float,
                                             C.c_uint32, C.c_uint32, VP, VP, VP])
        L.ref_render_frame.argtypes = [VP, C.POINTER(A.Camera), C.c_uint32, C.c_uint32, C.c_uint64,
                                       C.c_uint32, C.c_uint32, VP, VP, VP]
        L.ref_render_rows.argtypes = [VP, C.POINTER(A.Camera), C.c_uint32, C.c_uint32, C.c_uint64,
                                      C.c_uint32, C.c_uint32, C.c_uint32, C.c_uint32, VP, VP, VP]
        L.ref_distribution1d.argtypes = [VP, C.c_uint32, C.c_uint32, VP, VP, VP, VP, VP, VP]
        L.ref_ggx_helper.argtypes = [C.c_int, C.c_uint32, VP, VP, VP, VP]
        L.ref_lambert.argtypes = [VP, C.c_uint32, VP, VP]
        for n in ("ref_hair_kk_probe", "ref_hair_live_probe", "ref_ss_probe"):
            getattr(L, n).argtypes = [C.POINTER(A.Material), C.c_uint32, VP, VP]
        L.ref_ss_eval_ggx.argtypes = [C.c_uint32, VP, VP]
        L.ref_dome_sample.argtypes = [VP, C.c_uint32, VP, VP]
        L.ref_dome_eval.argtypes = [VP, C.c_uint32, VP, VP]
        L.ref_dome_tables.argtypes = [VP, VP, VP, VP, VP, VP]
        self._bind_probes()

    prefix = "ref_"
    dome_sample_name, dome_eval_name = "dome_sample", "dome_eval"

    def probe(self, mat, inp, flags=0):
        """The reference's material on probe rows: StandardSurface (opaque, constant parameters),
        or hair as it runs (Kajiya-Kay) / with its Marschner code reached (AHIP_FLAG_HAIR_MARSCHNER)."""
        inp = np.ascontiguousarray(inp, np.float32)
        out = np.zeros((len(inp), PROBE_OUT), np.float32)
        if mat.type == A.AHIP_MAT_STANDARD:
            fn = self.lib.ref_ss_probe
        elif mat.type == A.AHIP_MAT_HAIR:
            fn = self.lib.ref_hair_live_probe if flags & A.AHIP_FLAG_HAIR_MARSCHNER else self.lib.ref_hair_kk_probe
        else:
            raise ValueError("no reference probe for this material type")
        if fn(C.byref(mat), len(inp), P(inp), P(out)) != 0:
            raise ValueError("reference probe refused the material")
        return out

    def ss_eval_ggx(self, inp):
        inp = np.ascontiguousarray(inp, np.float32)
        out = np.zeros((len(inp), 5), np.float32)
        self.lib.ref_ss_eval_ggx(len(inp), P(inp), P(out))
        return out

    def dome_tables(self, dome):
        w, h = dome.w, dome.h
        mf, mc = np.zeros(h, np.float32), np.zeros(h + 1, np.float32)
        cf, cc = np.zeros((h, w), np.float32), np.zeros((h, w + 1), np.float32)
        mi = np.zeros(1, np.float32)
        self.lib.ref_dome_tables(dome.p, P(mf), P(mc), P(cf), P(cc), P(mi))
        return mf, mc, cf, cc, float(mi[0])

    def pcg(self, seed, seq, n):
        out = np.zeros(n, np.uint32)
        self.lib.ref_pcg_u32(seed, seq, n, P(out))
        return out

    def halton_seq(self, seed, px, py, idx, nd):
        out = np.zeros(nd, np.float32)
        self.lib.ref_halton_seq(seed, px, py, idx, nd, P(out))
        return out

    def halton_seq2d(self, seed, px, py, idx, npairs):
        out = np.zeros(2 * npairs, np.float32)
        self.lib.ref_halton_seq2d(seed, px, py, idx, npairs, P(out))
        return out

    def pinhole(self, frm, at, up, vfov, w, h):
        cam = A.Camera()
        f, a, u = (np.array(v, np.float32) for v in (frm, at, up))
        self.lib.ref_make_pinhole(P(f), P(a), P(u), vfov, w, h, C.byref(cam))
        return cam

    def generate_rays(self, cam, px, py, jx, jy):
        px, py = (np.ascontiguousarray(a, np.uint32) for a in (px, py))
        jx, jy = (np.ascontiguousarray(a, np.float32) for a in (jx, jy))
        out = np.zeros((len(px), 6), np.float32)
        self.lib.ref_generate_rays(C.byref(cam), len(px), P(px), P(py), P(jx), P(jy), P(out))
        return out

    def ggx_helper(self, which, a, b=None, c=None):
        a = np.ascontiguousarray(a, np.float32)
        b = np.ascontiguousarray(a if b is None else b, np.float32)
        c = np.ascontiguousarray(a if c is None else c, np.float32)
        out = np.zeros(len(a) * (3 if which == 6 else 1), np.float32)
        self.lib.ref_ggx_helper(which, len(a), P(a), P(b), P(c), P(out))
        return out

    def distribution1d(self, values, us):
        values = np.ascontiguousarray(values, np.float32)
        us = np.ascontiguousarray(us, np.float32)
        n = len(us)
        idx, pdf, ur = np.zeros(n, np.uint32), np.zeros(n, np.float32), np.zeros(n, np.float32)
        cdf, integ = np.zeros(len(values) + 1, np.float32), np.zeros(1, np.float32)
        self.lib.ref_distribution1d(P(values), len(values), n, P(us), P(idx), P(pdf), P(ur), P(cdf), P(integ))
        return idx, pdf, ur, cdf, float(integ[0])

    def scene(self, sc):
        return RefScene(self, sc)


class RefScene:
    def __init__(self, ref: Reference, sc):
        self.ref, self.sc = ref, sc
        self.desc, self._keep = sc.desc()
        self.h = ref.lib.ref_scene_create(C.byref(self.desc))
        if not self.h:
            raise RuntimeError("reference build cannot represent this scene")
        self.cam = ref.pinhole(sc.cam_from, sc.cam_at, sc.cam_up, sc.vfov, sc.width, sc.height)

    def __del__(self):
        if getattr(self, "h", None):
            self.ref.lib.ref_scene_destroy(self.h)
            self.h = None

    def bvh(self):
        nn, npr = np.zeros(1, np.uint32), np.zeros(1, np.uint32)
        self.ref.lib.ref_bvh_sizes(self.h, P(nn), P(npr))
        nodes = np.zeros((int(nn[0]), 8), np.uint32)
        prims = np.zeros(int(npr[0]), np.uint32)
        self.ref.lib.ref_bvh_dump(self.h, P(nodes), P(prims))
        return nodes, prims

    def trace(self, rays):
        rays = np.ascontiguousarray(rays, np.float32)
        out = np.zeros((len(rays), 12), np.float32)
        self.ref.lib.ref_trace(self.h, P(rays), len(rays), P(out))
        mesh = out[:, 1].view(np.uint32).copy()
        prim = out[:, 2].view(np.uint32).copy()
        return dict(t=out[:, 0].copy(), mesh_id=mesh, prim_id=prim, p=out[:, 3:6], n=out[:, 6:9], ng=out[:, 9:12])

    def occluded(self, rays):
        rays = np.ascontiguousarray(rays, np.float32)
        out = np.zeros(len(rays), np.uint8)
        self.ref.lib.ref_occluded(self.h, P(rays), len(rays), P(out))
        return out

    def render_tile(self, x0, y0, w, h, s0, ns, max_depth=None, seed=None, cam=None):
        out = np.zeros((h, w, 5), np.float32)
        sc = self.sc
        self.ref.lib.ref_render_tile(self.h, C.byref(cam or self.cam), x0, y0, w, h, s0, ns,
                                     sc.max_depth if max_depth is None else max_depth,
                                     sc.seed if seed is None else seed, P(out))
        return out

    def render_tile_aovs(self, x0, y0, w, h, s0, ns, max_depth=None, seed=None, cam=None):
        out = np.zeros((h, w, 5), np.float32)
        alb = np.zeros((h, w, 3), np.float32)
        nrm = np.zeros((h, w, 3), np.float32)
        sc = self.sc
        self.ref.lib.ref_render_tile_aovs(self.h, C.byref(cam or self.cam), x0, y0, w, h, s0, ns,
                                          sc.max_depth if max_depth is None else max_depth,
                                          sc.seed if seed is None else seed, P(out), P(alb), P(nrm))
        return out, alb, nrm

    def render_tile_bdpt(self, x0, y0, w, h, s0, ns, clamp=10.0, max_depth=None, seed=None, cam=None):
        """The reference's BDPTIntegrator(maxDepth, clamp)::renderTile: (accum, albedo sums, normal sums)."""
        out = np.zeros((h, w, 5), np.float32)
        alb = np.zeros((h, w, 3), np.float32)
        nrm = np.zeros((h, w, 3), np.float32)
        sc = self.sc
        self.ref.lib.ref_render_tile_bdpt(self.h, C.byref(cam or self.cam), x0, y0, w, h, s0, ns,
                                          sc.max_depth if max_depth is None else max_depth,
                                          sc.seed if seed is None else seed, clamp, P(out), P(alb), P(nrm))
        return out, alb, nrm

    def render_frame_bdpt(self, spp, clamp=10.0, threads=0, tile=64, max_depth=None, cam=None):
        """The frame through RenderSession's tile loop with BDPTIntegrator: (accum, seconds, threads)."""
        cam = cam or self.cam
        out = np.zeros((cam.height, cam.width, 5), np.float32)
        sec = np.zeros(1, np.float64)
        used = np.zeros(1, np.uint32)
        self.ref.lib.ref_render_frame_bdpt(self.h, C.byref(cam), spp, self.sc.max_depth if max_depth is None else max_depth,
                                           self.sc.seed, clamp, threads, tile, P(out), P(sec), P(used))
        return out, float(sec[0]), int(used[0])

    def render_frame(self, spp, threads=0, tile=64, max_depth=None, cam=None, row0=0, row1=None):
        """RenderSession's non-adaptive tile loop (64x64 tiles on the reference ThreadPool) over the
        rows [row0, row1) (the whole frame by default): (accum, seconds of the tile loop, threads)."""
        cam = cam or self.cam
        out = np.zeros((cam.height, cam.width, 5), np.float32)
        sec = np.zeros(1, np.float64)
        used = np.zeros(1, np.uint32)
        self.ref.lib.ref_render_rows(self.h, C.byref(cam), spp, self.sc.max_depth if max_depth is None else max_depth,
                                     self.sc.seed, row0, cam.height if row1 is None else row1, threads, tile,
                                     P(out), P(sec), P(used))
        return out, float(sec[0]), int(used[0])


def reference_available() -> bool:
    return os.path.exists(REF_SO)

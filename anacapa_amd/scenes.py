"""Procedural scenes for the five workloads, as flat arrays for the C ABI.

Each builder returns a ``Scene``: triangle meshes (positions / normals / uvs / local
indices, one range per mesh as GeometryPool holds them, include/anacapa/accel/GeometryPool.h),
per-mesh materials (SceneView::materials is indexed by meshID), lights, the HDRI of the
dome light and the camera parameters.  All randomness comes from ``numpy.random.default_rng``
seeded by the scene seed, so a scene is a pure function of its arguments.
"""
from __future__ import annotations

import ctypes as C
import math
from dataclasses import dataclass, field

import numpy as np

from . import _abi as A


@dataclass
class Scene:
    name: str
    positions: np.ndarray                 # [V,3] f32
    normals: np.ndarray | None            # [V,3] f32 (per-mesh counts in normal_count)
    uvs: np.ndarray | None                # [V,2] f32
    indices: np.ndarray                   # [I] u32, local to mesh
    vertex_offset: np.ndarray             # [M+1] u32
    index_offset: np.ndarray              # [M+1] u32
    normal_count: np.ndarray              # [M] u32
    uv_count: np.ndarray                  # [M] u32
    mesh_material: np.ndarray             # [M] i32
    materials: list = field(default_factory=list)     # list[A.Material]
    lights: list = field(default_factory=list)        # list[A.Light]
    env_light: int = -1
    env_radiance: tuple = (0.0, 0.0, 0.0)
    env_pixels: np.ndarray | None = None  # [H,W,3] f32
    textures: list = field(default_factory=list)      # list[np.ndarray [H,W] f32]
    strand_cv_offset: np.ndarray | None = None       # [S+1] u32
    curve_cvs: np.ndarray | None = None               # [CV,3] f32
    strand_width: np.ndarray | None = None            # [S,2] f32 root/tip diameter
    strand_mesh_index: np.ndarray | None = None       # [S] u32
    strand_color: np.ndarray | None = None            # [S,3] f32
    tri_albedo: np.ndarray | None = None              # [T,3] f32, commit order (per-triangle base colour)
    strand_width_offset: np.ndarray | None = None     # [S+1] u32 ranges of strand_widths
    strand_widths: np.ndarray | None = None           # per-CV diameters (replace strand_width)
    cam_from: tuple = (0.0, 0.0, 0.0)
    cam_at: tuple = (0.0, 0.0, 1.0)
    cam_up: tuple = (0.0, 1.0, 0.0)
    vfov: float = 40.0
    width: int = 64
    height: int = 64
    spp: int = 1
    max_depth: int = 8
    min_depth: int = 2
    seed: int = 0

    @property
    def num_meshes(self) -> int:
        return int(len(self.mesh_material))

    @property
    def num_triangles(self) -> int:
        return int(len(self.indices) // 3)

    def _with_sphere(self, light):
        """A dome or directional light left without a scene sphere gets the one its reference loader
        gives it: the triangle bounds' centroid, and the half diagonal x 2 for the dome
        (RenderSession.cpp:407-416) or x 1.5 for a directional light (USDLoader.cpp:1722-1744)."""
        if light.type not in (A.AHIP_LIGHT_DOME, A.AHIP_LIGHT_DIRECTIONAL) or light.scene_radius != 0 \
                or len(self.positions) == 0:
            return light
        p = np.asarray(self.positions, np.float32).reshape(-1, 3)
        lo, hi = p.min(0), p.max(0)
        dg = hi - lo
        half = np.float32(np.sqrt(np.float32(dg[0] * dg[0] + dg[1] * dg[1] + dg[2] * dg[2]))) * np.float32(0.5)
        l = A.Light.from_buffer_copy(light)
        l.scene_center = A.F3(*((lo + hi) * np.float32(0.5)).tolist())
        l.scene_radius = float(half * np.float32(2.0 if light.type == A.AHIP_LIGHT_DOME else 1.5))
        return l

    def desc(self):
        """Build an ahip_scene_desc; returns (desc, keepalive) — keep both alive while in use."""
        keep = []

        def arr(a, dt):
            if a is None:
                return None
            b = np.ascontiguousarray(a, dtype=dt)
            keep.append(b)
            return b

        pos = arr(self.positions.reshape(-1), np.float32)
        nrm = arr(None if self.normals is None else self.normals.reshape(-1), np.float32)
        uvs = arr(None if self.uvs is None else self.uvs.reshape(-1), np.float32)
        idx = arr(self.indices, np.uint32)
        vo = arr(self.vertex_offset, np.uint32)
        io = arr(self.index_offset, np.uint32)
        nc = arr(self.normal_count, np.uint32)
        uc = arr(self.uv_count, np.uint32)
        mm = arr(self.mesh_material, np.int32)
        mats = (A.Material * max(1, len(self.materials)))(*self.materials)
        lts = (A.Light * max(1, len(self.lights)))(*[self._with_sphere(l) for l in self.lights])
        keep += [mats, lts]
        d = A.SceneDesc()
        d.num_meshes = self.num_meshes
        d.mesh_vertex_offset = vo.ctypes.data_as(A.U32P)
        d.mesh_index_offset = io.ctypes.data_as(A.U32P)
        d.mesh_normal_count = nc.ctypes.data_as(A.U32P)
        d.mesh_uv_count = uc.ctypes.data_as(A.U32P)
        d.mesh_material = mm.ctypes.data_as(A.I32P)
        d.positions = pos.ctypes.data_as(A.F32P)
        d.normals = nrm.ctypes.data_as(A.F32P) if nrm is not None else C.cast(None, A.F32P)
        d.uvs = uvs.ctypes.data_as(A.F32P) if uvs is not None else C.cast(None, A.F32P)
        d.indices = idx.ctypes.data_as(A.U32P)
        d.num_vertices = len(self.positions)
        d.num_indices = len(self.indices)
        d.num_materials = len(self.materials)
        d.materials = C.cast(mats, C.POINTER(A.Material))
        d.num_lights = len(self.lights)
        d.lights = C.cast(lts, C.POINTER(A.Light))
        d.env_light = self.env_light
        d.env_radiance = A.F3(*self.env_radiance)
        if self.env_pixels is not None:
            env = arr(self.env_pixels.reshape(-1), np.float32)
            d.env_height, d.env_width = self.env_pixels.shape[:2]
            d.env_pixels = env.ctypes.data_as(A.F32P)
        d.num_textures = len(self.textures)
        if self.textures:
            tw = arr(np.array([t.shape[1] for t in self.textures]), np.uint32)
            th = arr(np.array([t.shape[0] for t in self.textures]), np.uint32)
            tps = [arr(t.reshape(-1), np.float32) for t in self.textures]
            tpa = (A.F32P * len(tps))(*[t.ctypes.data_as(A.F32P) for t in tps])
            keep.append(tpa)
            d.texture_width = tw.ctypes.data_as(A.U32P)
            d.texture_height = th.ctypes.data_as(A.U32P)
            d.texture_pixels = C.cast(tpa, C.POINTER(A.F32P))
        if self.strand_cv_offset is not None and len(self.strand_cv_offset) > 1:
            so = arr(self.strand_cv_offset, np.uint32)
            cv = arr(self.curve_cvs.reshape(-1), np.float32)
            sm = arr(self.strand_mesh_index, np.uint32)
            d.num_strands = len(so) - 1
            d.strand_cv_offset = so.ctypes.data_as(A.U32P)
            d.curve_cvs = cv.ctypes.data_as(A.F32P)
            if self.strand_width is not None:
                d.strand_width = arr(self.strand_width.reshape(-1), np.float32).ctypes.data_as(A.F32P)
            if self.strand_widths is not None:
                d.strand_width_offset = arr(self.strand_width_offset, np.uint32).ctypes.data_as(A.U32P)
                d.strand_widths = arr(self.strand_widths, np.float32).ctypes.data_as(A.F32P)
            d.strand_mesh_index = sm.ctypes.data_as(A.U32P)
            if self.strand_color is not None:
                d.strand_color = arr(self.strand_color.reshape(-1), np.float32).ctypes.data_as(A.F32P)
        if self.tri_albedo is not None:
            assert self.tri_albedo.size == 3 * self.num_triangles
            d.tri_albedo = arr(self.tri_albedo.reshape(-1), np.float32).ctypes.data_as(A.F32P)
        keep.append(d)
        return d, keep

    @property
    def num_strands(self) -> int:
        return 0 if self.strand_cv_offset is None else int(len(self.strand_cv_offset) - 1)


def lambert(rgb) -> A.Material:
    m = A.Material()
    m.type = A.AHIP_MAT_LAMBERT
    m.albedo = A.F3(*rgb)
    m.roughness_texture = -1
    return m


def emissive(albedo, emission) -> A.Material:
    m = lambert(albedo)
    m.type = A.AHIP_MAT_EMISSIVE
    m.emission = A.F3(*emission)
    return m


def standard_surface(base_color, metalness=0.0, roughness=0.5, specular=1.0, ior=1.5, coat=0.0,
                     coat_roughness=0.1, base=1.0, roughness_texture=-1) -> A.Material:
    """StandardSurfaceMaterial::Params defaults (src/shading/StandardSurface.h:117-148)."""
    m = A.Material()
    m.type = A.AHIP_MAT_STANDARD
    m.albedo = A.F3(*base_color)
    m.base = base
    m.metalness = metalness
    m.roughness = roughness
    m.specular = specular
    m.specular_color = A.F3(1.0, 1.0, 1.0)
    m.specular_ior = ior
    m.coat = coat
    m.coat_roughness = coat_roughness
    m.transmission = 0.0
    m.roughness_texture = roughness_texture
    return m


def glass(base_color=(1.0, 1.0, 1.0), roughness=0.0, ior=1.5, transmission=1.0) -> A.Material:
    """Dielectric glass: standard_surface with transmission, no metal (StandardSurface.h:248-347)."""
    m = standard_surface(base_color, metalness=0.0, roughness=roughness, specular=1.0, ior=ior)
    m.transmission = transmission
    return m


def alpha_masked(base_color, opacity=1.0, opacity_texture=-1, roughness=0.5) -> A.Material:
    """standard_surface with Params::alphaMask: opacity a constant or a texture (StandardSurface.h:141-143,
    197-217)."""
    m = standard_surface(base_color, roughness=roughness)
    m.alpha_mask = 1
    m.opacity = opacity
    m.opacity_texture = opacity_texture
    return m


def hair(sigma_a=(0.40, 0.65, 1.35), eta=1.55, beta_m=0.3, beta_n=0.3, alpha=2.0) -> A.Material:
    """MarschnerHairMaterial::Params defaults (src/shading/MarschnerHair.h:193-199)."""
    m = A.Material()
    m.type = A.AHIP_MAT_HAIR
    m.roughness_texture = -1
    m.hair_sigma_a = A.F3(*sigma_a)
    m.hair_eta = eta
    m.hair_beta_m = beta_m
    m.hair_beta_n = beta_n
    m.hair_alpha = alpha
    return m


def eumelanin_sigma_a(ce):
    """Absorption of eumelanin concentration ce (d'Eon et al. 2011 as used by PBRT-v4's
    HairBxDF::SigmaAFromConcentration with no pheomelanin)."""
    return np.outer(np.asarray(ce, np.float64), np.array([0.419, 0.697, 1.37]))


def area_light(center, u_axis, v_axis, Le) -> A.Light:
    l = A.Light()
    l.type = A.AHIP_LIGHT_AREA
    l.center = A.F3(*center)
    l.u_axis = A.F3(*u_axis)
    l.v_axis = A.F3(*v_axis)
    l.Le = A.F3(*Le)
    l.intensity = 1.0
    l.rot = (C.c_float * 9)(1, 0, 0, 0, 1, 0, 0, 0, 1)
    return l


def dome_light(intensity=1.0) -> A.Light:
    l = A.Light()
    l.type = A.AHIP_LIGHT_DOME
    l.intensity = intensity
    l.rot = (C.c_float * 9)(1, 0, 0, 0, 1, 0, 0, 0, 1)
    return l


def directional_light(direction, Le, half_angle=0.0) -> A.Light:
    """DirectionalLight(dirToLight, Le, ..., halfAngleDeg) (src/shading/lights/DirectionalLight.h:33-45):
    direction points from the scene toward the light."""
    l = A.Light()
    l.type = A.AHIP_LIGHT_DIRECTIONAL
    l.direction = A.F3(*direction)
    l.Le = A.F3(*Le)
    l.half_angle = half_angle
    l.intensity = 1.0
    l.rot = (C.c_float * 9)(1, 0, 0, 0, 1, 0, 0, 0, 1)
    return l


def with_sun(sc: Scene, direction=(0.35, 0.8, -0.5), Le=(3.0, 2.8, 2.5), half_angle=0.0) -> Scene:
    """The scene with a directional light added to its lights (the light pick is uniform over them,
    PathIntegrator.cpp:134-144)."""
    sc.lights = list(sc.lights) + [directional_light(direction, Le, half_angle)]
    return sc


class _MeshBuilder:
    def __init__(self):
        self.pos, self.nrm, self.uv, self.idx = [], [], [], []
        self.vo, self.io, self.nc, self.uc, self.mat = [0], [0], [], [], []

    def add(self, positions, indices, material, normals=None, uvs=None):
        positions = np.asarray(positions, np.float32).reshape(-1, 3)
        nv = len(positions)
        self.pos.append(positions)
        nrm = np.zeros((nv, 3), np.float32) if normals is None else np.asarray(normals, np.float32).reshape(-1, 3)
        uv = np.zeros((nv, 2), np.float32) if uvs is None else np.asarray(uvs, np.float32).reshape(-1, 2)
        self.nrm.append(nrm)
        self.uv.append(uv)
        self.nc.append(0 if normals is None else nv)
        self.uc.append(0 if uvs is None else nv)
        self.idx.append(np.asarray(indices, np.uint32).reshape(-1))
        self.vo.append(self.vo[-1] + nv)
        self.io.append(self.io[-1] + len(self.idx[-1]))
        self.mat.append(material)

    def scene(self, name, **kw) -> Scene:
        return Scene(
            name=name,
            positions=np.concatenate(self.pos).astype(np.float32),
            normals=np.concatenate(self.nrm).astype(np.float32),
            uvs=np.concatenate(self.uv).astype(np.float32),
            indices=np.concatenate(self.idx).astype(np.uint32),
            vertex_offset=np.array(self.vo, np.uint32),
            index_offset=np.array(self.io, np.uint32),
            normal_count=np.array(self.nc, np.uint32),
            uv_count=np.array(self.uc, np.uint32),
            mesh_material=np.array(self.mat, np.int32),
            **kw,
        )


def _quad(mb, v0, v1, v2, v3, normal, mat):
    n = np.asarray(normal, np.float32)
    n = n / np.linalg.norm(n)
    mb.add([v0, v1, v2, v3], [0, 1, 2, 0, 2, 3], mat, normals=[n, n, n, n],
           uvs=[(0, 0), (1, 0), (1, 1), (0, 1)])


def _block(mb, lo, hi, deg, mat):
    """Axis-aligned box rotated about the vertical axis through its centre: 12 triangles."""
    x0, y0, z0 = lo
    x1, y1, z1 = hi
    cx, cz = 0.5 * (x0 + x1), 0.5 * (z0 + z1)
    a = math.radians(deg)
    ca, sa = math.cos(a), math.sin(a)

    def rot(p):
        x, y, z = p
        dx, dz = x - cx, z - cz
        return (cx + ca * dx + sa * dz, y, cz - sa * dx + ca * dz)

    def rn(n):
        x, y, z = n
        return (ca * x + sa * z, y, -sa * x + ca * z)

    faces = [
        ([(x0, y0, z0), (x0, y0, z1), (x1, y0, z1), (x1, y0, z0)], (0, -1, 0)),
        ([(x0, y1, z0), (x1, y1, z0), (x1, y1, z1), (x0, y1, z1)], (0, 1, 0)),
        ([(x0, y0, z0), (x1, y0, z0), (x1, y1, z0), (x0, y1, z0)], (0, 0, -1)),
        ([(x1, y0, z1), (x0, y0, z1), (x0, y1, z1), (x1, y1, z1)], (0, 0, 1)),
        ([(x0, y0, z1), (x0, y0, z0), (x0, y1, z0), (x0, y1, z1)], (-1, 0, 0)),
        ([(x1, y0, z0), (x1, y0, z1), (x1, y1, z1), (x1, y1, z0)], (1, 0, 0)),
    ]
    pos, nrm, idx = [], [], []
    for quad, n in faces:
        b = len(pos)
        pos += [rot(p) for p in quad]
        nrm += [rn(n)] * 4
        idx += [b, b + 1, b + 2, b, b + 2, b + 3]
    mb.add(pos, idx, mat, normals=nrm)


def cornell(width=512, height=512, spp=64, max_depth=8, seed=0, second_light=False, glossy=False) -> Scene:
    """The Cornell box: unit-box scene, 36 triangles, one 0.25x0.25 quad light (17,12,4).
    second_light: a dimmer 0.1x0.1 quad (4,6,9) near the back of the ceiling as well (two lights of
    unequal power, for light selection).  glossy: the tall block rough gold metal, the short block a
    coated blue dielectric and a small mirror panel on the back wall (standard_surface, the mirror a
    delta BSDF)."""
    mats = [lambert((0.73, 0.73, 0.73)), lambert((0.63, 0.065, 0.05)), lambert((0.14, 0.45, 0.09)),
            emissive((0.0, 0.0, 0.0), (17.0, 12.0, 4.0))]
    tall, short = 0, 0
    if glossy:
        mats += [standard_surface((0.9, 0.75, 0.45), metalness=1.0, roughness=0.3),
                 standard_surface((0.2, 0.4, 0.8), roughness=0.25, coat=0.5, coat_roughness=0.1),
                 standard_surface((0.95, 0.95, 0.95), metalness=1.0, roughness=0.0)]
        tall, short = 4, 5
    mb = _MeshBuilder()
    _quad(mb, (0, 0, 0), (1, 0, 0), (1, 0, 1), (0, 0, 1), (0, 1, 0), 0)      # floor
    _quad(mb, (0, 1, 1), (1, 1, 1), (1, 1, 0), (0, 1, 0), (0, -1, 0), 0)     # ceiling
    _quad(mb, (0, 0, 1), (0, 1, 1), (1, 1, 1), (1, 0, 1), (0, 0, -1), 0)     # back
    _quad(mb, (0, 0, 1), (0, 1, 1), (0, 1, 0), (0, 0, 0), (1, 0, 0), 1)      # left (red)
    _quad(mb, (1, 0, 0), (1, 1, 0), (1, 1, 1), (1, 0, 1), (-1, 0, 0), 2)     # right (green)
    _block(mb, (0.13, 0.0, 0.12), (0.43, 0.3, 0.42), -18.0, short)          # short block
    _block(mb, (0.53, 0.0, 0.45), (0.83, 0.6, 0.75), 18.0, tall)            # tall block
    if glossy:   # mirror panel just in front of the back wall
        _quad(mb, (0.1, 0.35, 0.998), (0.1, 0.75, 0.998), (0.45, 0.75, 0.998), (0.45, 0.35, 0.998), (0, 0, -1), 6)
    y = 0.999
    _quad(mb, (0.375, y, 0.375), (0.625, y, 0.375), (0.625, y, 0.625), (0.375, y, 0.625), (0, -1, 0), 3)
    lights = [area_light((0.5, y, 0.5), (0.125, 0.0, 0.0), (0.0, 0.0, 0.125), (17.0, 12.0, 4.0))]
    if second_light:
        mats.append(emissive((0.0, 0.0, 0.0), (4.0, 6.0, 9.0)))
        _quad(mb, (0.2, y, 0.8), (0.3, y, 0.8), (0.3, y, 0.9), (0.2, y, 0.9), (0, -1, 0), len(mats) - 1)
        lights.append(area_light((0.25, y, 0.85), (0.05, 0.0, 0.0), (0.0, 0.0, 0.05), (4.0, 6.0, 9.0)))
    return mb.scene("cornell", materials=mats, lights=lights, cam_from=(0.5, 0.5, -1.4),
                    cam_at=(0.5, 0.5, 0.0), cam_up=(0.0, 1.0, 0.0), vfov=40.0, width=width,
                    height=height, spp=spp, max_depth=max_depth, seed=seed)


def cornell_glass(width=512, height=512, spp=64, max_depth=8, seed=0, alpha_texture=False) -> Scene:
    """The Cornell box with transparency: the tall block smooth glass (IOR 1.5, a delta BSDF), the short
    block rough tinted glass (roughness 0.3), and two alpha-masked panels between the blocks and the back
    wall: a partly opaque one (opacity 0.45, the stochastic pass-through) and a cut-out one (opacity 0).
    alpha_texture: the partly opaque panel reads its opacity from an 8x8 texture of 0 / 0.45 / 1 squares
    instead (texture 0; only this build's lookup reads it)."""
    mats = [lambert((0.73, 0.73, 0.73)), lambert((0.63, 0.065, 0.05)), lambert((0.14, 0.45, 0.09)),
            emissive((0.0, 0.0, 0.0), (17.0, 12.0, 4.0)),
            glass(roughness=0.0), glass((0.85, 0.95, 0.9), roughness=0.3, ior=1.45),
            alpha_masked((0.9, 0.6, 0.2), opacity=0.45, opacity_texture=0 if alpha_texture else -1),
            alpha_masked((0.2, 0.3, 0.9), opacity=0.0)]
    textures = []
    if alpha_texture:
        ii, jj = np.meshgrid(np.arange(8), np.arange(8), indexing="ij")
        textures.append(np.choose((ii + 2 * jj) % 3, [0.0, 0.45, 1.0]).astype(np.float32))
    mb = _MeshBuilder()
    _quad(mb, (0, 0, 0), (1, 0, 0), (1, 0, 1), (0, 0, 1), (0, 1, 0), 0)      # floor
    _quad(mb, (0, 1, 1), (1, 1, 1), (1, 1, 0), (0, 1, 0), (0, -1, 0), 0)     # ceiling
    _quad(mb, (0, 0, 1), (0, 1, 1), (1, 1, 1), (1, 0, 1), (0, 0, -1), 0)     # back
    _quad(mb, (0, 0, 1), (0, 1, 1), (0, 1, 0), (0, 0, 0), (1, 0, 0), 1)      # left (red)
    _quad(mb, (1, 0, 0), (1, 1, 0), (1, 1, 1), (1, 0, 1), (-1, 0, 0), 2)     # right (green)
    # the glass blocks float 2 mm above the floor: a bottom face on the floor plane would tie with the
    # floor for rays inside the glass, a tie decided by the last bit of each build's hit distance
    _block(mb, (0.13, 0.002, 0.12), (0.43, 0.3, 0.42), -18.0, 5)            # short block: rough glass
    _block(mb, (0.53, 0.002, 0.45), (0.83, 0.6, 0.75), 18.0, 4)             # tall block: smooth glass
    _quad(mb, (0.08, 0.3, 0.9), (0.08, 0.8, 0.9), (0.48, 0.8, 0.9), (0.48, 0.3, 0.9), (0, 0, -1), 6)
    _quad(mb, (0.55, 0.75, 0.95), (0.55, 0.95, 0.95), (0.9, 0.95, 0.95), (0.9, 0.75, 0.95), (0, 0, -1), 7)
    y = 0.999
    _quad(mb, (0.375, y, 0.375), (0.625, y, 0.375), (0.625, y, 0.625), (0.375, y, 0.625), (0, -1, 0), 3)
    lights = [area_light((0.5, y, 0.5), (0.125, 0.0, 0.0), (0.0, 0.0, 0.125), (17.0, 12.0, 4.0))]
    return mb.scene("cornell_glass", materials=mats, lights=lights, cam_from=(0.5, 0.5, -1.4),
                    cam_at=(0.5, 0.5, 0.0), cam_up=(0.0, 1.0, 0.0), vfov=40.0, width=width,
                    height=height, spp=spp, max_depth=max_depth, seed=seed, textures=textures)


def geodesic_sphere(freq: int):
    """Class-I geodesic sphere: each icosahedron face split into freq^2 triangles, outward CCW."""
    t = (1.0 + 5 ** 0.5) / 2.0
    v = np.array([(-1, t, 0), (1, t, 0), (-1, -t, 0), (1, -t, 0), (0, -1, t), (0, 1, t), (0, -1, -t),
                  (0, 1, -t), (t, 0, -1), (t, 0, 1), (-t, 0, -1), (-t, 0, 1)], np.float64)
    v /= np.linalg.norm(v, axis=1, keepdims=True)
    f = np.array([(0, 11, 5), (0, 5, 1), (0, 1, 7), (0, 7, 10), (0, 10, 11), (1, 5, 9), (5, 11, 4),
                  (11, 10, 2), (10, 7, 6), (7, 1, 8), (3, 9, 4), (3, 4, 2), (3, 2, 6), (3, 6, 8),
                  (3, 8, 9), (4, 9, 5), (2, 4, 11), (6, 2, 10), (8, 6, 7), (9, 8, 1)])
    pts, tris = [], []
    for a, b, c in f:
        A_, B_, C_ = v[a], v[b], v[c]
        base = len(pts)
        lid = {}
        for i in range(freq + 1):
            for j in range(freq + 1 - i):
                lid[(i, j)] = base + len(lid)
                pts.append(A_ + (B_ - A_) * (i / freq) + (C_ - A_) * (j / freq))
        for i in range(freq):
            for j in range(freq - i):
                tris.append((lid[(i, j)], lid[(i + 1, j)], lid[(i, j + 1)]))
                if i + j < freq - 1:
                    tris.append((lid[(i + 1, j)], lid[(i + 1, j + 1)], lid[(i, j + 1)]))
    pts = np.array(pts)
    pts /= np.linalg.norm(pts, axis=1, keepdims=True)
    key = np.round(pts * 1e6).astype(np.int64)
    _, first, inv = np.unique(key, axis=0, return_index=True, return_inverse=True)
    verts = pts[first]
    tris = inv.reshape(-1)[np.array(tris)]
    c = verts[tris].mean(axis=1)
    nrm = np.cross(verts[tris[:, 1]] - verts[tris[:, 0]], verts[tris[:, 2]] - verts[tris[:, 0]])
    flip = (nrm * c).sum(1) < 0
    tris[flip] = tris[flip][:, [0, 2, 1]]
    return verts, tris.astype(np.uint32)


def _vertex_normals(p, tris):
    fn = np.cross(p[tris[:, 1]] - p[tris[:, 0]], p[tris[:, 2]] - p[tris[:, 0]])
    vn = np.zeros_like(p)
    for k in range(3):
        np.add.at(vn, tris[:, k], fn)
    return vn / np.maximum(np.linalg.norm(vn, axis=1, keepdims=True), 1e-20)


def procedural_hdri(width=2048, height=1024, sun_peak=5e4, sun_sigma_deg=0.5,
                    sun_elev_deg=40.0, sun_azim_deg=60.0) -> np.ndarray:
    """Equirect RGB f32 (row 0 = zenith, DomeLight.h:63-69 convention): sky gradient + Gaussian sun."""
    r = (np.arange(height, dtype=np.float64) + 0.5) / height * math.pi
    c = (np.arange(width, dtype=np.float64) + 0.5) / width * 2 * math.pi
    th, ph = np.meshgrid(r, c, indexing="ij")
    d = np.stack([np.sin(th) * np.sin(ph), np.cos(th), np.sin(th) * np.cos(ph)], -1)
    y = d[..., 1]
    zen = np.array([0.22, 0.42, 0.95])
    hor = np.array([0.95, 0.95, 1.0])
    gnd = np.array([0.25, 0.22, 0.2])
    s = np.clip(y, 0, 1)[..., None] ** 0.5
    sky = np.where((y >= 0)[..., None], hor * (1 - s) + zen * s, gnd * 0.5)
    el, az = math.radians(sun_elev_deg), math.radians(sun_azim_deg)
    sd = np.array([math.cos(el) * math.sin(az), math.sin(el), math.cos(el) * math.cos(az)])
    ang = np.arccos(np.clip((d * sd).sum(-1), -1, 1))
    sig = math.radians(sun_sigma_deg)
    sun = sun_peak * np.exp(-0.5 * (ang / sig) ** 2)
    img = sky + sun[..., None] * np.array([1.0, 0.96, 0.88])
    return img.astype(np.float32)


def ggx_spheres(width=1920, height=1080, spp=16, max_depth=8, seed=0, grid=32, freq=7,
                hdri_size=(2048, 1024), ground_tilt_deg=0.0) -> Scene:
    """The GGX-spheres scene: grid x grid displaced geodesic spheres (20*freq^2 triangles each; freq=7 gives
    1,003,520 sphere triangles for a 32x32 grid) with per-sphere standard_surface on a 2-triangle
    ground, lit by a procedural HDRI dome with importance sampling.

    The flat ground at y = 0 sits alone in a BVH leaf whose box has no height, and the reference's
    slab test rejects such a box (t1 <= t0, BVHBackend.cpp:249-262): no ray hits the ground.
    ground_tilt_deg > 0 turns the ground about the grid's centre line (x axis) by that angle, so its
    box has height and the ground is seen, as the workload describes it."""
    rng = np.random.default_rng(seed)
    unit, tris = geodesic_sphere(freq)
    mb = _MeshBuilder()
    mats = [standard_surface((0.5, 0.5, 0.5), metalness=0.0, roughness=0.8)]
    spacing = 1.6
    for gi in range(grid):
        for gj in range(grid):
            k = len(mats)
            rad = rng.uniform(0.3, 0.6)
            cx = spacing * gi + rng.uniform(-0.15, 0.15)
            cz = spacing * gj + rng.uniform(-0.15, 0.15)
            noise = rng.uniform(-1.0, 1.0, size=len(unit)) * 0.05
            p = unit * (rad + noise)[:, None] + np.array([cx, rad, cz])
            metal = 1.0 if rng.uniform() < 0.3 else 0.0
            rough = rng.uniform(0.05, 1.0)
            base = rng.uniform(0.0, 1.0, size=3)
            coat = rng.uniform(0.0, 0.5)
            mats.append(standard_surface(tuple(base), metalness=metal, roughness=rough, ior=1.5, coat=coat))
            uv = np.stack([np.arctan2(unit[:, 0], unit[:, 2]) / (2 * math.pi) + 0.5,
                           np.arccos(np.clip(unit[:, 1], -1, 1)) / math.pi], -1)
            mb.add(p.astype(np.float32), tris.reshape(-1), k, normals=_vertex_normals(p, tris).astype(np.float32),
                   uvs=uv.astype(np.float32))
    g = 200.0
    ext = spacing * (grid - 1)
    c0 = ext * 0.5
    ty = math.tan(math.radians(ground_tilt_deg)) * g       # height of the ground's z edges
    gn = (0.0, math.cos(math.radians(ground_tilt_deg)), -math.sin(math.radians(ground_tilt_deg)))
    mb.add([(c0 - g, -ty, c0 - g), (c0 - g, ty, c0 + g), (c0 + g, ty, c0 + g), (c0 + g, -ty, c0 - g)],
           [0, 1, 2, 0, 2, 3], 0, normals=[gn] * 4, uvs=[(0, 0), (0, 1), (1, 1), (1, 0)])
    # the ground is the last mesh; material 0 is the ground's
    env = procedural_hdri(*hdri_size)
    lights = [dome_light(1.0)]
    cx = ext * 0.5
    return mb.scene("ggx_spheres", materials=mats, lights=lights, env_light=0, env_pixels=env,
                    cam_from=(cx, 10.0, -18.0), cam_at=(cx, 0.0, 22.0), cam_up=(0.0, 1.0, 0.0),
                    vfov=40.0, width=width, height=height, spp=spp, max_depth=max_depth, seed=seed)


def inverse_scene(width=512, height=512, spp=8, max_depth=4, seed=0, grid=32, freq=7, tex=1024,
                  hdri_size=(2048, 1024)) -> Scene:
    """The inverse-rendering scene at its ground truth: the GGX-spheres geometry with a per-triangle base colour
    (U[0.05, 0.95]^3 per triangle, the albedo parameters) and one tex x tex roughness texture
    (U[0.05, 1] per texel, sampled through each sphere's UVs) shared by every sphere material."""
    sc = ggx_spheres(width=width, height=height, spp=spp, max_depth=max_depth, seed=seed, grid=grid, freq=freq,
                     hdri_size=hdri_size)
    rng = np.random.default_rng(seed + 7919)
    sc.name = "inverse"
    sc.tri_albedo = rng.uniform(0.05, 0.95, size=(sc.num_triangles, 3)).astype(np.float32)
    sc.textures = [rng.uniform(0.05, 1.0, size=(tex, tex)).astype(np.float32)]
    for m in sc.materials[1:]:
        m.roughness_texture = 0
    return sc


def perturbed_params(sc: Scene, seed=1, sigma=0.1):
    """Initial parameters of the inverse-rendering scene: ground truth + N(0, sigma), clamped to [0, 1]."""
    rng = np.random.default_rng(seed)
    alb = np.clip(sc.tri_albedo + rng.normal(0.0, sigma, size=sc.tri_albedo.shape), 0.0, 1.0).astype(np.float32)
    tex = np.clip(sc.textures[0] + rng.normal(0.0, sigma, size=sc.textures[0].shape), 0.0, 1.0).astype(np.float32)
    return alb, tex


def small_inverse(width=32, height=24, spp=2, max_depth=4, seed=0) -> Scene:
    """A few-hundred-triangle instance of the inverse-rendering scene for oracle-speed gradient tests."""
    sc = inverse_scene(width=width, height=height, spp=spp, max_depth=max_depth, seed=seed, grid=3, freq=3,
                       tex=16, hdri_size=(256, 128))
    c = 1.6                                         # centre of the 3x3 grid
    sc.cam_from, sc.cam_at = (c, 2.2, c - 3.4), (c, 0.3, c)
    return sc


@dataclass
class ViewSpec:
    frm: tuple
    at: tuple
    width: int
    height: int
    spp: int


def batch_views(sc: Scene, n=64, seed=0, resolutions=((256, 256), (512, 512), (1280, 720), (1920, 1080)),
                spps=(1, 4, 16, 64)) -> list:
    """`n` views of a grid scene for the batched form: resolution and spp drawn from the given sets,
    cameras on a seeded orbit (radius U[28,40], height U[6,16]) looking at the grid centre."""
    rng = np.random.default_rng(seed)
    c = float(sc.cam_from[0])                      # the sphere grid is square: centre (c, 0, c)
    out = []
    for _ in range(n):
        w, h = resolutions[rng.integers(len(resolutions))]
        phi = rng.uniform(0.0, 2.0 * math.pi)
        r = rng.uniform(28.0, 40.0)
        frm = (c + r * math.cos(phi), rng.uniform(6.0, 16.0), c + r * math.sin(phi))
        out.append(ViewSpec(frm, (c, 0.0, c), int(w), int(h), int(spps[rng.integers(len(spps))])))
    return out


def small_ggx(width=96, height=64, spp=4, max_depth=8, seed=0, grid=3, freq=3) -> Scene:
    """A few-hundred-triangle instance of the GGX-spheres scene for oracle-speed parity tests."""
    return ggx_spheres(width=width, height=height, spp=spp, max_depth=max_depth, seed=seed, grid=grid,
                       freq=freq, hdri_size=(256, 128))


def triangle_soup(n=2000, seed=1) -> Scene:
    """Random triangles in a unit cube (BVH / traversal parity: many overlaps and near-ties)."""
    rng = np.random.default_rng(seed)
    c = rng.uniform(0, 1, size=(n, 1, 3))
    p = (c + rng.normal(0, 0.03, size=(n, 3, 3))).reshape(-1, 3).astype(np.float32)
    mb = _MeshBuilder()
    mb.add(p, np.arange(3 * n, dtype=np.uint32), 0)
    return mb.scene("soup", materials=[lambert((0.5, 0.5, 0.5))], lights=[], env_radiance=(1.0, 1.0, 1.0))


def hair_sphere(width=1024, height=1024, spp=32, max_depth=16, seed=0, strands=200_000, segments=4,
                hdri_size=(1024, 512), scalp_freq=8) -> Scene:
    """The hair sphere: `strands` cubic-Bezier strands of `segments` segments (shared end points) rooted
    at Fibonacci points of a unit sphere, grown along the normal with length U[0.3, 0.6] and a
    seed-driven curl, width tapering 0.004 -> 0.001; per-strand eumelanin U[0.2, 1.5] encoded as
    the strand colour the hair material turns back into sigma_a (MarschnerHair.h:430-450); one
    1x1 area light plus a dome."""
    rng = np.random.default_rng(seed)
    mb = _MeshBuilder()
    mats = [lambert((0.05, 0.04, 0.03)), emissive((0.0, 0.0, 0.0), (6.0, 6.0, 6.0)), hair()]
    unit, tris = geodesic_sphere(scalp_freq)
    mb.add(unit.astype(np.float32), tris.reshape(-1), 0, normals=unit.astype(np.float32))
    y = 2.5
    _quad(mb, (-0.5, y, -0.5), (0.5, y, -0.5), (0.5, y, 0.5), (-0.5, y, 0.5), (0, -1, 0), 1)
    mb.add(np.zeros((0, 3), np.float32), np.zeros(0, np.uint32), 2)      # mesh slot 2: the hair material
    n = strands
    i = np.arange(n) + 0.5
    polar = np.arccos(1 - 2 * i / n)
    azim = np.pi * (1 + 5 ** 0.5) * i
    root = np.stack([np.cos(azim) * np.sin(polar), np.cos(polar), np.sin(azim) * np.sin(polar)], 1)
    ref = np.where(np.abs(root[:, 1:2]) < 0.9, np.array([[0.0, 1.0, 0.0]]), np.array([[1.0, 0.0, 0.0]]))
    t1 = np.cross(root, ref)
    t1 /= np.linalg.norm(t1, axis=1, keepdims=True)
    t2 = np.cross(root, t1)
    length = rng.uniform(0.3, 0.6, n)
    freq = rng.uniform(1.0, 3.0, n)
    amp = rng.uniform(0.02, 0.08, n)
    ph1, ph2 = rng.uniform(0, 2 * np.pi, (2, n))
    s = np.linspace(0.0, 1.0, 3 * segments + 1)
    ang = 2 * np.pi * freq[:, None] * s[None, :]
    curl = (t1[:, None, :] * (np.sin(ang + ph1[:, None]) - np.sin(ph1[:, None]))[..., None]
            + t2[:, None, :] * (np.cos(ang + ph2[:, None]) - np.cos(ph2[:, None]))[..., None]) * amp[:, None, None]
    cvs = root[:, None, :] * (1.0 + length[:, None, None] * s[None, :, None]) + curl * s[None, :, None]
    ce = rng.uniform(0.2, 1.5, n)
    sig = eumelanin_sigma_a(ce)
    color = np.exp(-sig / 2.2)
    lights = [area_light((0.0, y, 0.0), (0.5, 0.0, 0.0), (0.0, 0.0, 0.5), (6.0, 6.0, 6.0)), dome_light(1.0)]
    env = procedural_hdri(*hdri_size, sun_peak=500.0, sun_sigma_deg=2.0)
    sc = mb.scene("hair_sphere", materials=mats, lights=lights, env_light=1, env_pixels=env,
                  cam_from=(0.0, 0.5, -3.6), cam_at=(0.0, 0.1, 0.0), cam_up=(0.0, 1.0, 0.0), vfov=40.0,
                  width=width, height=height, spp=spp, max_depth=max_depth, seed=seed)
    sc.strand_cv_offset = (np.arange(n + 1) * (3 * segments + 1)).astype(np.uint32)
    sc.curve_cvs = cvs.reshape(-1, 3).astype(np.float32)
    sc.strand_width = np.tile(np.array([[0.004, 0.001]], np.float32), (n, 1))
    sc.strand_mesh_index = np.full(n, 2, np.uint32)
    sc.strand_color = color.astype(np.float32)
    return sc


def small_hair(width=64, height=64, spp=2, max_depth=16, seed=0) -> Scene:
    """A few-hundred-strand instance of the hair sphere for oracle-speed parity tests."""
    return hair_sphere(width=width, height=height, spp=spp, max_depth=max_depth, seed=seed, strands=400,
                       hdri_size=(128, 64), scalp_freq=3)

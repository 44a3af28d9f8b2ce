"""USD-lite ingest (ahip_usd_load, anacapa_amd/csrc/usd_lite.cpp) against known answers.

The reference reads USD through OpenUSD (src/scene/usd/USDLoader.cpp), which is not installed
here, so the reader is pinned by what the USD specification and the reference loader's own rules
fix independently of this implementation:
  - the reference's scenes/cornell_box.usda (when the checkout is present): positions, normals
    and bindings re-read from the file text with regular expressions, fan triangulation, flat
    normals facing out of the blocks, the lights and the camera its doc string describes;
  - transform composition (xformOpOrder, right-handed rotations, parents, resetXformStack,
    pivots, quaternions, matrices, Z-up), each a point whose image is known in closed form;
  - the loader's rules (USDLoader.cpp): fan triangulation, face-varying primvars, GeomSubset
    splitting, displayColor fallback, binding inheritance, UsdPreviewSurface defaults, light
    intensities, camera selection, and loud refusal of what the device path lacks;
  - the committed tests/golden/usd_cornell.usda against the procedural Cornell box it encodes.
"""
import math
import os
import re

import numpy as np
import pytest

from anacapa_amd import _abi as A
from anacapa_amd import scenes
from anacapa_amd.usd import UsdLoadError, load_usda

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
REF_SCENES = "/root/reference/scenes"
GOLDEN_USDA = os.path.join(ROOT, "tests", "golden", "usd_cornell.usda")


def load(text, w=64, h=64, **kw):
    return load_usda(text=text, width=w, height=h, **kw)


def tris(sc, m):
    """World-space triangles of mesh m, [T,3,3]."""
    vo, i0, i1 = sc.vertex_offset[m], sc.index_offset[m], sc.index_offset[m + 1]
    return sc.positions[sc.indices[i0:i1].reshape(-1, 3) + vo]


def mesh_normals(sc, m):
    return sc.normals[sc.vertex_offset[m]:sc.vertex_offset[m + 1]]


def layer(body, meta='upAxis = "Y"'):
    return f"#usda 1.0\n(\n    {meta}\n)\n\n{body}\n"


def tri_mesh(name, pts=((1, 0, 0), (0, 1, 0), (0, 0, 1)), extra=""):
    p = ", ".join(f"({a}, {b}, {c})" for a, b, c in pts)
    return (f'def Mesh "{name}"\n{{\n    int[] faceVertexCounts = [3]\n    int[] faceVertexIndices = [0, 1, 2]\n'
            f"    point3f[] points = [{p}]\n{extra}}}\n")


def xform(name, ops, order, children="", meta=""):
    lines = "\n".join(f"    {o}" for o in ops)
    order_s = ", ".join(f'"{o}"' for o in order)
    return f'def Xform "{name}"{meta}\n{{\n{lines}\n    uniform token[] xformOpOrder = [{order_s}]\n{children}}}\n'


# ------------------------------------------------------------------------------------------------
# the reference's own scene file
# ------------------------------------------------------------------------------------------------
def _ref_cornell_text():
    path = os.path.join(REF_SCENES, "cornell_box.usda")
    if not os.path.exists(path):
        pytest.skip("reference checkout not present")
    return path, open(path).read()


def _floats(s):
    return [float(x) for x in re.findall(r"-?\d+(?:\.\d*)?(?:e-?\d+)?", s)]


def _regex_meshes(text):
    """(name, points, counts, indices, normals or None, bound material) per Mesh, file order."""
    out = []
    for m in re.finditer(r'def Mesh "(\w+)"(.*?)subdivisionScheme', text, re.S):
        body = m.group(2)
        pts = np.array(_floats(re.search(r"point3f\[\] points = \[(.*?)\]", body).group(1)), np.float32).reshape(-1, 3)
        cnt = [int(x) for x in _floats(re.search(r"faceVertexCounts = \[(.*?)\]", body).group(1))]
        idx = [int(x) for x in _floats(re.search(r"faceVertexIndices = \[(.*?)\]", body).group(1))]
        nm = re.search(r"normal3f\[\] normals = \[(.*?)\]", body)
        nrm = None if nm is None else np.array(_floats(nm.group(1)), np.float32).reshape(-1, 3)
        mat = re.search(r"material:binding = </World/(\w+)>", body).group(1)
        out.append((m.group(1), pts, cnt, idx, nrm, mat))
    return out


def _regex_materials(text):
    mats = {}
    for m in re.finditer(r'def Material "(\w+)"(.*?)token outputs:surface\n', text, re.S):
        body = m.group(2)
        d = re.search(r"diffuseColor = \((.*?)\)", body)
        e = re.search(r"emissiveColor = \((.*?)\)", body)
        mats[m.group(1)] = (_floats(d.group(1)), None if e is None else _floats(e.group(1)))
    return mats


def test_reference_cornell_box_geometry_and_materials():
    path, text = _ref_cornell_text()
    sc = load_usda(path, width=512, height=512)
    meshes = _regex_meshes(text)
    mats = _regex_materials(text)
    assert sc.mesh_names == [m[0] for m in meshes]
    assert sc.num_triangles == 36
    seen = []
    for mi, (name, pts, cnt, idx, nrm, mat) in enumerate(meshes):
        exp, exp_n, start = [], [], 0
        for c in cnt:  # fan from the face's first vertex
            for k in range(c - 2):
                exp.append(pts[[idx[start], idx[start + k + 1], idx[start + k + 2]]])
                if nrm is not None:
                    exp_n += [nrm[idx[start]], nrm[idx[start + k + 1]], nrm[idx[start + k + 2]]]
            start += c
        got = tris(sc, mi)
        assert np.array_equal(got, np.array(exp)), name          # identity transform: exact
        n = mesh_normals(sc, mi)
        if nrm is not None:
            assert np.array_equal(n, np.array(exp_n, np.float32)), name
        else:                                                     # flat normals, facing out
            t = got.astype(np.float64)
            fn = np.cross(t[:, 1] - t[:, 0], t[:, 2] - t[:, 0])
            fn /= np.linalg.norm(fn, axis=1, keepdims=True)
            assert np.allclose(n.reshape(-1, 3, 3), fn[:, None, :], atol=1e-6), name
            centre = pts.astype(np.float64).mean(0)
            assert (np.einsum("ij,ij->i", fn, t.mean(1) - centre) > 0).all(), name
        # material: default Lambert(0.5) first, then materials in first-binding order
        if mat not in seen:
            seen.append(mat)
        m = sc.materials[sc.mesh_material[mi]]
        assert sc.mesh_material[mi] == 1 + seen.index(mat)
        diffuse, emissive = mats[mat]
        assert m.type == A.AHIP_MAT_STANDARD
        assert np.array_equal(np.array(list(m.albedo), np.float32), np.array(diffuse, np.float32))
        assert (m.roughness, m.metalness, m.specular, m.specular_ior, m.coat) == (1.0, 0.0, 0.5, 1.5, 0.0)
        if emissive is None:
            assert m.emission_weight == 0.0
        else:
            assert m.emission_weight == 1.0 and list(m.emission) == emissive
    d = sc.materials[0]
    assert d.type == A.AHIP_MAT_LAMBERT and list(d.albedo) == [0.5, 0.5, 0.5]


def test_reference_cornell_box_lights_and_camera():
    path, _ = _ref_cornell_text()
    sc = load_usda(path, width=512, height=512)
    # RectLight 0.5 x 0.3 at (0, 0.999, 1) turned by rotateX(90): width along x, height along z,
    # then the emissive LightPatch registered over its bounding box (thin along y)
    assert len(sc.lights) == 2
    for l in sc.lights:
        assert l.type == A.AHIP_LIGHT_AREA
        assert np.allclose(list(l.center), [0, 0.999, 1.0], atol=1e-6)
        assert np.allclose(list(l.u_axis), [0.25, 0, 0], atol=1e-6)
        assert np.allclose(list(l.v_axis), [0, 0, 0.15], atol=1e-6)
        assert list(l.Le) == [15.0, 15.0, 15.0]
        n = np.cross(list(l.u_axis), list(l.v_axis))
        assert n[1] < 0 and abs(n[0]) < 1e-9 and abs(n[2]) < 1e-9    # faces down into the box
    # "Camera sits at z=-2.5 looking in +z, 50 deg vertical FOV" (the file's doc string):
    # translate (0,0,-2.5) after rotateY(180); vfov = 2 atan(aperture / 2 / focal)
    assert sc.has_camera
    assert sc.cam_from == (0.0, 0.0, -2.5)
    assert np.allclose(sc.cam_at, (0, 0, -1.5), atol=1e-6)
    assert np.allclose(sc.cam_up, (0, 1, 0), atol=1e-7)
    assert abs(sc.vfov - math.degrees(2 * math.atan(46.631 / 2 / 50))) < 1e-4
    assert abs(sc.vfov - 50.0) < 0.01
    # the vertical aperture is the horizontal one over the film aspect (USDLoader.cpp:1107-1110)
    wide = load_usda(path, width=1024, height=512)
    assert abs(wide.vfov - math.degrees(2 * math.atan(46.631 / 2 / 2 / 50))) < 1e-4


def test_reference_motion_and_composed_scenes_are_refused():
    for name, word in (("cornell_box_motion.usda", "animated"), ("kitchen_set_materials.usda", "references")):
        path = os.path.join(REF_SCENES, name)
        if not os.path.exists(path):
            pytest.skip("reference checkout not present")
        with pytest.raises(UsdLoadError) as e:
            load_usda(path, width=64, height=64)
        assert e.value.status == A.AHIP_E_UNSUPPORTED and word in str(e.value), str(e.value)


# ------------------------------------------------------------------------------------------------
# transforms: points with closed-form images
# ------------------------------------------------------------------------------------------------
def _first_vertex(body, meta='upAxis = "Y"'):
    sc = load(layer(body, meta))
    return tris(sc, 0)[0].astype(np.float64)


def test_xform_op_order_applies_last_op_first():
    # [translate, rotateZ, scale]: (1,0,0) -> scale 2 -> (2,0,0) -> rotateZ 90 -> (0,2,0) -> +(10,0,0)
    body = xform("X", ["double3 xformOp:translate = (10, 0, 0)", "float xformOp:rotateZ = 90",
                       "float3 xformOp:scale = (2, 2, 2)"],
                 ["xformOp:translate", "xformOp:rotateZ", "xformOp:scale"], tri_mesh("M"))
    assert np.allclose(_first_vertex(body)[0], (10, 2, 0), atol=1e-6)


def test_rotate_xyz_turns_about_x_first():
    # rotateXYZ (90, 0, 90): X first. (1,0,0) -> (1,0,0) -> (0,1,0); (0,1,0) -> (0,0,1) -> (0,0,1)
    body = xform("X", ["double3 xformOp:rotateXYZ = (90, 0, 90)"], ["xformOp:rotateXYZ"], tri_mesh("M"))
    v = _first_vertex(body)
    assert np.allclose(v[0], (0, 1, 0), atol=1e-6)
    assert np.allclose(v[1], (0, 0, 1), atol=1e-6)
    # rotateZYX with the same angles: Z first. (0,1,0) -> (-1,0,0) -> (-1,0,0)
    body = xform("X", ["double3 xformOp:rotateZYX = (90, 0, 90)"], ["xformOp:rotateZYX"], tri_mesh("M"))
    assert np.allclose(_first_vertex(body)[1], (-1, 0, 0), atol=1e-6)


def test_parent_transforms_and_reset_xform_stack():
    child = xform("C", ["float3 xformOp:scale = (2, 2, 2)"], ["xformOp:scale"], tri_mesh("M"))
    body = xform("P", ["double3 xformOp:translate = (0, 5, 0)"], ["xformOp:translate"], child)
    assert np.allclose(_first_vertex(body)[0], (2, 5, 0), atol=1e-6)      # child first, then parent
    child = xform("C", ["float3 xformOp:scale = (2, 2, 2)"], ["!resetXformStack!", "xformOp:scale"], tri_mesh("M"))
    body = xform("P", ["double3 xformOp:translate = (0, 5, 0)"], ["xformOp:translate"], child)
    assert np.allclose(_first_vertex(body)[0], (2, 0, 0), atol=1e-6)      # parent ignored


def test_orient_matrix_and_inverted_pivot():
    h = math.sqrt(0.5)
    body = xform("X", [f"quatf xformOp:orient = ({h}, 0, 0, {h})"], ["xformOp:orient"], tri_mesh("M"))
    assert np.allclose(_first_vertex(body)[0], (0, 1, 0), atol=1e-6)      # 90 deg about z
    body = xform("X", ["matrix4d xformOp:transform = ( (1, 0, 0, 0), (0, 1, 0, 0), (0, 0, 1, 0), (3, 4, 5, 1) )"],
                 ["xformOp:transform"], tri_mesh("M"))
    assert np.allclose(_first_vertex(body)[0], (4, 4, 5), atol=1e-6)      # translation in the last row
    # rotate 90 deg about z around the pivot (1,1,1): (1,0,0) - pivot = (0,-1,-1) -> (1,0,-1) -> (2,1,0)
    body = xform("X", ["double3 xformOp:translate:pivot = (1, 1, 1)", "float xformOp:rotateZ = 90"],
                 ["xformOp:translate:pivot", "xformOp:rotateZ", "!invert!xformOp:translate:pivot"], tri_mesh("M"))
    assert np.allclose(_first_vertex(body)[0], (2, 1, 0), atol=1e-6)


def test_z_up_stage_is_turned_to_y_up():
    v = _first_vertex(tri_mesh("M"), meta='upAxis = "Z"')
    assert np.allclose(v, [(1, 0, 0), (0, 0, -1), (0, 1, 0)], atol=0)     # (x, y, z) -> (x, z, -y)


def test_normals_use_the_inverse_transpose():
    n = math.sqrt(0.5)
    mesh = tri_mesh("M", extra=f"    normal3f[] normals = [({n}, {n}, 0), ({n}, {n}, 0), ({n}, {n}, 0)]\n")
    body = xform("X", ["float3 xformOp:scale = (2, 1, 1)"], ["xformOp:scale"], mesh)
    sc = load(layer(body))
    exp = np.array([0.5, 1.0, 0.0]) / math.sqrt(1.25)                     # (nx/sx, ny/sy, nz/sz), normalized
    assert np.allclose(mesh_normals(sc, 0), exp[None, :], atol=1e-6)


# ------------------------------------------------------------------------------------------------
# meshes: triangulation, primvars, subsets, display colours, bindings
# ------------------------------------------------------------------------------------------------
PENTA = [(0, 0, 0), (1, 0, 0), (1.5, 1, 0), (0.5, 2, 0), (-0.5, 1, 0)]


def _penta(extra="", order=(0, 1, 2, 3, 4)):
    p = ", ".join(f"({a}, {b}, {c})" for a, b, c in PENTA)
    return (f'def Mesh "P"\n{{\n    int[] faceVertexCounts = [5]\n    int[] faceVertexIndices = [{", ".join(map(str, order))}]\n'
            f"    point3f[] points = [{p}]\n{extra}}}\n")


def test_fan_triangulation_and_vertex_arrays():
    order = (4, 3, 2, 1, 0)
    sc = load(layer(_penta(order=order)))
    P = np.array(PENTA, np.float32)
    exp = np.array([P[[order[0], order[k + 1], order[k + 2]]] for k in range(3)])
    assert np.array_equal(tris(sc, 0), exp)
    assert list(sc.indices) == list(range(9))                  # three fresh vertices per triangle
    assert sc.normal_count[0] == 9 and sc.uv_count[0] == 9
    assert not sc.uvs.any()                                    # no st primvar: zeros
    assert np.allclose(sc.normals, [0, 0, -1], atol=1e-7)      # flat normal of a clockwise fan


def test_vertex_and_face_varying_normals_and_uvs():
    order = (4, 3, 2, 1, 0)
    nrm = [(0, 0, 1), (0, 1, 0), (1, 0, 0), (0, 0, -1), (0, -1, 0)]
    ns = ", ".join(f"({a}, {b}, {c})" for a, b, c in nrm)
    fan = [(0, k + 1, k + 2) for k in range(3)]                # face-vertex positions per triangle
    for interp in ("vertex", "faceVarying"):
        extra = (f'    normal3f[] normals = [{ns}] (\n        interpolation = "{interp}"\n    )\n'
                 '    texCoord2f[] primvars:st = [(0, 0), (1, 0), (1, 1)] (\n        interpolation = "faceVarying"\n    )\n'
                 "    int[] primvars:st:indices = [2, 1, 0, 1, 2]\n")
        sc = load(layer(_penta(extra, order)))
        for t, f in enumerate(fan):
            for k in range(3):
                ni = f[k] if interp == "faceVarying" else order[f[k]]
                assert np.array_equal(sc.normals[3 * t + k], np.array(nrm[ni], np.float32)), (interp, t, k)
                st = [(0, 0), (1, 0), (1, 1)][[2, 1, 0, 1, 2][f[k]]]
                assert tuple(sc.uvs[3 * t + k]) == st


def _looks(*mats):
    out = []
    for name, inputs in mats:
        ins = "\n".join(f"            {i}" for i in inputs)
        out.append(f'    def Material "{name}"\n    {{\n'
                   f"        token outputs:surface.connect = </Looks/{name}/S.outputs:surface>\n"
                   f'        def Shader "S"\n        {{\n            uniform token info:id = "UsdPreviewSurface"\n'
                   f"{ins}\n            token outputs:surface\n        }}\n    }}\n")
    return 'def Scope "Looks"\n{\n' + "".join(out) + "}\n"


def _grid_mesh(name, binding=None, subsets=(), display=None):
    """Three unit quads along x: faces 0, 1, 2."""
    pts = [(x, y, 0) for x in range(4) for y in (0, 1)]
    p = ", ".join(f"({a}, {b}, {c})" for a, b, c in pts)
    s = [f'def Mesh "{name}"', "{", "    int[] faceVertexCounts = [4, 4, 4]",
         "    int[] faceVertexIndices = [0, 2, 3, 1, 2, 4, 5, 3, 4, 6, 7, 5]", f"    point3f[] points = [{p}]"]
    if binding:
        s.append(f"    rel material:binding = </Looks/{binding}>")
    if display:
        s.append(f"    color3f[] primvars:displayColor = [({display[0]}, {display[1]}, {display[2]})]")
    for sname, faces, mat in subsets:
        s += [f'    def GeomSubset "{sname}"', "    {", '        uniform token elementType = "face"',
              f"        int[] indices = [{', '.join(map(str, faces))}]"]
        if mat:
            s.append(f"        rel material:binding = </Looks/{mat}>")
        s.append("    }")
    s.append("}")
    return "\n".join(s) + "\n"


def test_geom_subsets_split_meshes():
    looks = _looks(("A", ["color3f inputs:diffuseColor = (1, 0, 0)"]), ("B", ["color3f inputs:diffuseColor = (0, 1, 0)"]))
    sc = load(layer(looks + _grid_mesh("G", binding="A", subsets=[("S", [2, 0], "B")])))
    # the subset becomes mesh 1 (faces in subset order), the mesh keeps face 1; B resolved first
    assert sc.num_meshes == 2 and sc.mesh_names == ["G", "G"]
    assert list(sc.mesh_material) == [2, 1]
    assert list(sc.materials[1].albedo) == [0, 1, 0] and list(sc.materials[2].albedo) == [1, 0, 0]
    assert np.allclose(tris(sc, 0)[:, :, 0].mean(), 1.5)            # face 1 spans x in [1, 2]
    t1 = tris(sc, 1)
    assert len(t1) == 4 and np.allclose(t1[:2, :, 0].mean(), 2.5) and np.allclose(t1[2:, :, 0].mean(), 0.5)
    # a subset with no binding of its own takes the mesh's (bindings are inherited from ancestors)
    sc = load(layer(looks + _grid_mesh("G", binding="A", subsets=[("S", [1], None)])))
    assert sc.num_meshes == 2 and list(sc.mesh_material) == [1, 1]
    # a subset covering every face leaves the mesh empty
    sc = load(layer(looks + _grid_mesh("G", binding="A", subsets=[("S", [0, 1, 2], "B")])))
    assert list(sc.index_offset) == [0, 0, 18]


def test_display_color_fallback_is_shared_lambert():
    body = (_grid_mesh("G1", display=(0.2, 0.4, 0.6)) + _grid_mesh("G2", display=(0.2, 0.4, 0.6)) +
            _grid_mesh("G3", display=(0.9, 0.1, 0.1)) + _grid_mesh("G4"))
    sc = load(layer(body))
    assert list(sc.mesh_material) == [1, 1, 2, 0]
    assert sc.materials[1].type == A.AHIP_MAT_LAMBERT
    assert np.allclose(list(sc.materials[1].albedo), [0.2, 0.4, 0.6], atol=1e-7)


def test_binding_inheritance_and_strength():
    looks = _looks(("A", ["color3f inputs:diffuseColor = (1, 0, 0)"]), ("B", ["color3f inputs:diffuseColor = (0, 1, 0)"]))
    parent = ('def Xform "P"\n{{\n    rel material:binding = </Looks/A>{meta}\n{kids}}}\n')
    kids = _grid_mesh("Inherits") + _grid_mesh("Own", binding="B")
    sc = load(layer(looks + parent.format(meta="", kids=kids)))
    assert [sc.materials[i].albedo[0] for i in sc.mesh_material] == [1, 0]
    stronger = ' (\n        bindMaterialAs = "strongerThanDescendants"\n    )'
    sc = load(layer(looks + parent.format(meta=stronger, kids=kids)))
    assert [sc.materials[i].albedo[0] for i in sc.mesh_material] == [1, 1]


# ------------------------------------------------------------------------------------------------
# materials (USDLoader.cpp:386-405, 526-779)
# ------------------------------------------------------------------------------------------------
def _one_material(inputs, name="M", shader_id="UsdPreviewSurface", extra=""):
    ins = "\n".join(f"        {i}" for i in inputs)
    return layer(f'def Material "{name}"\n{{\n    token outputs:surface.connect = </{name}/S.outputs:surface>\n'
                 f'    def Shader "S"\n    {{\n        uniform token info:id = "{shader_id}"\n{ins}\n'
                 f"        token outputs:surface\n    }}\n{extra}}}\n" +
                 tri_mesh("T", extra=f"    rel material:binding = </{name}>\n"))


def _mat(text):
    sc = load(text)
    return sc.materials[sc.mesh_material[0]]


def test_preview_surface_defaults():
    m = _mat(_one_material(["color3f inputs:diffuseColor = (0.1, 0.2, 0.3)"]))
    assert m.type == A.AHIP_MAT_STANDARD
    assert np.allclose(list(m.albedo), [0.1, 0.2, 0.3], atol=1e-8)
    assert (m.base, m.roughness, m.metalness, m.specular, m.specular_ior) == (1.0, 1.0, 0.0, 0.5, 1.5)
    assert (m.coat, m.transmission, m.emission_weight, m.roughness_texture) == (0.0, 0.0, 0.0, -1)
    assert abs(m.coat_roughness - 0.1) < 1e-8 and list(m.specular_color) == [1, 1, 1]
    m = _mat(_one_material([]))                                   # no diffuseColor: 0.5 grey
    assert list(m.albedo) == [0.5, 0.5, 0.5]


def test_preview_surface_rules():
    m = _mat(_one_material(["float inputs:metallic = 0.8"]))
    assert m.specular == 0.0                                      # metals: no dielectric layer
    m = _mat(_one_material(["float inputs:opacity = 0.25"]))
    assert m.transmission == 0.75 and m.specular == 1.0           # opacity -> transmission, full Fresnel
    m = _mat(_one_material(["float inputs:specular = 0.3", "float inputs:clearcoat = 0.4",
                            "float inputs:clearcoatRoughness = 0.2", "float inputs:ior = 1.33",
                            "float inputs:transmission = 0.1", "float inputs:opacity = 1"]))
    assert np.allclose([m.specular, m.coat, m.coat_roughness, m.specular_ior, m.transmission],
                       [0.3, 0.4, 0.2, 1.33, 0.1], atol=1e-7)
    m = _mat(_one_material(["color3f inputs:emissiveColor = (2, 3, 4)"]))
    assert list(m.emission) == [2, 3, 4] and m.emission_weight == 1.0


def test_materials_without_preview_surface():
    m = _mat(_one_material(["color3f inputs:base_color = (1, 0, 0)"], shader_id="ND_standard_surface_surfaceshader"))
    assert m.type == A.AHIP_MAT_LAMBERT and list(m.albedo) == [0.5, 0.5, 0.5]
    glass = layer('def Material "WindowGlass"\n{\n}\n' + tri_mesh("T", extra="    rel material:binding = </WindowGlass>\n"))
    m = _mat(glass)
    assert m.type == A.AHIP_MAT_STANDARD and m.transmission == 1.0 and m.roughness == 0.0 and m.specular == 1.0
    assert list(m.albedo) == [1, 1, 1]
    plain = layer('def Material "Wood"\n{\n}\n' + tri_mesh("T", extra="    rel material:binding = </Wood>\n"))
    assert _mat(plain).type == A.AHIP_MAT_LAMBERT


def test_node_graph_indirection_and_texture_inputs():
    ng = ('def Material "M"\n{\n    token outputs:surface.connect = </M/NG.outputs:surf>\n'
          '    def NodeGraph "NG"\n    {\n        token outputs:surf.connect = </M/NG/S.outputs:surface>\n'
          '        def Shader "S"\n        {\n            uniform token info:id = "UsdPreviewSurface"\n'
          "            color3f inputs:diffuseColor = (0.25, 0.5, 0.75)\n            token outputs:surface\n"
          "        }\n    }\n}\n" + tri_mesh("T", extra="    rel material:binding = </M>\n"))
    assert list(_mat(layer(ng)).albedo) == [0.25, 0.5, 0.75]
    tex = ('    def Shader "Tex"\n    {{\n        uniform token info:id = "UsdUVTexture"\n{file}'
           "        float3 outputs:rgb\n    }}\n")
    inputs = ["color3f inputs:diffuseColor = (0.2, 0.3, 0.4)", "color3f inputs:diffuseColor.connect = </M/Tex.outputs:rgb>"]
    # a texture node without a file resolves to nothing: the authored constant stays
    m = _mat(_one_material(inputs, extra=tex.format(file="")))
    assert np.allclose(list(m.albedo), [0.2, 0.3, 0.4], atol=1e-8)
    with pytest.raises(UsdLoadError) as e:
        load(_one_material(inputs, extra=tex.format(file="        asset inputs:file = @wood.png@\n")))
    assert e.value.status == A.AHIP_E_UNSUPPORTED and "texture" in str(e.value)


# ------------------------------------------------------------------------------------------------
# lights (USDLoader.cpp:1433-1545, 1748-1826)
# ------------------------------------------------------------------------------------------------
def _lights(body):
    return load(layer(body)).lights


def test_sphere_disk_and_rect_light_intensities():
    c = (1.0, 0.5, 0.25)
    sphere = ('def SphereLight "S"\n{\n    color3f inputs:color = (1, 0.5, 0.25)\n    float inputs:intensity = 5\n'
              "    bool inputs:normalize = 1\n    float inputs:exposure = 3\n    float inputs:radius = 0.2\n"
              '    double3 xformOp:translate = (1, 2, 3)\n    uniform token[] xformOpOrder = ["xformOp:translate"]\n}\n')
    disk = ('def DiskLight "D"\n{\n    color3f inputs:color = (1, 0.5, 0.25)\n    float inputs:intensity = 2\n'
            "    float inputs:exposure = 1\n    float inputs:radius = 0.4\n    double3 xformOp:rotateXYZ = (90, 0, 0)\n"
            '    uniform token[] xformOpOrder = ["xformOp:rotateXYZ"]\n}\n')
    rect = ('def RectLight "R"\n{\n    float inputs:intensity = 7\n    float inputs:exposure = 2\n'
            "    float inputs:width = 2\n    float inputs:height = 4\n}\n")
    s, d, r = _lights(sphere + disk + rect)
    # sphere: normalize -> / pi, exposure not applied; a square of half-size r facing -y
    assert np.allclose(list(s.Le), [x * np.float32(5) * np.float32(1 / 3.14159265) for x in c], rtol=1e-6)
    assert list(s.center) == [1, 2, 3] and list(s.u_axis) == [0.2 * 1.0, 0, 0] or np.allclose(list(s.u_axis), [0.2, 0, 0])
    assert np.allclose(list(s.v_axis), [0, 0, 0.2])
    # disk: intensity * 2^exposure, a square of the disk's area on its local x/y axes (y turned to z)
    assert np.allclose(list(d.Le), [x * 4 for x in c], rtol=1e-6)
    h = 0.4 * 0.8862
    assert np.allclose(list(d.u_axis), [h, 0, 0], atol=1e-6) and np.allclose(list(d.v_axis), [0, 0, h], atol=1e-6)
    # rect: exposure not applied (USDLoader.cpp:1443-1444), white by default, half extents
    assert list(r.Le) == [7, 7, 7]
    assert list(r.u_axis) == [1, 0, 0] and list(r.v_axis) == [0, 2, 0]


def test_emissive_meshes_become_area_lights():
    looks = _looks(("E", ["color3f inputs:emissiveColor = (3, 2, 1)"]))
    wall = ('def Mesh "W"\n{\n    int[] faceVertexCounts = [4]\n    int[] faceVertexIndices = [0, 1, 2, 3]\n'
            "    point3f[] points = [(2, 0, 0), (2, 1, 0), (2, 1, 3), (2, 0, 3)]\n"
            "    rel material:binding = </Looks/E>\n}\n")
    (l,) = _lights(looks + wall)
    assert list(l.center) == [2, 0.5, 1.5] and list(l.Le) == [3, 2, 1]
    assert list(l.u_axis) == [0, 0.5, 0] and list(l.v_axis) == [0, 0, 1.5]   # thin along x


def test_distant_light_is_directional():
    """DistantLight (USDLoader.cpp:1553-1575): the direction toward the light is the prim's local +Z in
    world space (rotated here by 90 degrees about x: +Z turns to -Y), radiance colour x intensity x
    2^exposure, a hard direction; with upAxis Z the direction is turned to Y-up like every light axis."""
    body = ('def DistantLight "Sun"\n{\n    color3f inputs:color = (1, 0.5, 0.25)\n    float inputs:intensity = 3\n'
            "    float inputs:exposure = 1\n    double3 xformOp:rotateXYZ = (90, 0, 0)\n"
            '    uniform token[] xformOpOrder = ["xformOp:rotateXYZ"]\n}\n')
    (l,) = _lights(body)
    assert l.type == A.AHIP_LIGHT_DIRECTIONAL and l.half_angle == 0.0
    assert np.allclose(list(l.Le), [6, 3, 1.5], rtol=1e-6)
    assert np.allclose(list(l.direction), [0, -1, 0], atol=1e-6)
    (z,) = load(layer('def DistantLight "Sun"\n{\n    float inputs:intensity = 2\n}\n', meta='upAxis = "Z"')).lights
    assert list(z.Le) == [2, 2, 2]
    assert np.allclose(list(z.direction), [0, 1, 0], atol=1e-6)   # local +Z is up in a Z-up layer


def test_distant_light_scene_sphere():
    """Once every mesh is loaded the loader gives its directional lights the scene's bounding sphere
    (USDLoader.cpp:1730-1744): the bounds' centroid and 1.5 x the half diagonal, where BDPT's light
    subpaths start."""
    body = tri_mesh("T", pts=((0, 0, 0), (4, 0, 0), (0, 2, 6))) + 'def DistantLight "Sun"\n{\n}\n'
    (l,) = load(layer(body)).lights
    assert list(l.scene_center) == [2, 1, 3]
    assert l.scene_radius == pytest.approx(0.5 * 1.5 * np.sqrt(16 + 4 + 36), rel=1e-6)


@pytest.mark.parametrize("kind", ["DomeLight"])
def test_unsupported_lights_are_refused(kind):
    with pytest.raises(UsdLoadError) as e:
        load(layer(f'def {kind} "L"\n{{\n    float inputs:intensity = 1\n}}\n'))
    assert e.value.status == A.AHIP_E_UNSUPPORTED and kind in str(e.value)


# ------------------------------------------------------------------------------------------------
# camera selection (USDLoader.cpp:1635-1722)
# ------------------------------------------------------------------------------------------------
def _cam(name, z, extra=""):
    return (f'def Camera "{name}"\n{{\n    float focalLength = 50\n    float horizontalAperture = 20\n{extra}'
            f'    double3 xformOp:translate = (0, 0, {z})\n    uniform token[] xformOpOrder = ["xformOp:translate"]\n}}\n')


def test_camera_selection_order():
    two = _cam("A", 1) + _cam("B", 2)
    assert load(layer(two)).cam_from == (0, 0, 1)                                 # first camera
    rs = 'def RenderSettings "RS"\n{\n    rel camera = </B>\n}\n'
    assert load(layer(two + rs)).cam_from == (0, 0, 2)                           # RenderSettings
    assert load(layer(two + rs), camera_path="/A").cam_from == (0, 0, 1)         # explicit path
    sc = load(layer(two), w=200, h=100)
    assert abs(sc.vfov - math.degrees(2 * math.atan(10 / 2 / 50))) < 1e-4
    assert sc.cam_at == (0, 0, 0) and sc.cam_up == (0, 1, 0)                      # looks down -z
    nocam = load(layer(tri_mesh("T")))
    assert not nocam.has_camera


def test_camera_under_a_parent_and_thin_lens():
    body = xform("Rig", ["double3 xformOp:translate = (5, 0, 0)"], ["xformOp:translate"], _cam("C", 3))
    sc = load(layer(body))
    assert sc.cam_from == (5, 0, 3) and sc.cam_at == (5, 0, 2)
    with pytest.raises(UsdLoadError) as e:
        load(layer(_cam("C", 0, "    float fStop = 2.8\n    float focusDistance = 3\n")))
    assert e.value.status == A.AHIP_E_UNSUPPORTED and "thin-lens" in str(e.value)


# ------------------------------------------------------------------------------------------------
# traversal and errors
# ------------------------------------------------------------------------------------------------
def test_traversal_skips_over_class_and_inactive():
    body = (tri_mesh("Kept") + "over " + tri_mesh("Over")[4:] + "class " + tri_mesh("Abstract")[4:] +
            'def Xform "Off" (\n    active = false\n)\n{\n' + tri_mesh("Inside") + "}\n")
    sc = load(layer(body))
    assert sc.mesh_names == ["Kept"]


@pytest.mark.parametrize("text,status,word", [
    ("#usda 1.0\ndef Mesh \"M\" {\n    point3f[] points = [(0, 0, 0)\n}\n", A.AHIP_E_INVALID_ARG, "line"),
    ("def Mesh \"M\" {\n}\n", A.AHIP_E_INVALID_ARG, "#usda"),
    (layer('def Xform "R" (\n    prepend references = @other.usda@\n)\n{\n}\n'), A.AHIP_E_UNSUPPORTED, "references"),
    (layer('def Xform "V" (\n    variants = {\n        string look = "red"\n    }\n)\n{\n}\n'), A.AHIP_E_UNSUPPORTED, "variants"),
    (layer('def Xform "V"\n{\n    variantSet "look" = {\n    }\n}\n'), A.AHIP_E_UNSUPPORTED, "variant"),
    ("#usda 1.0\n(\n    subLayers = [@base.usda@]\n)\n", A.AHIP_E_UNSUPPORTED, "subLayers"),
    (layer('def Mesh "M"\n{\n    int[] faceVertexCounts = [3]\n    int[] faceVertexIndices = [0, 1, 7]\n'
           "    point3f[] points = [(0, 0, 0), (1, 0, 0), (0, 1, 0)]\n}\n"), A.AHIP_E_INVALID_ARG, "out of range"),
])
def test_malformed_and_unsupported_layers(text, status, word):
    with pytest.raises(UsdLoadError) as e:
        load(text)
    assert e.value.status == status and word in str(e.value), str(e.value)


def test_animation_is_refused_unless_static():
    anim = ('def Xform "X"\n{\n    double3 xformOp:translate.timeSamples = {\n        1: (0, 0, 0),\n'
            '        24: (1, 0, 0),\n    }\n    uniform token[] xformOpOrder = ["xformOp:translate"]\n' + tri_mesh("M") + "}\n")
    timed = 'upAxis = "Y"\n    startTimeCode = 1\n    endTimeCode = 24'
    with pytest.raises(UsdLoadError) as e:
        load(layer(anim, meta=timed))
    assert e.value.status == A.AHIP_E_UNSUPPORTED and "motion" in str(e.value)
    held = anim.replace("24: (1, 0, 0)", "24: (0, 0, 0)")
    assert load(layer(held, meta=timed)).num_triangles == 1
    with pytest.raises(UsdLoadError):
        load(layer(tri_mesh("M").replace("point3f[] points =", "point3f[] points.timeSamples = { 0:") + "}\n"))


def test_missing_file():
    with pytest.raises(UsdLoadError) as e:
        load_usda("/nonexistent/scene.usda", width=8, height=8)
    assert e.value.status == A.AHIP_E_INVALID_ARG


# ------------------------------------------------------------------------------------------------
# the committed scene against the procedural box it encodes
# ------------------------------------------------------------------------------------------------
def test_golden_usd_cornell_matches_procedural_box():
    sc = load_usda(GOLDEN_USDA, width=64, height=64)
    ref = scenes.cornell(width=64, height=64)
    assert sc.mesh_names == ["Floor", "Floor", "Ceiling", "BackWall", "LeftWall", "RightWall", "Cube", "Cube",
                             "LampPatch"]
    # walls and lamp: same quads, same fan
    for got_m, ref_m in ((2, 1), (3, 2), (4, 3), (5, 4), (8, 7)):
        assert np.array_equal(tris(sc, got_m), tris(ref, ref_m)), sc.mesh_names[got_m]
    floor = np.concatenate([tris(sc, 0), tris(sc, 1)])
    assert np.allclose(floor[..., 1], 0) and floor[..., 0].min() == 0 and floor[..., 0].max() == 1
    # blocks: unit cube under translate/rotateY/scale, each face (a, d, c, b) where the procedural
    # box has (a, b, c, d): fan triangles (a, d, c), (a, c, b) are its (a, c, d), (a, b, c) reversed
    for got_m, ref_m in ((6, 5), (7, 6)):
        g, r = tris(sc, got_m), tris(ref, ref_m)
        swap = np.arange(12) ^ 1
        assert np.allclose(g, r[swap][:, [0, 2, 1]], atol=1e-6, rtol=0), sc.mesh_names[got_m]
        # flat (short block) and face-varying (tall block, non-uniform scale) normals both give
        # the procedural box's outward face normals
        gn = mesh_normals(sc, got_m).reshape(-1, 3, 3)
        rn = ref.normals[ref.vertex_offset[ref_m]:ref.vertex_offset[ref_m + 1]]
        ri = ref.indices[ref.index_offset[ref_m]:ref.index_offset[ref_m + 1]].reshape(-1, 3)[:, 0]
        assert np.allclose(gn, rn[ri][swap][:, None, :], atol=1e-6)
    assert sc.cam_from == (0.5, 0.5, np.float32(-1.4)) and np.allclose(sc.cam_at, (0.5, 0.5, -0.4), atol=1e-6)
    assert abs(sc.vfov - 40.0) < 1e-4
    # materials: Steel (subset, first bound), White, Lambert ceiling, Red, Lambert green wall, Lamp
    kinds = [sc.materials[i].type for i in sc.mesh_material]
    assert kinds == [2, 2, 0, 2, 2, 0, 2, 2, 2]
    assert list(sc.materials[sc.mesh_material[8]].emission) == [17, 12, 4]
    # lights: the RectLight then the lamp patch over its bounding box
    assert len(sc.lights) == 2
    assert np.allclose(list(sc.lights[0].Le), [10, 9, 8], atol=1e-6)
    assert np.allclose(list(sc.lights[1].u_axis), [0.125, 0, 0]) and list(sc.lights[1].Le) == [17, 12, 4]


def test_golden_usd_cornell_is_current():
    """The committed file is what tests/golden/make_usda.py writes."""
    import importlib.util
    spec = importlib.util.spec_from_file_location("make_usda", os.path.join(ROOT, "tests", "golden", "make_usda.py"))
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    assert open(GOLDEN_USDA).read() == mod.build()


def test_oracle_matches_reference_on_the_usd_box_geometry(oracle, reference):
    """The loaded geometry and both of its area lights (one without geometry) through the oracle
    and the reference's PathIntegrator, with every material made Lambertian / emissive (the
    reference build links no StandardSurface)."""
    sc = load_usda(GOLDEN_USDA, width=32, height=32, spp=4)
    mats = []
    for m in sc.materials:
        if m.emission_weight > 0:
            mats.append(scenes.emissive((0, 0, 0), tuple(x * m.emission_weight for x in m.emission)))
        else:
            mats.append(scenes.lambert(tuple(m.albedo)))
    sc.materials = mats
    o = oracle.scene(sc).render_tile(0, 0, 32, 32, 0, 4)
    r = reference.scene(sc).render_tile(0, 0, 32, 32, 0, 4)
    assert np.array_equal(o[..., 3], r[..., 3])
    g, e = o[..., :3].astype(np.float64), r[..., :3].astype(np.float64)
    err = np.abs(g - e).max(-1) / np.maximum(np.abs(e).max(-1), 1e-3 * e.mean())
    assert (err <= 1e-4).mean() >= 0.99 and abs(g.mean() / e.mean() - 1) < 1e-4

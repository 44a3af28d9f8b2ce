"""Writes tests/golden/usd_cornell.usda, the text USD scene of the USD-lite ingest tests.

The unit Cornell box of ``anacapa_amd.scenes.cornell`` (same wall quads, same two blocks, same
camera) authored the way a DCC exports it, so that loading it exercises the reader's paths:
  - walls as Meshes with vertex normals; two walls bound to UsdPreviewSurface materials, two
    carrying only ``primvars:displayColor`` (the Lambert fallback), one with a GeomSubset that
    gives part of its faces a second material;
  - blocks as a unit cube under Xforms with translate / rotateY / scale: the short block without
    normals (flat normals from the winding), the tall one with face-varying normals under a
    non-uniform scale and its material bound on the parent Xform;
  - a RectLight (rotateX 90 to face down) and an emissive UsdPreviewSurface patch, which the
    loader also registers as an area light;
  - a Camera under translate + rotateXYZ(0, 180, 0) with the aperture of a 40 degree field.
Run: python tests/golden/make_usda.py
"""
from __future__ import annotations

import math
import os

HERE = os.path.dirname(os.path.abspath(__file__))


def tup(v):
    return "(" + ", ".join(f"{x:.9g}" for x in v) + ")"


def arr(vs):
    return "[" + ", ".join(tup(v) for v in vs) + "]"


def material(name, diffuse, roughness=None, metallic=None, emissive=None):
    lines = [f'        def Material "{name}"', "        {",
             f"            token outputs:surface.connect = </World/Looks/{name}/Surface.outputs:surface>",
             "", '            def Shader "Surface"', "            {",
             '                uniform token info:id = "UsdPreviewSurface"',
             f"                color3f inputs:diffuseColor = {tup(diffuse)}"]
    if roughness is not None:
        lines.append(f"                float inputs:roughness = {roughness:.9g}")
    if metallic is not None:
        lines.append(f"                float inputs:metallic = {metallic:.9g}")
    if emissive is not None:
        lines.append(f"                color3f inputs:emissiveColor = {tup(emissive)}")
    lines += ["                token outputs:surface", "            }", "        }", ""]
    return "\n".join(lines)


def quad_mesh(name, pts, normal, binding=None, display=None, subset=None):
    lines = [f'    def Mesh "{name}"' + (' (\n        prepend apiSchemas = ["MaterialBindingAPI"]\n    )' if binding else ""),
             "    {", "        int[] faceVertexCounts = [4]", "        int[] faceVertexIndices = [0, 1, 2, 3]"]
    if binding:
        lines.append(f"        rel material:binding = </World/Looks/{binding}>")
    lines.append(f"        normal3f[] normals = {arr([normal] * 4)}")
    lines.append(f"        point3f[] points = {arr(pts)}")
    if display:
        lines.append(f"        color3f[] primvars:displayColor = [{tup(display)}] (")
        lines.append('            interpolation = "constant"')
        lines.append("        )")
    lines.append('        uniform token subdivisionScheme = "none"')
    if subset:
        sname, faces, mat = subset
        lines += ["", f'        def GeomSubset "{sname}" (', '            prepend apiSchemas = ["MaterialBindingAPI"]',
                  "        )", "        {", '            uniform token elementType = "face"',
                  '            uniform token familyName = "materialBind"',
                  f"            int[] indices = [{', '.join(str(f) for f in faces)}]",
                  f"            rel material:binding = </World/Looks/{mat}>", "        }"]
    lines += ["    }", ""]
    return "\n".join(lines)


def split_quad_mesh(name, x0, x1, z0, z1, y, normal, binding, subset):
    """A 2x1 grid of quads (two faces) so that a GeomSubset can take one of them."""
    xm = 0.5 * (x0 + x1)
    pts = [(x0, y, z0), (xm, y, z0), (x1, y, z0), (x1, y, z1), (xm, y, z1), (x0, y, z1)]
    sname, faces, mat = subset
    return "\n".join([
        f'    def Mesh "{name}" (', '        prepend apiSchemas = ["MaterialBindingAPI"]', "    )", "    {",
        "        int[] faceVertexCounts = [4, 4]",
        "        int[] faceVertexIndices = [0, 1, 4, 5, 1, 2, 3, 4]",
        f"        rel material:binding = </World/Looks/{binding}>",
        f"        normal3f[] normals = {arr([normal] * 6)}",
        f"        point3f[] points = {arr(pts)}",
        '        uniform token subdivisionScheme = "none"', "",
        f'        def GeomSubset "{sname}" (', '            prepend apiSchemas = ["MaterialBindingAPI"]', "        )",
        "        {", '            uniform token elementType = "face"', '            uniform token familyName = "materialBind"',
        f"            int[] indices = [{', '.join(str(f) for f in faces)}]",
        f"            rel material:binding = </World/Looks/{mat}>", "        }", "    }", ""])


# unit cube [-0.5, 0.5]^3, faces in the order of scenes._block (bottom, top, -z, +z, -x, +x), each
# wound counter-clockwise seen from outside (scenes._block lists the same corners the other way
# round, starting from the same corner), with the outward face normals for face-varying use
CUBE_PTS = [(-0.5, -0.5, -0.5), (0.5, -0.5, -0.5), (0.5, -0.5, 0.5), (-0.5, -0.5, 0.5),
            (-0.5, 0.5, -0.5), (0.5, 0.5, -0.5), (0.5, 0.5, 0.5), (-0.5, 0.5, 0.5)]
CUBE_FACES = [0, 1, 2, 3,  4, 7, 6, 5,  0, 4, 5, 1,  2, 6, 7, 3,  3, 7, 4, 0,  1, 5, 6, 2]
CUBE_NORMALS = [(0, -1, 0), (0, 1, 0), (0, 0, -1), (0, 0, 1), (-1, 0, 0), (1, 0, 0)]


def block(name, lo, hi, deg, binding_on_xform=None, binding_on_mesh=None, face_normals=False):
    c = tuple(0.5 * (a + b) for a, b in zip(lo, hi))
    s = tuple(b - a for a, b in zip(lo, hi))
    xf_meta = ' (\n        prepend apiSchemas = ["MaterialBindingAPI"]\n    )' if binding_on_xform else ""
    lines = [f'    def Xform "{name}"{xf_meta}', "    {"]
    if binding_on_xform:
        lines.append(f"        rel material:binding = </World/Looks/{binding_on_xform}>")
    lines += [f"        double3 xformOp:translate = {tup(c)}",
              f"        float xformOp:rotateY = {deg:.9g}",
              f"        float3 xformOp:scale = {tup(s)}",
              '        uniform token[] xformOpOrder = ["xformOp:translate", "xformOp:rotateY", "xformOp:scale"]', "",
              '        def Mesh "Cube"' + (' (\n            prepend apiSchemas = ["MaterialBindingAPI"]\n        )'
                                          if binding_on_mesh else ""),
              "        {", "            int[] faceVertexCounts = [4, 4, 4, 4, 4, 4]",
              f"            int[] faceVertexIndices = [{', '.join(str(i) for i in CUBE_FACES)}]"]
    if binding_on_mesh:
        lines.append(f"            rel material:binding = </World/Looks/{binding_on_mesh}>")
    if face_normals:
        lines += [f"            normal3f[] normals = {arr([n for n in CUBE_NORMALS for _ in range(4)])} (",
                  '                interpolation = "faceVarying"', "            )"]
    lines += [f"            point3f[] points = {arr(CUBE_PTS)}", '            uniform token subdivisionScheme = "none"',
              "        }", "    }", ""]
    return "\n".join(lines)


def build() -> str:
    y = 0.999
    vfov = 40.0
    focal = 50.0
    aperture = 2.0 * focal * math.tan(math.radians(vfov) / 2.0)
    out = ["#usda 1.0", "(", '    defaultPrim = "World"', "    metersPerUnit = 1", '    upAxis = "Y"', ")", "",
           'def Xform "World"', "{", '    def Scope "Looks"', "    {",
           material("White", (0.73, 0.73, 0.73), roughness=1.0),
           material("Red", (0.63, 0.065, 0.05), roughness=0.9),
           material("Steel", (0.8, 0.8, 0.85), roughness=0.35, metallic=0.9),
           material("Lamp", (0.0, 0.0, 0.0), emissive=(17.0, 12.0, 4.0)),
           "    }", "",
           split_quad_mesh("Floor", 0.0, 1.0, 0.0, 1.0, 0.0, (0, 1, 0), "White", ("Tiles", [1], "Steel")),
           quad_mesh("Ceiling", [(0, 1, 1), (1, 1, 1), (1, 1, 0), (0, 1, 0)], (0, -1, 0),
                     display=(0.73, 0.73, 0.73)),
           quad_mesh("BackWall", [(0, 0, 1), (0, 1, 1), (1, 1, 1), (1, 0, 1)], (0, 0, -1), binding="White"),
           quad_mesh("LeftWall", [(0, 0, 1), (0, 1, 1), (0, 1, 0), (0, 0, 0)], (1, 0, 0), binding="Red"),
           quad_mesh("RightWall", [(1, 0, 0), (1, 1, 0), (1, 1, 1), (1, 0, 1)], (-1, 0, 0),
                     display=(0.14, 0.45, 0.09)),
           block("ShortBlock", (0.13, 0.0, 0.12), (0.43, 0.3, 0.42), -18.0, binding_on_mesh="White"),
           block("TallBlock", (0.53, 0.0, 0.45), (0.83, 0.6, 0.75), 18.0, binding_on_xform="Steel",
                 face_normals=True),
           quad_mesh("LampPatch", [(0.375, y, 0.375), (0.625, y, 0.375), (0.625, y, 0.625), (0.375, y, 0.625)],
                     (0, -1, 0), binding="Lamp"),
           '    def RectLight "KeyLight"', "    {", "        color3f inputs:color = (1, 0.9, 0.8)",
           "        float inputs:height = 0.2", "        float inputs:intensity = 10", "        bool inputs:normalize = 0",
           "        float inputs:width = 0.3", "        double3 xformOp:rotateXYZ = (90, 0, 0)",
           f"        double3 xformOp:translate = (0.5, {y:.9g}, 0.5)",
           '        uniform token[] xformOpOrder = ["xformOp:translate", "xformOp:rotateXYZ"]', "    }", "",
           '    def Camera "Camera"', "    {", "        float2 clippingRange = (0.1, 1000)",
           f"        float focalLength = {focal:.9g}", f"        float horizontalAperture = {aperture:.9g}",
           f"        float verticalAperture = {aperture:.9g}", "        double3 xformOp:rotateXYZ = (0, 180, 0)",
           "        double3 xformOp:translate = (0.5, 0.5, -1.4)",
           '        uniform token[] xformOpOrder = ["xformOp:translate", "xformOp:rotateXYZ"]', "    }", "}", ""]
    return "\n".join(out)


if __name__ == "__main__":
    path = os.path.join(HERE, "usd_cornell.usda")
    with open(path, "w") as f:
        f.write(build())
    print("wrote", path)

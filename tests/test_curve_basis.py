"""Curve-basis conversion (ahip_curves_to_bezier, anacapa_amd/csrc/curve_basis.cpp) and per-CV
strand widths, against what the mathematics fixes.

The reference's converter lives in its Alembic reader (src/scene/alembic/AlembicLoader.cpp:58-306),
which needs the Alembic SDK and is not built here; the conversion is pinned instead by identities:
each Bezier span must trace the same cubic as the uniform B-spline / Catmull-Rom span it replaces
(evaluated with the textbook basis functions, not the conversion matrices), a linear curve must
stay the straight segment with a uniform parameter, spans must share end points, and the width,
skip, colour and transform rules are checked on known answers.  StrandDesc::widthAt
(CurvePool.h:45-54) is pinned on the oracle through its closed-form values; the product's copy is
compared with the oracle through the GPU curve tests.
"""
import numpy as np
import pytest

from anacapa_amd import _abi as A
from anacapa_amd import curves, scenes

T = np.linspace(0.0, 1.0, 9)


def bezier(b, t):
    t = t[:, None]
    return ((1 - t) ** 3) * b[0] + 3 * ((1 - t) ** 2) * t * b[1] + 3 * (1 - t) * t * t * b[2] + t ** 3 * b[3]


def bspline(p, t):   # uniform cubic B-spline span over p[0..3]
    t = t[:, None]
    return (((1 - t) ** 3) * p[0] + (3 * t ** 3 - 6 * t ** 2 + 4) * p[1] + (-3 * t ** 3 + 3 * t ** 2 + 3 * t + 1) * p[2]
            + t ** 3 * p[3]) / 6.0


def catmull_rom(p, t):   # uniform Catmull-Rom span from p[1] to p[2]
    t = t[:, None]
    return 0.5 * ((-t ** 3 + 2 * t ** 2 - t) * p[0] + (3 * t ** 3 - 5 * t ** 2 + 2) * p[1]
                  + (-3 * t ** 3 + 4 * t ** 2 + t) * p[2] + (t ** 3 - t ** 2) * p[3])


def spans(cvs):
    n = (len(cvs) - 1) // 3
    return [cvs[3 * i:3 * i + 4].astype(np.float64) for i in range(n)]


@pytest.mark.parametrize("basis,fn", [("bspline", bspline), ("catmullrom", catmull_rom)])
def test_cubic_bases_trace_the_same_curve(basis, fn):
    rng = np.random.default_rng(5)
    p = rng.normal(size=(7, 3)).astype(np.float32)
    w = rng.uniform(0.001, 0.01, 7).astype(np.float32)
    st = curves.to_bezier(basis, [7], p, widths=w)
    cvs, bw = st["curve_cvs"], st["strand_widths"]
    assert len(cvs) == 3 * 4 + 1 and len(bw) == len(cvs)                 # N-3 spans, shared end points
    for i, b in enumerate(spans(cvs)):
        assert np.allclose(bezier(b, T), fn(p[i:i + 4].astype(np.float64), T), atol=2e-6), i
        wb = bw[3 * i:3 * i + 4].astype(np.float64)[:, None]
        assert np.allclose(bezier(wb, T), fn(w[i:i + 4].astype(np.float64)[:, None], T), atol=1e-8), i
    if basis == "catmullrom":                                              # interpolates the inner CVs
        assert np.array_equal(cvs[::3], p[1:-1])


def test_linear_stays_straight_with_uniform_speed():
    p = np.array([[0, 0, 0], [3, 0, 0], [3, 3, 3]], np.float32)
    st = curves.to_bezier("linear", [3], p, widths=[1.0, 4.0, 7.0])
    for i, b in enumerate(spans(st["curve_cvs"])):
        lin = p[i].astype(np.float64) + T[:, None] * (p[i + 1] - p[i]).astype(np.float64)
        assert np.allclose(bezier(b, T), lin, atol=1e-6)
    assert np.allclose(st["strand_widths"], [1, 2, 3, 4, 5, 6, 7], atol=1e-6)


def test_bezier_passthrough_and_skips():
    p = np.arange(3 * (7 + 5 + 1 + 3), dtype=np.float32).reshape(-1, 3)
    st = curves.to_bezier("bezier", [7, 5, 1, 3], p)
    assert list(st["strand_cv_offset"]) == [0, 7]                          # 3n+1 kept as is
    assert np.array_equal(st["curve_cvs"], p[:7])
    assert st["num_skipped"] == 2                                          # 5 and 3 CVs; 1 CV dropped silently
    st = curves.to_bezier("bspline", [3, 4], np.zeros((7, 3), np.float32))
    assert st["num_skipped"] == 1 and len(st["strand_cv_offset"]) == 2


def test_width_scopes():
    p = np.zeros((5, 3), np.float32)
    p[:, 0] = np.arange(5)
    kw = dict(default_width=0.5, width_scale=2.0)
    per_vertex = curves.to_bezier("bezier", [4, 1], p, widths=[1, 2, 3, 4, 9], **kw)
    assert list(per_vertex["strand_widths"]) == [2, 4, 6, 8]
    per_curve = curves.to_bezier("bezier", [4, 1], p, widths=[3, 9], **kw)
    assert list(per_curve["strand_widths"]) == [6, 6, 6, 6]
    const = curves.to_bezier("bezier", [4, 1], p, widths=[0.25], **kw)
    assert list(const["strand_widths"]) == [0.5] * 4
    for w in (None, [1.0, 2.0, 3.0]):                                     # none, or an unmatched count
        st = curves.to_bezier("bezier", [4, 1], p, widths=w, **kw)
        assert list(st["strand_widths"]) == [0.5] * 4                      # default, unscaled


def test_world_matrix_colors_and_material():
    p = np.array([[0, 0, 0], [1, 0, 0], [1, 1, 0], [0, 1, 0]], np.float32)
    m = np.array([[1, 0, 0, 5], [0, 1, 0, 6], [0, 0, 1, 7], [0, 0, 0, 2]], np.float32)   # column vectors
    st = curves.to_bezier("bezier", [4], p, world=m, colors=[[0.1, 0.2, 0.3]], material_index=7)
    assert np.allclose(st["curve_cvs"], (p + [5, 6, 7]) / 2)
    assert list(st["strand_mesh_index"]) == [7]
    assert np.allclose(st["strand_color"], [[0.1, 0.2, 0.3]])
    st = curves.to_bezier("bezier", [4], p)
    assert list(st["strand_color"][0]) == [1, 1, 1]


def test_unknown_basis_is_refused():
    with pytest.raises(curves.CurveError) as e:
        curves.to_bezier(9, [4], np.zeros((4, 3), np.float32))
    assert e.value.status == A.AHIP_E_UNSUPPORTED


def test_oracle_width_at_closed_forms(oracle):
    f32 = np.float32
    assert oracle.width_at([], 0.3) == f32(0.001)
    assert oracle.width_at([0.7], 0.3) == f32(0.7)
    assert oracle.width_at([2.0, 4.0], 0.25) == f32(2.5)
    w = np.array([1.0, 3.0, 2.0, 6.0, 4.0], np.float32)
    assert oracle.width_at(w, 0.0) == w[0] and oracle.width_at(w, 1.0) == w[-1]
    for k in range(4):                                                      # midpoints of each piece
        assert abs(oracle.width_at(w, (k + 0.5) / 4) - 0.5 * (w[k] + w[k + 1])) < 1e-6
    assert oracle.width_at(w, 0.25) == w[1] and oracle.width_at(w, 0.75) == w[3]


def test_oracle_per_cv_two_widths_equal_root_tip(oracle):
    """Two per-CV diameters are the root/tip form: the same image, bit for bit."""
    sc = scenes.small_hair(width=32, height=32, spp=1)
    a = oracle.scene(sc).render_tile(0, 0, 32, 32, 0, 1)
    n = sc.num_strands
    sc.strand_widths = np.ascontiguousarray(sc.strand_width.reshape(-1))
    sc.strand_width_offset = (2 * np.arange(n + 1)).astype(np.uint32)
    sc.strand_width = None
    b = oracle.scene(sc).render_tile(0, 0, 32, 32, 0, 1)
    assert np.array_equal(a, b)

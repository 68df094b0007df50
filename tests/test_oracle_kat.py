"""The oracle pinned against the reference's own known-answer tests, closed
forms and an independent exact-area computation (CPU only)."""
import ctypes as C
import math

import numpy as np
import pytest

import _oracle as O
from _paths import OCTAGON_EVENODD, fill_rule_shapes, octagon, path_size, rect, squirrel
from rasterize_amd import PathBuilder, Transform, workload_cubics64

EPS = 2.220446049250313e-16
I = Transform.identity()


# ---- src/rasterize.rs:945-1003 test_signed_difference_line -----------------
def test_signed_difference_line_kat():
    img = O.signed_difference_line(5, 2, (0.5, 1.0, 3.5, 0.0))
    a0 = (0.5 * (1.0 / 6.0)) / 2.0
    a1 = ((1.0 / 6.0) + (3.0 / 6.0)) / 2.0
    a2 = ((3.0 / 6.0) + (5.0 / 6.0)) / 2.0
    for x, v in enumerate([-a0, a0 - a1, a1 - a2, a0 - a1, -a0]):
        assert abs(img[0, x] - v) < EPS
    assert abs(img.sum() - (-1.0)) < EPS

    img = O.signed_difference_line(5, 2, (-1.0, 0.0, 1.0, 1.0))
    assert abs(img[0, 0] - 3.0 / 4.0) < EPS and abs(img[0, 1] - 1.0 / 4.0) < EPS

    img = O.signed_difference_line(5, 2, (0.0, -0.5, 2.0, 1.5))
    assert abs(img[0, 0] - 1.0 / 8.0) < EPS
    assert abs(img[0, 1] - (1.0 - 2.0 / 8.0)) < EPS
    assert abs(img[0, 2] - 1.0 / 8.0) < EPS
    assert abs(img[1, 1] - 1.0 / 8.0) < EPS
    assert abs(img[1, 2] - (0.5 - 1.0 / 8.0)) < EPS

    img = O.signed_difference_line(5, 2, (0.1, 0.1, 1.9, 0.9))
    for x, v in enumerate([0.18, 0.44, 0.18]):
        assert abs(img[0, x] - v) < EPS

    img = O.signed_difference_line(5, 2, (0.1, 0.1, 0.9, 0.9))
    assert abs(img[0, 0] - 0.4) < EPS and abs(img[0, 1] - (0.8 - 0.4)) < EPS

    img = O.signed_difference_line(5, 2, (0.5, 0.5, 0.5, 1.75))
    assert abs(img[0, 0] - 0.25) < EPS and abs(img[0, 1] - 0.25) < EPS
    assert abs(img[1, 0] - 3.0 / 8.0) < EPS and abs(img[1, 1] - 3.0 / 8.0) < EPS


def test_total_deposit_is_signed_height():
    """Closed form: a line's deposits in a row sum to dir * dy of that row."""
    rng = np.random.default_rng(7)
    for _ in range(200):
        x0, x1 = rng.uniform(0.0, 30.0, 2)
        y0, y1 = rng.uniform(0.0, 8.0, 2)
        img = O.signed_difference_line(40, 8, (x0, y0, x1, y1))
        assert abs(img.sum() - (y1 - y0)) < 1e-12


# ---- src/rasterize.rs:1005-1063 test_edge_iter (through ActiveEdgeIter) -----
def test_edge_iter_values():
    # Edge (0,0)->(6,2): row 0 yields 1/6, 2/6, 2/6, 1/6 from column 0, so the
    # accumulated winding of row 0 is their running sum, and 1 to the right.
    cov = O.ae_lines([(0.0, 0.0, 6.0, 2.0)], 10, 2, 0)
    exp = np.cumsum([1 / 6, 2 / 6, 2 / 6, 1 / 6])
    assert np.allclose(cov[0, :4], exp, atol=1e-15)
    assert np.allclose(cov[0, 4:], 1.0)
    # second row starts at column 3 and sums to 1
    assert abs(cov[1, -1] - 1.0) < 1e-15 and np.all(cov[1, :3] == 0.0)
    # Edge (1,1)->(4,0): -1/6, -2/6, -2/6, -1/6 from column 1 (NonZero takes |w|)
    cov = O.ae_lines([(1.0, 1.0, 4.0, 0.0)], 10, 1, 0)
    assert np.allclose(cov[0, 1:5], np.cumsum([1 / 6, 2 / 6, 2 / 6, 1 / 6]), atol=1e-15)
    # vertical (0.5,0.5)->(0.5,1.5): 1/4, 1/4 in row 0, sum 0.5 in row 1
    cov = O.ae_lines([(0.5, 0.5, 0.5, 1.5)], 4, 2, 0)
    assert np.allclose(cov[0, :2], [0.25, 0.5], atol=1e-15)
    assert abs(cov[1, -1] - 0.5) < 1e-15


# ---- src/rasterize.rs:1065-1120 test_rasterizer (both rasterizers) ---------
def test_rasterizer_octagon_evenodd():
    p = octagon()
    sd = O.sd_mask(p, I, 9, 4, 1)
    ae = O.ae_mask(p, I, 9, 4, 1)
    assert np.all(np.abs(sd - OCTAGON_EVENODD) < EPS)
    assert np.all(np.abs(ae - OCTAGON_EVENODD) < EPS)


# ---- src/rasterize.rs:1122-1161 test_fill_rule -------------------------------
@pytest.mark.parametrize("which", ["sd", "ae"])
def test_fill_rule(which):
    path = fill_rule_shapes()
    w, h, tr = path_size(path, I, O.path_bbox(path, I))
    mask = O.sd_mask if which == "sd" else O.ae_mask
    y, x0, x1, x2 = 50, 50, 150, 250
    eo = mask(path, tr, w, h, 1)
    assert abs(eo[y, x0]) < 1e-6 and abs(eo[y, x1]) < 1e-6 and abs(eo[y, x2]) < 1e-6
    assert abs(eo.sum() - 13130.0) < 1.0
    nz = mask(path, tr, w, h, 0)
    assert abs(nz[y, x0] - 1.0) < 1e-6 and abs(nz[y, x1] - 1.0) < 1e-6 and abs(nz[y, x2]) < 1e-6
    assert abs(nz.sum() - 16492.5) < 1.0


# ---- src/path.rs:1191-1211 test_flatten; src/path.rs:1097-1105 test_bbox -----
def test_flatten_connected_and_subpath_concat():
    path = squirrel()
    tr = Transform.new_rotate(math.pi / 3.0).pre_translate(-10.0, -20.0)
    lines = O.flatten(path, tr, 0.05, True)
    ref = []
    for i in range(len(path)):
        b, e = path.subpaths[i], path.subpaths[i + 1]
        sub = PathBuilder()
        # the same segments as one closed subpath
        from rasterize_amd import Path
        k0, k1 = path.point_offsets[b], path.point_offsets[e]
        sp = Path(path.kinds[b:e], path.point_offsets[b:e + 1] - k0, path.points[2 * k0:2 * k1],
                  [0, e - b], [1])
        sl = O.flatten(sp, tr, 0.05, False)
        assert np.all(np.abs(sl[1:, :2] - sl[:-1, 2:]) < EPS)  # connected
        ref.append(sl)
    ref = np.concatenate(ref)
    assert ref.shape == lines.shape
    assert np.array_equal(ref, lines)


def test_bbox_squirrel():
    bb = O.path_bbox(squirrel(), I)
    assert abs(bb[0] - 0.25) < EPS and abs(bb[1] - 1.0) < EPS
    assert abs((bb[2] - bb[0]) - 15.75) < EPS and abs((bb[3] - bb[1]) - 14.0) < EPS


# ---- src/curve.rs:1569-1582 test_split ----------------------------------------
def _split(kind, pts, fast):
    pts = np.array(pts, float).reshape(-1)
    out = np.zeros(2 * pts.size)
    O.lib().orc_split(kind, O._dp(pts), fast, O._dp(out))
    return out


def test_split_matches_optimised_split():
    q = [(0.0, 0.0), (8.0, 5.0), (4.0, 0.0)]
    c = [(3.0, 7.0), (2.0, 8.0), (0.0, 3.0), (6.0, 5.0)]
    assert np.array_equal(_split(1, q, 0), _split(1, q, 1))
    assert np.array_equal(_split(2, c, 0), _split(2, c, 1))
    rng = np.random.default_rng(3)
    for _ in range(1000):
        c = rng.uniform(-1e3, 1e3, 8)
        assert np.array_equal(_split(2, c, 0), _split(2, c, 1))


def test_quad_leaf_count_is_power_of_two():
    """Closed form: children of a quad have exactly 1/16 of its flatness when
    the arithmetic is exact, so a quad yields 2^k lines, k = min{k: F/16^k < T}."""
    for scale in [1.0, 4.0, 16.0, 64.0, 256.0, 1024.0]:
        p = PathBuilder().move_to((0.0, 0.0)).quad_to((scale, 2 * scale), (2 * scale, 0.0)).build()
        F = (2 * 2 * scale) ** 2  # |2p1 - p0 - p2|^2
        T = 16 * 0.05 * 0.05
        k = 0
        while F / 16 ** k >= T:
            k += 1
        lines = O.flatten(p, I, 0.05, False)
        assert len(lines) == 2 ** k


def test_cubic_leaf_chords_within_tolerance():
    """Invariant: every flattened line is within the tolerance of the curve
    (sampled), and line endpoints lie on the curve."""
    ws = workload_cubics64(1, 512.0)
    path = ws.paths[0]
    lines = O.flatten(path, I, 0.05, True)
    assert np.all(np.abs(lines[1:, :2] - lines[:-1, 2:]) == 0.0)
    assert np.array_equal(lines[0, :2], lines[-1, 2:])


# ---- src/grad.rs:514-602, src/color.rs:538-554, src/utils.rs:458-491 ----------
def test_spread():
    sp = O.lib().orc_spread_at
    assert abs(sp(2, 0.3) - 0.3) < 1e-6 and abs(sp(2, -0.3) - 0.3) < 1e-6
    assert abs(sp(2, 1.3) - 0.7) < 1e-6 and abs(sp(2, -1.3) - 0.7) < 1e-6
    assert abs(sp(1, 0.3) - 0.3) < EPS and abs(sp(1, -0.3) - 0.7) < EPS


def _paint(kind, stops, **kw):
    pos = np.array([s[0] for s in stops], float)
    cols = np.array([c for s in stops for c in s[1]], np.float32)
    p = O.orc_paint()
    p.kind = kind
    p.units = kw.get("units", 0)
    p.linear_colors = int(kw.get("linear_colors", True))
    p.spread = kw.get("spread", 0)
    for i, v in enumerate(kw.get("tr", (1, 0, 0, 0, 1, 0))):
        p.tr[i] = v
    for f in ("start", "end", "center", "fcenter"):
        if f in kw:
            getattr(p, f)[0], getattr(p, f)[1] = kw[f]
    p.radius = kw.get("radius", 0.0)
    p.fradius = kw.get("fradius", 0.0)
    p.n_stops = len(stops)
    p.stop_positions = pos.ctypes.data_as(C.POINTER(C.c_double))
    p.stop_colors = cols.ctypes.data_as(C.POINTER(C.c_float))
    p._keep = (pos, cols)
    return p


def _at(p, x, y):
    out = (C.c_float * 4)()
    O.lib().orc_paint_at(C.byref(p), x, y, out)
    return list(out)


def test_grad_stops():
    stops = [(0.0, (1, 0, 0, 1)), (0.5, (0, 1, 0, 1)), (1.0, (0, 0, 1, 1))]
    p = _paint(1, stops, start=(0.0, 0.0), end=(1.0, 0.0))
    assert _at(p, -1.0, 0.0) == [1, 0, 0, 1]
    assert _at(p, 0.25, 0.0) == [0.5, 0.5, 0, 1]
    assert _at(p, 0.75, 0.0) == [0, 0.5, 0.5, 1]
    assert _at(p, 2.0, 0.0) == [0, 0, 1, 1]


def test_radial_grad_offset():
    p = _paint(2, [], center=(0.5, 0.0), radius=0.5, fcenter=(0.25, 0.0), fradius=0.1, units=1)
    off = C.c_double()
    O.lib().orc_radial_offset(C.byref(p), 0.25, 0.0, C.byref(off))
    assert off.value < 0.0
    O.lib().orc_radial_offset(C.byref(p), 0.675, 0.0, C.byref(off))
    assert abs(off.value - 0.5) < EPS
    O.lib().orc_radial_offset(C.byref(p), 1.0, 0.0, C.byref(off))
    assert abs(off.value - 1.0) < EPS


def _rgba_lin(rgba):
    out = (C.c_float * 4)()
    O.lib().orc_rgba_to_lin((C.c_uint8 * 4)(*rgba), out)
    return list(out)


def test_linear_grad():
    c0 = _rgba_lin((0x89, 0x15, 0x51, 0x80))
    c1 = _rgba_lin((0xff, 0x27, 0x2d, 0x80))
    c2 = _rgba_lin((0xff, 0x27, 0x2d, 0x00))
    p = _paint(1, [(0.0, c0), (0.5, c1), (1.0, c2)], start=(0.0, 0.0), end=(1.0, 1.0))
    assert _at(p, -0.5, -0.5) == c0 and _at(p, 0.0, 0.0) == c0
    assert _at(p, 1.0, 0.0) == c1 and _at(p, 0.0, 1.0) == c1
    assert _at(p, 1.0, 1.0) == c2 and _at(p, 1.5, 1.5) == c2


def test_lin_and_srgb_round_trip():
    L = O.lib()
    for i in range(255):
        v = i / 255.0
        assert abs(v - L.orc_linear_to_srgb(L.orc_srgb_to_linear(v))) < 1e-4
        assert abs(v - L.orc_srgb_to_linear(L.orc_linear_to_srgb(v))) < 1e-4


def test_rgba_conversion_round_trip():
    lin = (C.c_float * 4)(*_rgba_lin((0xff, 0x80, 0x40, 0x10)))
    out = (C.c_uint8 * 4)()
    O.lib().orc_lin_to_rgba(lin, out)
    assert list(out) == [0xff, 0x80, 0x40, 0x10]


def test_simd_srgb_alpha_lane():
    """x86+simd semantics: l2s/s2l apply to the alpha lane as well."""
    one = (C.c_float * 4)(1.0, 1.0, 1.0, 1.0)
    o = (C.c_float * 4)()
    O.lib().orc_l2s(one, o)
    assert abs(o[3] - 0.99996865) < 2e-7
    O.lib().orc_s2l(one, o)
    assert abs(o[3] - 1.0008736) < 2e-7


def test_quadratic_solve():
    out = np.zeros(2)
    n = O.lib().orc_quadratic_solve(1.0, -5.0, 6.0, O._dp(out))
    assert n == 2 and sorted(out) == [2.0, 3.0]
    n = O.lib().orc_quadratic_solve(1.0, -6.0, 9.0, O._dp(out))
    assert n == 1 and out[0] == 3.0
    assert O.lib().orc_quadratic_solve(1.0, 3.0, 5.0, O._dp(out)) == 0
    n = O.lib().orc_quadratic_solve(0.0, 5.0, 10.0, O._dp(out))
    assert n == 1 and out[0] == -2.0


def test_fill_rule_alpha():
    a = O.lib().orc_alpha_from_winding
    assert a(0, 0.5) == 0.5 and a(0, -2.0) == 1.0 and a(0, 5e-7) == 0.0
    assert a(1, 0.25) == 0.25 and a(1, 1.75) == 0.25 and a(1, 2.0) == 0.0 and a(1, -1.0) == 1.0


# ---- independent exact area --------------------------------------------------
@pytest.mark.parametrize("seed", [1, 2, 3])
def test_accumulation_equals_exact_winding_area(seed):
    """The restated accumulation (before the fill rule) equals the exact
    winding-weighted area per pixel (Sutherland-Hodgman + shoelace)."""
    ws = workload_cubics64(seed, 40.0)
    path = ws.paths[0]
    tr = Transform.new_translate(4.0, 4.0)
    W = H = 48
    acc = O.sd_mask(path, tr, W, H, -1)
    ex = O.exact_winding_area(path, tr, W, H)
    assert np.max(np.abs(acc - ex)) < 1e-9


def test_nonzero_simple_polygon_equals_exact_coverage():
    """For a simple polygon NonZero coverage is the exact filled area."""
    b = PathBuilder().move_to((3.3, 2.1)).line_to((20.7, 5.2)).line_to((17.1, 19.9))
    path = b.line_to((4.2, 14.4)).close().build()
    m = O.sd_mask(path, I, 24, 24, 0)
    ex = O.exact_fill_coverage(path, I, 24, 24, 0)
    assert np.max(np.abs(m - ex)) < 1e-9
    assert abs(m.sum() - ex.sum()) < 1e-9


def test_translation_by_whole_pixels_shifts_mask():
    ws = workload_cubics64(4, 30.0)
    p = ws.paths[0]
    a = O.sd_mask(p, Transform.new_translate(2.0, 3.0), 40, 40, 0)
    b = O.sd_mask(p, Transform.new_translate(7.0, 8.0), 40, 40, 0)
    assert np.max(np.abs(a[3:33, 2:32] - b[8:38, 7:37])) < 1e-12


def test_nan_is_an_error():
    p = PathBuilder().move_to((0.0, 0.0)).line_to((float("nan"), 1.0)).line_to((2.0, 2.0)).close().build()
    with pytest.raises(O.OracleError):
        O.sd_mask(p, I, 4, 4, 0)


def test_arc_expansion_product_matches_oracle():
    from rasterize_amd import _abi
    import ctypes
    rng = np.random.default_rng(11)
    for _ in range(300):
        src, dst = rng.uniform(0, 100, 2), rng.uniform(0, 100, 2)
        radii = rng.uniform(1, 60, 2)
        rot, large, sweep = rng.uniform(0, 360), int(rng.integers(2)), int(rng.integers(2))
        n, ref = O.arc_to(src, radii, rot, large, sweep, dst)
        out = np.zeros(32)
        dp = lambda a: a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))
        m = _abi.lib().rzw_arc_to(dp(np.array(src)), dp(np.array(radii)), rot, large, sweep,
                                  dp(np.array(dst)), dp(out))
        assert m == n
        assert np.array_equal(out[: 8 * max(n, 0)], ref[: 8 * max(n, 0)])

"""Pins for the oracle functions the reference's own tests leave open (CPU):
the sRGB polynomials against the exact sRGB transfer functions (and s2l's
coefficients against the fit the reference's comment describes), the
elliptic-arc expansion against the SVG endpoint-to-centre conversion, and
Scene::render's layer geometry against the reference's test_scene_view.
Also: the oracle's multi-threaded scene and atlas forms against its
single-threaded ones, and its forward-mode derivative against its own
finite differences."""
import ctypes as C
import math

import numpy as np
import pytest

import _oracle as O
import rasterize_amd as rz
from rasterize_amd import PathBuilder, Transform

I = Transform.identity()


def _lanes(fn, xs):
    out = np.zeros(len(xs), np.float32)
    for i in range(0, len(xs), 4):
        chunk = [float(v) for v in xs[i:i + 4]]
        a = (C.c_float * 4)(*(chunk + [0.0] * (4 - len(chunk))))
        b = (C.c_float * 4)()
        fn(a, b)
        out[i:i + len(chunk)] = np.array(b[:len(chunk)])
    return out


# ---- sRGB: src/simd/x86.rs:196-244 --------------------------------------------
def test_s2l_is_the_documented_cubic_fit():
    """s2l above 0.04045 is the cubic np.polynomial.Polynomial.fit of the exact
    sRGB decode over linspace(0.04045, 1, 16000) (the fit src/simd/x86.rs
    states in its comment); below, x / 12.92."""
    xs = np.linspace(0.04045, 1.0, 16000)
    fit = np.polynomial.Polynomial.fit(xs, ((xs + 0.055) / 1.055) ** 2.4, 3)
    x = np.linspace(0.0, 1.0, 40001).astype(np.float32)
    got = _lanes(O.lib().orc_s2l, x).astype(np.float64)
    xd = x.astype(np.float64)
    hi = xd > 0.04045
    assert np.abs(got[hi] - fit(xd[hi])).max() < 1e-6
    assert np.abs(got[~hi] - xd[~hi] / 12.92).max() < 1e-9
    # and so within the fit's own error of the exact decode
    exact = np.where(xd <= 0.04045, xd / 12.92, ((xd + 0.055) / 1.055) ** 2.4)
    assert np.abs(got - exact).max() < 1.3e-3


def test_l2s_tracks_the_exact_srgb_encode():
    """l2s is a fit of the exact sRGB encode in x, x^1/2, x^1/4, x^1/8 above
    0.0031308 and 12.92 x below: within 1e-4 of the exact encode on [0, 1]
    (a sign, a dropped term, a swapped or wrong coefficient or power is off
    by far more), monotone, and continuous at the break."""
    x = np.linspace(0.0, 1.0, 40001).astype(np.float32)
    got = _lanes(O.lib().orc_l2s, x).astype(np.float64)
    xd = x.astype(np.float64)
    exact = np.where(xd <= 0.0031308, 12.92 * xd, 1.055 * xd ** (1 / 2.4) - 0.055)
    err = np.abs(got - exact)
    assert err.max() < 1e-4, err.max()
    lo = xd <= 0.0031308
    assert np.abs(got[lo] - (xd[lo] * np.float32(12.92)).astype(np.float32)).max() < 1e-9
    assert np.all(np.diff(got) >= -1e-7)
    b = np.float32(0.0031308)
    below = _lanes(O.lib().orc_l2s, [b])[0]
    above = _lanes(O.lib().orc_l2s, [np.nextafter(b, np.float32(1))])[0]
    assert abs(float(above) - float(below)) < 1e-4


# ---- EllipArc: src/ellipse.rs:40-96, 159-214 vs the SVG conversion -------------
def _svg_center(src, dst, rx, ry, phi_deg, large, sweep):
    """SVG 1.1 implementation notes F.6.5-F.6.6, written from the spec."""
    phi = math.radians(phi_deg)
    c, s = math.cos(phi), math.sin(phi)
    dx, dy = (src[0] - dst[0]) / 2, (src[1] - dst[1]) / 2
    x1, y1 = c * dx + s * dy, -s * dx + c * dy
    rx, ry = abs(rx), abs(ry)
    lam = x1 * x1 / (rx * rx) + y1 * y1 / (ry * ry)
    if lam > 1:
        rx, ry = rx * math.sqrt(lam), ry * math.sqrt(lam)
    num = rx * rx * ry * ry - rx * rx * y1 * y1 - ry * ry * x1 * x1
    den = rx * rx * y1 * y1 + ry * ry * x1 * x1
    k = math.sqrt(max(num / den, 0.0)) * (-1 if large == sweep else 1)
    cx1, cy1 = k * rx * y1 / ry, -k * ry * x1 / rx
    cx = c * cx1 - s * cy1 + (src[0] + dst[0]) / 2
    cy = s * cx1 + c * cy1 + (src[1] + dst[1]) / 2

    def ang(ux, uy, vx, vy):
        a = math.atan2(ux * vy - uy * vx, ux * vx + uy * vy)
        return a
    t1 = ang(1, 0, (x1 - cx1) / rx, (y1 - cy1) / ry)
    dt = ang((x1 - cx1) / rx, (y1 - cy1) / ry, (-x1 - cx1) / rx, (-y1 - cy1) / ry)
    if not sweep and dt > 0:
        dt -= 2 * math.pi
    elif sweep and dt < 0:
        dt += 2 * math.pi
    return (cx, cy), rx, ry, phi, t1, dt


def _ellipse_at(center, rx, ry, phi, t):
    c, s = math.cos(phi), math.sin(phi)
    x, y = rx * math.cos(t), ry * math.sin(t)
    return np.array([c * x - s * y + center[0], s * x + c * y + center[1]])


def _ellipse_d(rx, ry, phi, t):
    c, s = math.cos(phi), math.sin(phi)
    x, y = -rx * math.sin(t), ry * math.cos(t)
    return np.array([c * x - s * y, s * x + c * y])


def test_arc_expansion_against_svg_center_parameterisation():
    rng = np.random.default_rng(21)
    worst = 0.0
    checked = 0
    for _ in range(400):
        src, dst = rng.uniform(0, 300, 2), rng.uniform(0, 300, 2)
        radii = rng.uniform(1, 200, 2)
        rot, large, sweep = float(rng.uniform(0, 360)), int(rng.integers(2)), int(rng.integers(2))
        n, out = O.arc_to(src, radii, rot, large, sweep, dst)
        center, rx, ry, phi, t1, dt = _svg_center(src, dst, radii[0], radii[1], rot, large, sweep)
        q = abs(dt) / (math.pi / 2)
        if abs(q - round(q)) < 1e-7:
            # a span at a multiple of 90 degrees (e.g. every scaled-up arc is a
            # half ellipse): the count is a rounding tie
            assert n in (round(q), round(q) + 1), (n, q)
        else:
            assert n == math.ceil(q), (n, q)
        cub = out[: 8 * n].reshape(n, 4, 2)
        scale = max(rx, ry)
        tol = 5e-8 * (1 + scale)  # sqrt(eps) amplification where the radii were scaled up
        assert np.abs(cub[0, 0] - src).max() < tol
        assert np.abs(cub[-1, 3] - dst).max() < tol
        for k in range(n):
            a1 = t1 + (dt / n) * k
            a2 = a1 + dt / n
            p0, p1, p2, p3 = cub[k]
            # the cubic Maisonobe's construction gives (L. Maisonobe, "Drawing an
            # elliptical arc using polylines, quadratic or cubic Bezier curves",
            # 2003): ends on the ellipse, arms alpha times the derivative there,
            # alpha = sin(D) (sqrt(4 + 3 tan^2(D / 2)) - 1) / 3
            D = a2 - a1
            alpha = math.sin(D) * (math.sqrt(4 + 3 * math.tan(D / 2) ** 2) - 1) / 3
            e1, e2 = _ellipse_at(center, rx, ry, phi, a1), _ellipse_at(center, rx, ry, phi, a2)
            d1, d2 = _ellipse_d(rx, ry, phi, a1), _ellipse_d(rx, ry, phi, a2)

            assert np.abs(p0 - e1).max() < tol and np.abs(p3 - e2).max() < tol
            assert np.abs(p1 - (e1 + alpha * d1)).max() < tol
            assert np.abs(p2 - (e2 - alpha * d2)).max() < tol
            # the approximation error at the middle (diagnostic)
            mid = 0.125 * p0 + 0.375 * p1 + 0.375 * p2 + 0.125 * p3
            e = np.linalg.norm(mid - _ellipse_at(center, rx, ry, phi, 0.5 * (a1 + a2)))
            worst = max(worst, e / scale)
        checked += 1
    assert checked == 400
    print(f"arcs checked {checked}, worst midpoint error / radius {worst:.2e}")


def test_arc_circle_midpoint_follows_maisonobe():
    """On a circle the cubic of a span D, with control arms alpha times the
    tangent (L. Maisonobe, "Drawing an elliptical arc using polylines,
    quadratic or cubic Bezier curves", 2003: alpha = sin(D) (sqrt(4 + 3
    tan^2(D / 2)) - 1) / 3), has its t = 1/2 point at distance
    r (cos(D / 2) + 3/4 alpha sin(D / 2)) from the centre, on the bisector."""
    rng = np.random.default_rng(5)
    for _ in range(200):
        r = float(rng.uniform(2, 150))
        c = rng.uniform(-50, 50, 2)
        a0, span = rng.uniform(0, 2 * math.pi), rng.uniform(0.2, 1.95 * math.pi)
        sweep = int(rng.integers(2))
        sgn = 1 if sweep else -1
        src = c + r * np.array([math.cos(a0), math.sin(a0)])
        dst = c + r * np.array([math.cos(a0 + sgn * span), math.sin(a0 + sgn * span)])
        large = int(span > math.pi)
        n, out = O.arc_to(src, (r, r), float(rng.uniform(0, 360)), large, sweep, dst)
        D = span / n
        alpha = math.sin(D) * (math.sqrt(4 + 3 * math.tan(D / 2) ** 2) - 1) / 3
        want = r * (math.cos(D / 2) + 0.75 * alpha * math.sin(D / 2))
        for k, q in enumerate(out[: 8 * n].reshape(n, 4, 2)):
            mid = 0.125 * q[0] + 0.375 * q[1] + 0.375 * q[2] + 0.125 * q[3]
            ang = a0 + sgn * D * (k + 0.5)
            assert abs(np.linalg.norm(mid - c) - want) < 1e-9 * r
            assert abs(math.remainder(math.atan2(*(mid - c)[::-1]) - ang, 2 * math.pi)) < 1e-9


def test_arc_radii_scaled_up_when_too_small():
    """Radii too small to span the endpoints are scaled up (src/ellipse.rs:56-62):
    the centre is then the chord's midpoint and the arc is a half ellipse."""
    src, dst = np.array([10.0, 20.0]), np.array([90.0, 60.0])
    n, out = O.arc_to(src, (5.0, 3.0), 30.0, 0, 1, dst)
    center, rx, ry, phi, t1, dt = _svg_center(src, dst, 5.0, 3.0, 30.0, 0, 1)
    assert np.allclose(center, (src + dst) / 2, atol=1e-6)
    assert abs(abs(dt) - math.pi) < 1e-6  # sqrt(eps)-level: the scaled radii just span
    assert n in (2, 3)
    cub = out[: 8 * n].reshape(n, 4, 2)
    for k in range(n):  # every cubic's ends lie on the scaled-up ellipse
        assert np.abs(cub[k, 3] - _ellipse_at(center, rx, ry, phi, t1 + dt * (k + 1) / n)).max() < 1e-5


# ---- Scene::render layer geometry: src/scene.rs:648-693 test_scene_view ---------
def _scene_view_fill():
    """The reference test's scene: transform "translate(7, 7) rotate(45, 7, 7)
    scale(10)" of a fill of "M0,0 h1 v1 h-1 z", as (path, transform)."""
    def m(a):
        return np.array([[a[0], a[1], a[2]], [a[3], a[4], a[5]], [0, 0, 1.0]])
    c, s = math.cos(math.radians(45)), math.sin(math.radians(45))
    T = m([1, 0, 7, 0, 1, 7])
    R = m([1, 0, 7, 0, 1, 7]) @ m([c, -s, 0, s, c, 0]) @ m([1, 0, -7, 0, 1, -7])
    S = m([10, 0, 0, 0, 10, 0])
    M = T @ R @ S
    tr = Transform(M[0, 0], M[0, 1], M[0, 2], M[1, 0], M[1, 1], M[1, 2])
    path = PathBuilder().move_to((0, 0)).line_to((1, 0)).line_to((1, 1)).line_to((0, 1)).close().build()
    return path, tr


def test_scene_view_layer_geometry_oracle():
    path, tr = _scene_view_fill()
    paint = rz.LinColor(1.0, 0.2158605, 0.05126946, 1.0)  # #ff8040, any colour will do
    fills = [(path, paint.abi(), 0)]
    img, (x, y) = O.scene_render(fills, tr, view=(1.0, 2.0, 23.0, 23.0))
    assert (x, y, img.shape[1], img.shape[0]) == (1, 2, 22, 21)
    img, (x, y) = O.scene_render(fills, tr)
    assert (x, y, img.shape[1], img.shape[0]) == (6, 4, 16, 15)


# ---- the oracle's parallel forms: bit for bit its single-threaded ones -----------
@pytest.mark.parametrize("threads", [1, 3, 8])
def test_scene_render_mt_equals_scene_render(threads):
    ws = rz.workload_scene(3, 300, 512.0)
    fills = [(p, q.abi(), r) for p, q, r in zip(ws.paths, ws.paints, ws.rules)]
    tr = Transform.new_scale(0.5, 0.5)
    for view, bg in (((0.0, 0.0, 256.0, 256.0), None), ((13.5, 7.25, 200.0, 251.0), (0.1, 0.2, 0.0, 0.3))):
        a, oa = O.scene_render(fills, tr, view, bg)
        b, ob = O.scene_render_mt(fills, tr, view, bg, threads=threads)
        assert oa == ob and a.shape == b.shape
        assert np.array_equal(a, b)


def test_mask_atlas_mt_equals_per_glyph_masks():
    gs = rz.workload_glyphs(4, 300, 1024)
    W, H = gs.atlas_size
    trs = [list(s.tr.m) for s in gs.slots]
    q = [[s.x, s.y, s.width, s.height] for s in gs.slots]
    at = O.mask_atlas(gs.paths, trs, q, W, H, threads=5)
    for p, s in zip(gs.paths[:60], gs.slots[:60]):
        m = O.sd_mask(p, Transform(*s.tr.m), s.width, s.height, 0)
        assert np.array_equal(at[s.y:s.y + s.height, s.x:s.x + s.width], m.astype(np.float32))


# ---- the forward-mode derivative (CPU FP64 reference gradient) -------------------
@pytest.mark.parametrize("rule", [0, 1])
def test_dual_mask_iter_values_and_derivative(rule):
    from _diffscene import blob_scene
    ds = blob_scene(3, 2, 6, 64)
    path = ds.paths()[0]
    tr = Transform(0.9, 0.2, 3.0, -0.1, 1.1, -1.0)
    W = H = 72
    n_pts = int(path.point_offsets[-1])
    rng = np.random.default_rng(7)
    # a direction in the unique points (shared endpoints move together, so
    # the outline stays closed and the interior's winding stays put)
    uniq = rng.normal(0, 1, ds.points.shape)
    tan = uniq[ds.index[:n_pts]]
    a, da = O.sd_mask_iter_dual(path, tan, tr, W, H, rule)
    assert np.array_equal(a, O.sd_mask_iter(path, tr, W, H, rule))
    wgt = rng.uniform(-1, 1, (H, W))
    h = 1e-6
    pts = path.points.reshape(-1, 2)

    def f(eps):
        p = rz.Path(path.kinds, path.point_offsets, (pts + eps * tan).reshape(-1), path.subpaths,
                    path.closed)
        assert len(O.flatten(p, tr)) == len(O.flatten(path, tr))  # same subdivision
        return float((wgt * O.sd_mask_iter(p, tr, W, H, rule)).sum())
    fd = (f(h) - f(-h)) / (2 * h)
    ad = float((wgt * da).sum())
    print(f"rule {rule}: fd {fd:.9e} dual {ad:.9e}")
    assert abs(fd - ad) <= 1e-6 * max(1.0, abs(fd))

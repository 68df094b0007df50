"""Rasterizer::fill on the GPU against the oracle's restatement of the
reference's fill (src/rasterize.rs:69-115): the mask_iter coverage of the path
composites solid, linear- and radial-gradient paint source-over into a
premultiplied linear RGBA f32 image. Tolerances are per channel."""
import sys

import numpy as np
import pytest

import _oracle as O
from _paths import squirrel
import rasterize_amd as rz
from rasterize_amd import (ActiveEdgeRasterizer, FillRule, GradLinear, GradRadial, GradSpread,
                           LinColor, PathBuilder, SignedDifferenceRasterizer, Transform, Units,
                           workload_cubics64)

pytestmark = pytest.mark.gpu

MAX_ABS = 2e-3
MEAN_ABS = 1e-5
I = Transform.identity()


def _say(msg):
    print(msg, file=sys.stderr, flush=True)


def check_close(gpu, ref, what):
    d = np.abs(gpu.astype(np.float64) - ref.astype(np.float64))
    _say(f"{what}: max|d|={d.max():.3e} mean|d|={d.mean():.3e}")
    assert d.max() <= MAX_ABS, f"{what}: max |d| {d.max()}"
    assert d.mean() <= MEAN_ABS, f"{what}: mean |d| {d.mean()}"


def background(H, W, seed):
    """Random premultiplied background (alpha in [0, 1], colour <= alpha)."""
    r = np.random.default_rng(seed)
    a = r.uniform(0.0, 1.0, (H, W, 1)).astype(np.float32)
    c = (r.uniform(0.0, 1.0, (H, W, 3)) * a).astype(np.float32)
    return np.concatenate([c, a], axis=2).astype(np.float32)


STOPS = [(0.0, LinColor(0.8, 0.1, 0.1, 0.9)), (0.35, LinColor(0.05, 0.5, 0.1, 0.6)),
         (0.7, LinColor(0.0, 0.0, 0.0, 0.0)), (1.0, LinColor(0.1, 0.2, 0.7, 1.0))]


def paints():
    rot = Transform.new_rotate(0.4)
    return {
        "solid": LinColor(0.2, 0.4, 0.1, 0.7),
        "linear_pad_srgb": GradLinear(STOPS, start=(40, 30), end=(400, 300)),
        "linear_repeat_lin": GradLinear(STOPS, linear_colors=True, spread=GradSpread.Repeat,
                                        start=(100, 100), end=(180, 140), tr=rot),
        "linear_reflect_bbox": GradLinear(STOPS, units=Units.BoundingBox, spread=GradSpread.Reflect,
                                          start=(0.1, 0.2), end=(0.4, 0.3)),
        "radial_focal": GradRadial(STOPS, center=(256, 256), radius=200, fcenter=(200, 220),
                                   fradius=20),
        "radial_cone_reflect": GradRadial(STOPS, spread=GradSpread.Reflect, center=(300, 250),
                                          radius=40, fcenter=(150, 150), fradius=60),
        "radial_bbox_repeat": GradRadial(STOPS, units=Units.BoundingBox, spread=GradSpread.Repeat,
                                         linear_colors=True, center=(0.5, 0.5), radius=0.3),
    }


@pytest.fixture(scope="module")
def c1():
    return workload_cubics64(1, 512.0).paths[0]


@pytest.mark.parametrize("name", sorted(paints()))
@pytest.mark.parametrize("rule", [FillRule.NonZero, FillRule.EvenOdd])
def test_fill_c1(c1, name, rule):
    paint = paints()[name]
    _say(f"fill c1 {name} rule={rule}")
    bg = background(512, 512, 7)
    got = SignedDifferenceRasterizer().fill(c1, I, rule, paint, bg.copy())
    ref = O.fill(c1, I, paint.abi(), rule, bg)
    check_close(got, ref, f"fill c1 {name} rule={rule}")


def test_fill_edges_and_transform():
    """Path crossing all four image edges under a rotation (left split and
    right-edge column in the (W+1)-wide mask_iter semantics)."""
    path = squirrel()
    tr = Transform.new_translate(-30.0, -25.0) * Transform.new_scale(9.0, 9.0) * Transform.new_rotate(0.2)
    W, H = 161, 143
    bg = background(H, W, 3)
    for name, paint in sorted(paints().items()):
        _say(f"fill edges {name}")
        got = SignedDifferenceRasterizer().fill(path, tr, FillRule.NonZero, paint, bg.copy())
        ref = O.fill(path, tr, paint.abi(), FillRule.NonZero, bg)
        check_close(got, ref, f"fill edges {name}")


def test_fill_strided_view(c1):
    """A sub-view of a larger image: only the view changes."""
    big = background(600, 700, 11)
    view = big[40:552, 60:572]
    before = big.copy()
    paint = paints()["linear_pad_srgb"]
    ref = O.fill(c1, I, paint.abi(), FillRule.NonZero, np.ascontiguousarray(view))
    SignedDifferenceRasterizer().fill(c1, I, FillRule.NonZero, paint, view)
    check_close(big[40:552, 60:572], ref, "strided view")
    outside = np.ones(big.shape[:2], bool)
    outside[40:552, 60:572] = False
    assert np.array_equal(big[outside], before[outside])


def test_fill_active_edge():
    """ActiveEdgeRasterizer::fill = its mask_iter (no right clip) + the same
    compositing; expected image from the oracle's active-edge mask and paint."""
    path = workload_cubics64(3, 60.0).paths[0]
    W, H = 48, 56
    bg = background(H, W, 5)
    paint = paints()["radial_focal"]
    pa = GradRadial(STOPS, center=(30, 30), radius=25, fcenter=(20, 25), fradius=3)
    for p in (paint, pa):
        got = ActiveEdgeRasterizer().fill(path, I, FillRule.NonZero, p, bg.copy())
        alpha = O.ae_mask(path, I, W, H, FillRule.NonZero)
        ref = bg.copy()
        abi = p.abi()
        for y in range(H):
            for x in range(W):
                a = alpha[y, x]
                if abs(a) < 1e-6:
                    continue
                c = O.paint_at(abi, x + 0.5, y + 0.5) * np.float32(a)
                ref[y, x] = c + ref[y, x] * (np.float32(1.0) - c[3])
        check_close(got, ref, "active-edge fill")


def test_fill_degenerate_gradient_draws_nothing(c1):
    """A gradient whose unit transform is singular draws nothing (the
    reference's `invert()` returns None)."""
    bg = background(64, 64, 1)
    flat = Transform(1.0, 0.0, 0.0, 0.0, 0.0, 0.0)
    paint = GradLinear(STOPS, start=(0, 0), end=(10, 0), tr=flat)
    got = SignedDifferenceRasterizer().fill(c1, I, FillRule.NonZero, paint, bg.copy())
    assert np.array_equal(got, bg)


def test_fill_nan_is_error():
    p = PathBuilder().move_to((0, 0)).line_to((float("nan"), 3)).line_to((5, 5)).close().build()
    with pytest.raises(rz.RzError) as e:
        SignedDifferenceRasterizer().fill(p, I, FillRule.NonZero, LinColor(1, 0, 0, 1),
                                          np.zeros((8, 8, 4), np.float32))
    assert e.value.status == rz._abi.RZ_E_NAN

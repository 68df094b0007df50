"""rz_render_tree (Scene::render of Fill / Group / Transform / Opacity /
Clip trees, src/scene.rs:186-199, 268-459) on the GPU against the oracle's
restatement of the same tree, on the reference's scene files and on
generated trees: nested transforms, opacities, clips in user space and in
the child's bounding box, views, backgrounds, fill rules."""
import os
import sys

import numpy as np
import pytest

import _oracle as O
from _paths import rect
from rasterize_amd import (ActiveEdgeRasterizer, FillRule, GradLinear, GradRadial, GradSpread, LinColor,
                           PathBuilder, Scene, SignedDifferenceRasterizer, Transform, Units,
                           workload_scene)

pytestmark = pytest.mark.gpu

MAX_ABS = 2e-3
MEAN_ABS = 1e-5
GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
I = Transform.identity()


def _say(msg):
    print(msg, file=sys.stderr, flush=True)


def check(scene, tr, view=None, bg=None, what="", rast=None, max_abs=MAX_ABS):
    _say(f"tree {what}: view={view} bg={bg is not None}")
    got = scene.render(rast or SignedDifferenceRasterizer(), tr, view, bg)
    img, (x, y) = O.scene_tree_render(scene, tr, view, bg,
                                      active_edge=isinstance(rast, ActiveEdgeRasterizer))
    assert (got.x, got.y) == (x, y), f"{what}: layer origin {(got.x, got.y)} != {(x, y)}"
    assert got.image.shape == img.shape, f"{what}: layer shape {got.image.shape} != {img.shape}"
    if img.size == 0:
        return got
    d = np.abs(got.image.astype(np.float64) - img.astype(np.float64))
    _say(f"tree {what}: {img.shape[1]}x{img.shape[0]} max|d|={d.max():.3e} mean|d|={d.mean():.3e}")
    assert d.max() <= max_abs, f"{what}: max |d| {d.max()} at {np.unravel_index(d.argmax(), d.shape)}"
    assert d.mean() <= MEAN_ABS, f"{what}: mean |d| {d.mean()}"
    return got


def load(name):
    with open(os.path.join(GOLDEN, name)) as f:
        return Scene.from_json(f.read())


@pytest.mark.parametrize("name", ["grad.scene", "linear-colors.scene", "firefox.scene"])
def test_reference_scene_files(name):
    check(load(name), I, what=name)


def test_reference_scene_files_in_views():
    check(load("firefox.scene"), Transform.new_scale(0.5, 0.5), (40.0, 30.0, 200.5, 180.25),
          (0.9, 0.9, 0.85, 1.0), what="firefox view+bg")
    check(load("grad.scene"), Transform.new_translate(-20, 5), (0.0, 0.0, 128.0, 96.0), what="grad view")


def _blob(rng, cx, cy, r, n=5):
    a = np.sort(rng.uniform(0, 2 * np.pi, n))
    pts = [(cx + r * rng.uniform(0.4, 1.0) * np.cos(t), cy + r * rng.uniform(0.4, 1.0) * np.sin(t))
           for t in a]
    pb = PathBuilder().move_to(pts[0])
    for k in range(1, n + 1):
        p0, p1 = pts[k - 1], pts[k % n]
        m = ((p0[0] + p1[0]) / 2 + rng.uniform(-r, r) * 0.3, (p0[1] + p1[1]) / 2 + rng.uniform(-r, r) * 0.3)
        pb.quad_to(m, p1)
    return pb.close().build()


def _random_tree(rng, depth, size):
    if depth == 0 or rng.uniform() < 0.25:
        c = rng.uniform(0, 1, 4).astype(np.float32)
        c[:3] *= c[3]
        if rng.uniform() < 0.3:
            stops = [(0.0, LinColor(*c)), (1.0, LinColor(0.1, 0.5, 0.2, 1.0))]
            paint = GradLinear(stops, units=Units.BoundingBox, spread=GradSpread.Reflect,
                               start=(0.2, 0.1), end=(0.6, 0.9))
        else:
            paint = LinColor(*c)
        rule = FillRule.EvenOdd if rng.uniform() < 0.3 else FillRule.NonZero
        return Scene.fill(_blob(rng, *rng.uniform(0.2 * size, 0.8 * size, 2), rng.uniform(8, size / 3)),
                          paint, rule)
    k = rng.integers(0, 4)
    child = Scene(Scene.GROUP, children=[_random_tree(rng, depth - 1, size)
                                         for _ in range(rng.integers(1, 4))])
    if k == 0:
        return child
    if k == 1:
        T = (Transform.new_translate(*rng.uniform(-10, 10, 2)) *
             Transform.new_rotate(rng.uniform(-0.4, 0.4)) * Transform.new_scale(*rng.uniform(0.7, 1.2, 2)))
        return child.with_transform(T)
    if k == 2:
        return Scene(Scene.OPACITY, opacity=float(rng.uniform(0.1, 1.0)), child=child)
    units = Units.BoundingBox if rng.uniform() < 0.5 else Units.UserSpaceOnUse
    clip = (_blob(rng, 0.5, 0.5, 0.45) if units == Units.BoundingBox
            else _blob(rng, *rng.uniform(0.3 * size, 0.7 * size, 2), size / 2.5))
    rule = FillRule.EvenOdd if rng.uniform() < 0.3 else FillRule.NonZero
    return child.with_clip(clip, units, rule)


@pytest.mark.parametrize("seed", range(6))
def test_generated_trees(seed):
    rng = np.random.default_rng(seed)
    sc = Scene(Scene.GROUP, children=[_random_tree(rng, 4, 192.0) for _ in range(3)])
    check(sc, I, what=f"seed {seed}")
    check(sc, Transform.new_translate(3.5, -2.25) * Transform.new_scale(1.3, 1.3),
          (16.0, 8.0, 200.0, 230.0), (0.0, 0.0, 0.0, 1.0), what=f"seed {seed} view")


def test_nested_opacity_and_clip():
    a = rect(10, 10, 90, 70)
    b = _blob(np.random.default_rng(4), 60, 50, 40)
    inner = Scene.group([Scene.fill(a, LinColor(0.4, 0.2, 0.1, 0.8)),
                         Scene.fill(b, LinColor(0.0, 0.3, 0.5, 0.6), FillRule.EvenOdd)])
    sc = (inner.with_clip(rect(0.1, 0.1, 0.9, 0.7), Units.BoundingBox)
          .with_opacity(0.6)
          .with_transform(Transform.new_rotate(0.2))
          .with_clip(_blob(np.random.default_rng(5), 50, 50, 45), Units.UserSpaceOnUse, FillRule.EvenOdd)
          .with_opacity(0.9))
    check(sc, I, what="nested")
    check(sc, I, (0.0, 0.0, 100.0, 100.0), (1.0, 1.0, 1.0, 1.0), what="nested view+bg")


def test_empty_results():
    far = Scene.fill(rect(0, 0, 8, 8), LinColor(1, 1, 1, 1)).with_clip(rect(20, 20, 30, 30))
    got = check(far, I, what="disjoint clip")
    assert got.image.size == 0 and (got.x, got.y) == (0, 0)
    out = Scene.fill(rect(0, 0, 8, 8), LinColor(1, 1, 1, 1)).with_opacity(0.5)
    got = check(out, I, (100.0, 100.0, 120.0, 110.0), what="out of view")
    assert got.image.size == 0


def test_many_fills_under_one_transform():
    """A run of fills in one layer goes through one batched render, each fill
    under its own transform (a Transform node per fill, and a shared one)."""
    ws = workload_scene(9, 120, 512.0)
    fills = [Scene.fill(p, q, r) for p, q, r in zip(ws.paths, ws.paints, ws.rules)]
    rng = np.random.default_rng(1)
    parts = []
    for k, f in enumerate(fills):
        if k % 3 == 0:
            f = f.with_transform(Transform.new_translate(*rng.uniform(-20, 20, 2)))
        parts.append(f)
    sc = Scene(Scene.GROUP, children=parts).with_transform(Transform.new_scale(0.5, 0.5))
    check(sc, I, (0.0, 0.0, 256.0, 256.0), what="120 fills")


def test_active_edge_rasterizer():
    """ActiveEdgeRasterizer (RZ_EDGE_NONE) for fills and clip masks, against
    the oracle's tree with its active-edge restatement."""
    check(load("grad.scene"), I, what="grad active-edge", rast=ActiveEdgeRasterizer())
    check(load("firefox.scene"), Transform.new_scale(0.5, 0.5), what="firefox active-edge",
          rast=ActiveEdgeRasterizer())
    rng = np.random.default_rng(11)
    sc = Scene(Scene.GROUP, children=[_random_tree(rng, 4, 192.0) for _ in range(3)])
    check(sc, I, (0.0, 0.0, 192.0, 192.0), what="generated active-edge", rast=ActiveEdgeRasterizer())


def test_stroke_nodes():
    """Stroke nodes (src/scene.rs:297-306): each path stroked by rz_path_stroke
    and filled non-zero, inside transforms, opacities and a bounding-box clip
    (whose box is the unstroked path's, src/scene.rs:158)."""
    from rasterize_amd import LineCap, LineJoin, StrokeStyle
    a = PathBuilder().move_to((6, 6)).cubic_to((60, 2), (64, 60), (10, 56)).line_to((30, 30)).build()
    b = PathBuilder().move_to((20, 70)).quad_to((50, 40), (90, 80)).line_to((70, 20)).close().build()
    sa = Scene.stroke(a, LinColor(0.8, 0.2, 0.1, 1.0), StrokeStyle(4.0, LineJoin.Round, LineCap.Round))
    sb = Scene.stroke(b, GradLinear([(0.0, LinColor(0, 0, 1, 1)), (1.0, LinColor(0, 1, 0, 1))],
                                    start=(20, 20), end=(90, 80)),
                      StrokeStyle(2.5, LineJoin.Miter, LineCap.Square, miter_limit=10.0))
    fill = Scene.fill(rect(30, 30, 70, 70), LinColor(0.3, 0.3, 0.3, 0.5))
    tree = Scene.group([fill, sa.with_opacity(0.6), sb.with_transform(Transform.new_rotate(0.1))])
    check(tree, Transform.new_scale(2.0, 2.0), what="strokes")
    check(sa.with_clip(rect(0.1, 0.1, 0.7, 0.9), Units.BoundingBox), I, what="stroke bbox clip")
    check(Scene.stroke(a, LinColor(1, 1, 1, 1), 3.0), I, (0.0, 0.0, 50.0, 40.0), (0, 0, 0, 1),
          what="stroke view")

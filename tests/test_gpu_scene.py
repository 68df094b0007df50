"""Scene::render on the GPU (batched and plain forms) against the oracle's
restatement of Scene::group(fills).render(SignedDifferenceRasterizer, tr,
view, bg) (src/scene.rs:186-199, 268-357, 384-459): per fill the bbox ∩ view,
its sub-view and alignment, mask_iter + paint + source-over in path order."""
import sys

import numpy as np
import pytest

import _oracle as O
import rasterize_amd as rz
from rasterize_amd import (FillRule, GradLinear, GradRadial, GradSpread, LinColor, PathBuilder, Scene,
                           SignedDifferenceRasterizer, Transform, Units, workload_circles,
                           workload_scene)

pytestmark = pytest.mark.gpu

MAX_ABS = 2e-3
MEAN_ABS = 1e-5
I = Transform.identity()


def _say(msg):
    print(msg, file=sys.stderr, flush=True)


def check_layer(got, ref, what):
    img, (x, y) = ref
    assert (got.x, got.y) == (x, y), f"{what}: layer origin {(got.x, got.y)} != {(x, y)}"
    assert got.image.shape == img.shape, f"{what}: layer shape {got.image.shape} != {img.shape}"
    if img.size == 0:
        return
    d = np.abs(got.image.astype(np.float64) - img.astype(np.float64))
    _say(f"{what}: max|d|={d.max():.3e} mean|d|={d.mean():.3e}")
    assert d.max() <= MAX_ABS, f"{what}: max |d| {d.max()} at {np.unravel_index(d.argmax(), d.shape)}"
    assert d.mean() <= MEAN_ABS, f"{what}: mean |d| {d.mean()}"


def fills_of(ws):
    return [(p, q.abi(), r) for p, q, r in zip(ws.paths, ws.paints, ws.rules)]


@pytest.fixture(scope="module")
def small_scene():
    """150 paths generated on a 1024 canvas, drawn at a quarter scale (S) so
    that the 256 x 256 view holds paths of 2 to 256 px."""
    return workload_scene(3, 150, 1024.0)


S = Transform.new_scale(0.25, 0.25)


@pytest.mark.parametrize("batched", [True, False])
def test_scene_small(small_scene, batched):
    view = (0.0, 0.0, 256.0, 256.0)
    _say(f"scene small batched={batched}")
    got = small_scene.scene().render(SignedDifferenceRasterizer(), S, view, batched=batched)
    ref = O.scene_render(fills_of(small_scene), S, view)
    check_layer(got, ref, f"scene small batched={batched}")


@pytest.mark.parametrize("batched", [True, False])
def test_scene_view_offset_bg_transform(small_scene, batched):
    """Fractional view corners, a background colour and a rotated, scaled
    transform: sub-views land at every tile phase."""
    view = (10.5, 20.25, 230.7, 241.1)
    bg = (0.1, 0.05, 0.2, 0.4)
    tr = Transform.new_translate(128, 128) * Transform.new_rotate(0.3) * Transform.new_scale(1.3, 0.8) * \
        Transform.new_translate(-128, -128) * S
    _say(f"scene offset batched={batched}")
    got = small_scene.scene().render(SignedDifferenceRasterizer(), tr, view, bg=bg, batched=batched)
    ref = O.scene_render(fills_of(small_scene), tr, view, bg=bg)
    check_layer(got, ref, f"scene offset batched={batched}")


def test_scene_without_view(small_scene):
    """view = None: the layer is the union of the fills' bboxes."""
    got = small_scene.scene().render(SignedDifferenceRasterizer(), S)
    ref = O.scene_render(fills_of(small_scene), S)
    check_layer(got, ref, "scene no view")


def test_scene_circles():
    """The reference's own scene benchmark shape (benches/scene_bench.rs), scaled down."""
    ws = workload_circles(96, 256.0)
    view = (0.0, 0.0, 256.0, 256.0)
    got = ws.scene().render(SignedDifferenceRasterizer(), I, view)
    ref = O.scene_render(fills_of(ws), I, view)
    check_layer(got, ref, "circles")


def test_scene_gradients_and_rules():
    """Bounding-box and user-space gradients of every spread, both fill rules,
    overlapping paths (order matters) and paths partly outside the view."""
    stops = [(0.0, LinColor(0.9, 0.1, 0.1, 0.9)), (0.5, LinColor(0.0, 0.3, 0.1, 0.5)),
             (1.0, LinColor(0.1, 0.1, 0.6, 0.8))]
    p1 = rz.workload_cubics64(7, 180.0).paths[0]
    p2 = PathBuilder().move_to((-30, 40)).cubic_to((100, -20), (220, 300), (150, 100)).close().build()
    p3 = PathBuilder().move_to((60, 60)).circle(70.0).build()
    scene = Scene.group([
        Scene.fill(p1, GradLinear(stops, units=Units.BoundingBox, spread=GradSpread.Reflect,
                                  start=(0.2, 0.1), end=(0.5, 0.4)), FillRule.EvenOdd),
        Scene.fill(p2, GradRadial(stops, spread=GradSpread.Repeat, center=(90, 90), radius=30,
                                  fcenter=(70, 80), fradius=5), FillRule.NonZero),
        Scene.fill(p3, LinColor(0.2, 0.5, 0.1, 0.6), FillRule.NonZero),
        Scene.fill(p1, GradRadial(stops, units=Units.BoundingBox, linear_colors=True,
                                  center=(0.5, 0.5), radius=0.4), FillRule.NonZero),
    ])
    view = (0.0, 0.0, 200.0, 190.0)
    fills = [(p, q.abi(), r) for p, q, r in scene.fills()]
    ref = O.scene_render(fills, I, view)
    for batched in (True, False):
        got = scene.render(SignedDifferenceRasterizer(), I, view, batched=batched)
        check_layer(got, ref, f"gradients batched={batched}")


def test_scene_nothing_in_view(small_scene):
    got = small_scene.scene().render(SignedDifferenceRasterizer(), I, (5000.0, 5000.0, 5100.0, 5100.0))
    assert got.image.size == 0 and (got.x, got.y) == (0, 0)


def test_scene_batched_is_deterministic(small_scene):
    view = (0.0, 0.0, 256.0, 256.0)
    a = small_scene.scene().render(SignedDifferenceRasterizer(), S, view).image
    b = small_scene.scene().render(SignedDifferenceRasterizer(), S, view).image
    assert np.array_equal(a.view(np.uint32), b.view(np.uint32))


@pytest.mark.parametrize("batched", [True, False])
def test_scene_empty_outside_and_clipped_paths(small_scene, batched):
    """Fills without segments, fills entirely outside the view and fills cut by
    the view's edges, between ordinary ones: the per-segment box reduction
    must leave an empty path undrawn and give clipped paths their sub-views."""
    empty = PathBuilder().build()
    far = PathBuilder().move_to((5000.0, 5000.0)).line_to((5100.0, 5000.0)).line_to((5050.0, 5080.0)).close().build()
    edge = (PathBuilder().move_to((-40.0, 30.0)).cubic_to((600.0, -50.0), (1200.0, 1300.0), (1100.0, 90.0))
            .line_to((120.0, 1070.0)).close().build())
    ws = small_scene
    paint = ws.paints[1]
    order = list(range(10)) + ["e", "f"] + list(range(10, 25)) + ["c", "e"] + list(range(25, 40))
    pick = {"e": (empty, paint, 0), "f": (far, paint, 0), "c": (edge, paint, 1)}
    items = [(ws.paths[k], ws.paints[k], ws.rules[k]) if isinstance(k, int) else pick[k] for k in order]
    scene = Scene.group([Scene.fill(p, q, r) for p, q, r in items])
    view = (0.0, 0.0, 256.0, 256.0)
    _say(f"scene empty/outside/clipped batched={batched}")
    got = scene.render(SignedDifferenceRasterizer(), S, view, batched=batched)
    ref = O.scene_render([(p, q.abi(), r) for p, q, r in items], S, view)
    check_layer(got, ref, f"scene empty/outside/clipped batched={batched}")


@pytest.mark.parametrize("batched", [True, False])
def test_scene_view_layer_geometry(batched):
    """src/scene.rs:669-693 test_scene_view: the layer of a view is the view's
    pixel box (1, 2, 22 x 21); without a view it is the fill's box
    (6, 4, 16 x 15). The scene: transform "translate(7, 7) rotate(45, 7, 7)
    scale(10)" of a fill of the unit square, ActiveEdgeRasterizer."""
    from test_oracle_pins import _scene_view_fill
    path, tr = _scene_view_fill()
    sc = Scene.fill(path, LinColor(1.0, 0.2158605, 0.05126946, 1.0))
    r = rz.ActiveEdgeRasterizer()
    lay = sc.render(r, tr, view=(1.0, 2.0, 23.0, 23.0), batched=batched)
    assert (lay.x, lay.y, lay.image.shape[1], lay.image.shape[0]) == (1, 2, 22, 21)
    lay = sc.render(r, tr, batched=batched)
    assert (lay.x, lay.y, lay.image.shape[1], lay.image.shape[0]) == (6, 4, 16, 15)
    # and the pixels agree with the oracle's render of the same layer
    ref, (x, y) = O.scene_render([(path, LinColor(1.0, 0.2158605, 0.05126946, 1.0).abi(), 0)], tr)
    assert (x, y) == (6, 4)
    d = np.abs(lay.image.astype(np.float64) - ref)
    assert d.max() <= 2e-3


@pytest.mark.parametrize("mode", ["scene", "atlas"])
def test_fused_flatten_and_count_batched_forms(small_scene, mode):
    """RZ_PARAM_FUSED 1 (flatten and count pass as one kernel) against the two
    kernels for the batched forms: the same pixels, bit for bit."""
    from rasterize_amd import mask_atlas, workload_glyphs
    L = rz._abi.lib()
    out = []
    gs = workload_glyphs(4, 300, 1024) if mode == "atlas" else None
    for fused in (0, 1):
        ctx = rz.Context(0)
        try:
            ctx.check(L.rz_set_param(ctx.h, rz._abi.RZ_PARAM_FUSED, fused))
            _say(f"fused={fused} {mode}")
            if mode == "scene":
                lay = small_scene.scene().render(SignedDifferenceRasterizer(ctx=ctx), S,
                                                 (0.0, 0.0, 256.0, 256.0))
                out.append(np.array(lay.image))
            else:
                W, H = gs.atlas_size
                atlas = np.zeros((H, W), np.float32)
                out.append(mask_atlas(gs.paths, gs.slots, atlas, ctx=ctx))
        finally:
            ctx.close()
    assert out[0].shape == out[1].shape and np.array_equal(out[0], out[1])

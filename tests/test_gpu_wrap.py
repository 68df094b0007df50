"""Lines that start exactly on the right clip abscissa (W - 1 for a mask, W for
mask_iter and scenes) and end to its right. The reference keeps such a line
whole (src/rasterize.rs:373-384), and its deposits past a row's end land in the
next row's first pixels; the GPU flatten replaces it by the vertical lines
rzg::wrap_lines makes (pinned against the oracle's deposits on the CPU in
test_wrap_cpu.py). Here the whole pipeline against the oracle: dense masks,
a mask into a wider device image, mask_iter, and a scene fill."""
import ctypes as C

import numpy as np
import pytest
import torch

import _oracle as O
import rasterize_amd as rz
from rasterize_amd import FillRule, PathBuilder, Transform

pytestmark = pytest.mark.gpu

I = Transform.identity()


def _say(*a):
    print(*a, flush=True)


def _check(gpu, ref, what):
    d = np.abs(np.asarray(gpu, np.float64) - ref)
    _say(f"{what}: max|d|={d.max():.3e} mean|d|={d.mean():.3e}")
    assert d.max() <= 2e-3 and d.mean() <= 1e-5, what


def _wrap_path(wc, H, shift=0.0):
    """Vertices exactly on x = wc (+ shift), each followed by a point beyond
    it: a small overshoot, one of several pixels, a long line down the side, an
    upward one, and one into the last row."""
    x = wc + shift
    pb = PathBuilder().move_to((2.0, 1.0))
    for a, b in (((x, 1.5), (x + 0.4, 3.2)), ((x, 6.25), (x + 3.7, 8.9)),
                 ((x, 10.0), (x + 0.8, 30.0)), ((x, 40.0), (x + 2.5, 33.0)),
                 ((x, H - 0.5), (x + 2.0, H + 2.0))):
        pb = pb.line_to(a).line_to(b)
    return pb.line_to((1.0, H - 1.0)).close().build()


def _oracle_mask_strided(path, W, H, rs, rule):
    img = np.zeros((H, rs))
    p, t = O.opath(path), O._tr(I)
    O._check(O.lib().orc_sd_mask(C.byref(p), O._dp(t), 0.05, rule, O._dp(img), img.size, W, H, 0,
                                 rs, 1))
    return img[:, :W]


@pytest.mark.parametrize("rule", [FillRule.NonZero, FillRule.EvenOdd])
def test_wrap_dense_mask_and_iter(rule):
    W, H = 48, 48
    for mode, wc, ref_fn in ((rz._abi.RZ_EDGE_MASK, W - 1.0, O.sd_mask),
                             (rz._abi.RZ_EDGE_ITER, float(W), O.sd_mask_iter)):
        path = _wrap_path(wc, H)
        ref = ref_fn(path, I, W, H, rule)
        # the quirk matters: the same path with its vertices 1e-9 left of the clip
        near = ref_fn(_wrap_path(wc, H, -1e-9), I, W, H, rule)
        assert np.abs(ref - near).max() > 0.05
        _say(f"wrap mask mode={mode} rule={rule}")
        g = rz.SignedDifferenceRasterizer().mask_f32(path, I, W, H, rule, edge_mode=mode)
        _check(g, ref, f"mode {mode}")


def test_wrap_mask_into_wider_image():
    W, H, rs = 48, 48, 61
    path = _wrap_path(W - 1.0, H)
    ref = _oracle_mask_strided(path, W, H, rs, 0)
    out = torch.zeros((H, rs), dtype=torch.float32, device="cuda")
    ctx = rz.Context(0)
    try:
        rz.SignedDifferenceRasterizer(ctx=ctx).mask_f32(path, I, W, H, FillRule.NonZero)  # sizing
        dp = rz.DevicePath(path, ctx)
        _say("wrap mask_device row_stride 61")
        dp.mask_device(I, W, H, out.data_ptr(), rs)
        ctx.sync()
        dp.close()
    finally:
        ctx.close()
    g = out.cpu().numpy()
    _check(g[:, :W], ref, "row stride 61")
    assert not g[:, W:].any()


def test_wrap_scene_fill():
    # the path runs past the view on the right, so its sub-view is clipped at
    # x = 40 and its vertices on x = 40 are on the clip of mask_iter
    W, H = 40, 48
    path = _wrap_path(float(W), H)
    sc = rz.Scene.fill(path, rz.LinColor(0.1, 0.2, 0.3, 0.9))
    view = (0.0, 0.0, float(W), float(H))
    _say("wrap scene fill")
    lay = sc.render(rz.SignedDifferenceRasterizer(), I, view)
    img, org = O.scene_tree_render(sc, I, view)
    assert org == (0, 0)
    _check(lay.image, img, "scene fill")


def test_wrap_too_far_is_refused():
    # a line from the clip running 5000 px right within half a row would need
    # deposits of ~1250 per pixel: refused with RZ_E_DEPTH, not expanded
    W, H = 48, 48
    path = PathBuilder().move_to((2.0, 2.0)).line_to((W - 1.0, 5.0)).line_to((W + 5000.0, 5.5)) \
        .line_to((2.0, 30.0)).close().build()
    ctx = rz.Context(0)
    try:
        _say("wrap refused")
        with pytest.raises(rz.RzError) as e:
            rz.SignedDifferenceRasterizer(ctx=ctx).mask_f32(path, I, W, H, FillRule.NonZero)
        assert e.value.status == rz._abi.RZ_E_DEPTH
        # the context stays usable
        g = rz.SignedDifferenceRasterizer(ctx=ctx).mask_f32(_wrap_path(W - 1.0, H), I, W, H,
                                                           FillRule.NonZero)
        _check(g, O.sd_mask(_wrap_path(W - 1.0, H), I, W, H, 0), "after refusal")
    finally:
        ctx.close()

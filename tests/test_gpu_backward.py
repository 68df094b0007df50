"""Adjoint of the batched scene render (rz_render_backward_device) against
FP64 central finite differences of the loss, the scene rendered in f64 from
the oracle's per-path coverage (SignedDifferenceRasterizer::mask_iter) and an
f64 source-over. Parameters: unique control points (shared endpoints move
together) and solid colours."""
import sys

import numpy as np
import pytest

import _oracle as O
from _diffscene import blob_scene, render_f64
import rasterize_amd as rz
from rasterize_amd import DeviceScene, Transform, render_torch

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, file=sys.stderr, flush=True)


def test_gradients_match_finite_differences():
    import torch
    W = H = 96
    NP, NC = 6, 6
    ds = blob_scene(7, NP, NC, W)
    uniq, colors = ds.points, ds.colors.astype(np.float64)
    rng = np.random.default_rng(1)
    tp = uniq + rng.normal(0, 2.0, uniq.shape)
    tc = np.clip(colors + rng.normal(0, 0.05, colors.shape), 0, 1)
    target = render_f64(O, ds, tp, tc, W, H)
    dev = DeviceScene(ds.scene())
    idx = torch.tensor(ds.index, device="cuda")
    U = torch.tensor(uniq, dtype=torch.float64, device="cuda", requires_grad=True)
    Cc = torch.tensor(colors, dtype=torch.float32, device="cuda", requires_grad=True)
    view = (0.0, 0.0, float(W), float(H))
    out = render_torch(dev, U[idx], Cc, Transform.identity(), view)
    T = torch.tensor(target, dtype=torch.float32, device="cuda")
    loss = ((out - T) ** 2).sum()
    loss.backward()
    gU = U.grad.cpu().numpy()
    gC = Cc.grad.cpu().numpy()
    # the forward agrees with the f64 render
    f64 = render_f64(O, ds, uniq, colors, W, H)
    assert np.abs(out.detach().cpu().numpy() - f64).max() < 2e-3

    def loss64(u, c):
        return float(((render_f64(O, ds, u, c, W, H) - target) ** 2).sum())

    rng = np.random.default_rng(5)
    errs = []
    h = 1e-3
    for _ in range(24):
        i, d = rng.integers(len(uniq)), rng.integers(2)
        up, um = uniq.copy(), uniq.copy()
        up[i, d] += h
        um[i, d] -= h
        fd = (loss64(up, colors) - loss64(um, colors)) / (2 * h)
        errs.append((fd, float(gU[i, d]), f"point {i},{'xy'[d]}"))
    hc = 1e-4
    for _ in range(8):
        k, ch = rng.integers(NP), rng.integers(4)
        cp, cm = colors.copy(), colors.copy()
        cp[k, ch] += hc
        cm[k, ch] -= hc
        fd = (loss64(uniq, cp) - loss64(uniq, cm)) / (2 * hc)
        errs.append((fd, float(gC[k, ch]), f"colour {k},{ch}"))
    scale = max(abs(e[0]) for e in errs)
    rel = []
    for fd, g, what in errs:
        r = abs(g - fd) / max(abs(fd), 1e-2 * scale)
        rel.append(r)
        _say(f"{what}: fd={fd:+.6e} adjoint={g:+.6e} rel={r:.2e}")
    rel = np.array(rel)
    _say(f"relative error: median {np.median(rel):.2e} max {rel.max():.2e}")
    assert np.median(rel) < 1e-4
    assert rel.max() < 1e-3


def test_transform_gradient_matches_finite_differences():
    """dL/d(transform) of a rotated, scaled render against f64 finite differences."""
    import torch
    W = H = 80
    ds = blob_scene(11, 4, 5, W)
    tr0 = np.array([0.9, 0.15, 3.0, -0.1, 1.05, -2.0])
    colors = ds.colors.astype(np.float64)
    rng = np.random.default_rng(4)
    target = rng.uniform(0, 0.5, (H, W, 4))
    dev = DeviceScene(ds.scene())
    view = (0.0, 0.0, float(W), float(H))
    tr = Transform(*tr0)
    out = torch.empty((H, W, 4), dtype=torch.float32, device="cuda")
    dev.render_device(tr, view, out.data_ptr(), W)
    dev.ctx.check(rz._abi.lib().rz_sync(dev.ctx.h))
    g = (2.0 * (out - torch.tensor(target, dtype=torch.float32, device="cuda"))).contiguous()
    gp = torch.zeros((dev.n_points(), 2), dtype=torch.float32, device="cuda")
    gc = torch.zeros((4, 4), dtype=torch.float32, device="cuda")
    gtr = torch.zeros(6, dtype=torch.float64, device="cuda")
    dev.backward_device(tr, view, g.data_ptr(), W, gp.data_ptr(), gc.data_ptr(), gtr.data_ptr())
    dev.ctx.check(rz._abi.lib().rz_sync(dev.ctx.h))
    gtr = gtr.cpu().numpy()

    def loss64(m):
        dst = np.zeros((H, W, 4))
        for p, c in zip(ds.paths(), colors):
            a = O.sd_mask_iter(p, Transform(*m), W, H, 0)
            s = c[None, None, :] * a[..., None]
            dst = s + dst * (1.0 - s[..., 3:4])
        return float(((dst - target) ** 2).sum())

    for k, h in enumerate([1e-6, 1e-6, 1e-4, 1e-6, 1e-6, 1e-4]):
        mp, mm = tr0.copy(), tr0.copy()
        mp[k] += h
        mm[k] -= h
        fd = (loss64(mp) - loss64(mm)) / (2 * h)
        r = abs(gtr[k] - fd) / max(abs(fd), 1e-3)
        _say(f"transform m[{k}]: fd={fd:+.6e} adjoint={gtr[k]:+.6e} rel={r:.2e}")
        assert r < 1e-3


def test_zero_upstream_gradient_gives_zero():
    import torch
    ds = DeviceScene(blob_scene(3, 3, 4, 64).scene())
    n = ds.n_points()
    gp = torch.zeros((n, 2), dtype=torch.float32, device="cuda")
    gc = torch.zeros((3, 4), dtype=torch.float32, device="cuda")
    gtr = torch.zeros(6, dtype=torch.float64, device="cuda")
    g = torch.zeros((64, 64, 4), dtype=torch.float32, device="cuda")
    out = torch.empty((64, 64, 4), dtype=torch.float32, device="cuda")
    ds.render_device(Transform.identity(), (0, 0, 64, 64), out.data_ptr(), 64)
    ds.backward_device(Transform.identity(), (0, 0, 64, 64), g.data_ptr(), 64, gp.data_ptr(),
                       gc.data_ptr(), gtr.data_ptr())
    ds.ctx.check(rz._abi.lib().rz_sync(ds.ctx.h))
    assert float(gp.abs().sum()) == 0.0 and float(gc.abs().sum()) == 0.0
    assert float(gtr.abs().sum()) == 0.0


def _busy_torch_stream(torch):
    """Keep torch's current stream busy for a few milliseconds; returns a
    zero-valued scalar tensor that is ready only when that work is done."""
    x = torch.ones(1 << 26, device="cuda")
    for _ in range(40):
        x = x * 1.0000001
    return x[:1] * 0.0


def test_render_torch_orders_after_torch_stream():
    """Inputs produced by kernels still queued on torch's stream, and gradient
    buffers torch has just cleared, are seen by the library's stream."""
    import torch
    W = H = 96
    ds = blob_scene(17, 5, 6, W)
    dev = DeviceScene(ds.scene())
    idx = torch.tensor(ds.index, device="cuda")
    view = (0.0, 0.0, float(W), float(H))
    U0 = torch.tensor(ds.points, dtype=torch.float64, device="cuda")
    C0 = torch.tensor(ds.colors, dtype=torch.float32, device="cuda")
    # reference: everything settled before each call
    Ur = U0.clone().requires_grad_(True)
    Cr = C0.clone().requires_grad_(True)
    ref = render_torch(dev, Ur[idx], Cr, Transform.identity(), view)
    (ref ** 2).sum().backward()
    torch.cuda.synchronize()
    for trial in range(3):
        z = _busy_torch_stream(torch)
        U = (U0 + z.to(torch.float64)).requires_grad_(True)
        Cc = (C0 + z).detach().requires_grad_(True)
        out = render_torch(dev, U[idx], Cc, Transform.identity(), view)
        z2 = _busy_torch_stream(torch)
        ((out + z2) ** 2).sum().backward()
        torch.cuda.synchronize()
        assert torch.equal(out, ref), f"trial {trial}: forward differs"
        assert torch.allclose(U.grad, Ur.grad, rtol=1e-4, atol=1e-4), f"trial {trial}: dL/dpoints"
        assert torch.allclose(Cc.grad, Cr.grad, rtol=1e-4, atol=1e-4), f"trial {trial}: dL/dcolors"


def test_backward_uses_its_forward_state():
    """A second forward with other points before loss.backward() of the first
    must not change the first's gradients."""
    import torch
    W = H = 96
    ds = blob_scene(23, 4, 6, W)
    dev = DeviceScene(ds.scene())
    idx = torch.tensor(ds.index, device="cuda")
    view = (0.0, 0.0, float(W), float(H))
    rng = np.random.default_rng(2)
    target = torch.tensor(rng.uniform(0, 0.4, (H, W, 4)), dtype=torch.float32, device="cuda")

    def grads(second):
        U = torch.tensor(ds.points, dtype=torch.float64, device="cuda", requires_grad=True)
        Cc = torch.tensor(ds.colors, dtype=torch.float32, device="cuda", requires_grad=True)
        out = render_torch(dev, U[idx], Cc, Transform.identity(), view)
        loss = ((out - target) ** 2).sum()
        if second:
            with torch.no_grad():
                moved = torch.tensor(ds.points + 7.0, dtype=torch.float64, device="cuda")
                render_torch(dev, moved[idx], Cc * 0.5, Transform.identity(), view)
        loss.backward()
        return U.grad.cpu().numpy(), Cc.grad.cpu().numpy()

    gu0, gc0 = grads(False)
    gu1, gc1 = grads(True)
    _say(f"max |dgrad| points {np.abs(gu1 - gu0).max():.3e} colours {np.abs(gc1 - gc0).max():.3e}")
    assert np.allclose(gu1, gu0, rtol=1e-4, atol=1e-3 * np.abs(gu0).max())
    assert np.allclose(gc1, gc0, rtol=1e-4, atol=1e-3 * np.abs(gc0).max())


def test_transform_tensor_gradient_through_render_torch():
    """render_torch with a 6-element transform tensor returns dL/d(transform),
    the same values rz_render_backward_device adds to d_grad_tr."""
    import torch
    W = H = 80
    ds = blob_scene(11, 4, 5, W)
    dev = DeviceScene(ds.scene())
    idx = torch.tensor(ds.index, device="cuda")
    view = (0.0, 0.0, float(W), float(H))
    m = [0.9, 0.15, 3.0, -0.1, 1.05, -2.0]
    trt = torch.tensor(m, dtype=torch.float64, device="cuda", requires_grad=True)
    U = torch.tensor(ds.points, dtype=torch.float64, device="cuda")
    Cc = torch.tensor(ds.colors, dtype=torch.float32, device="cuda")
    out = render_torch(dev, U[idx], Cc, trt, view)
    target = torch.full_like(out, 0.2)
    ((out - target) ** 2).sum().backward()
    g = (2.0 * (out.detach() - target)).contiguous()
    gp = torch.zeros((dev.n_points(), 2), dtype=torch.float32, device="cuda")
    gc = torch.zeros((4, 4), dtype=torch.float32, device="cuda")
    gtr = torch.zeros(6, dtype=torch.float64, device="cuda")
    dev.update_device(U[idx].contiguous().data_ptr(), Cc.data_ptr())
    dev.backward_device(Transform(*m), view, g.data_ptr(), W, gp.data_ptr(), gc.data_ptr(),
                        gtr.data_ptr())
    dev.ctx.check(rz._abi.lib().rz_sync(dev.ctx.h))
    _say(f"render_torch dL/dtr {trt.grad.cpu().numpy()} raw {gtr.cpu().numpy()}")
    assert torch.allclose(trt.grad, gtr, rtol=1e-5, atol=1e-6 * float(gtr.abs().max()))


def test_backward_capacity_miss_leaves_transform_gradient():
    """A backward on a fresh context outgrows its buffers (RZ_E_CAPACITY):
    d_grad_tr must be left as it was; the re-run then adds the gradient once."""
    import torch
    W = H = 512
    ds, _, _ = rz.workload_diff(9, 128, 32, 512.0)  # about 200k lines: more than a first call holds
    ctx = rz.Context(0)
    dev = DeviceScene(ds.scene(), ctx=ctx)
    view = (0.0, 0.0, float(W), float(H))
    g = torch.full((H, W, 4), 0.1, dtype=torch.float32, device="cuda")
    gp = torch.zeros((dev.n_points(), 2), dtype=torch.float32, device="cuda")
    gc = torch.zeros((128, 4), dtype=torch.float32, device="cuda")
    gtr = torch.zeros(6, dtype=torch.float64, device="cuda")
    torch.cuda.synchronize()
    L = rz._abi.lib()
    dev.backward_device(Transform.identity(), view, g.data_ptr(), W, gp.data_ptr(), gc.data_ptr(),
                        gtr.data_ptr())
    s = L.rz_sync(ctx.h)
    _say(f"first backward status {s}, |gtr| {float(gtr.abs().sum()):.3e}")
    assert s == rz._abi.RZ_E_CAPACITY
    assert float(gtr.abs().sum()) == 0.0
    assert float(gp.abs().sum()) == 0.0 and float(gc.abs().sum()) == 0.0
    for _ in range(3):
        dev.backward_device(Transform.identity(), view, g.data_ptr(), W, gp.data_ptr(),
                            gc.data_ptr(), gtr.data_ptr())
        s = L.rz_sync(ctx.h)
        if s != rz._abi.RZ_E_CAPACITY:
            break
    ctx.check(s)
    once = gtr.clone()
    gtr.zero_()
    gp.zero_()
    gc.zero_()
    dev.backward_device(Transform.identity(), view, g.data_ptr(), W, gp.data_ptr(), gc.data_ptr(),
                        gtr.data_ptr())
    ctx.check(L.rz_sync(ctx.h))
    assert float(once.abs().sum()) > 0.0
    assert torch.allclose(once, gtr, rtol=1e-6, atol=1e-9 * float(gtr.abs().max()))
    dev.close()
    ctx.close()


def _gradient_scene(stops):
    """Three gradient fills (linear with linear colours and unsorted stops,
    radial and reflected linear interpolating sRGB colours) under a solid
    one; `stops` (7, 4) premultiplied linear RGBA in the scene's stop order."""
    from rasterize_amd import (GradLinear, GradRadial, GradSpread, LinColor, PathBuilder, Scene,
                               Units)
    from _paths import rect
    s = [tuple(float(v) for v in c) for c in stops]
    p1 = rect(8, 8, 70, 60)
    g1 = GradLinear([(0.8, s[0]), (0.1, s[1]), (0.5, s[2])], units=Units.BoundingBox,
                    linear_colors=True, start=(0.0, 0.0), end=(1.0, 0.3))
    p2 = PathBuilder().move_to((50, 50)).circle(30).build()
    g2 = GradRadial([(0.0, s[3]), (1.0, s[4])], center=(50, 45), radius=35, fcenter=(45, 40))
    p3 = rect(30, 20, 90, 80)
    g3 = GradLinear([(0.2, s[5]), (0.9, s[6])], spread=GradSpread.Reflect, start=(30, 20),
                    end=(50, 40))
    p4 = rect(20, 50, 60, 90)
    return Scene.group([Scene.fill(p1, g1), Scene.fill(p2, g2), Scene.fill(p3, g3),
                        Scene.fill(p4, LinColor(0.2, 0.3, 0.1, 0.5))])


def test_gradient_stop_colours_match_finite_differences():
    """dL/d(stop colour) through GradStops::at and the sRGB conversions, for
    stops given out of order, against central differences of the oracle's
    render (loss linear in the layer, so the differences see only the paint)."""
    import torch
    W = H = 96
    rng = np.random.default_rng(8)
    a = rng.uniform(0.35, 0.95, 7)
    stops = np.concatenate([rng.uniform(0.1, 0.9, (7, 3)) * a[:, None], a[:, None]], 1)
    stops = stops.astype(np.float32)
    weights = rng.uniform(-1.0, 1.0, (H, W, 4)).astype(np.float32)
    sc = _gradient_scene(stops)
    dev = DeviceScene(sc)
    assert dev.n_stops() == 7
    view = (0.0, 0.0, float(W), float(H))
    fl = sc.fills()
    pts = torch.tensor(np.concatenate([f[0].points for f in fl]).reshape(-1, 2), dtype=torch.float64,
                       device="cuda")
    cols = torch.tensor([f[1] if isinstance(f[1], tuple) else (0.0, 0.0, 0.0, 0.0) for f in fl],
                        dtype=torch.float32, device="cuda")
    S = torch.tensor(stops, device="cuda", requires_grad=True)
    out = render_torch(dev, pts, cols, Transform.identity(), view, stops=S)
    ref, _ = O.scene_render([(f[0], f[1].abi(), f[2]) for f in fl], Transform.identity(), view)
    assert np.abs(out.detach().cpu().numpy() - ref).max() <= 2e-3
    loss = (out * torch.tensor(weights, device="cuda")).sum()
    loss.backward()
    g = S.grad.cpu().numpy()

    def loss_ref(st):
        f2 = _gradient_scene(st).fills()
        img, _ = O.scene_render([(f[0], f[1].abi(), f[2]) for f in f2], Transform.identity(), view)
        return float((img.astype(np.float64) * weights).sum())

    errs = []
    for i in range(7):
        for c in range(4):
            h = 1e-2 if i < 3 else 2e-3  # the first paint interpolates linear colours: exactly linear
            sp, sm = stops.copy(), stops.copy()
            sp[i, c] += h
            sm[i, c] -= h
            fd = (loss_ref(sp) - loss_ref(sm)) / (2 * h)
            errs.append((fd, float(g[i, c]), f"stop {i} channel {c}"))
    scale = max(abs(e[0]) for e in errs)
    bad = [(fd, ad, w) for fd, ad, w in errs if abs(fd - ad) > 2e-3 * max(abs(fd), 0.05 * scale)]
    _say("stop gradients: " + ", ".join(f"{w}: fd {fd:.4g} adj {ad:.4g}" for fd, ad, w in errs[:8]))
    assert not bad, bad
    assert np.count_nonzero(g) >= 20

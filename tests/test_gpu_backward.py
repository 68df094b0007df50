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

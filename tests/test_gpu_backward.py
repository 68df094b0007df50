"""Adjoint of the batched scene render (rz_render_backward_device) against
FP64 central finite differences of the loss, the scene rendered in f64 from
the oracle's per-path coverage (SignedDifferenceRasterizer::mask_iter) and an
f64 source-over. Parameters: unique control points (shared endpoints move
together) and solid colours."""
import sys

import numpy as np
import pytest

import _oracle as O
from _diffscene import blob_scene, gather_index, render_f64, scene_from
import rasterize_amd as rz
from rasterize_amd import DeviceScene, Transform, render_torch

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, file=sys.stderr, flush=True)


def test_gradients_match_finite_differences():
    import torch
    W = H = 96
    NP, NC = 6, 6
    uniq, colors = blob_scene(7, NP, NC, W)
    target = render_f64(O, uniq + np.random.default_rng(1).normal(0, 2.0, uniq.shape),
                        np.clip(colors + np.random.default_rng(2).normal(0, 0.05, colors.shape), 0, 1),
                        W, H)
    ds = DeviceScene(scene_from(uniq, colors))
    idx = torch.tensor(gather_index(NP, NC), device="cuda")
    U = torch.tensor(uniq.reshape(-1, 2), dtype=torch.float64, device="cuda", requires_grad=True)
    Cc = torch.tensor(colors, dtype=torch.float32, device="cuda", requires_grad=True)
    view = (0.0, 0.0, float(W), float(H))
    out = render_torch(ds, U[idx], Cc, Transform.identity(), view)
    T = torch.tensor(target, dtype=torch.float32, device="cuda")
    loss = ((out - T) ** 2).sum()
    loss.backward()
    gU = U.grad.cpu().numpy().reshape(uniq.shape)
    gC = Cc.grad.cpu().numpy()
    # forward agrees with the f64 render
    f64 = render_f64(O, uniq, colors, W, H)
    assert np.abs(out.detach().cpu().numpy() - f64).max() < 2e-3

    def loss64(u, c):
        return float(((render_f64(O, u, c, W, H) - target) ** 2).sum())

    rng = np.random.default_rng(5)
    errs = []
    h = 1e-3
    for _ in range(24):
        k, i, d = rng.integers(NP), rng.integers(3 * NC), rng.integers(2)
        up, um = uniq.copy(), uniq.copy()
        up[k, i, d] += h
        um[k, i, d] -= h
        fd = (loss64(up, colors) - loss64(um, colors)) / (2 * h)
        errs.append((fd, float(gU[k, i, d]), f"point {k},{i},{'xy'[d]}"))
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
    assert np.median(rel) < 2e-3
    assert rel.max() < 5e-2


def test_zero_upstream_gradient_gives_zero():
    import torch
    uniq, colors = blob_scene(3, 3, 4, 64)
    ds = DeviceScene(scene_from(uniq, colors))
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

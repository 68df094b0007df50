"""HIP-graph replay of repeated frames (RZ_PARAM_GRAPH): a captured frame
must give exactly what the eager calls give, on every replay, with the
program's own eager work and synchronisations between the replays (the
replays must not depend on state an earlier replay left), and must follow
device data the frame reads (rz_scene_update_device) between replays."""
import ctypes as C
import sys

import numpy as np
import pytest

import _oracle as O
import rasterize_amd as rz
from rasterize_amd import Transform, _abi

pytestmark = pytest.mark.gpu
I = Transform.identity()


def _say(msg):
    print(msg, file=sys.stderr, flush=True)


def _ctx(graph):
    ctx = rz.Context(0)
    ctx.check(_abi.lib().rz_set_param(ctx.h, _abi.RZ_PARAM_GRAPH, int(graph)))
    return ctx


def _replays(ctx):
    v = C.c_int()
    ctx.check(_abi.lib().rz_get_param(ctx.h, _abi.RZ_PARAM_GRAPH_REPLAYS, C.byref(v)))
    return v.value


def test_mask_frames_replay_exactly():
    import torch
    L = _abi.lib()
    ws = rz.workload_random_walk(2, 20_000, 1024.0)
    path, tr = ws.paths[0], ws.transform
    outs = {}
    for graph in (False, True):
        ctx = _ctx(graph)
        d = C.c_void_p()
        ctx.check(L.rz_path_upload(ctx.h, C.byref(path.abi()), C.byref(d)))
        out = torch.zeros((1024, 1024), dtype=torch.float32, device="cuda")
        frames = []
        for f in range(6):
            out.fill_(0.5)  # eager torch work between frames
            torch.cuda.synchronize()
            s = rz._abi.RZ_E_CAPACITY
            for _ in range(3):
                ctx.check(L.rz_mask_device(ctx.h, d, tr.abi(), 0.05, f % 2, _abi.RZ_EDGE_MASK, 1024,
                                           1024, C.c_void_p(out.data_ptr()), 1024))
                s = L.rz_sync(ctx.h)
                if s != rz._abi.RZ_E_CAPACITY:
                    break
            ctx.check(s)
            frames.append(out.cpu().numpy().copy())
        outs[graph] = frames
        _say(f"graph={graph}: replays {_replays(ctx)}")
        if graph:
            assert _replays(ctx) >= 2
        L.rz_path_free(ctx.h, d)
        ctx.close()
    for f, (a, b) in enumerate(zip(outs[False], outs[True])):
        assert np.array_equal(a, b), f"frame {f}"
    ref = O.sd_mask(path, tr, 1024, 1024, 1)
    assert np.abs(outs[True][-1].astype(np.float64) - ref).max() <= 2e-3


def test_scene_frames_follow_updated_points():
    """The differentiable render's loop: new points each frame, one captured
    forward (and backward) replayed."""
    import torch
    from _diffscene import blob_scene
    W = H = 128
    ds = blob_scene(31, 5, 6, W)
    idx = torch.tensor(ds.index, device="cuda")
    view = (0.0, 0.0, float(W), float(H))
    res = {}
    for graph in (False, True):
        ctx = _ctx(graph)
        dev = rz.DeviceScene(ds.scene(), ctx=ctx)
        out = torch.zeros((H, W, 4), dtype=torch.float32, device="cuda")
        g = torch.full((H, W, 4), 0.25, dtype=torch.float32, device="cuda")
        gp = torch.zeros((dev.n_points(), 2), dtype=torch.float32, device="cuda")
        gc = torch.zeros((5, 4), dtype=torch.float32, device="cuda")
        gt = torch.zeros(6, dtype=torch.float64, device="cuda")
        L = _abi.lib()
        frames = []
        for f in range(7):
            U = torch.tensor(ds.points + 1.5 * f, dtype=torch.float64, device="cuda")[idx].contiguous()
            cols = torch.tensor(ds.colors * (1.0 - 0.05 * f), dtype=torch.float32, device="cuda")
            gp.zero_()
            gc.zero_()
            gt.zero_()
            torch.cuda.synchronize()
            for _ in range(3):
                dev.update_device(U.data_ptr(), cols.data_ptr())
                dev.render_device(I, view, out.data_ptr(), W)
                dev.backward_device(I, view, g.data_ptr(), W, gp.data_ptr(), gc.data_ptr(),
                                    gt.data_ptr())
                s = L.rz_sync(ctx.h)
                if s != _abi.RZ_E_CAPACITY:
                    break
                gp.zero_()
                gc.zero_()
                gt.zero_()
            ctx.check(s)
            frames.append((out.cpu().numpy().copy(), gp.cpu().numpy().copy(),
                           gc.cpu().numpy().copy(), gt.cpu().numpy().copy()))
        res[graph] = frames
        _say(f"graph={graph}: replays {_replays(ctx)}")
        if graph:
            assert _replays(ctx) >= 8
        dev.close()
        ctx.close()
    for f, (a, b) in enumerate(zip(res[False], res[True])):
        assert np.array_equal(a[0], b[0]), f"frame {f}: layer"
        for k in (1, 2, 3):  # gradients: float atomics, so equal up to order
            assert np.allclose(a[k], b[k], rtol=1e-5, atol=1e-6 * max(1.0, np.abs(a[k]).max()))


def test_atlas_and_scene_frames_replay():
    import torch
    L = _abi.lib()
    gs = rz.workload_glyphs(4, 500, 2048)
    W, H = gs.atlas_size
    ws = rz.workload_scene(3, 300, 512.0)
    res = {}
    for graph in (False, True):
        ctx = _ctx(graph)
        dg = rz.DeviceGlyphs(gs.paths, gs.slots, ctx=ctx)
        dsn = rz.DeviceScene(ws.scene(), ctx=ctx)
        atlas = torch.zeros((H, W), dtype=torch.float32, device="cuda")
        lay = torch.zeros((512, 512, 4), dtype=torch.float32, device="cuda")
        frames = []
        for f in range(5):
            atlas.fill_(-1.0)
            lay.fill_(-1.0)
            torch.cuda.synchronize()
            for _ in range(4):
                dg.mask_device(atlas.data_ptr(), W)
                dsn.render_device(I, (0.0, 0.0, 512.0, 512.0), lay.data_ptr(), 512, bg=(0.1, 0.0, 0.2, 0.5))
                s = L.rz_sync(ctx.h)
                if s != _abi.RZ_E_CAPACITY:
                    break
            ctx.check(s)
            frames.append((atlas.cpu().numpy().copy(), lay.cpu().numpy().copy()))
        res[graph] = frames
        _say(f"graph={graph}: replays {_replays(ctx)}")
        if graph:
            assert _replays(ctx) >= 4
        dg.close()
        dsn.close()
        ctx.close()
    for f, (a, b) in enumerate(zip(res[False], res[True])):
        assert np.array_equal(a[0], b[0]), f"frame {f}: atlas"
        assert np.array_equal(a[1], b[1]), f"frame {f}: scene"

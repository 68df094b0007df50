"""Launch tuning (rz_set_param's launch choices and the rz_tune_* search):
every choice gives the same output bit for bit, the search never picks a
slower choice than the context's, its entry is used by later calls of the
same workload, and the cache file round-trips."""
import os
import sys

import numpy as np
import pytest

import rasterize_amd as rz
from rasterize_amd import Transform, _abi

pytestmark = pytest.mark.gpu
I = Transform.identity()


def _say(msg):
    print(msg, file=sys.stderr, flush=True)


def _mask(dp, ctx, tr, out, W, H, rule=0):
    for _ in range(4):
        dp.mask_device(tr, W, H, out.data_ptr(), W, rule)
        s = _abi.lib().rz_sync(ctx.h)
        if s != _abi.RZ_E_CAPACITY:
            break
    ctx.check(s)
    return out.cpu().numpy().copy()


SETTINGS = ([(_abi.RZ_PARAM_SPLIT_QUAD, v) for v in (0, 2, 6)] +
            [(_abi.RZ_PARAM_SPLIT_CUBIC, v) for v in (0, 3, 6)] +
            [(_abi.RZ_PARAM_FLAT_LDS, v) for v in (0, 16)] +
            [(_abi.RZ_PARAM_BIN_LDS, v) for v in (8, 24)] +
            [(_abi.RZ_PARAM_FINE_LDS, v) for v in (16, 32)] +
            [(_abi.RZ_PARAM_SCATTER_WAVES, v) for v in (8, 12, 18)])


def test_launch_params_do_not_change_masks():
    import torch
    ctx = rz.Context(0)
    ws = rz.workload_random_walk(2, 20_000, 1024.0)
    dp = rz.DevicePath(ws.paths[0], ctx)
    W = H = 1024
    out = torch.zeros((H, W), dtype=torch.float32, device="cuda")
    for rule in (0, 1):
        ref = _mask(dp, ctx, ws.transform, out, W, H, rule)
        for p, v in SETTINGS:
            ctx.set_param(p, v)
            got = _mask(dp, ctx, ws.transform, out, W, H, rule)
            assert np.array_equal(got, ref), f"rule {rule}, param {p} = {v}"
        ctx.set_param(_abi.RZ_PARAM_SPLIT_AUTO, 1)
        for p, v in ((_abi.RZ_PARAM_FLAT_LDS, -1), (_abi.RZ_PARAM_BIN_LDS, 0),
                     (_abi.RZ_PARAM_FINE_LDS, 0), (_abi.RZ_PARAM_SCATTER_WAVES, 15)):
            ctx.set_param(p, v)
    for p, v in ((_abi.RZ_PARAM_FLAT_LDS, 33), (_abi.RZ_PARAM_SCATTER_WAVES, 7),
                 (_abi.RZ_PARAM_TUNED, 2)):
        with pytest.raises(rz.RzError):
            ctx.set_param(p, v)
    dp.close()
    ctx.close()


def test_tuner_mask_entry_is_used_and_cached(tmp_path):
    import torch
    ctx = rz.Context(0)
    ws = rz.workload_random_walk(2, 20_000, 1024.0)
    dp = rz.DevicePath(ws.paths[0], ctx)
    W = H = 1024
    out = torch.zeros((H, W), dtype=torch.float32, device="cuda")
    ref = _mask(dp, ctx, ws.transform, out, W, H)
    r = dp.tune(ws.transform, W, H, out.data_ptr(), W, budget_ms=600.0)
    _say(f"tuned mask: {r}")
    assert r["trials"] >= 2
    assert r["ms_tuned"] <= r["ms_default"]
    assert ctx.tune_entries() == 1
    assert np.array_equal(out.cpu().numpy(), ref)  # the search's own frames
    assert np.array_equal(_mask(dp, ctx, ws.transform, out, W, H), ref)  # with the entry
    # another workload (size) has no entry
    small = torch.zeros((256, 256), dtype=torch.float32, device="cuda")
    _mask(dp, ctx, Transform.new_scale(0.25, 0.25), small, 256, 256)
    assert ctx.tune_entries() == 1
    f = str(tmp_path / "tune.txt")
    with open(f, "w") as fh:
        fh.write("# a comment\nothergpu/1cu 0 0 20 14 1 2 0 3 4 -1 0 0 15 6 0 1.0 0.9\nnot a line\n")
    ctx.tune_save(f)
    text = open(f).read()
    assert "othergpu/1cu" in text and "not a line" not in text
    ctx.tune_clear()
    assert ctx.tune_entries() == 0
    assert ctx.tune_load(f) == 1
    assert ctx.tune_entries() == 1
    ctx.set_param(_abi.RZ_PARAM_TUNED, 0)
    assert np.array_equal(_mask(dp, ctx, ws.transform, out, W, H), ref)
    dp.close()
    ctx.close()


def test_tuner_scene_and_atlas_keep_their_output():
    import torch
    ctx = rz.Context(0)
    ws = rz.workload_scene(3, 400, 512.0)
    ds = rz.DeviceScene(ws.scene(), ctx)
    W = H = 512
    view = (0.0, 0.0, float(W), float(H))
    out = torch.zeros((H, W, 4), dtype=torch.float32, device="cuda")

    def render():
        for _ in range(4):
            ds.render_device(I, view, out.data_ptr(), W)
            s = _abi.lib().rz_sync(ctx.h)
            if s != _abi.RZ_E_CAPACITY:
                break
        ctx.check(s)
        return out.cpu().numpy().copy()
    ref = render()
    r = ds.tune(I, view, out.data_ptr(), W, budget_ms=500.0)
    _say(f"tuned scene: {r}")
    assert r["ms_tuned"] <= r["ms_default"]
    assert np.array_equal(render(), ref)
    # the scene fine kernel's register budgets give the same image
    ctx.set_param(_abi.RZ_PARAM_TUNED, 0)
    for v in (1, 8, 6):
        ctx.set_param(_abi.RZ_PARAM_SCENE_WAVES, v)
        assert np.array_equal(render(), ref), f"scene waves {v}"
    with pytest.raises(rz.RzError):
        ctx.set_param(_abi.RZ_PARAM_SCENE_WAVES, 4)
    # the queued (persistent, costliest tiles first) fine stage too
    for v in (8, 6):
        ctx.set_param(_abi.RZ_PARAM_SCENE_WAVES, v)
        ctx.set_param(_abi.RZ_PARAM_SCENE_QUEUE, 1)
        assert np.array_equal(render(), ref), f"queued, scene waves {v}"
    ctx.set_param(_abi.RZ_PARAM_SCENE_QUEUE, 0)
    ctx.set_param(_abi.RZ_PARAM_TUNED, 1)

    gs = rz.workload_glyphs(4, 2000, 2048)
    AW, AH = gs.atlas_size
    dg = rz.DeviceGlyphs(gs.paths, gs.slots, ctx)
    atlas = torch.zeros((AH, AW), dtype=torch.float32, device="cuda")

    def masks():
        for _ in range(4):
            dg.mask_device(atlas.data_ptr(), AW)
            s = _abi.lib().rz_sync(ctx.h)
            if s != _abi.RZ_E_CAPACITY:
                break
        ctx.check(s)
        return atlas.cpu().numpy().copy()
    aref = masks()
    r = dg.tune(atlas.data_ptr(), AW, budget_ms=500.0)
    _say(f"tuned atlas: {r}")
    assert r["ms_tuned"] <= r["ms_default"]
    assert np.array_equal(masks(), aref)
    assert ctx.tune_entries() == 2
    ds.close()
    dg.close()
    ctx.close()

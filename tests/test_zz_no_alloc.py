"""rz_mask_device_noalloc (the entry point that allocates nothing) against
rz_mask_device on the same inputs, through the C ABI: the masks must be
identical (both pipelines deposit 2^-20 fixed-point coverage with integer
atomics, so buffer placement cannot change a bit), on consecutive calls whose
second input is smaller, and within the mask tolerance of the oracle."""
import ctypes as C
import math

import numpy as np
import pytest

import _oracle as O
from _paths import squirrel
import rasterize_amd as rz
from rasterize_amd import Transform, workload_cubics64, workload_random_walk

pytestmark = pytest.mark.gpu
I = Transform.identity()


def _say(*a):
    print(*a, flush=True)


def _upload(L, ctx, path):
    d = C.c_void_p()
    ctx.check(L.rz_path_upload(ctx.h, C.byref(path.abi()), C.byref(d)))
    return d


def _mask(L, ctx, fn, d, tr, W, H, rule, out):
    ctx.check(fn(ctx.h, d, tr.abi(), 0.05, rule, rz._abi.RZ_EDGE_MASK, W, H,
                 C.c_void_p(out.data_ptr()), W))


def test_noalloc_matches_mask_device():
    import torch
    L = rz._abi.lib()
    ctx = rz.Context(0)
    c1 = workload_cubics64(1, 512.0).paths[0]
    sq = squirrel()
    tr_sq = Transform.new_rotate(math.pi / 5.0).pre_translate(-4.0, -6.0)
    cases = [(c1, I, 512, 512, 0), (sq, tr_sq, 200, 150, 1)]  # the second is smaller
    ds = [_upload(L, ctx, p) for p, *_ in cases]
    try:
        outs = []
        for (p, tr, W, H, rule), d in zip(cases, ds):
            _say(f"no-alloc mask {W}x{H} rule {rule}")
            ref = torch.zeros((H, W), dtype=torch.float32, device="cuda")
            got = torch.full((H, W), 7.0, dtype=torch.float32, device="cuda")
            torch.cuda.synchronize()
            _mask(L, ctx, L.rz_mask_device, d, tr, W, H, rule, ref)
            ctx.check(L.rz_sync(ctx.h))
            _mask(L, ctx, L.rz_mask_device_noalloc, d, tr, W, H, rule, got)
            ctx.check(L.rz_sync(ctx.h))
            g, r = got.cpu().numpy(), ref.cpu().numpy()
            assert np.array_equal(g, r), f"{W}x{H}: no-alloc mask differs from rz_mask_device"
            o = O.sd_mask(p, tr, W, H, rule)
            dm = np.abs(g.astype(np.float64) - o)
            _say(f"  identical to rz_mask_device; vs oracle max {dm.max():.2e} mean {dm.mean():.2e}")
            assert dm.max() <= 2e-3 and dm.mean() <= 1e-5
            outs.append((g, got))
        # both queued before one sync, the larger first: the second run reuses
        # the scratch only after the first run's kernels are done (stream order)
        for (_, got), (p, tr, W, H, rule) in zip(outs, cases):
            got.fill_(3.0)
        torch.cuda.synchronize()
        for (_, got), (p, tr, W, H, rule), d in zip(outs, cases, ds):
            _mask(L, ctx, L.rz_mask_device_noalloc, d, tr, W, H, rule, got)
        ctx.check(L.rz_sync(ctx.h))
        for (g0, got), (p, tr, W, H, rule) in zip(outs, cases):
            assert np.array_equal(got.cpu().numpy(), g0), f"{W}x{H}: queued no-alloc call differs"
    finally:
        for d in ds:
            L.rz_path_free(ctx.h, d)
        ctx.close()


def test_noalloc_refuses_what_does_not_fit():
    import torch
    L = rz._abi.lib()
    ctx = rz.Context(0)
    c1 = workload_cubics64(1, 512.0).paths[0]
    ws = workload_random_walk(2, 200_000, 4096.0)
    big, tr2 = ws.paths[0], ws.transform
    d1, d2 = _upload(L, ctx, c1), _upload(L, ctx, big)
    try:
        # a 32768^2 mask: its per-tile arrays alone exceed the scratch; refused
        # before anything is enqueued (the output could hold it anyway)
        huge = torch.empty((32768, 32768), dtype=torch.float32, device="cuda")
        s = L.rz_mask_device_noalloc(ctx.h, d1, I.abi(), 0.05, 0, rz._abi.RZ_EDGE_MASK, 32768,
                                     32768, C.c_void_p(huge.data_ptr()), 32768)
        msg = L.rz_last_error(ctx.h).decode()
        _say(f"32768^2: status {s} ({msg})")
        assert s == rz._abi.RZ_E_CAPACITY and "scratch" in msg
        del huge
        # 6.6 M lines: enqueued, reported at rz_sync, nothing grown
        out = torch.zeros((4096, 4096), dtype=torch.float32, device="cuda")
        torch.cuda.synchronize()
        _mask(L, ctx, L.rz_mask_device_noalloc, d2, tr2, 4096, 4096, 0, out)
        s = L.rz_sync(ctx.h)
        msg = L.rz_last_error(ctx.h).decode()
        _say(f"200k-segment path: status {s} ({msg})")
        assert s == rz._abi.RZ_E_CAPACITY and "no-alloc" in msg
        assert float(out.abs().sum()) == 0.0
        # the next call fits and is right; the context's own path still works
        small = torch.zeros((512, 512), dtype=torch.float32, device="cuda")
        ref = torch.zeros((512, 512), dtype=torch.float32, device="cuda")
        torch.cuda.synchronize()
        _mask(L, ctx, L.rz_mask_device_noalloc, d1, I, 512, 512, 0, small)
        ctx.check(L.rz_sync(ctx.h))
        _mask(L, ctx, L.rz_mask_device, d1, I, 512, 512, 0, ref)
        ctx.check(L.rz_sync(ctx.h))
        assert torch.equal(small, ref)
    finally:
        L.rz_path_free(ctx.h, d1)
        L.rz_path_free(ctx.h, d2)
        ctx.close()

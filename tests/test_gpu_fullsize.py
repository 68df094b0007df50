"""GPU parity at the full configured sizes of the batched workloads, the
sizes the benchmark times:

- the 10,000-path scene on a 2048x2048 RGBA layer, batched and plain forms,
  against the oracle's Scene::render (all pixels, all channels);
- the 65,536-glyph atlas (409 Mpx of masks) against the oracle's per-glyph
  SignedDifferenceRasterizer::mask;
- the differentiable render, 256 paths x 32 cubics on 1024x1024: the
  adjoint's gradients on 512 sampled parameters against FP64 central finite
  differences (h = 1e-3 px for coordinates, 1e-4 for colours) and against
  the oracle's forward-mode FP64 reference gradient.

Tolerances (written in the tests): coverage and colour max |d| <= 2e-3 and
mean |d| <= 1e-5 per pixel and channel; gradients relative error
|g - ref| / max(|ref|, 1e-2 max|ref|) <= 1e-3 for every sampled parameter.
The oracle runs on the host's cores (bands of rows, glyphs, or parameters
shared out over threads)."""
import concurrent.futures as cf
import sys
import time

import numpy as np
import pytest

import _oracle as O
from _diffscene import local_index, path_of
import rasterize_amd as rz
from rasterize_amd import Transform

pytestmark = pytest.mark.gpu
I = Transform.identity()
MAX_ABS, MEAN_ABS = 2e-3, 1e-5


def _say(msg):
    print(msg, file=sys.stderr, flush=True)


def _settle(L, ctx, run):
    for _ in range(4):
        run()
        s = L.rz_sync(ctx.h)
        if s != rz._abi.RZ_E_CAPACITY:
            break
    ctx.check(s)


@pytest.fixture(scope="module")
def c3():
    ws = rz.workload_scene(3, 10_000, 2048.0)
    fills = [(p, q.abi(), r) for p, q, r in zip(ws.paths, ws.paints, ws.rules)]
    t0 = time.perf_counter()
    ref, org = O.scene_render_mt(fills, I, (0.0, 0.0, 2048.0, 2048.0))
    _say(f"c3 oracle: {time.perf_counter() - t0:.1f} s on {O.host_threads()} threads")
    assert org == (0, 0) and ref.shape == (2048, 2048, 4)
    return ws, ref


@pytest.mark.parametrize("batched", [True, False])
def test_c3_scene_full_size(c3, batched):
    import torch
    ws, ref = c3
    L = rz._abi.lib()
    ds = rz.DeviceScene(ws.scene())
    out = torch.full((2048, 2048, 4), -1.0, dtype=torch.float32, device="cuda")
    torch.cuda.synchronize()
    _settle(L, ds.ctx, lambda: ds.render_device(I, (0.0, 0.0, 2048.0, 2048.0), out.data_ptr(), 2048,
                                                batched=batched))
    nl, nr, ne = (rz._abi.C.c_size_t() for _ in range(3))
    ds.ctx.check(L.rz_last_counts(ds.ctx.h, rz._abi.C.byref(nl), rz._abi.C.byref(nr),
                                  rz._abi.C.byref(ne)))
    got = out.cpu().numpy()
    ds.close()
    d = np.abs(got.astype(np.float64) - ref)
    mx, mean = float(d.max()), float(d.mean())
    worst = np.unravel_index(int(d.argmax()), d.shape)
    _say(f"c3 {'batched' if batched else 'plain'}: {nl.value} lines, {nr.value} records; "
         f"max|d|={mx:.3e} at {worst} mean|d|={mean:.3e}")
    assert mx <= MAX_ABS and mean <= MEAN_ABS


def test_c4_atlas_full_size():
    import torch
    gs = rz.workload_glyphs(4, 65_536, 16384)
    W, H = gs.atlas_size
    L = rz._abi.lib()
    dg = rz.DeviceGlyphs(gs.paths, gs.slots)
    out = torch.full((H, W), -1.0, dtype=torch.float32, device="cuda")
    torch.cuda.synchronize()
    _settle(L, dg.ctx, lambda: dg.mask_device(out.data_ptr(), W))
    got = out.cpu().numpy()
    del out
    dg.close()
    trs = np.array([list(s.tr.m) for s in gs.slots])
    q = np.array([[s.x, s.y, s.width, s.height] for s in gs.slots], np.uint32)
    t0 = time.perf_counter()
    ref = np.full((H, W), -1.0, np.float32)  # pixels outside every view stay as they were
    O.mask_atlas(gs.paths, trs, q, W, H, out=ref)
    _say(f"c4 oracle: {time.perf_counter() - t0:.1f} s on {O.host_threads()} threads")
    mx, tot = 0.0, 0.0
    for r0 in range(0, H, 1024):
        d = np.abs(got[r0:r0 + 1024].astype(np.float64) - ref[r0:r0 + 1024])
        mx = max(mx, float(d.max()))
        tot += float(d.sum())
    px = int((q[:, 2].astype(np.int64) * q[:, 3]).sum())
    _say(f"c4 atlas {W}x{H}, {len(gs.paths)} glyphs, {px} mask pixels: max|d|={mx:.3e} "
         f"mean|d|={tot / px:.3e}")
    assert mx <= MAX_ABS and tot / px <= MEAN_ABS


def test_c5_gradients_full_size():
    """Adjoint vs FP64 finite differences and the FP64 forward-mode gradient on
    512 sampled parameters (448 control-point coordinates of 64 paths, 64
    colour channels). A coordinate whose +-h moves change the path's
    subdivision (a leaf count differs) is redrawn: the adjoint, like the
    forward-mode gradient, differentiates at fixed subdivision."""
    import torch
    W = H = 1024
    NP, NC = 256, 32
    ds, tp, tc = rz.workload_diff(5, NP, NC, 1024.0)
    U0 = ds.points.copy()
    cols = ds.colors.astype(np.float64)
    pool = cf.ThreadPoolExecutor(O.host_threads())

    def alpha(u, k):
        return O.sd_mask_iter(path_of(ds, u, k), I, W, H, 0)

    t0 = time.perf_counter()
    # target: the perturbed scene, f64
    T = np.zeros((H, W, 4))
    for k, a in enumerate(pool.map(lambda k: alpha(tp, k), range(NP))):
        s = tc[k][None, None, :] * a[..., None]
        T = s + T * (1.0 - s[..., 3:4])
    rng = np.random.default_rng(12)
    sampled = sorted(rng.choice(NP, 64, replace=False).tolist())
    # reverse sweep: C_final and, for sampled paths, the transmittance above
    Q = np.ones((H, W))
    P = np.zeros((H, W, 4))
    Qk = {}
    for k, a in zip(range(NP - 1, -1, -1), pool.map(lambda k: alpha(U0, k), range(NP - 1, -1, -1))):
        if k in sampled:
            Qk[k] = Q.copy()
        P += Q[..., None] * (cols[k][None, None, :] * a[..., None])
        Q *= 1.0 - cols[k][3] * a
    E = P - T  # C_final - target
    _say(f"c5 oracle setup {time.perf_counter() - t0:.1f} s")

    # the adjoint on the GPU
    dev = rz.DeviceScene(ds.scene())
    idx = torch.tensor(ds.index, device="cuda")
    Ug = torch.tensor(U0, dtype=torch.float64, device="cuda", requires_grad=True)
    Cg = torch.tensor(ds.colors, dtype=torch.float32, device="cuda", requires_grad=True)
    out = rz.render_torch(dev, Ug[idx], Cg, I, (0.0, 0.0, float(W), float(H)))
    Tg = torch.tensor(T, dtype=torch.float32, device="cuda")
    ((out - Tg) ** 2).sum().backward()
    gU = Ug.grad.cpu().numpy()
    gC = Cg.grad.cpu().numpy()
    fwd = np.abs(out.detach().cpu().numpy() - P)
    _say(f"c5 forward vs f64: max|d|={fwd.max():.3e} mean|d|={fwd.mean():.3e}")
    assert fwd.max() <= MAX_ABS and fwd.mean() <= MEAN_ABS

    h, hc = 1e-3, 1e-4
    rows = []  # (kind, adjoint, fd, forward-mode)
    redrawn = 0
    B = np.zeros((H, W, 4))
    for k, ak in zip(range(NP), pool.map(lambda k: alpha(U0, k), range(NP))):
        if k in Qk:
            c = cols[k]
            Qp = Qk[k]
            G = 2.0 * E * Qp[..., None]            # dL/dC_k
            A = c[None, None, :] - c[3] * B        # dC_k/dalpha
            base = len(O.flatten(path_of(ds, U0, k), I))
            picks = []
            while len(picks) < 7:
                i, j = int(rng.integers(3 * NC)), int(rng.integers(2))
                g = 3 * NC * k + i
                up, um = U0.copy(), U0.copy()
                up[g, j] += h
                um[g, j] -= h
                if (len(O.flatten(path_of(ds, up, k), I)) != base or
                        len(O.flatten(path_of(ds, um, k), I)) != base):
                    redrawn += 1
                    continue
                picks.append((i, j, up, um))
            li = local_index(ds, k)

            def dual(ij):
                i, j = ij
                tan = np.zeros((len(li), 2))
                tan[li == i, j] = 1.0
                return O.sd_mask_iter_dual(path_of(ds, U0, k), tan, I, W, H, 0)[1]
            plus = pool.map(lambda t: alpha(t[2], k), picks)
            minus = pool.map(lambda t: alpha(t[3], k), picks)
            duals = pool.map(dual, [(t[0], t[1]) for t in picks])

            def dl(da):
                nz = da != 0.0
                dC = A[nz] * da[nz][:, None]
                q = Qp[nz][:, None]
                return float((2.0 * E[nz] * q * dC + (q * dC) ** 2).sum())
            for (i, j, _, _), ap, am, dd in zip(picks, plus, minus, duals):
                fd = (dl(ap - ak) - dl(am - ak)) / (2 * h)
                fm = float((G * A * dd[..., None]).sum())
                rows.append(("point", float(gU[3 * NC * k + i, j]), fd, fm))
            # one colour channel: C_k is linear in it, the loss quadratic
            ch = int(rng.integers(4))

            def ck(cc):
                return cc[None, None, :] * ak[..., None] + B * (1.0 - cc[3] * ak[..., None])
            C0 = ck(c)
            dls = []
            for s in (1.0, -1.0):
                cc = c.copy()
                cc[ch] += s * hc
                dC = ck(cc) - C0
                dls.append(float((2.0 * E * Qp[..., None] * dC + (Qp[..., None] * dC) ** 2).sum()))
            fd = (dls[0] - dls[1]) / (2 * hc)
            e = np.zeros(4)
            e[ch] = 1.0
            dCdc = e[None, None, :] * ak[..., None] - (B * ak[..., None] if ch == 3 else 0.0)
            fm = float((G * dCdc).sum())
            rows.append(("colour", float(gC[k, ch]), fd, fm))
        B = cols[k][None, None, :] * ak[..., None] + B * (1.0 - cols[k][3] * ak[..., None])
    pool.shutdown()
    _say(f"c5: {len(rows)} parameters ({redrawn} redrawn for a subdivision change), "
         f"{time.perf_counter() - t0:.1f} s")
    assert len(rows) == 512
    for kind in ("point", "colour"):
        rs = [r for r in rows if r[0] == kind]
        sfd = max(abs(r[2]) for r in rs)
        sfm = max(abs(r[3]) for r in rs)
        e_fd = np.array([abs(r[1] - r[2]) / max(abs(r[2]), 1e-2 * sfd) for r in rs])
        e_fm = np.array([abs(r[1] - r[3]) / max(abs(r[3]), 1e-2 * sfm) for r in rs])
        e_or = np.array([abs(r[2] - r[3]) / max(abs(r[3]), 1e-2 * sfm) for r in rs])
        _say(f"c5 {kind} ({len(rs)}): adjoint vs FD median {np.median(e_fd):.2e} max {e_fd.max():.2e}; "
             f"adjoint vs forward-mode median {np.median(e_fm):.2e} max {e_fm.max():.2e}; "
             f"FD vs forward-mode max {e_or.max():.2e}")
        assert e_fd.max() <= 1e-3 and e_fm.max() <= 1e-3
    dev.close()

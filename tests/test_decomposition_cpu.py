"""CPU check of the tile decomposition: the binning replay's records, band
backdrops and events, accumulated the way the fine stage does (here in f64
numpy), must reproduce the oracle's reference mask. This validates the
decomposition itself (closed-chain backdrops, left-edge crossing terms,
free-endpoint events, right clip) without a GPU."""
import numpy as np
import pytest

import _oracle as O
from _paths import fill_rule_shapes, octagon, path_size, squirrel
from rasterize_amd import PathBuilder, Transform, workload_cubics64, workload_random_walk

I = Transform.identity()


def V(y):
    return np.clip(y - np.arange(16), 0.0, 1.0)


def fine_emulation(counts, recs, bd, events, redge, W, H, rule, mode, phase=(0, 0), arith="f64"):
    """arith: "f64" accumulates the tile's local winding in f64; "int32" wraps
    it to int32 in 2^-20 fixed point, as a fine stage that kept every tile in
    int32 would. In both, every tile the fine stage keeps in int32 (3 records
    + band events < 2000) is checked to stay inside int32."""
    phx, phy = phase
    tx_n, ty_n = (W + phx + 15) // 16, (H + phy + 15) // 16
    out = np.zeros((ty_n * 16, tx_n * 16))
    ev_band = events[:, 0].view(np.int32) if len(events) else np.zeros(0, np.int32)
    ev_col = events[:, 1].view(np.int32) if len(events) else np.zeros(0, np.int32)
    off = np.concatenate([[0], np.cumsum(counts)]).astype(np.int64)
    for t in range(tx_n * ty_n):
        tx, ty = t % tx_n, t // tx_n
        acc = np.zeros((16, 17))
        base = np.zeros(16)
        for r in recs[off[t]:off[t + 1]]:
            x0, y0, x1, y1 = r
            if np.signbit(x0):
                base += V(y0)
            if np.signbit(x1):
                base -= V(y1)
            x0, x1 = abs(x0), abs(x1)
            if y0 == y1:
                continue
            d_ = 1.0
            if y0 > y1:
                d_ = -1.0
                x0, x1, y0, y1 = x1, x0, y1, y0
            dxdy = (x1 - x0) / (y1 - y0)
            xn = x0
            for y in range(int(y0), min(16, int(np.ceil(y1)))):
                x = xn
                dy = min(y + 1.0, y1) - max(float(y), y0)
                d = d_ * dy
                xn = x + dxdy * dy
                xa, xb = min(x, xn), max(x, xn)
                xaf0 = max(np.floor(xa), 0.0)
                xai, xbc = int(xaf0), np.ceil(xb)
                xbi = int(xbc)

                def dep(c, v):
                    if c < 16:
                        acc[y, c] += v
                if xbi <= xai + 1:
                    xmf = 0.5 * (x + xn) - xaf0
                    dep(xai, d * (1.0 - xmf))
                    dep(xai + 1, d * xmf)
                else:
                    s = 1.0 / (xb - xa)
                    xf = xa - xaf0
                    xbf = xb - xbc + 1.0
                    a0 = 0.5 * s * (1.0 - xf) ** 2
                    am = 0.5 * s * xbf * xbf
                    dep(xai, d * a0)
                    if xbi == xai + 2:
                        dep(xai + 1, d * (1.0 - a0 - am))
                    else:
                        a1 = s * (1.5 - xf)
                        dep(xai + 1, d * (a1 - a0))
                        for xi in range(xai + 2, xbi - 1):
                            dep(xi, d * s)
                        a2 = a1 + (xbi - xai - 3) * s
                        dep(xbi - 1, d * (1.0 - a2 - am))
                    dep(xbi, d * am)
        sel = (ev_band == ty) & (ev_col < tx)
        for e in events[sel]:
            base += e[3] * V(e[2])
        local = base[:, None] + np.cumsum(acc[:, :16], axis=1)
        lf = np.rint(local * float(1 << 20)).astype(np.int64)
        if 3 * int(off[t + 1] - off[t]) + int((ev_band == ty).sum()) < 2000:
            assert np.abs(lf).max() < 2 ** 31, "an int32 tile's local winding leaves int32"
        if arith == "int32":
            local = (((lf + 2 ** 31) % 2 ** 32) - 2 ** 31) / float(1 << 20)
        w = bd[t] + local
        if mode != 2:
            # right-edge column: lines beyond the clip, as per-row sums R
            X = (W - 1.0 if mode == 0 else float(W)) - 0.001
            xr = int(np.floor(X))
            rows = np.arange(ty * 16, ty * 16 + 16) - phy
            ok = (rows >= 0) & (rows < H)
            R = np.where(ok, redge[np.clip(rows, 0, H - 1)], 0) / float(1 << 20)
            cols = tx * 16 + np.arange(16) - phx
            wcol = np.where(cols > xr, 1.0, np.where(cols == xr, 1.0 - (X - xr), 0.0))
            w = w + R[:, None] * wcol[None, :]
        if rule == 0:
            a = np.abs(w)
            a = np.where(a >= 1.0, 1.0, np.where(a < 1e-6, 0.0, a))
        else:
            a = np.abs(np.mod(w + 1.0, 2.0) - 1.0)
        out[ty * 16:ty * 16 + 16, tx * 16:tx * 16 + 16] = a
    return out[phy:phy + H, phx:phx + W]


CASES = {
    "octagon": (lambda: (octagon(), I), 9, 4),
    "squirrel": (lambda: (squirrel(), Transform.new_scale(6.0, 6.0).pre_rotate(0.3)), 100, 110),
    "c1_small": (lambda: (workload_cubics64(1, 96.0).paths[0], I), 96, 96),
    "edges": (lambda: (workload_cubics64(5, 120.0).paths[0], Transform.new_translate(-20, -20)), 81, 70),
    "narrow": (lambda: (workload_cubics64(6, 60.0).paths[0], Transform.new_translate(-10, -5)), 17, 40),
    "walk": (lambda: (workload_random_walk(2, 400, 160.0).paths[0], I), 160, 150),
    "fill_rule": (None, None, None),
}


@pytest.mark.parametrize("name", sorted(CASES))
@pytest.mark.parametrize("mode", [0, 1, 2])
@pytest.mark.parametrize("rule", [0, 1])
def test_decomposition_reproduces_reference(name, mode, rule):
    if name == "fill_rule":
        path = fill_rule_shapes()
        W, H, tr = path_size(path, I, O.path_bbox(path, I))
    else:
        mk, W, H = CASES[name]
        path, tr = mk()
    counts, recs, bd, ev, redge = O.bin_replay(path, tr, W, H, mode)
    got = fine_emulation(counts, recs.astype(np.float64), bd, ev, redge, W, H, rule, mode)
    if mode == 0:
        ref = O.sd_mask(path, tr, W, H, rule)
    elif mode == 1:
        ref = O.sd_mask_iter(path, tr, W, H, rule)
        got = np.where(np.abs(got) >= 1e-6, got, 0.0)
    else:
        # ActiveEdgeIter's left-split quirk (a split point at x = -1e-16 moves
        # row pieces one column right) is reproduced by rzg::ae_left_shift
        ref = O.ae_mask(path, tr, W, H, rule)
    d = np.abs(got - ref)
    assert d.max() <= 2e-4, f"max {d.max()} at {np.unravel_index(d.argmax(), d.shape)}"
    assert d.mean() <= 2e-6


@pytest.mark.parametrize("phase", [(5, 0), (0, 11), (15, 1), (7, 9)])
@pytest.mark.parametrize("mode", [0, 1])
def test_decomposition_with_grid_phase(phase, mode):
    """A scene path drawn into a sub-view at (ox, oy) bins on the canvas tile
    grid: tile columns start at 16k - (ox mod 16) in the sub-view's frame. The
    decomposition must still reproduce the reference mask of the sub-view."""
    path = workload_cubics64(5, 120.0).paths[0]
    tr = Transform.new_translate(-20, -20)
    W, H = 81, 70
    for rule in (0, 1):
        counts, recs, bd, ev, redge = O.bin_replay(path, tr, W, H, mode, phase=phase)
        got = fine_emulation(counts, recs.astype(np.float64), bd, ev, redge, W, H, rule, mode, phase)
        ref = O.sd_mask(path, tr, W, H, rule) if mode == 0 else O.sd_mask_iter(path, tr, W, H, rule)
        if mode == 1:
            got = np.where(np.abs(got) >= 1e-6, got, 0.0)
        d = np.abs(got - ref)
        assert d.max() <= 2e-4, f"rule {rule}: max {d.max()} at {np.unravel_index(d.argmax(), d.shape)}"
        assert d.mean() <= 2e-6


def _rect_copies(n, x0, y0, x1, y1):
    pb = PathBuilder()
    for _ in range(n):
        pb.move_to((x0, y0)).line_to((x1, y0)).line_to((x1, y1)).line_to((x0, y1)).close()
    return pb.build()


def test_many_coincident_edges_need_64_bit_tiles():
    """2,100 copies of one rectangle: in the tile holding their right edges the
    backdrop is +2100 and the local sum -2100 windings, -2100 * 2^20 in fixed
    point, outside int32. Kept in int32 it wraps to +1996 and NonZero draws
    coverage 1 right of the rectangle, where the reference has 0; the fine
    stage keeps such a tile in int64 (3 x 2100 records >= 2000)."""
    path = _rect_copies(2100, 10.5, 10.5, 40.5, 40.5)
    W = H = 64
    counts, recs, bd, ev, redge = O.bin_replay(path, I, W, H, 0)
    ref = O.sd_mask(path, I, W, H, 0)
    wrapped = fine_emulation(counts, recs.astype(np.float64), bd, ev, redge, W, H, 0, 0, arith="int32")
    assert np.abs(wrapped - ref).max() > 0.5  # the wrap is real
    exact = fine_emulation(counts, recs.astype(np.float64), bd, ev, redge, W, H, 0, 0)
    assert np.abs(exact - ref).max() <= 2e-4


def _zero_crossing_path(seed, n=400, W=64, H=96):
    """Closed random walks that cross x = 0 and touch it: lines with
    non-dyadic ends left and right of the image's left edge, and vertices
    exactly on x = 0."""
    rng = np.random.default_rng(seed)
    pb = PathBuilder()
    for sub in range(4):
        pts = [(float(rng.uniform(-3.0, 6.0)), float(rng.uniform(0.0, H)))]
        for k in range(n // 4):
            x = 0.0 if rng.uniform() < 0.15 else float(rng.uniform(-3.0, 6.0)) / 3.0
            pts.append((x, float(np.clip(pts[-1][1] + rng.normal(0.0, 7.3), -4.0, H + 4.0))))
        pb.move_to(pts[0])
        for p in pts[1:]:
            pb.line_to(p)
        pb.close()
    return pb.build(), W, H


@pytest.mark.parametrize("seed", range(4))
@pytest.mark.parametrize("rule", [0, 1])
def test_active_edge_left_edge_chains(seed, rule):
    """The replay with ActiveEdgeRasterizer's chains reproduces the oracle's
    ActiveEdgeRasterizer::mask on paths that cross and touch x = 0, where
    the exact-area mask differs from it by whole row pieces."""
    path, W, H = _zero_crossing_path(seed)
    counts, recs, bd, ev, redge = O.bin_replay(path, I, W, H, 2)
    got = fine_emulation(counts, recs.astype(np.float64), bd, ev, redge, W, H, rule, 2)
    ref = O.ae_mask(path, I, W, H, rule)
    d = np.abs(got - ref)
    assert d.max() <= 2e-4, f"max {d.max()} at {np.unravel_index(d.argmax(), d.shape)}"
    # the quirk is there to reproduce: the exact-area mask differs
    exact = O.sd_mask_iter(path, I, W, H, rule)
    assert np.abs(exact - ref).max() > 0.01

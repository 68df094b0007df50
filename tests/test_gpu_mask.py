"""GPU parity of the mask pipeline (flatten -> bin -> fine) against the oracle.

Tolerance for coverage (f32 output vs the reference's f64): max |d| <= 2e-3
and mean |d| <= 1e-5 per pixel. Flattened lines, tile records, backdrops and
events are compared bit for bit.
"""
import ctypes as C
import math
import sys

import numpy as np
import pytest

import _oracle as O
from _paths import OCTAGON_EVENODD, fill_rule_shapes, octagon, path_size, rect, squirrel
import rasterize_amd as rz
from rasterize_amd import (ActiveEdgeRasterizer, FillRule, PathBuilder, SignedDifferenceRasterizer,
                           Transform, workload_cubics64, workload_random_walk)

pytestmark = pytest.mark.gpu

MAX_ABS = 2e-3
MEAN_ABS = 1e-5
I = Transform.identity()


def _say(*a):
    print(*a, flush=True)


def check_close(gpu, ref, what):
    d = np.abs(gpu.astype(np.float64) - ref)
    mx, mean = float(d.max()) if d.size else 0.0, float(d.mean()) if d.size else 0.0
    _say(f"{what}: max|d|={mx:.3e} mean|d|={mean:.3e}")
    assert mx <= MAX_ABS, f"{what}: max |d| {mx}"
    assert mean <= MEAN_ABS, f"{what}: mean |d| {mean}"


@pytest.fixture(scope="module")
def c1():
    return workload_cubics64(1, 512.0).paths[0]


@pytest.fixture(scope="module")
def c2():
    ws = workload_random_walk(2, 200_000, 4096.0)
    return ws.paths[0], ws.transform


# ---- flatten: bit-exact line lists -------------------------------------------
def test_flatten_bit_exact_c1(c1):
    _say("flatten c1")
    g = rz.flatten(c1, I, 0.05, True)
    r = O.flatten(c1, I, 0.05, True)
    assert g.shape == r.shape and np.array_equal(g, r)


def test_flatten_bit_exact_squirrel_rotated():
    p = squirrel()
    tr = Transform.new_rotate(math.pi / 3.0).pre_translate(-10.0, -20.0)
    for close in (True, False):
        g = rz.flatten(p, tr, 0.05, close)
        r = O.flatten(p, tr, 0.05, close)
        assert np.array_equal(g, r)
    # a much finer tolerance deepens every tree
    g = rz.flatten(p, tr.pre_scale(40.0, 40.0), 0.001, True)
    r = O.flatten(p, tr.pre_scale(40.0, 40.0), 0.001, True)
    assert np.array_equal(g, r)


def test_flatten_bit_exact_c2(c2):
    path, tr = c2
    _say("flatten c2", path.segments_count())
    g = rz.flatten(path, tr, 0.05, True)
    r = O.flatten(path, tr, 0.05, True)
    _say("lines", len(g))
    assert g.shape == r.shape and np.array_equal(g, r)


def test_flatten_nan_is_error():
    p = PathBuilder().move_to((0, 0)).cubic_to((1, 1), (float("nan"), 2), (3, 0)).close().build()
    with pytest.raises(rz.RzError) as e:
        rz.flatten(p, I)
    assert e.value.status == rz._abi.RZ_E_NAN


# ---- binning: bit-exact against the sequential replay -------------------------
def _sort_tiles(counts, recs):
    """Sort records within each tile by their bit patterns."""
    bits = recs.view(np.uint32)
    tile = np.repeat(np.arange(counts.size), counts)
    order = np.lexsort((bits[:, 3], bits[:, 2], bits[:, 1], bits[:, 0], tile))
    return bits[order]


def _sort_events(ev):
    b = ev.view(np.uint32)
    return b[np.lexsort((b[:, 3], b[:, 2], b[:, 1], b[:, 0]))]


@pytest.mark.parametrize("case", ["c1", "squirrel", "edges", "c2"])
@pytest.mark.parametrize("mode", [rz._abi.RZ_EDGE_MASK, rz._abi.RZ_EDGE_ITER, rz._abi.RZ_EDGE_NONE])
def test_bin_bit_exact(case, mode, c1, c2):
    if case == "c1":
        path, tr, W, H = c1, I, 512, 512
    elif case == "squirrel":
        path, tr, W, H = squirrel(), Transform.new_scale(20.0, 20.0), 330, 310
    elif case == "edges":
        # crosses all four image edges, W-1 a multiple of 16 (1-column last tile)
        path, tr, W, H = workload_cubics64(5, 120.0).paths[0], Transform.new_translate(-20, -20), 81, 70
    else:
        path, tr = c2
        W = H = 4096
    _say(f"bin {case} mode={mode}")
    gc, gr, gb, ge, gx = rz.bin_stage(path, tr, W, H, mode)
    rc, rr, rb, re, rx = O.bin_replay(path, tr, W, H, mode)
    assert np.array_equal(gc, rc), "per-tile record counts differ"
    assert np.array_equal(gb, rb), "band backdrops differ"
    assert np.array_equal(gx, rx), "right-edge column differs"
    assert np.array_equal(_sort_tiles(gc, gr), _sort_tiles(rc, rr)), "records differ"
    assert np.array_equal(_sort_events(ge), _sort_events(re)), "events differ"


def _spokes(n, size):
    """n thin triangles from the centre almost to the border, the last ones
    crossing the right edge: nearly every line spans many cells."""
    pb = PathBuilder()
    c, r = size / 2.0, size * 0.55
    for k in range(n):
        a0 = 2.0 * math.pi * k / n
        a1 = a0 + math.pi / n
        pb.move_to((c, c))
        pb.line_to((c + r * math.cos(a0), c + r * math.sin(a0)))
        pb.line_to((c + r * math.cos(a1), c + r * math.sin(a1)))
        pb.close()
    return pb.build()


@pytest.mark.parametrize("mode", [rz._abi.RZ_EDGE_MASK, rz._abi.RZ_EDGE_ITER])
def test_bin_long_lines_fresh_context(mode):
    """Long lines only, on a fresh context: the first run outgrows the record
    and long-line buffers and the re-run must still be bit-exact."""
    path = _spokes(3000, 4096)
    ctx = rz.Context(0)
    try:
        _say(f"bin spokes mode={mode}")
        gc, gr, gb, ge, gx = rz.bin_stage(path, I, 4096, 4096, mode, ctx=ctx)
    finally:
        ctx.close()
    rc, rr, rb, re, rx = O.bin_replay(path, I, 4096, 4096, mode)
    _say(f"records {len(rr)} events {len(re)}")
    assert len(rr) > 500_000
    assert np.array_equal(gc, rc), "per-tile record counts differ"
    assert np.array_equal(gb, rb), "band backdrops differ"
    assert np.array_equal(gx, rx), "right-edge column differs"
    assert np.array_equal(_sort_tiles(gc, gr), _sort_tiles(rc, rr)), "records differ"
    assert np.array_equal(_sort_events(ge), _sort_events(re)), "events differ"


@pytest.mark.parametrize("dq,dc", [(0, 0), (1, 2), (3, 5), (6, 6)])
def test_bin_split_depths_agree(dq, dc, c2):
    """The virtual-subtree depths only spread the flatten work over threads:
    the binning must come out bit-identical for any of them (and for the
    depths chosen automatically by the default context)."""
    path, tr = c2
    rc, rr, rb, re, rx = rz.bin_stage(path, tr, 4096, 4096)
    ctx = rz.Context(0)
    try:
        L = rz._abi.lib()
        ctx.check(L.rz_set_param(ctx.h, rz._abi.RZ_PARAM_SPLIT_QUAD, dq))
        ctx.check(L.rz_set_param(ctx.h, rz._abi.RZ_PARAM_SPLIT_CUBIC, dc))
        v = C.c_int()
        ctx.check(L.rz_get_param(ctx.h, rz._abi.RZ_PARAM_SPLIT_AUTO, C.byref(v)))
        assert v.value == 0
        _say(f"bin c2 split dq={dq} dc={dc}")
        gc, gr, gb, ge, gx = rz.bin_stage(path, tr, 4096, 4096, ctx=ctx)
    finally:
        ctx.close()
    assert np.array_equal(gc, rc) and np.array_equal(gb, rb) and np.array_equal(gx, rx)
    assert np.array_equal(_sort_tiles(gc, gr), _sort_tiles(rc, rr))
    assert np.array_equal(_sort_events(ge), _sort_events(re))


@pytest.mark.parametrize("mode", [rz._abi.RZ_EDGE_MASK, rz._abi.RZ_EDGE_NONE])
@pytest.mark.parametrize("dq,dc", [(None, None), (0, 0), (6, 6)])
def test_bin_fused_and_two_kernels_agree(mode, dq, dc, c2):
    """Flatten and count pass as one kernel (lines from LDS, RZ_PARAM_FUSED 1)
    and as two (lines through HBM, the default): bit-identical bins, with
    subtrees of few lines (all kept in LDS) and of many (through the line
    buffer), and with ActiveEdge chains."""
    path, tr = c2
    out = []
    for fused in (0, 1):
        ctx = rz.Context(0)
        try:
            L = rz._abi.lib()
            ctx.check(L.rz_set_param(ctx.h, rz._abi.RZ_PARAM_FUSED, fused))
            if dq is not None:
                ctx.check(L.rz_set_param(ctx.h, rz._abi.RZ_PARAM_SPLIT_QUAD, dq))
                ctx.check(L.rz_set_param(ctx.h, rz._abi.RZ_PARAM_SPLIT_CUBIC, dc))
            _say(f"bin c2 fused={fused} mode={mode} dq={dq} dc={dc}")
            out.append(rz.bin_stage(path, tr, 4096, 4096, edge_mode=mode, ctx=ctx))
        finally:
            ctx.close()
    (rc, rr, rb, re, rx), (gc, gr, gb, ge, gx) = out
    assert np.array_equal(gc, rc) and np.array_equal(gb, rb) and np.array_equal(gx, rx)
    assert np.array_equal(_sort_tiles(gc, gr), _sort_tiles(rc, rr))
    assert np.array_equal(_sort_events(ge), _sort_events(re))


@pytest.mark.parametrize("direct_log2", [0, 3])
def test_bin_scan_forms_agree(direct_log2, c2):
    """The tile scan's second pass sums the block totals itself up to
    2^RZ_PARAM_SCAN_DIRECT blocks and has them scanned by a kernel of their
    own beyond: both forms (C2 has 32 blocks), with the band scan and the
    backdrop / right-edge row prefixes riding along, must bin identically."""
    path, tr = c2
    rc, rr, rb, re, rx = rz.bin_stage(path, tr, 4096, 4096)
    ctx = rz.Context(0)
    try:
        L = rz._abi.lib()
        ctx.check(L.rz_set_param(ctx.h, rz._abi.RZ_PARAM_SCAN_DIRECT, direct_log2))
        v = C.c_int()
        ctx.check(L.rz_get_param(ctx.h, rz._abi.RZ_PARAM_SCAN_DIRECT, C.byref(v)))
        assert v.value == direct_log2
        _say(f"bin c2 scan direct up to 2^{direct_log2} blocks")
        gc, gr, gb, ge, gx = rz.bin_stage(path, tr, 4096, 4096, ctx=ctx)
    finally:
        ctx.close()
    assert np.array_equal(gc, rc) and np.array_equal(gb, rb) and np.array_equal(gx, rx)
    assert np.array_equal(_sort_tiles(gc, gr), _sort_tiles(rc, rr))
    assert np.array_equal(_sort_events(ge), _sort_events(re))


@pytest.mark.parametrize("rule", [FillRule.NonZero, FillRule.EvenOdd])
def test_mask_long_lines(rule):
    path = _spokes(3000, 4096)
    g = SignedDifferenceRasterizer().mask_f32(path, I, 4096, 4096, rule)
    r = O.sd_mask(path, I, 4096, 4096, rule)
    check_close(g, r, f"spokes rule={rule}")


# ---- masks: reference known answers on the GPU --------------------------------
@pytest.mark.parametrize("R", [SignedDifferenceRasterizer, ActiveEdgeRasterizer])
def test_rasterizer_octagon_evenodd(R):
    img = np.zeros((4, 9))
    R().mask(octagon(), I, img, FillRule.EvenOdd)
    assert np.max(np.abs(img - OCTAGON_EVENODD)) < 1e-6


@pytest.mark.parametrize("R", [SignedDifferenceRasterizer, ActiveEdgeRasterizer])
def test_fill_rule_areas(R):
    path = fill_rule_shapes()
    w, h, tr = path_size(path, I, O.path_bbox(path, I))
    img = np.zeros((h, w))
    R().mask(path, tr, img, FillRule.EvenOdd)
    y, x0, x1, x2 = 50, 50, 150, 250
    assert abs(img[y, x0]) < 1e-6 and abs(img[y, x1]) < 1e-6 and abs(img[y, x2]) < 1e-6
    assert abs(img.sum() - 13130.0) < 1.0
    img = np.zeros((h, w))
    R().mask(path, tr, img, FillRule.NonZero)
    assert abs(img[y, x0] - 1.0) < 1e-6 and abs(img[y, x1] - 1.0) < 1e-6 and abs(img[y, x2]) < 1e-6
    assert abs(img.sum() - 16492.5) < 1.0


# ---- masks: parity with the oracle --------------------------------------------
@pytest.mark.parametrize("rule", [FillRule.NonZero, FillRule.EvenOdd])
def test_mask_c1(rule, c1):
    g = SignedDifferenceRasterizer().mask_f32(c1, I, 512, 512, rule)
    r = O.sd_mask(c1, I, 512, 512, rule)
    check_close(g, r, f"c1 rule={rule}")


@pytest.mark.parametrize("rule", [FillRule.NonZero, FillRule.EvenOdd])
def test_mask_c2(rule, c2):
    path, tr = c2
    g = SignedDifferenceRasterizer().mask_f32(path, tr, 4096, 4096, rule)
    r = O.sd_mask(path, tr, 4096, 4096, rule)
    check_close(g, r, f"c2 rule={rule}")


@pytest.mark.parametrize("W,H", [(81, 70), (17, 33), (1, 5), (16, 16), (15, 47), (100, 1)])
@pytest.mark.parametrize("rule", [FillRule.NonZero, FillRule.EvenOdd])
def test_mask_edges(W, H, rule):
    """Paths crossing every image edge; image widths with W-1 = 16k (one-column
    last tile), W < 16, single rows/columns."""
    path = workload_cubics64(7, 1.6 * max(W, H)).paths[0]
    tr = Transform.new_translate(-0.3 * W, -0.3 * H)
    g = SignedDifferenceRasterizer().mask_f32(path, tr, W, H, rule)
    r = O.sd_mask(path, tr, W, H, rule)
    check_close(g, r, f"edges {W}x{H} rule={rule}")


@pytest.mark.parametrize("rule", [FillRule.NonZero, FillRule.EvenOdd])
def test_mask_iter_matches_oracle(rule):
    path = workload_cubics64(8, 200.0).paths[0]
    tr = Transform.new_translate(-20.0, 10.0)
    W, H = 160, 150
    g = SignedDifferenceRasterizer().mask_f32(path, tr, W, H, rule, rz._abi.RZ_EDGE_ITER)
    r = O.sd_mask_iter(path, tr, W, H, rule)
    g = np.where(np.abs(g) >= 1e-6, g, 0.0)
    check_close(g, r, f"mask_iter rule={rule}")


@pytest.mark.parametrize("rule", [FillRule.NonZero, FillRule.EvenOdd])
def test_active_edge_matches_oracle(rule):
    # crosses the right and bottom edges
    path = workload_cubics64(9, 300.0).paths[0]
    tr = Transform.new_translate(2.0, -10.0)
    W, H = 200, 230
    g = ActiveEdgeRasterizer().mask_f32(path, tr, W, H, rule)
    r = O.ae_mask(path, tr, W, H, rule)
    check_close(g, r, f"active-edge rule={rule}")


@pytest.mark.parametrize("seed", range(3))
@pytest.mark.parametrize("rule", [FillRule.NonZero, FillRule.EvenOdd])
def test_active_edge_left_edge(seed, rule):
    """Paths that cross and touch x = 0, where ActiveEdgeIter moves row pieces
    one column right (a split point or a row's walk a rounding below x = 0,
    src/rasterize.rs:733-737, 812-871): the GPU's chains reproduce it."""
    from test_decomposition_cpu import _zero_crossing_path
    path, W, H = _zero_crossing_path(seed)
    _say(f"active-edge left edge seed={seed} rule={rule}")
    g = ActiveEdgeRasterizer().mask_f32(path, I, W, H, rule)
    r = O.ae_mask(path, I, W, H, rule)
    check_close(g, r, f"active-edge left edge seed={seed} rule={rule}")
    # and the binning, bit for bit
    gc, gr, gb, ge, gx = rz.bin_stage(path, I, W, H, rz._abi.RZ_EDGE_NONE)
    rc, rr, rb, re, rx = O.bin_replay(path, I, W, H, rz._abi.RZ_EDGE_NONE)
    assert np.array_equal(gc, rc) and np.array_equal(gb, rb)
    assert np.array_equal(_sort_tiles(gc, gr), _sort_tiles(rc, rr))


def test_strided_view_output():
    """mask into a strided f64 view (ImageMut view_mut, src/image.rs:300-311)."""
    path = workload_cubics64(10, 60.0).paths[0]
    big = np.full((100, 120), -7.0)
    view = big[10:80, 20:90]
    SignedDifferenceRasterizer().mask(path, I, view, FillRule.NonZero)
    r = O.sd_mask(path, I, 70, 70, 0)
    check_close(view, r, "strided view")
    assert np.all(big[:10] == -7.0) and np.all(big[:, :20] == -7.0) and np.all(big[80:] == -7.0)


def test_empty_and_zero_size():
    empty = PathBuilder().build()
    img = np.zeros((8, 8))
    SignedDifferenceRasterizer().mask(empty, I, img, FillRule.NonZero)
    assert np.all(img == 0.0)
    p = rect(1, 1, 5, 5)
    z = np.zeros((0, 0))
    SignedDifferenceRasterizer().mask(p, I, np.zeros((0, 4)), FillRule.NonZero)
    assert SignedDifferenceRasterizer().mask_iter(p, I, (0, 0), FillRule.NonZero) == []
    del z


def test_mask_nan_is_error():
    p = PathBuilder().move_to((0, 0)).line_to((float("nan"), 1)).line_to((2, 2)).close().build()
    with pytest.raises(rz.RzError) as e:
        SignedDifferenceRasterizer().mask(p, I, np.zeros((4, 4)), FillRule.NonZero)
    assert e.value.status == rz._abi.RZ_E_NAN


def test_deterministic(c1):
    a = SignedDifferenceRasterizer().mask_f32(c1, I, 512, 512, FillRule.EvenOdd)
    b = SignedDifferenceRasterizer().mask_f32(c1, I, 512, 512, FillRule.EvenOdd)
    assert np.array_equal(a, b)


def test_error_vs_exact_area(c1):
    """Error of the GPU mask against the FP64 exact coverage of the flattened
    path (reported, and bounded: box-filtering the winding before the fill
    rule is the reference's method, so the reference has the same error)."""
    path = workload_cubics64(12, 100.0).paths[0]
    tr = Transform.new_translate(6.0, 6.0)
    W = H = 112
    for rule in (0, 1):
        g = SignedDifferenceRasterizer().mask_f32(path, tr, W, H, rule).astype(np.float64)
        r = O.sd_mask(path, tr, W, H, rule)
        ex = O.exact_fill_coverage(path, tr, W, H, rule)
        dg, dr = np.abs(g - ex), np.abs(r - ex)
        _say(f"rule={rule}: gpu vs exact max {dg.max():.3e} mean {dg.mean():.3e}; "
             f"reference vs exact max {dr.max():.3e} mean {dr.mean():.3e}")
        assert abs(dg.mean() - dr.mean()) < 1e-5


def test_queued_capacity_miss_is_reported(c1, c2):
    """Two asynchronous masks queued before one rz_sync on a fresh context: the
    first (200k-segment path, 6.6 M lines) outgrows the first-call buffers,
    the second (64 cubics) fits. rz_sync must still report RZ_E_CAPACITY
    (the first call's mask was not written); after a re-run of both, both
    masks are right."""
    import torch
    L = rz._abi.lib()
    big, tr2 = c2
    ctx = rz.Context(0)
    d_big, d_small = C.c_void_p(), C.c_void_p()
    try:
        ctx.check(L.rz_path_upload(ctx.h, C.byref(big.abi()), C.byref(d_big)))
        ctx.check(L.rz_path_upload(ctx.h, C.byref(c1.abi()), C.byref(d_small)))
        ob = torch.zeros((4096, 4096), dtype=torch.float32, device="cuda")
        os_ = torch.zeros((512, 512), dtype=torch.float32, device="cuda")
        torch.cuda.synchronize()

        def run():
            ctx.check(L.rz_mask_device(ctx.h, d_big, tr2.abi(), 0.05, 0, rz._abi.RZ_EDGE_MASK, 4096,
                                       4096, C.c_void_p(ob.data_ptr()), 4096))
            ctx.check(L.rz_mask_device(ctx.h, d_small, I.abi(), 0.05, 0, rz._abi.RZ_EDGE_MASK, 512,
                                       512, C.c_void_p(os_.data_ptr()), 512))
        _say("queued big + small masks on a fresh context")
        run()
        s = L.rz_sync(ctx.h)
        assert s == rz._abi.RZ_E_CAPACITY, f"expected RZ_E_CAPACITY, got {s}"
        for _ in range(3):
            run()
            s = L.rz_sync(ctx.h)
            if s != rz._abi.RZ_E_CAPACITY:
                break
        ctx.check(s)
        check_close(os_.cpu().numpy(), O.sd_mask(c1, I, 512, 512, 0), "small mask after re-run")
        check_close(ob.cpu().numpy(), O.sd_mask(big, tr2, 4096, 4096, 0), "big mask after re-run")
    finally:
        L.rz_path_free(ctx.h, d_big)
        L.rz_path_free(ctx.h, d_small)
        ctx.close()


def _rect_copies(n, x0, y0, x1, y1, pb=None):
    pb = pb or PathBuilder()
    for _ in range(n):
        pb.move_to((x0, y0)).line_to((x1, y0)).line_to((x1, y1)).line_to((x0, y1)).close()
    return pb


@pytest.mark.parametrize("rule", [0, 1])
def test_many_coincident_edges(rule):
    """2,100 copies of a rectangle inside the image and 2,100 of one that
    crosses its right edge: windings of +-2100, beyond int32 in 2^-20 fixed
    point, in the tiles holding their edges and in the right-edge column.
    Mask, batched scene and atlas against the oracle."""
    pb = _rect_copies(2100, 10.5, 10.5, 40.5, 40.5)
    path = _rect_copies(2100, 20.25, 44.5, 100.0, 60.75, pb).build()
    W = H = 64
    for mode in (rz._abi.RZ_EDGE_MASK, rz._abi.RZ_EDGE_ITER):
        got = SignedDifferenceRasterizer().mask_f32(path, I, W, H, rule, mode)
        ref = O.sd_mask(path, I, W, H, rule) if mode == 0 else O.sd_mask_iter(path, I, W, H, rule)
        check_close(got, ref, f"2x2100 rectangles, mode {mode}, rule {rule}")
    paint = rz.LinColor(0.2, 0.3, 0.1, 0.6)
    lay = rz.Scene.fill(path, paint, rule).render(SignedDifferenceRasterizer(), I, (0.0, 0.0, 64.0, 64.0))
    sref, _ = O.scene_render([(path, paint.abi(), rule)], I, (0.0, 0.0, 64.0, 64.0))
    d = np.abs(lay.image.astype(np.float64) - sref)
    _say(f"scene: max|d|={d.max():.3e}")
    assert d.max() <= MAX_ABS and d.mean() <= MEAN_ABS
    atlas = np.full((80, 80), -1.0, np.float32)
    rz.mask_atlas([path], [rz.atlas_slot(7, 9, W, H)], atlas, rule)
    check_close(atlas[9:9 + H, 7:7 + W], O.sd_mask(path, I, W, H, rule), "atlas view")
    assert (atlas[:9] == -1.0).all() and (atlas[:, :7] == -1.0).all()


@pytest.mark.parametrize("name", ["squirrel.path", "tv.path", "rust.path"])
def test_real_world_paths_from_svg(name):
    """Path files the reference ships (data/*.path, copied under
    tests/golden), parsed by rz_svg_path_parse and fitted into 512 px as the
    reference's rasterize example fits them (Transform::fit_size, src/
    geometry.rs:470-516), masked with both rules against the oracle."""
    import os
    text = open(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", name)).read()
    path = rz.Path.from_svg(text)
    bb = O.path_bbox(path, I)
    x0, y0, x1, y1 = np.floor(bb[0]), np.floor(bb[1]), np.ceil(bb[2]), np.ceil(bb[3])
    W = 512
    H = int(np.ceil((y1 - y0) * W / (x1 - x0)))
    s = min(H / (y1 - y0), W / (x1 - x0))
    tr = Transform(s, 0.0, (W - (x1 - x0) * s) / 2.0 - x0 * s, 0.0, s, (H - (y1 - y0) * s) / 2.0 - y0 * s)
    for rule in (0, 1):
        got = SignedDifferenceRasterizer().mask_f32(path, tr, W, H, rule)
        check_close(got, O.sd_mask(path, tr, W, H, rule), f"{name} {W}x{H} rule {rule}")

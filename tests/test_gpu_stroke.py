"""Path::stroke on the GPU (rz_path_stroke: offset curves, joins and caps as
work items, rasterize_amd/csrc/rz_stroke_kernels.hpp) against the oracle's
sequential restatement (oracle/orc_stroke.hpp, pinned in test_stroke_cpu.py).

On well-conditioned paths (the reference's test paths, the squirrel, the tv,
degenerate segments, random curves that keep their control points near the
chord) segment kinds, subpath structure and closed flags must match exactly and
points within 1e-9 of the coordinate scale: the same f64 arithmetic in the same
order, with hypot correctly rounded on the GPU (the host libm's differs from it
in ~0.1% of calls by one ulp) and sincos, tan, acos from the device library.
Random curves with cusps and loops make the offset construction ill-conditioned
(a one-ulp difference in a normal can add or drop a sub-nanometre piece); there
the stroked outlines are compared as masks, against the coverage bar, and the
structure is reported."""
import json
import os

import numpy as np
import pytest

import _oracle as O
import _paths
import rasterize_amd as rz
from rasterize_amd import LineCap, LineJoin, Path, PathBuilder, StrokeStyle

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
GOLDEN = json.load(open(os.path.join(HERE, "golden", "stroke_cases.json")))
JOIN = {"miter": LineJoin.Miter, "bevel": LineJoin.Bevel, "round": LineJoin.Round}
CAP = {"butt": LineCap.Butt, "square": LineCap.Square, "round": LineCap.Round}


def _say(*a):
    print(*a, flush=True)


def _svg(text):
    return Path(*O.svg_parse(text))


def _paths_sample():
    out = [("squirrel", _paths.squirrel())]
    for c in GOLDEN["cases"]:
        out.append((c["path"], _svg(c["path"])))
    with open(os.path.join(HERE, "golden", "tv.path")) as f:
        out.append(("tv", _svg(f.read())))
    out.append(("degenerate", _svg("M0,0 C0,0 0,0 5,5 Q5,5 10,0 C20,10 0,10 10,0 L10,0 Z "
                                   "M30,30 C40,40 20,40 30,30 Z M50,50 L50,50.0000000000001")))
    # random curves whose control points stay near the chord: every join,
    # cap, split and close path, no cusps
    rnd = np.random.default_rng(11)
    b = PathBuilder().move_to((0.0, 0.0))
    pos = np.zeros(2)
    for _ in range(2000):
        k = rnd.integers(0, 3)
        nxt = pos + rnd.uniform(-20, 20, 2)
        d = nxt - pos
        nrm = np.array([-d[1], d[0]])
        c1 = pos + d / 3 + nrm * rnd.uniform(-0.25, 0.25)
        c2 = pos + 2 * d / 3 + nrm * rnd.uniform(-0.25, 0.25)
        if k == 0:
            b = b.line_to(tuple(nxt))
        elif k == 1:
            b = b.quad_to(tuple(c1), tuple(nxt))
        else:
            b = b.cubic_to(tuple(c1), tuple(c2), tuple(nxt))
        pos = nxt
        if rnd.random() < 0.02:
            b = b.close()
            pos = rnd.uniform(-200, 200, 2)
            b = b.move_to(tuple(pos))
    out.append(("gentle random", b.build()))
    return out


def _chaotic(seed, n):
    """Random control points anywhere in a box: cusps, loops, reversals."""
    rnd = np.random.default_rng(seed)
    b = PathBuilder().move_to((0.0, 0.0))
    for _ in range(n):
        k = rnd.integers(0, 3)
        q = [tuple(v) for v in rnd.uniform(-50, 50, (3, 2))]
        b = b.line_to(q[0]) if k == 0 else (b.quad_to(q[0], q[1]) if k == 1 else b.cubic_to(*q))
        if rnd.random() < 0.02:
            b = b.close().move_to(tuple(rnd.uniform(-50, 50, 2)))
    return b.build()


def _same_structure(got, ref):
    """Subpaths (of the stroked outline) whose segment kinds agree."""
    rk, ro, rp, rs, rc = ref
    n = min(len(got), len(rs) - 1 if len(rs) else 0)
    same = 0
    for i in range(n):
        a = got.kinds[got.subpaths[i]:got.subpaths[i + 1]]
        b = rk[rs[i]:rs[i + 1]]
        same += a.size == b.size and np.array_equal(a, b)
    return same, max(len(got), len(rs) - 1 if len(rs) else 0)


def _mask_pair(got, ref, tr, W, H):
    gpu = rz.SignedDifferenceRasterizer().mask_f32(got, tr, W, H, rz.FillRule.NonZero)
    cpu = O.sd_mask(Path(*ref), tr, W, H, rz.FillRule.NonZero)
    return np.abs(np.asarray(gpu, np.float64) - cpu)


def _compare(name, got, ref, scale):
    rk, ro, rp, rs, rc = ref
    assert np.array_equal(got.kinds, rk), (name, "kinds", got.kinds.size, rk.size)
    assert np.array_equal(got.point_offsets, ro), (name, "offsets")
    assert np.array_equal(got.subpaths, rs) and np.array_equal(got.closed, rc), (name, "subpaths")
    d = np.abs(got.points - rp)
    finite = np.isfinite(rp)
    assert np.array_equal(np.isfinite(got.points), finite), (name, "non-finite points differ")
    worst = float(d[finite].max()) if finite.any() else 0.0
    assert worst <= 1e-9 * scale, (name, worst)
    return worst


@pytest.mark.parametrize("join", ["miter", "bevel", "round"])
@pytest.mark.parametrize("cap", ["butt", "square", "round"])
def test_stroke_matches_oracle(join, cap):
    ctx = rz.Context(0)
    for name, p in _paths_sample():
        scale = max(1.0, float(np.abs(p.points).max()) if p.points.size else 1.0)
        for width, limit in ((1.0, 4.0), (3.5, 10.0)):
            _say("stroke", name, join, cap, width)
            got = p.stroke(StrokeStyle(width, JOIN[join], CAP[cap], limit), ctx)
            ref = O.path_stroke(p, width, join, limit, cap)
            _compare(name, got, ref, scale)


@pytest.mark.parametrize("case", range(len(GOLDEN["cases"])))
def test_stroke_golden(case):
    c = GOLDEN["cases"][case]
    got = _svg(c["path"]).stroke(StrokeStyle(c["width"], JOIN[c["join"]], CAP[c["cap"]],
                                             c["miter_limit"]))
    ref = _svg(c["expected"])
    assert np.array_equal(got.kinds, ref.kinds) and np.array_equal(got.subpaths, ref.subpaths)
    np.testing.assert_allclose(got.points, ref.points, rtol=0, atol=6e-6)


def test_stroke_empty_and_bad_style():
    empty = Path(np.zeros(0, np.uint8), np.zeros(1, np.uint32), np.zeros(0), np.zeros(0, np.uint32),
                 np.zeros(0, np.uint8))
    assert empty.stroke(1.0).segments_count() == 0
    p = _paths.squirrel()
    with pytest.raises(rz.RzError):
        p.stroke(StrokeStyle(1.0, 7, LineCap.Butt))


@pytest.mark.parametrize("join", ["miter", "round"])
def test_stroke_chaotic_paths_as_masks(join):
    # cusps and loops: the outlines agree as coverage (the mask bar), and
    # nearly all subpaths piece for piece
    p = _chaotic(11, 2000)
    got = p.stroke(StrokeStyle(1.0, JOIN[join], LineCap.Butt))
    ref = O.path_stroke(p, 1.0, join, 4.0, "butt")
    same, n = _same_structure(got, ref)
    assert len(got) == len(ref[4])
    tr = rz.Transform(4.0, 0.0, 256.0, 0.0, 4.0, 256.0)
    d = _mask_pair(got, ref, tr, 512, 512)
    _say(f"chaotic {join}: {got.segments_count()} vs {ref[0].size} segments, {same}/{n} subpaths "
         f"alike; mask max|d| {d.max():.3e} mean {d.mean():.3e}")
    assert d.max() <= 2e-3 and d.mean() <= 1e-5
    assert abs(got.segments_count() - ref[0].size) <= 0.01 * ref[0].size


def test_stroke_random_walk_full_size_and_mask():
    # the 200k-segment random walk (the mask benchmark's path) stroked, then
    # masked: the GPU stroke against the oracle's, and the GPU mask of the
    # GPU stroke against the oracle's mask of the oracle's stroke
    ws = rz.workload_random_walk(2, 200_000, 4096.0)
    path, tr = ws.paths[0], ws.transform
    style = StrokeStyle(2.0, LineJoin.Round, LineCap.Round)
    got = path.stroke(style)
    ref = O.path_stroke(path, 2.0, "round", 4.0, "round")
    same, n = _same_structure(got, ref)
    _say(f"random walk: {path.segments_count()} segments -> {got.segments_count()} "
         f"(oracle {ref[0].size}), {same}/{n} subpaths alike")
    assert len(got) == len(ref[4]) and abs(got.segments_count() - ref[0].size) <= 0.01 * ref[0].size
    small = rz.Transform(*(v / 4 for v in tr.m))
    d = _mask_pair(got, ref, small, 1024, 1024)
    _say(f"stroked mask: max|d| {d.max():.3e} mean {d.mean():.3e}")
    assert d.max() <= 2e-3 and d.mean() <= 1e-5

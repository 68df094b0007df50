"""ctypes loader for the oracle (oracle/_build/liboracle.so).

The oracle is the CPU restatement of the reference used as the checker; only
tests/, __graft_entry__.smoke() and bench.py's CPU-baseline leg load it.
"""
import ctypes as C
import os
import subprocess

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_SO = os.path.join(ROOT, "oracle", "_build", "liboracle.so")


class orc_path(C.Structure):
    _fields_ = [
        ("n_segments", C.c_uint32), ("n_subpaths", C.c_uint32),
        ("kinds", C.POINTER(C.c_uint8)), ("point_offsets", C.POINTER(C.c_uint32)),
        ("points", C.POINTER(C.c_double)), ("subpaths", C.POINTER(C.c_uint32)),
        ("closed", C.POINTER(C.c_uint8)),
    ]


class orc_paint(C.Structure):
    _fields_ = [
        ("kind", C.c_int32), ("color", C.c_float * 4), ("units", C.c_int32),
        ("linear_colors", C.c_int32), ("spread", C.c_int32), ("tr", C.c_double * 6),
        ("start", C.c_double * 2), ("end", C.c_double * 2), ("center", C.c_double * 2),
        ("radius", C.c_double), ("fcenter", C.c_double * 2), ("fradius", C.c_double),
        ("n_stops", C.c_uint32), ("stop_positions", C.POINTER(C.c_double)),
        ("stop_colors", C.POINTER(C.c_float)),
    ]


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(ORACLE_SO):
            subprocess.run([os.path.join(ROOT, "oracle", "build.sh")], check=True)
        _lib = C.CDLL(ORACLE_SO)
        P = C.POINTER
        d, sz, i = C.c_double, C.c_size_t, C.c_int
        protos = {
            "orc_flatten": (i, [P(orc_path), P(d), d, i, P(d), sz, P(sz)]),
            "orc_signed_difference_line": (i, [P(d), sz, sz, P(d)]),
            "orc_sd_mask": (i, [P(orc_path), P(d), d, i, P(d), sz, sz, sz, sz, sz, sz]),
            "orc_sd_mask_iter": (i, [P(orc_path), P(d), d, i, sz, sz, P(d)]),
            "orc_ae_mask": (i, [P(orc_path), P(d), d, i, sz, sz, P(d)]),
            "orc_ae_lines": (i, [P(d), sz, sz, sz, i, P(d)]),
            "orc_exact_winding_area": (i, [P(orc_path), P(d), d, sz, sz, P(d)]),
            "orc_exact_fill_coverage": (i, [P(orc_path), P(d), d, i, sz, sz, P(d)]),
            "orc_paint_at": (i, [P(orc_paint), d, d, P(C.c_float)]),
            "orc_radial_offset": (i, [P(orc_paint), d, d, P(d)]),
            "orc_l2s": (i, [P(C.c_float), P(C.c_float)]),
            "orc_s2l": (i, [P(C.c_float), P(C.c_float)]),
            "orc_linear_to_srgb": (C.c_float, [C.c_float]),
            "orc_srgb_to_linear": (C.c_float, [C.c_float]),
            "orc_rgba_to_lin": (i, [P(C.c_uint8), P(C.c_float)]),
            "orc_lin_to_rgba": (i, [P(C.c_float), P(C.c_uint8)]),
            "orc_spread_at": (d, [i, d]),
            "orc_quadratic_solve": (i, [d, d, d, P(d)]),
            "orc_alpha_from_winding": (d, [i, d]),
            "orc_split": (i, [i, P(d), i, P(d)]),
            "orc_flatness": (d, [i, P(d)]),
            "orc_arc_to": (i, [P(d), P(d), d, i, i, P(d), P(d)]),
            "orc_path_bbox": (i, [P(orc_path), P(d), P(d)]),
            "orc_time_sd_mask": (d, [P(orc_path), P(d), d, i, sz, sz, i, i]),
            "orc_time_flatten": (d, [P(orc_path), P(d), d, i, P(sz)]),
            "orc_fill": (i, [P(orc_path), P(d), d, i, P(orc_paint), P(C.c_float), sz, sz, sz, sz,
                             sz]),
            "orc_scene_render": (i, [sz, P(orc_path), P(orc_paint), P(C.c_int32), P(d), P(d),
                                     P(C.c_float), d, P(C.c_float), sz, P(C.c_int64)]),
            "orc_scene_render_mt": (i, [sz, P(orc_path), P(orc_paint), P(C.c_int32), P(d), P(d),
                                        P(C.c_float), d, P(C.c_float), sz, P(C.c_int64), i]),
            "orc_scene_tree_render": (i, [P(orc_path), sz, P(orc_paint), sz, C.c_void_p,
                                          P(C.c_uint32), C.c_uint32, P(d), P(d), P(C.c_float), d,
                                          i, P(C.c_float), sz, P(C.c_int64)]),
            "orc_mask_atlas_mt": (i, [sz, P(orc_path), P(d), P(C.c_uint32), d, i, P(C.c_float), sz,
                                      sz, i]),
            "orc_sd_mask_iter_dual": (i, [P(orc_path), P(d), P(d), d, i, sz, sz, P(d), P(d)]),
            "orc_svg_transform": (i, [C.c_char_p, sz, P(d)]),
            "orc_svg_parse": (i, [C.c_char_p, sz, P(C.c_uint8), P(C.c_uint32), P(d), P(C.c_uint32),
                                  P(C.c_uint8), sz, sz, sz, P(sz), P(sz), P(sz)]),
            "orc_path_stroke": (i, [P(orc_path), d, i, d, i, P(C.c_uint8), P(C.c_uint32), P(d),
                                    P(C.c_uint32), P(C.c_uint8), sz, sz, sz, P(sz), P(sz), P(sz)]),
            "orc_time_stroke": (d, [P(orc_path), d, i, d, i, i, P(sz)]),
            "orc_bin_replay": (i, [P(orc_path), P(d), d, i, sz, sz, P(C.c_uint32), P(C.c_float),
                                   sz, P(C.c_int32), P(C.c_float), sz, P(C.c_int32), P(sz),
                                   P(sz), i, i]),
        }
        for name, (res, args) in protos.items():
            f = getattr(_lib, name)
            f.restype = res
            f.argtypes = args
    return _lib


def _dp(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def opath(path):
    """orc_path view of a rasterize_amd.Path (same layout as rz_path)."""
    p = orc_path()
    p.n_segments = path.kinds.size
    p.n_subpaths = max(0, path.subpaths.size - 1)
    p.kinds = path.kinds.ctypes.data_as(C.POINTER(C.c_uint8))
    p.point_offsets = path.point_offsets.ctypes.data_as(C.POINTER(C.c_uint32))
    p.points = path.points.ctypes.data_as(C.POINTER(C.c_double))
    p.subpaths = path.subpaths.ctypes.data_as(C.POINTER(C.c_uint32))
    p.closed = path.closed.ctypes.data_as(C.POINTER(C.c_uint8))
    return p


def _tr(tr):
    return np.array(tr.m if hasattr(tr, "m") else tr, dtype=np.float64)


class OracleError(RuntimeError):
    pass


def _check(rc):
    if rc == 1:
        raise OracleError("cannot flatten segment with NaN")
    if rc != 0:
        raise OracleError(f"oracle error {rc}")


def flatten(path, tr, flatness=0.05, close=True):
    p, t = opath(path), _tr(tr)
    n = C.c_size_t()
    rc = lib().orc_flatten(C.byref(p), _dp(t), flatness, int(close), None, 0, C.byref(n))
    if rc == 1:
        raise OracleError("cannot flatten segment with NaN")
    out = np.zeros((max(n.value, 1), 4))
    _check(lib().orc_flatten(C.byref(p), _dp(t), flatness, int(close), _dp(out), n.value, C.byref(n)))
    return out[: n.value]


def sd_mask(path, tr, W, H, rule, flatness=0.05):
    """SignedDifferenceRasterizer::mask into a zeroed W x H f64 image. rule=-1: raw accumulation."""
    img = np.zeros((H, W))
    p, t = opath(path), _tr(tr)
    _check(lib().orc_sd_mask(C.byref(p), _dp(t), flatness, rule, _dp(img), img.size, W, H, 0, W, 1))
    return img


def sd_mask_iter(path, tr, W, H, rule, flatness=0.05):
    out = np.zeros((H, W))
    p, t = opath(path), _tr(tr)
    _check(lib().orc_sd_mask_iter(C.byref(p), _dp(t), flatness, rule, W, H, _dp(out)))
    return out


def ae_mask(path, tr, W, H, rule, flatness=0.05):
    out = np.zeros((H, W))
    p, t = opath(path), _tr(tr)
    _check(lib().orc_ae_mask(C.byref(p), _dp(t), flatness, rule, W, H, _dp(out)))
    return out


def ae_lines(lines, W, H, rule):
    lines = np.ascontiguousarray(lines, dtype=np.float64)
    out = np.zeros((H, W))
    _check(lib().orc_ae_lines(_dp(lines), len(lines), W, H, rule, _dp(out)))
    return out


def signed_difference_line(W, H, line):
    img = np.zeros((H, W))
    l = np.array(line, dtype=np.float64)
    lib().orc_signed_difference_line(_dp(img), W, H, _dp(l))
    return img


def exact_winding_area(path, tr, W, H, flatness=0.05):
    out = np.zeros((H, W))
    p, t = opath(path), _tr(tr)
    _check(lib().orc_exact_winding_area(C.byref(p), _dp(t), flatness, W, H, _dp(out)))
    return out


def exact_fill_coverage(path, tr, W, H, rule, flatness=0.05):
    out = np.zeros((H, W))
    p, t = opath(path), _tr(tr)
    _check(lib().orc_exact_fill_coverage(C.byref(p), _dp(t), flatness, rule, W, H, _dp(out)))
    return out


def path_bbox(path, tr):
    out = np.zeros(4)
    p, t = opath(path), _tr(tr)
    rc = lib().orc_path_bbox(C.byref(p), _dp(t), _dp(out))
    return None if rc else out


def arc_to(src, radii, rot, large, sweep, dst):
    out = np.zeros(32)
    n = lib().orc_arc_to(_dp(np.array(src, float)), _dp(np.array(radii, float)), rot, int(large),
                         int(sweep), _dp(np.array(dst, float)), _dp(out))
    return n, out


def bin_replay(path, tr, W, H, mode, flatness=0.05, phase=(0, 0)):
    """Sequential replay of the binning stage: (tile_counts, records[n,4] f32,
    backdrop[tiles], events[n,4], right_edge[H] int32). `phase` = (phx, phy)
    shifts the tile grid (tile column k spans x in [16k - phx, 16k + 16 - phx))."""
    p, t = opath(path), _tr(tr)
    nr, ne = C.c_size_t(), C.c_size_t()
    L = lib()
    phx, phy = phase
    _check(L.orc_bin_replay(C.byref(p), _dp(t), flatness, mode, W, H, None, None, 0, None, None,
                            0, None, C.byref(nr), C.byref(ne), phx, phy))
    tiles = ((W + phx + 15) // 16) * ((H + phy + 15) // 16)
    counts = np.zeros(tiles, np.uint32)
    recs = np.zeros((max(nr.value, 1), 4), np.float32)
    bd = np.zeros(tiles, np.int32)
    ev = np.zeros((max(ne.value, 1), 4), np.float32)
    redge = np.zeros(max(H, 1), np.int32)
    _check(L.orc_bin_replay(C.byref(p), _dp(t), flatness, mode, W, H,
                            counts.ctypes.data_as(C.POINTER(C.c_uint32)),
                            recs.ctypes.data_as(C.POINTER(C.c_float)), recs.shape[0],
                            bd.ctypes.data_as(C.POINTER(C.c_int32)),
                            ev.ctypes.data_as(C.POINTER(C.c_float)), ev.shape[0],
                            redge.ctypes.data_as(C.POINTER(C.c_int32)), C.byref(nr), C.byref(ne),
                            phx, phy))
    return counts, recs[: nr.value], bd, ev[: ne.value], redge[:H]


def fill(path, tr, paint_abi, rule, img, flatness=0.05):
    """SignedDifferenceRasterizer::fill into a copy of `img` ((H, W, 4) float32,
    premultiplied linear RGBA). `paint_abi` is an rz_paint (same layout as orc_paint)."""
    out = np.ascontiguousarray(img, dtype=np.float32).copy()
    H, W = out.shape[:2]
    p, t = opath(path), _tr(tr)
    op = orc_paint.from_buffer_copy(paint_abi)
    _check(lib().orc_fill(C.byref(p), _dp(t), flatness, rule, C.byref(op),
                          out.ctypes.data_as(C.POINTER(C.c_float)), W * H, W, H, 0, W))
    return out


def paint_at(paint_abi, x, y):
    op = orc_paint.from_buffer_copy(paint_abi)
    out = (C.c_float * 4)()
    lib().orc_paint_at(C.byref(op), x, y, out)
    return np.array(out, dtype=np.float32)


def scene_render(fills, tr, view=None, bg=None, flatness=0.05):
    """Scene::group(fills).render(SignedDifferenceRasterizer, tr, view, bg):
    (image (H, W, 4) float32, (x, y)). `fills` = [(path, paint_abi, rule)]."""
    n = len(fills)
    paths = (orc_path * max(n, 1))(*[opath(f[0]) for f in fills])
    paints = (orc_paint * max(n, 1))(*[orc_paint.from_buffer_copy(f[1]) for f in fills])
    rules = (C.c_int32 * max(n, 1))(*[int(f[2]) for f in fills])
    t = _tr(tr)
    v = (C.c_double * 4)(*view) if view is not None else None
    b = (C.c_float * 4)(*bg) if bg is not None else None
    xywh = (C.c_int64 * 4)()
    rc = lib().orc_scene_render(n, paths, paints, rules, _dp(t), v, b, flatness, None, 0, xywh)
    x, y, w, h = list(xywh)
    img = np.zeros((h, w, 4), np.float32)
    if w and h:
        _check(lib().orc_scene_render(n, paths, paints, rules, _dp(t), v, b, flatness,
                                      img.ctypes.data_as(C.POINTER(C.c_float)), w * h, xywh))
    elif rc not in (0, 2):
        _check(rc)
    return img, (x, y)


def scene_tree_render(scene, tr, view=None, bg=None, flatness=0.05, active_edge=False):
    """Scene::render(SignedDifferenceRasterizer, tr, view, bg) of a
    rasterize_amd.Scene tree (its node arrays, the rz_scene_node layout), or
    with ActiveEdgeRasterizer: (image (H, W, 4) float32, (x, y))."""
    ps, qs, nodes, children, root = scene.node_arrays()
    paths = (orc_path * max(len(ps), 1))(*[opath(p) for p in ps])
    keep = [q.abi() for q in qs]  # owns the stop arrays the copies point at
    paints = (orc_paint * max(len(qs), 1))(*[orc_paint.from_buffer_copy(q) for q in keep])
    na = (type(nodes[0]) * len(nodes))(*nodes)
    ca = (C.c_uint32 * max(len(children), 1))(*children)
    t = _tr(tr)
    v = (C.c_double * 4)(*view) if view is not None else None
    b = (C.c_float * 4)(*bg) if bg is not None else None
    xywh = (C.c_int64 * 4)()
    L = lib()
    args = (paths, len(ps), paints, len(qs), C.cast(na, C.c_void_p), ca, root, _dp(t), v, b, flatness,
            int(active_edge))
    rc = L.orc_scene_tree_render(*args, None, 0, xywh)
    x, y, w, h = list(xywh)
    img = np.zeros((h, w, 4), np.float32)
    if w and h:
        _check(L.orc_scene_tree_render(*args, img.ctypes.data_as(C.POINTER(C.c_float)), w * h, xywh))
    elif rc not in (0, 2):
        _check(rc)
    del keep
    return img, (x, y)


def host_threads():
    """Host threads for the oracle's parallel checks: this process's CPU share,
    at most 16."""
    try:
        n = len(os.sched_getaffinity(0))
    except AttributeError:
        n = os.cpu_count() or 1
    return max(1, min(16, n))


def scene_render_mt(fills, tr, view=None, bg=None, flatness=0.05, threads=None):
    """scene_render on several host threads (bands of layer rows): the same
    image, bit for bit."""
    n = len(fills)
    paths = (orc_path * max(n, 1))(*[opath(f[0]) for f in fills])
    paints = (orc_paint * max(n, 1))(*[orc_paint.from_buffer_copy(f[1]) for f in fills])
    rules = (C.c_int32 * max(n, 1))(*[int(f[2]) for f in fills])
    t = _tr(tr)
    v = (C.c_double * 4)(*view) if view is not None else None
    b = (C.c_float * 4)(*bg) if bg is not None else None
    xywh = (C.c_int64 * 4)()
    th = threads or host_threads()
    L = lib()
    rc = L.orc_scene_render_mt(n, paths, paints, rules, _dp(t), v, b, flatness, None, 0, xywh, th)
    x, y, w, h = list(xywh)
    img = np.zeros((h, w, 4), np.float32)
    if w and h:
        _check(L.orc_scene_render_mt(n, paths, paints, rules, _dp(t), v, b, flatness,
                                     img.ctypes.data_as(C.POINTER(C.c_float)), w * h, xywh, th))
    elif rc not in (0, 2):
        _check(rc)
    return img, (x, y)


def mask_atlas(paths, trs, xywh, atlas_w, atlas_h, rule=0, flatness=0.05, threads=None, out=None):
    """Per glyph SignedDifferenceRasterizer::mask into a zeroed w x h image,
    copied into a float32 atlas (atlas_h, atlas_w) at (x, y)."""
    n = len(paths)
    keep = [opath(p) for p in paths]
    arr = (orc_path * max(n, 1))(*keep)
    t = np.ascontiguousarray(np.asarray(trs, np.float64).reshape(-1))
    q = np.ascontiguousarray(np.asarray(xywh, np.uint32).reshape(-1))
    atlas = out if out is not None else np.zeros((atlas_h, atlas_w), np.float32)
    _check(lib().orc_mask_atlas_mt(n, arr, _dp(t), q.ctypes.data_as(C.POINTER(C.c_uint32)),
                                   flatness, rule, atlas.ctypes.data_as(C.POINTER(C.c_float)),
                                   atlas_w, atlas_h, threads or host_threads()))
    return atlas


def sd_mask_iter_dual(path, tan, tr, W, H, rule=0, flatness=0.05):
    """(alpha, dalpha): mask_iter coverage (H, W) and its derivative along
    `tan` ((n_points, 2): d(point)/d(parameter) per stored point)."""
    p, t = opath(path), _tr(tr)
    tn = np.ascontiguousarray(np.asarray(tan, np.float64).reshape(-1))
    a = np.zeros((H, W))
    da = np.zeros((H, W))
    _check(lib().orc_sd_mask_iter_dual(C.byref(p), _dp(tn), _dp(t), flatness, rule, W, H, _dp(a),
                                       _dp(da)))
    return a, da


def svg_parse(text):
    """Path::read_svg_path (src/path.rs:528-534) as the oracle restates it:
    (kinds, point_offsets, points, subpaths, closed) arrays, or raises
    OracleError with the byte offset on a syntax error."""
    data = text.encode() if isinstance(text, str) else bytes(text)
    L = lib()
    ns, npt, nsub = C.c_size_t(), C.c_size_t(), C.c_size_t()
    rc = L.orc_svg_parse(data, len(data), None, None, None, None, None, 0, 0, 0, C.byref(ns),
                         C.byref(npt), C.byref(nsub))
    if rc == 3:
        raise OracleError(f"invalid SVG path data at byte {ns.value}")
    kinds = np.zeros(ns.value, np.uint8)
    offs = np.zeros(ns.value + 1, np.uint32)
    pts = np.zeros(2 * npt.value)
    subs = np.zeros(nsub.value + 1 if nsub.value else 0, np.uint32)
    closed = np.zeros(nsub.value, np.uint8)
    u8, u32 = C.POINTER(C.c_uint8), C.POINTER(C.c_uint32)
    _check(L.orc_svg_parse(data, len(data), kinds.ctypes.data_as(u8), offs.ctypes.data_as(u32),
                           _dp(pts), subs.ctypes.data_as(u32), closed.ctypes.data_as(u8),
                           ns.value, npt.value, len(subs), C.byref(ns), C.byref(npt),
                           C.byref(nsub)))
    return kinds, offs, pts, subs, closed


JOINS = {"miter": 0, "bevel": 1, "round": 2}
CAPS = {"butt": 0, "square": 1, "round": 2}


def path_stroke(path, width, join="miter", miter_limit=4.0, cap="butt"):
    """Path::stroke (src/path.rs:370-415) as the oracle restates it:
    (kinds, point_offsets, points, subpaths, closed) arrays."""
    L = lib()
    op = opath(path)
    j, c = JOINS[join], CAPS[cap]
    ns, npt, nsub = C.c_size_t(), C.c_size_t(), C.c_size_t()
    _check(L.orc_path_stroke(C.byref(op), width, j, miter_limit, c, None, None, None, None, None,
                             0, 0, 0, C.byref(ns), C.byref(npt), C.byref(nsub)))
    kinds = np.zeros(ns.value, np.uint8)
    offs = np.zeros(ns.value + 1, np.uint32)
    pts = np.zeros(2 * npt.value)
    subs = np.zeros(nsub.value + 1 if nsub.value else 0, np.uint32)
    closed = np.zeros(nsub.value, np.uint8)
    u8, u32 = C.POINTER(C.c_uint8), C.POINTER(C.c_uint32)
    _check(L.orc_path_stroke(C.byref(op), width, j, miter_limit, c, kinds.ctypes.data_as(u8),
                             offs.ctypes.data_as(u32), _dp(pts), subs.ctypes.data_as(u32),
                             closed.ctypes.data_as(u8), ns.value, npt.value, len(subs),
                             C.byref(ns), C.byref(npt), C.byref(nsub)))
    return kinds, offs, pts, subs, closed


def time_stroke(path, width, join="miter", miter_limit=4.0, cap="butt", reps=1):
    """Seconds per Path::stroke call of the oracle, and the output's segment count."""
    n = C.c_size_t()
    t = lib().orc_time_stroke(C.byref(opath(path)), width, JOINS[join], miter_limit, CAPS[cap],
                              reps, C.byref(n))
    return t, n.value


def svg_transform(text):
    """FromStr for Transform as the oracle restates it: the 6 entries."""
    data = text.encode() if isinstance(text, str) else bytes(text)
    out = np.zeros(6)
    if lib().orc_svg_transform(data, len(data), _dp(out)) != 0:
        raise OracleError("invalid SVG transform")
    return out

"""rasterize_amd -- MI355X coverage rasterizer with the API of aslpavel/rasterize.

Python mirror of the crate's host API over the C ABI (include/rz.h):
Transform, Path / PathBuilder, FillRule, SignedDifferenceRasterizer and
ActiveEdgeRasterizer (`mask`, `mask_iter`), plus the flatten/bin stage hooks.
All rasterization runs in the HIP kernels of librz.so; nothing here computes
coverage on the CPU.
"""
import ctypes as C
import math
import os

import numpy as np

from . import _abi
from ._abi import lib

DEFAULT_FLATNESS = 0.05  # src/path.rs:16
EPSILON = 2.220446049250313e-16


_libm = C.CDLL("libm.so.6")
_libm.sincos.argtypes = [C.c_double, C.POINTER(C.c_double), C.POINTER(C.c_double)]
_libm.sincos.restype = None


def sin_cos(a):
    """f64::sin_cos: one glibc sincos() call, as LLVM lowers it on x86_64-linux-gnu."""
    s, c = C.c_double(), C.c_double()
    _libm.sincos(float(a), C.byref(s), C.byref(c))
    return s.value, c.value


class RzError(RuntimeError):
    def __init__(self, status, msg):
        super().__init__(f"{_abi.STATUS_NAMES.get(status, status)}: {msg}")
        self.status = status


class FillRule:
    NonZero = _abi.RZ_NONZERO
    EvenOdd = _abi.RZ_EVENODD


class LineJoin:
    """LineJoin (src/path.rs:70-93): Miter(limit) is LineJoin.Miter with
    StrokeStyle.miter_limit."""
    Miter = _abi.RZ_JOIN_MITER
    Bevel = _abi.RZ_JOIN_BEVEL
    Round = _abi.RZ_JOIN_ROUND


class LineCap:
    """LineCap (src/path.rs:95-116)."""
    Butt = _abi.RZ_CAP_BUTT
    Square = _abi.RZ_CAP_SQUARE
    Round = _abi.RZ_CAP_ROUND


class StrokeStyle:
    """StrokeStyle (src/path.rs:118-136); the default join is Miter(4.0), the
    default cap Butt."""

    __slots__ = ("width", "line_join", "miter_limit", "line_cap")

    def __init__(self, width, line_join=LineJoin.Miter, line_cap=LineCap.Butt, miter_limit=4.0):
        self.width = float(width)
        self.line_join = int(line_join)
        self.line_cap = int(line_cap)
        self.miter_limit = float(miter_limit)

    def abi(self):
        return _abi.rz_stroke_style(self.width, self.miter_limit, self.line_join, self.line_cap)

    def __repr__(self):
        return (f"StrokeStyle(width={self.width}, line_join={self.line_join}, "
                f"line_cap={self.line_cap}, miter_limit={self.miter_limit})")


class Transform:
    """2D affine [m00, m01, m02, m10, m11, m12] (src/geometry.rs:316-539)."""

    __slots__ = ("m",)

    def __init__(self, m00=1.0, m01=0.0, m02=0.0, m10=0.0, m11=1.0, m12=0.0):
        self.m = (float(m00), float(m01), float(m02), float(m10), float(m11), float(m12))

    @staticmethod
    def identity():
        return Transform()

    @staticmethod
    def new_translate(tx, ty):
        return Transform(1.0, 0.0, tx, 0.0, 1.0, ty)

    @staticmethod
    def new_scale(sx, sy):
        return Transform(sx, 0.0, 0.0, 0.0, sy, 0.0)

    @staticmethod
    def new_rotate(a):
        s, c = sin_cos(a)
        return Transform(c, -s, 0.0, s, c, 0.0)

    def __mul__(self, o):
        s, q = self.m, o.m
        return Transform(s[0] * q[0] + s[1] * q[3], s[0] * q[1] + s[1] * q[4],
                         s[0] * q[2] + s[1] * q[5] + s[2], s[3] * q[0] + s[4] * q[3],
                         s[3] * q[1] + s[4] * q[4], s[3] * q[2] + s[4] * q[5] + s[5])

    def pre_translate(self, tx, ty):
        return self * Transform.new_translate(tx, ty)

    def pre_scale(self, sx, sy):
        return self * Transform.new_scale(sx, sy)

    def pre_rotate(self, a):
        return self * Transform.new_rotate(a)

    @staticmethod
    def from_svg(text):
        """FromStr for Transform (src/svg.rs:591-601): an SVG transform list
        (rz_svg_transform_parse). Raises ValueError with the byte offset on a
        syntax error."""
        data = text.encode() if isinstance(text, str) else bytes(text)
        t = _abi.rz_transform()
        err = C.c_size_t()
        if lib().rz_svg_transform_parse(data, len(data), C.byref(t), C.byref(err)) != _abi.RZ_OK:
            raise ValueError(f"invalid SVG transform at byte {err.value}")
        return Transform(*t.m)

    @staticmethod
    def fit_bbox(src, dst, align=1):
        """Transform::fit_bbox (src/geometry.rs:470-487); boxes are (x0, y0,
        x1, y1), align 0 Min, 1 Mid, 2 Max."""
        sw, sh = src[2] - src[0], src[3] - src[1]
        dw, dh = dst[2] - dst[0], dst[3] - dst[1]
        scale = min(dh / sh, dw / sw)
        base = (Transform.new_translate(dst[0], dst[1]).pre_scale(scale, scale)
                .pre_translate(-src[0], -src[1]))
        if align == 0:
            a = Transform.identity()
        elif align == 1:
            a = Transform.new_translate((dw - sw * scale) / 2.0, (dh - sh * scale) / 2.0)
        else:
            a = Transform.new_translate(dw - sw * scale, dh - sh * scale)
        return a * base

    @staticmethod
    def fit_size(src, width, height, align=1):
        """Transform::fit_size (src/geometry.rs:490-516): ((width, height),
        transform) fitting the whole-pixel box of src into an image of that
        size; a 0 dimension follows the other (both 0: src's size)."""
        x0, y0 = math.floor(src[0]), math.floor(src[1])
        x1, y1 = math.ceil(src[2]), math.ceil(src[3])
        box = (float(x0), float(y0), float(x1), float(y1))
        sw, sh = box[2] - box[0], box[3] - box[1]
        if height == 0 and width == 0:
            h, w = sh, sw
        elif width == 0:
            h = float(height)
            w = float(math.ceil(sw * h / sh))
        elif height == 0:
            w = float(width)
            h = float(math.ceil(sh * w / sw))
        else:
            h, w = float(height), float(width)
        return (int(w), int(h)), Transform.fit_bbox(box, (0.0, 0.0, w, h), align)

    def apply(self, p):
        m = self.m
        return (p[0] * m[0] + p[1] * m[1] + m[2], p[0] * m[3] + p[1] * m[4] + m[5])

    def abi(self):
        t = _abi.rz_transform()
        for i, v in enumerate(self.m):
            t.m[i] = v
        return t


def _is_close(a, b):
    return abs(a[0] - b[0]) < EPSILON and abs(a[1] - b[1]) < EPSILON


class Path:
    """Path {segments, subpaths, closed} (src/path.rs:228-233) as SoA arrays."""

    def __init__(self, kinds, point_offsets, points, subpaths, closed):
        self.kinds = np.ascontiguousarray(kinds, dtype=np.uint8)
        self.point_offsets = np.ascontiguousarray(point_offsets, dtype=np.uint32)
        self.points = np.ascontiguousarray(points, dtype=np.float64).reshape(-1)
        self.subpaths = np.ascontiguousarray(subpaths, dtype=np.uint32)
        self.closed = np.ascontiguousarray(closed, dtype=np.uint8)
        self._abi = None

    @staticmethod
    def builder():
        return PathBuilder()

    @staticmethod
    def from_abi(p):
        """Copy an rz_path view into a Path."""
        ns, nsub = p.n_segments, p.n_subpaths
        kinds = np.ctypeslib.as_array(p.kinds, (ns,)).copy() if ns else np.zeros(0, np.uint8)
        offs = np.ctypeslib.as_array(p.point_offsets, (ns + 1,)).copy()
        npts = int(offs[-1])
        pts = np.ctypeslib.as_array(p.points, (2 * npts,)).copy() if npts else np.zeros(0)
        subs = np.ctypeslib.as_array(p.subpaths, (nsub + 1,)).copy() if nsub else np.zeros(0, np.uint32)
        closed = np.ctypeslib.as_array(p.closed, (nsub,)).copy() if nsub else np.zeros(0, np.uint8)
        return Path(kinds, offs, pts, subs, closed)

    @staticmethod
    def from_svg(text):
        """FromStr for Path (src/path.rs:1058-1064): SVG path data to a Path
        (rz_svg_path_parse). Raises ValueError with the byte offset on a
        syntax error."""
        data = text.encode() if isinstance(text, str) else bytes(text)
        h = C.c_void_p()
        err = C.c_size_t()
        L = lib()
        if L.rz_svg_path_parse(data, len(data), C.byref(h), C.byref(err)) != _abi.RZ_OK:
            raise ValueError(f"invalid SVG path data at byte {err.value}")
        try:
            v = _abi.rz_path()
            L.rz_path_buf_view(h, C.byref(v))
            return Path.from_abi(v) if v.n_segments else Path(
                np.zeros(0, np.uint8), np.zeros(1, np.uint32), np.zeros(0), np.zeros(0, np.uint32),
                np.zeros(0, np.uint8))
        finally:
            L.rz_path_buf_free(h)

    def segments_count(self):
        return int(self.kinds.size)

    def extend(self, other):
        """Extend<SubPath> for Path (src/path.rs:683-689): append the
        subpaths of `other` (in place; returns self)."""
        if other.kinds.size == 0:
            return self
        base_seg, base_pt = self.kinds.size, int(self.point_offsets[-1]) if self.kinds.size else 0
        subs = self.subpaths if self.subpaths.size else np.zeros(1, np.uint32)
        self.kinds = np.concatenate([self.kinds, other.kinds]).astype(np.uint8)
        self.point_offsets = np.concatenate(
            [self.point_offsets[:base_seg], other.point_offsets + base_pt]).astype(np.uint32)
        self.points = np.concatenate([self.points[:2 * base_pt], other.points])
        self.subpaths = np.concatenate([subs, other.subpaths[1:] + base_seg]).astype(np.uint32)
        self.closed = np.concatenate([self.closed, other.closed]).astype(np.uint8)
        self._abi = None
        return self

    def stroke(self, style, ctx=None):
        """Path::stroke (src/path.rs:370-415) on the GPU (rz_path_stroke): a
        new path, the outline of this one offset by width / 2 on both sides,
        joined and capped as `style` (a StrokeStyle, or a width) says."""
        if not isinstance(style, StrokeStyle):
            style = StrokeStyle(style)
        ctx = ctx or default_context()
        h = C.c_void_p()
        L = lib()
        ctx.check(L.rz_path_stroke(ctx.h, C.byref(self.abi()), style.abi(), C.byref(h)))
        try:
            v = _abi.rz_path()
            L.rz_path_buf_view(h, C.byref(v))
            return Path.from_abi(v) if v.n_segments else Path(
                np.zeros(0, np.uint8), np.zeros(1, np.uint32), np.zeros(0), np.zeros(0, np.uint32),
                np.zeros(0, np.uint8))
        finally:
            L.rz_path_buf_free(h)

    def __len__(self):
        return max(0, int(self.subpaths.size) - 1)

    def abi(self):
        if self._abi is None:
            p = _abi.rz_path()
            p.n_segments = self.kinds.size
            p.n_subpaths = max(0, self.subpaths.size - 1)
            p.kinds = self.kinds.ctypes.data_as(C.POINTER(C.c_uint8))
            p.point_offsets = self.point_offsets.ctypes.data_as(C.POINTER(C.c_uint32))
            p.points = self.points.ctypes.data_as(C.POINTER(C.c_double))
            p.subpaths = self.subpaths.ctypes.data_as(C.POINTER(C.c_uint32))
            p.closed = self.closed.ctypes.data_as(C.POINTER(C.c_uint8))
            self._abi = p
        return self._abi

    def bbox(self, tr=None):
        """Path::bbox (src/path.rs:428-431): (x0, y0, x1, y1) of the segments
        under tr (curve extrema), or None for an empty path."""
        has = C.c_int()
        out = (C.c_double * 4)()
        s = lib().rz_path_bbox(C.byref(self.abi()), (tr or Transform.identity()).abi(), C.byref(has), out)
        if s != _abi.RZ_OK:
            raise RzError(s, "rz_path_bbox failed")
        return tuple(out) if has.value else None

    def transform_points(self, tr):
        """Path::transform (src/path.rs:359-363) on a copy."""
        m = tr.m
        xs, ys = self.points[0::2], self.points[1::2]
        out = np.empty_like(self.points)
        out[0::2] = xs * m[0] + ys * m[1] + m[2]
        out[1::2] = xs * m[3] + ys * m[4] + m[5]
        return Path(self.kinds, self.point_offsets, out, self.subpaths, self.closed)

    def mask(self, rasterizer, tr, fill_rule, img):
        """Path::mask (src/path.rs:513-520)."""
        rasterizer.mask(self, tr, img, fill_rule)
        return img


class PathBuilder:
    """PathBuilder (src/path.rs:799-1056)."""

    def __init__(self):
        self.position = (0.0, 0.0)
        self._kinds, self._pts, self._subpaths, self._closed = [], [], [], []

    def _push(self, kind, pts):
        self._kinds.append(kind)
        self._pts.append(tuple(float(v) for p in pts for v in p))

    def _finish(self, close):
        n = len(self._kinds)
        if n == 0 or (self._subpaths and self._subpaths[-1] == n):
            return
        if not self._subpaths:
            self._subpaths.append(0)
        if close:
            first = self._pts[self._subpaths[-1]]
            self.position = (first[0], first[1])
        self._subpaths.append(n)
        self._closed.append(1 if close else 0)

    def move_to(self, p):
        self._finish(False)
        self.position = (float(p[0]), float(p[1]))
        return self

    def close(self):
        self._finish(True)
        return self

    def line_to(self, p):
        p = (float(p[0]), float(p[1]))
        if not _is_close(self.position, p):
            self._push(_abi.RZ_LINE, [self.position, p])
            self.position = p
        return self

    def quad_to(self, p1, p2):
        self._push(_abi.RZ_QUAD, [self.position, p1, p2])
        self.position = (float(p2[0]), float(p2[1]))
        return self

    def cubic_to(self, p1, p2, p3):
        self._push(_abi.RZ_CUBIC, [self.position, p1, p2, p3])
        self.position = (float(p3[0]), float(p3[1]))
        return self

    def arc_to(self, radii, x_axis_rot, large, sweep, p):
        out = (C.c_double * 32)()
        src = (C.c_double * 2)(*self.position)
        rad = (C.c_double * 2)(float(radii[0]), float(radii[1]))
        dst = (C.c_double * 2)(float(p[0]), float(p[1]))
        n = lib().rzw_arc_to(src, rad, float(x_axis_rot), int(bool(large)), int(bool(sweep)), dst, out)
        if n < 0:
            return self.line_to(p)
        for i in range(n):
            o = out[8 * i: 8 * i + 8]
            self._push(_abi.RZ_CUBIC, [(o[0], o[1]), (o[2], o[3]), (o[4], o[5]), (o[6], o[7])])
        self.position = (float(p[0]), float(p[1]))
        return self

    def circle(self, radius):
        off = 0.5522847498307935 * radius
        cx, cy = self.position
        p0, p1 = (cx - radius, cy), (cx, cy - radius)
        p2, p3 = (cx + radius, cy), (cx, cy + radius)
        return (self.move_to(p0)
                .cubic_to((p0[0], p0[1] - off), (p1[0] - off, p1[1]), p1)
                .cubic_to((p1[0] + off, p1[1]), (p2[0], p2[1] - off), p2)
                .cubic_to((p2[0], p2[1] + off), (p3[0] + off, p3[1]), p3)
                .cubic_to((p3[0] - off, p3[1]), (p0[0], p0[1] + off), p0)
                .close()
                .move_to((cx, cy)))

    def checkerboard(self, bbox, cell_size):
        """PathBuilder::checkerboard (src/path.rs:1031-1050): cells of the
        checkerboard over bbox = (x0, y0, x1, y1), to be filled EvenOdd."""
        x0, y0, x1, y1 = bbox
        y = y0
        while y < y1:
            x = x0
            while x < x1:
                self.move_to((x, y))
                self.line_to((x + cell_size, y))
                self.line_to((x + cell_size, y + 2.0 * cell_size))
                self.line_to((x + 2.0 * cell_size, y + 2.0 * cell_size))
                self.line_to((x + 2.0 * cell_size, y + cell_size))
                self.line_to((x, y + cell_size))
                self.close()
                x += 2.0 * cell_size
            y += 2.0 * cell_size
        return self.move_to((x0, y0))

    def build(self):
        self._finish(False)
        kinds = np.array(self._kinds, dtype=np.uint8)
        offs = np.zeros(len(kinds) + 1, dtype=np.uint32)
        offs[1:] = np.cumsum(kinds.astype(np.uint32) + 2)
        pts = np.array([v for seg in self._pts for v in seg], dtype=np.float64)
        path = Path(kinds, offs, pts, np.array(self._subpaths, np.uint32),
                    np.array(self._closed, np.uint8))
        self.__init__()
        return path


class Context:
    """One rz_ctx (device scratch, stream). Not thread-safe."""

    def __init__(self, device=0):
        h = C.c_void_p()
        s = lib().rz_ctx_create(C.byref(h), device)
        if s != _abi.RZ_OK:
            raise RzError(s, "rz_ctx_create failed (no usable GPU?)")
        self.h = h

    def check(self, s):
        if s != _abi.RZ_OK:
            raise RzError(s, lib().rz_last_error(self.h).decode())

    def sync(self):
        """rz_sync: wait for the context's work; raises the first error of the
        calls enqueued since the last sync."""
        self.check(lib().rz_sync(self.h))

    def set_param(self, param, value):
        """rz_set_param: launch choices (RZ_PARAM_*); results never depend on them."""
        self.check(lib().rz_set_param(self.h, int(param), int(value)))

    def get_param(self, param):
        v = C.c_int()
        self.check(lib().rz_get_param(self.h, int(param), C.byref(v)))
        return v.value

    def tune_save(self, path):
        """Write this context's tuned launch choices to a cache file (per GPU)."""
        self.check(lib().rz_tune_save(self.h, os.fsencode(path)))

    def tune_load(self, path):
        """Read the entries of this GPU from a cache file; returns how many."""
        n = C.c_size_t()
        self.check(lib().rz_tune_load(self.h, os.fsencode(path), C.byref(n)))
        return n.value

    def tune_clear(self):
        self.check(lib().rz_tune_clear(self.h))

    def tune_entries(self):
        n = C.c_size_t()
        self.check(lib().rz_tune_entries(self.h, C.byref(n)))
        return n.value

    def close(self):
        if self.h:
            lib().rz_ctx_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


_default_ctx = None


def default_context():
    global _default_ctx
    if _default_ctx is None:
        _default_ctx = Context(0)
    return _default_ctx


def _as_f64_view(img):
    """(data array, rz_shape) for a numpy f64 2D array or view of one."""
    if img.dtype != np.float64 or img.ndim != 2:
        raise TypeError("mask image must be a 2D float64 array")
    base = img
    while base.base is not None and isinstance(base.base, np.ndarray):
        base = base.base
    item = img.itemsize
    start = (img.__array_interface__["data"][0] - base.__array_interface__["data"][0]) // item
    shape = _abi.rz_shape(start, img.shape[1], img.shape[0], img.strides[0] // item,
                          img.strides[1] // item)
    return base, shape


# ---- paint (src/color.rs, src/grad.rs) --------------------------------------

_libm.powf.argtypes = [C.c_float, C.c_float]
_libm.powf.restype = C.c_float


def _srgb_to_linear(v):
    """Exact sRGB -> linear of one f32 channel (src/color.rs:479-486)."""
    v = np.float32(v)
    if v <= np.float32(0.04045):
        return np.float32(v / np.float32(12.92))
    return np.float32(_libm.powf(np.float32((v + np.float32(0.055)) / np.float32(1.055)), 2.4))


class LinColor(tuple):
    """Premultiplied linear RGBA colour (src/color.rs:271); also a solid Paint."""

    def __new__(cls, r, g, b, a):
        return tuple.__new__(cls, (float(np.float32(r)), float(np.float32(g)),
                                   float(np.float32(b)), float(np.float32(a))))

    @staticmethod
    def from_rgba8(r, g, b, a):
        """From<RGBA> (src/color.rs:394-406): sRGB 8-bit to premultiplied linear."""
        al = np.float32(a) / np.float32(255.0)
        ch = [_srgb_to_linear(np.float32(c) / np.float32(255.0)) * al for c in (r, g, b)]
        return LinColor(ch[0], ch[1], ch[2], al)

    def abi(self):
        p = _abi.rz_paint()
        p.kind = _abi.RZ_PAINT_SOLID
        for i in range(4):
            p.color[i] = self[i]
        return p


class Units:
    UserSpaceOnUse = 0
    BoundingBox = 1


class GradSpread:
    Pad = 0
    Repeat = 1
    Reflect = 2


class GradStop:
    def __init__(self, position, color):
        self.position = float(position)
        self.color = LinColor(*color)


class _Grad:
    kind = None

    def _abi_common(self):
        p = _abi.rz_paint()
        p.kind = self.kind
        p.units = self.units
        p.linear_colors = int(bool(self.linear_colors))
        p.spread = self.spread
        for i, v in enumerate(self.tr.m):
            p.tr[i] = v
        pos = np.array([s.position for s in self.stops], dtype=np.float64)
        cols = np.array([c for s in self.stops for c in s.color], dtype=np.float32)
        p.n_stops = len(self.stops)
        p.stop_positions = pos.ctypes.data_as(C.POINTER(C.c_double))
        p.stop_colors = cols.ctypes.data_as(C.POINTER(C.c_float))
        p._keep = (pos, cols)
        return p


class GradLinear(_Grad):
    """GradLinear (src/grad.rs:150-211)."""

    kind = _abi.RZ_PAINT_LINEAR

    def __init__(self, stops, units=Units.UserSpaceOnUse, linear_colors=False,
                 spread=GradSpread.Pad, tr=None, start=(0.0, 0.0), end=(1.0, 0.0)):
        self.stops = [s if isinstance(s, GradStop) else GradStop(*s) for s in stops]
        self.units, self.linear_colors, self.spread = units, linear_colors, spread
        self.tr = tr or Transform.identity()
        self.start, self.end = tuple(map(float, start)), tuple(map(float, end))

    def abi(self):
        p = self._abi_common()
        p.start[0], p.start[1] = self.start
        p.end[0], p.end[1] = self.end
        return p


class GradRadial(_Grad):
    """GradRadial (src/grad.rs:307-411): two-circle gradient."""

    kind = _abi.RZ_PAINT_RADIAL

    def __init__(self, stops, units=Units.UserSpaceOnUse, linear_colors=False,
                 spread=GradSpread.Pad, tr=None, center=(0.5, 0.5), radius=0.5,
                 fcenter=None, fradius=0.0):
        self.stops = [s if isinstance(s, GradStop) else GradStop(*s) for s in stops]
        self.units, self.linear_colors, self.spread = units, linear_colors, spread
        self.tr = tr or Transform.identity()
        self.center = tuple(map(float, center))
        self.radius = float(radius)
        self.fcenter = tuple(map(float, fcenter if fcenter is not None else center))
        self.fradius = float(fradius)

    def abi(self):
        p = self._abi_common()
        p.center[0], p.center[1] = self.center
        p.radius = self.radius
        p.fcenter[0], p.fcenter[1] = self.fcenter
        p.fradius = self.fradius
        return p


def _as_rgba_view(img):
    """(base float32 array, rz_shape in pixels) for an (H, W, 4) float32 array or view."""
    if img.dtype != np.float32 or img.ndim != 3 or img.shape[2] != 4 or img.strides[2] != 4:
        raise TypeError("colour image must be an (H, W, 4) float32 array with packed pixels")
    if img.strides[0] % 16 or img.strides[1] % 16:
        raise TypeError("pixel strides must be whole pixels")
    base = img
    while base.base is not None and isinstance(base.base, np.ndarray):
        base = base.base
    if not base.flags.c_contiguous:
        raise TypeError("the image's base buffer must be contiguous")
    start = (img.__array_interface__["data"][0] - base.__array_interface__["data"][0]) // 16
    shape = _abi.rz_shape(start, img.shape[1], img.shape[0], img.strides[0] // 16,
                          img.strides[1] // 16)
    return base, shape


class _RasterizerBase:
    edge_mode = None

    def __init__(self, flatness=DEFAULT_FLATNESS, ctx=None):
        self.flatness = float(flatness)
        self.ctx = ctx

    def _ctx(self):
        return self.ctx or default_context()

    def mask(self, path, tr, img, fill_rule):
        """Rasterizer::mask into a 2D float64 array (or strided view of one)."""
        ctx = self._ctx()
        base, shape = _as_f64_view(img)
        if not base.flags.c_contiguous:
            raise TypeError("the image's base buffer must be contiguous")
        data = base.ctypes.data_as(C.POINTER(C.c_double))
        ctx.check(lib().rz_mask(ctx.h, C.byref(path.abi()), tr.abi(), self.flatness, fill_rule,
                                self.edge_mode, data, base.size, shape))
        return img

    def mask_f32(self, path, tr, width, height, fill_rule, edge_mode=None):
        ctx = self._ctx()
        out = np.zeros((height, width), dtype=np.float32)
        mode = self.edge_mode if edge_mode is None else edge_mode
        ctx.check(lib().rz_mask_f32(ctx.h, C.byref(path.abi()), tr.abi(), self.flatness, fill_rule,
                                    mode, width, height,
                                    out.ctypes.data_as(C.POINTER(C.c_float)), width))
        return out

    def fill(self, path, tr, fill_rule, paint, img):
        """Rasterizer::fill: composite `paint` through the path's coverage into an
        (H, W, 4) float32 premultiplied linear RGBA image (or a view of one)."""
        ctx = self._ctx()
        base, shape = _as_rgba_view(img)
        p = paint.abi()
        ctx.check(lib().rz_fill(ctx.h, C.byref(path.abi()), tr.abi(), self.flatness, fill_rule,
                                self.iter_mode, C.byref(p),
                                base.ctypes.data_as(C.POINTER(C.c_float)), base.size, shape))
        return img

    def mask_iter(self, path, tr, size, fill_rule):
        """Rasterizer::mask_iter: list of (x, y, alpha) in row-major order."""
        width, height = size
        dense = self.mask_f32(path, tr, width, height, fill_rule, self.iter_mode)
        ys, xs = np.nonzero(np.abs(dense) >= 1e-6)
        return [(int(x), int(y), float(dense[y, x])) for y, x in zip(ys, xs)]


class SignedDifferenceRasterizer(_RasterizerBase):
    """SignedDifferenceRasterizer (src/rasterize.rs:278-360)."""

    edge_mode = _abi.RZ_EDGE_MASK
    iter_mode = _abi.RZ_EDGE_ITER

    def name(self):
        return "signed-difference"


class ActiveEdgeRasterizer(_RasterizerBase):
    """ActiveEdgeRasterizer (src/rasterize.rs:517-575); same coverage, no right clip."""

    edge_mode = _abi.RZ_EDGE_NONE
    iter_mode = _abi.RZ_EDGE_NONE

    def name(self):
        return "active-edge"


def flatten(path, tr, flatness=DEFAULT_FLATNESS, close=True, ctx=None):
    """Path::flatten on the GPU: (n, 4) float64 array of lines (x0, y0, x1, y1)."""
    ctx = ctx or default_context()
    n = C.c_size_t()
    ctx.check(lib().rz_flatten(ctx.h, C.byref(path.abi()), tr.abi(), flatness, int(close), None, 0,
                               C.byref(n)))
    out = np.zeros((max(n.value, 1), 4), dtype=np.float64)
    ctx.check(lib().rz_flatten(ctx.h, C.byref(path.abi()), tr.abi(), flatness, int(close),
                               out.ctypes.data_as(C.POINTER(C.c_double)), n.value, C.byref(n)))
    return out[: n.value]


def bin_stage(path, tr, width, height, edge_mode=_abi.RZ_EDGE_MASK, flatness=DEFAULT_FLATNESS,
              ctx=None):
    """Binning stage hook: (tile_counts, records[n,4] f32, backdrop, events[n,4],
    right_edge[height] int32)."""
    ctx = ctx or default_context()
    nr, ne = C.c_size_t(), C.c_size_t()
    ctx.check(lib().rz_bin(ctx.h, C.byref(path.abi()), tr.abi(), flatness, edge_mode, width, height,
                           None, None, 0, None, None, 0, None, C.byref(nr), C.byref(ne)))
    tx, ty = (width + 15) // 16, (height + 15) // 16
    counts = np.zeros(tx * ty, np.uint32)
    recs = np.zeros((max(nr.value, 1), 4), np.float32)
    bd = np.zeros(tx * ty, np.int32)
    ev = np.zeros((max(ne.value, 1), 4), np.float32)
    redge = np.zeros(max(height, 1), np.int32)
    ctx.check(lib().rz_bin(ctx.h, C.byref(path.abi()), tr.abi(), flatness, edge_mode, width, height,
                           counts.ctypes.data_as(C.POINTER(C.c_uint32)),
                           recs.ctypes.data_as(C.POINTER(C.c_float)), recs.shape[0],
                           bd.ctypes.data_as(C.POINTER(C.c_int32)),
                           ev.ctypes.data_as(C.POINTER(C.c_float)), ev.shape[0],
                           redge.ctypes.data_as(C.POINTER(C.c_int32)),
                           C.byref(nr), C.byref(ne)))
    return counts, recs[: nr.value], bd, ev[: ne.value], redge[:height]


# ---- workloads ------------------------------------------------------------

# ---- scenes (src/scene.rs) ---------------------------------------------------

class RawPaint:
    """A paint given directly as an rz_paint (stop arrays owned here)."""

    def __init__(self, p):
        self.p = _abi.rz_paint()
        C.pointer(self.p)[0] = p
        n = p.n_stops
        self.pos = np.ctypeslib.as_array(p.stop_positions, (n,)).copy() if n else np.zeros(1)
        self.col = np.ctypeslib.as_array(p.stop_colors, (4 * n,)).copy() if n else np.zeros(4, np.float32)
        self.p.stop_positions = self.pos.ctypes.data_as(C.POINTER(C.c_double))
        self.p.stop_colors = self.col.ctypes.data_as(C.POINTER(C.c_float))

    def abi(self):
        return self.p


def encode_rgba8(image, bottom_up=False, ctx=None):
    """RGBA::from(LinColor) (src/color.rs:164-175) of every pixel of an
    (H, W, 4) float32 premultiplied linear image, on the GPU: (H, W, 4) uint8
    sRGB with straight alpha (rows last to first when bottom_up)."""
    ctx = ctx or default_context()
    img = np.ascontiguousarray(image, dtype=np.float32)
    if img.ndim != 3 or img.shape[2] != 4:
        raise TypeError("image must be (H, W, 4)")
    h, w = img.shape[:2]
    out = np.zeros((h, w, 4), np.uint8)
    if w and h:
        ctx.check(lib().rz_encode_rgba8(ctx.h, img.ctypes.data_as(C.POINTER(C.c_float)), w, h, w,
                                        out.ctypes.data_as(C.POINTER(C.c_uint8)), 4 * w,
                                        int(bottom_up)))
    return out


def bmp_bytes(image, ctx=None):
    """ImageRead::write_bmp (src/image.rs:164-213) of a LinColor image."""
    h, w = image.shape[:2]
    hdr = (C.c_uint8 * 122)()
    if lib().rz_bmp_header(w, h, hdr) != _abi.RZ_OK:
        raise RzError(_abi.RZ_E_ARG, "image too large for BMP")
    return bytes(hdr) + encode_rgba8(image, bottom_up=True, ctx=ctx).tobytes()


def png_bytes(image, ctx=None):
    """ImageRead::write_png (src/image.rs:215-235) of a LinColor image (stored
    deflate blocks)."""
    px = encode_rgba8(image, ctx=ctx)
    h, w = px.shape[:2]
    n = C.c_size_t()
    L = lib()
    src = px.ctypes.data_as(C.POINTER(C.c_uint8))
    if L.rz_encode_png(src, w, h, 4 * w, None, 0, C.byref(n)) != _abi.RZ_OK:
        raise RzError(_abi.RZ_E_ARG, "image too large for PNG")
    buf = (C.c_uint8 * n.value)()
    if L.rz_encode_png(src, w, h, 4 * w, buf, n.value, C.byref(n)) != _abi.RZ_OK:
        raise RzError(_abi.RZ_E_ARG, "PNG encoding failed")
    return bytes(buf)


class Layer:
    """Rendered layer (src/scene.rs:463-520): `image` (H, W, 4) float32
    premultiplied linear RGBA whose pixel (0, 0) is at (x, y)."""

    def __init__(self, image, x, y):
        self.image, self.x, self.y = image, x, y

    def rgba8(self):
        return encode_rgba8(self.image)

    def write_bmp(self, path):
        with open(path, "wb") as f:
            f.write(bmp_bytes(self.image))

    def write_png(self, path):
        with open(path, "wb") as f:
            f.write(png_bytes(self.image))


class Scene:
    """A scene tree (SceneInner, src/scene.rs:19-63): Fill, Stroke, Group,
    Transform, Opacity and Clip nodes, built with the constructors of
    src/scene.rs:87-151 or read from the reference's JSON scene format
    (Scene.from_json)."""

    FILL, GROUP, TRANSFORM, OPACITY, CLIP, STROKE = (
        _abi.RZ_NODE_FILL, _abi.RZ_NODE_GROUP, _abi.RZ_NODE_TRANSFORM, _abi.RZ_NODE_OPACITY,
        _abi.RZ_NODE_CLIP, _abi.RZ_NODE_STROKE)

    def __init__(self, kind, path=None, paint=None, fill_rule=FillRule.NonZero, children=(),
                 child=None, tr=None, opacity=1.0, units=0, style=None):
        self.kind = kind
        self.path, self.paint, self.fill_rule = path, paint, int(fill_rule)
        self.children = list(children)
        self.child = child
        self.tr = tr
        self.opacity = float(opacity)
        self.units = int(units)
        self.style = style

    @staticmethod
    def fill(path, paint, fill_rule=FillRule.NonZero):
        return Scene(Scene.FILL, path=path, paint=paint, fill_rule=fill_rule)

    @staticmethod
    def stroke(path, paint, style):
        """Scene::stroke (src/scene.rs:98-101): the path stroked with `style`
        (a StrokeStyle, or a width) and filled with the non-zero rule."""
        if not isinstance(style, StrokeStyle):
            style = StrokeStyle(style)
        return Scene(Scene.STROKE, path=path, paint=paint, style=style)

    @staticmethod
    def group(children):
        """A group of one child is that child (src/scene.rs:104-109)."""
        children = list(children)
        if len(children) == 1:
            return children[0]
        return Scene(Scene.GROUP, children=children)

    def with_opacity(self, opacity):
        """Scene::opacity (src/scene.rs:112-125): clamped to [0, 1]; nested
        opacities multiply."""
        o = min(max(float(opacity), 0.0), 1.0)
        if self.kind == Scene.OPACITY:
            return self.child.with_opacity(o * self.opacity)
        return Scene(Scene.OPACITY, child=self, opacity=o)

    def with_clip(self, clip, units=0, fill_rule=FillRule.NonZero):
        """Scene::clip (src/scene.rs:128-136): clip path in user space or in
        the child's bounding box (Units.BoundingBox)."""
        return Scene(Scene.CLIP, path=clip, fill_rule=fill_rule, units=units, child=self)

    def with_transform(self, tr):
        """Scene::transform (src/scene.rs:139-151): nested transforms compose."""
        if self.kind == Scene.TRANSFORM:
            return self.child.with_transform(tr * self.tr)
        return Scene(Scene.TRANSFORM, child=self, tr=tr)

    # the reference's method names
    opacity_of = with_opacity
    clip_by = with_clip
    transform_by = with_transform

    @staticmethod
    def from_json(doc, colors=None):
        """Deserialize for Scene (src/scene.rs:13-63, 587-640): a JSON text
        (or its parsed value) of {"type": "fill" | "group" | "transform" |
        "opacity" | "clip", ...} nodes. Paths and transforms are SVG strings,
        colours "#rrggbb[aa][/alpha]" or names (`colors` adds names),
        gradients {"type": "linear-gradient" | "radial-gradient", ...}; a
        "stroke" node carries its StrokeStyle flattened: "width", "line_join"
        ({"miter": limit} | "bevel" | "round", default {"miter": 4}),
        "line_cap" ("butt" | "square" | "round", default "butt")."""
        import json
        if isinstance(doc, (str, bytes)):
            doc = json.loads(doc)
        try:
            return _scene_of_json(doc, colors)
        except KeyError as e:
            raise ValueError(f"missing field {e.args[0]!r}") from None

    def is_flat(self):
        """Only fills and groups: renders as its fills in order."""
        if self.kind == Scene.FILL:
            return True
        return self.kind == Scene.GROUP and all(c.is_flat() for c in self.children)

    def fills(self):
        """The fills of a flat scene in drawing order, (path, paint, rule)."""
        if self.kind == Scene.FILL:
            return [(self.path, self.paint, self.fill_rule)]
        if self.kind != Scene.GROUP:
            raise ValueError("fills(): the scene has transform, opacity or clip nodes")
        return [f for c in self.children for f in c.fills()]

    def _abi_arrays(self):
        fl = self.fills()
        n = len(fl)
        keep = [f[0].abi() for f in fl]
        paths = (C.POINTER(_abi.rz_path) * max(n, 1))(*[C.pointer(p) for p in keep])
        pk = [f[1].abi() for f in fl]
        paints = (_abi.rz_paint * max(n, 1))(*pk)
        rules = (C.c_int32 * max(n, 1))(*[f[2] for f in fl])
        return paths, paints, rules, (keep, pk)

    def node_arrays(self):
        """The tree as rz_render_tree's arrays: (paths, paints, nodes,
        children, root). A path or paint object used by several nodes is
        stored once."""
        paths, paints, nodes, children = [], [], [], []
        pidx, qidx = {}, {}

        def path_id(p):
            if id(p) not in pidx:
                pidx[id(p)] = len(paths)
                paths.append(p)
            return pidx[id(p)]

        def paint_id(q):
            if id(q) not in qidx:
                qidx[id(q)] = len(paints)
                paints.append(q)
            return qidx[id(q)]

        def visit(sc, depth):
            if depth > 256:
                raise ValueError("scene tree deeper than 256")
            n = _abi.rz_scene_node()
            n.kind = sc.kind
            n.fill_rule = sc.fill_rule
            n.units = sc.units
            n.opacity = sc.opacity
            if sc.kind == Scene.FILL:
                n.path, n.paint = path_id(sc.path), paint_id(sc.paint)
            elif sc.kind == Scene.STROKE:
                n.path, n.paint = path_id(sc.path), paint_id(sc.paint)
                n.stroke_width, n.miter_limit = sc.style.width, sc.style.miter_limit
                n.line_join, n.line_cap = sc.style.line_join, sc.style.line_cap
            elif sc.kind == Scene.GROUP:
                ids = [visit(c, depth + 1) for c in sc.children]
                n.first, n.count = len(children), len(ids)
                children.extend(ids)
            else:
                if sc.kind == Scene.TRANSFORM:
                    for i, v in enumerate(sc.tr.m):
                        n.tr[i] = v
                if sc.kind == Scene.CLIP:
                    n.path = path_id(sc.path)
                n.child = visit(sc.child, depth + 1)
            nodes.append(n)
            return len(nodes) - 1

        root = visit(self, 0)
        return paths, paints, nodes, children, root

    def render(self, rasterizer, tr, view=None, bg=None, batched=True):
        """Scene::render (src/scene.rs:186-199) -> Layer. view = (x0, y0, x1, y1).
        A flat scene renders through rz_render (`batched` picks its form),
        any other tree through rz_render_tree."""
        ctx = rasterizer._ctx()
        v = (C.c_double * 4)(*view) if view is not None else None
        b = (C.c_float * 4)(*bg) if bg is not None else None
        xywh = (C.c_int32 * 4)()
        L = lib()
        if self.is_flat():
            paths, paints, rules, keep = self._abi_arrays()
            n = len(keep[0])

            def call(img, cap):
                return L.rz_render(ctx.h, paths, paints, rules, n, tr.abi(), rasterizer.flatness,
                                   rasterizer.iter_mode, v, b, int(batched), img, cap, xywh)
        else:
            ps, qs, nodes, children, root = self.node_arrays()
            keep = [p.abi() for p in ps]
            pk = [q.abi() for q in qs]
            paths = (C.POINTER(_abi.rz_path) * max(len(ps), 1))(*[C.pointer(p) for p in keep])
            paints = (_abi.rz_paint * max(len(qs), 1))(*pk)
            na = (_abi.rz_scene_node * len(nodes))(*nodes)
            ca = (C.c_uint32 * max(len(children), 1))(*children)

            def call(img, cap):
                return L.rz_render_tree(ctx.h, paths, len(ps), paints, len(qs), na, len(nodes), ca,
                                        root, tr.abi(), rasterizer.flatness, rasterizer.iter_mode,
                                        v, b, img, cap, xywh)
        ctx.check(call(None, 0))
        x, y, w, h = list(xywh)
        img = np.zeros((h, w, 4), np.float32)
        if w and h:
            ctx.check(call(img.ctypes.data_as(C.POINTER(C.c_float)), w * h))
        return Layer(img, x, y)


def _json_units(v):
    if v is None or v == "userSpaceOnUse":
        return Units.UserSpaceOnUse
    if v == "objectBoundingBox":
        return Units.BoundingBox
    raise ValueError(f"unknown units: {v!r}")


def _json_rule(v):
    if v is None or v == "nonzero":
        return FillRule.NonZero
    if v == "evenodd":
        return FillRule.EvenOdd
    raise ValueError(f"unknown fill rule: {v!r}")


def _json_point(v):
    if not isinstance(v, (list, tuple)) or len(v) != 2:
        raise ValueError(f"point [x, y] expected, got {v!r}")
    return (float(v[0]), float(v[1]))


def _json_color(text, colors):
    from .color import parse_rgba
    return LinColor.from_rgba8(*parse_rgba(text, colors))


def _json_paint(v, colors):
    """serde_paint::deserialize (src/scene.rs:604-640), GradLinearSerialize /
    GradRadialSerialize (src/grad.rs:249-290, 449-492)."""
    if isinstance(v, str):
        return _json_color(v, colors)
    if not isinstance(v, dict) or "type" not in v:
        raise ValueError(f"failed to parse paint: {v!r}")
    if v["type"] not in ("linear-gradient", "radial-gradient"):
        raise ValueError(f"unknown paint type: {v['type']}")
    spreads = {"pad": GradSpread.Pad, "repeat": GradSpread.Repeat, "reflect": GradSpread.Reflect}
    if v.get("spread", "pad") not in spreads:
        raise ValueError(f"unknown spread: {v['spread']!r}")
    spread = spreads[v.get("spread", "pad")]
    stops = [GradStop(float(p), _json_color(c, colors)) for p, c in v["stops"]]
    common = dict(units=_json_units(v.get("units")), linear_colors=bool(v.get("linear_colors", False)),
                  spread=spread)
    if v["type"] == "linear-gradient":
        tr = Transform.from_svg(v["tr"]) if "tr" in v else Transform.identity()
        return GradLinear(stops, tr=tr, start=_json_point(v["start"]), end=_json_point(v["end"]),
                          **common)
    if v["type"] == "radial-gradient":
        fc = v.get("fcenter")
        return GradRadial(stops, tr=Transform.from_svg(v["tr"]), center=_json_point(v["center"]),
                          radius=float(v["radius"]), fcenter=_json_point(fc) if fc is not None else None,
                          fradius=float(v.get("fradius", 0.0)), **common)


def _scene_of_json(v, colors, depth=0):
    if depth > 256:
        raise ValueError("scene tree deeper than 256")
    t = v.get("type")
    if t == "fill":
        return Scene(Scene.FILL, path=Path.from_svg(v["path"]), paint=_json_paint(v["paint"], colors),
                     fill_rule=_json_rule(v.get("fill_rule")))
    if t == "group":
        return Scene(Scene.GROUP, children=[_scene_of_json(c, colors, depth + 1) for c in v["children"]])
    if t == "transform":
        return Scene(Scene.TRANSFORM, tr=Transform.from_svg(v["tr"]),
                     child=_scene_of_json(v["child"], colors, depth + 1))
    if t == "opacity":
        return Scene(Scene.OPACITY, opacity=float(v["opacity"]),
                     child=_scene_of_json(v["child"], colors, depth + 1))
    if t == "clip":
        return Scene(Scene.CLIP, path=Path.from_svg(v["clip"]), fill_rule=_json_rule(v.get("fill_rule")),
                     units=_json_units(v.get("units")), child=_scene_of_json(v["child"], colors, depth + 1))
    if t == "stroke":
        return Scene(Scene.STROKE, path=Path.from_svg(v["path"]), paint=_json_paint(v["paint"], colors),
                     style=_json_stroke_style(v))
    raise ValueError(f"unknown scene node type: {t!r}")


def _json_stroke_style(v):
    """StrokeStyle as serde reads it (src/path.rs:70-136, flattened into the
    Stroke node): LineJoin externally tagged, lower case."""
    j = v.get("line_join", {"miter": 4.0})
    if isinstance(j, dict) and list(j) == ["miter"]:
        join, limit = LineJoin.Miter, float(j["miter"])
    elif j == "bevel":
        join, limit = LineJoin.Bevel, 4.0
    elif j == "round":
        join, limit = LineJoin.Round, 4.0
    else:
        raise ValueError(f"unknown line join: {j!r}")
    caps = {"butt": LineCap.Butt, "square": LineCap.Square, "round": LineCap.Round}
    c = v.get("line_cap", "butt")
    if c not in caps:
        raise ValueError(f"unknown line cap: {c!r}")
    return StrokeStyle(float(v["width"]), join, caps[c], limit)


def _tune_dict(r):
    return {f: getattr(r, f) for f, _ in _abi.rz_tune_result._fields_}


class DevicePath:
    """A path uploaded once (rz_path_upload) for repeated masks into device
    memory (rz_mask_device)."""

    def __init__(self, path, ctx=None):
        self.ctx = ctx or default_context()
        self._abi_path = path.abi()
        h = C.c_void_p()
        self.ctx.check(lib().rz_path_upload(self.ctx.h, C.byref(self._abi_path), C.byref(h)))
        self.h = h

    def mask_device(self, tr, width, height, d_out, row_stride, fill_rule=FillRule.NonZero,
                    flatness=DEFAULT_FLATNESS, edge_mode=_abi.RZ_EDGE_MASK):
        """Asynchronous; errors surface at rz_sync (Context.sync)."""
        self.ctx.check(lib().rz_mask_device(self.ctx.h, self.h, tr.abi(), flatness, int(fill_rule),
                                            edge_mode, width, height, C.c_void_p(d_out), row_stride))

    def tune(self, tr, width, height, d_out, row_stride, fill_rule=FillRule.NonZero,
             flatness=DEFAULT_FLATNESS, edge_mode=_abi.RZ_EDGE_MASK, budget_ms=2000.0):
        """rz_tune_mask_device: search the launch choices for this workload
        (writing the mask meanwhile); returns the choice and both timings."""
        r = _abi.rz_tune_result()
        self.ctx.check(lib().rz_tune_mask_device(self.ctx.h, self.h, tr.abi(), flatness,
                                                 int(fill_rule), edge_mode, width, height,
                                                 C.c_void_p(d_out), row_stride, float(budget_ms),
                                                 C.byref(r)))
        return _tune_dict(r)

    def close(self):
        if self.h:
            lib().rz_path_free(self.ctx.h, self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class DeviceScene:
    """A scene uploaded once (rz_scene_upload) for repeated renders into device
    memory (rz_render_device)."""

    def __init__(self, scene, ctx=None):
        self.ctx = ctx or default_context()
        paths, paints, rules, keep = scene._abi_arrays()
        self.n = len(scene.fills())
        h = C.c_void_p()
        self.ctx.check(lib().rz_scene_upload(self.ctx.h, paths, paints, rules, self.n, C.byref(h)))
        self.h = h

    def render_device(self, tr, view, d_rgba, row_stride, bg=None, batched=True,
                      flatness=DEFAULT_FLATNESS, edge_mode=_abi.RZ_EDGE_ITER):
        v = (C.c_double * 4)(*view)
        b = (C.c_float * 4)(*bg) if bg is not None else None
        self.ctx.check(lib().rz_render_device(self.ctx.h, self.h, tr.abi(), flatness, edge_mode, v, b,
                                              int(batched), C.c_void_p(d_rgba), row_stride))

    def tune(self, tr, view, d_rgba, row_stride, bg=None, flatness=DEFAULT_FLATNESS,
             edge_mode=_abi.RZ_EDGE_ITER, budget_ms=2000.0):
        """rz_tune_render_device (batched form): search the launch choices for
        this scene and view; returns the choice and both timings."""
        v = (C.c_double * 4)(*view)
        b = (C.c_float * 4)(*bg) if bg is not None else None
        r = _abi.rz_tune_result()
        self.ctx.check(lib().rz_tune_render_device(self.ctx.h, self.h, tr.abi(), flatness, edge_mode,
                                                   v, b, C.c_void_p(d_rgba), row_stride,
                                                   float(budget_ms), C.byref(r)))
        return _tune_dict(r)

    def n_points(self):
        n = C.c_size_t()
        self.ctx.check(lib().rz_scene_points(self.ctx.h, self.h, C.byref(n)))
        return n.value

    def n_stops(self):
        """Gradient stops of the scene's paints, in path order, each paint's
        stops as given (solid paints have none)."""
        n = C.c_size_t()
        self.ctx.check(lib().rz_scene_stops(self.ctx.h, self.h, C.byref(n)))
        return n.value

    def update_stops_device(self, d_stops):
        """Replace the gradient stops' colours (device f32 premultiplied
        linear RGBA per stop, n_stops of them)."""
        self.ctx.check(lib().rz_scene_update_stops_device(self.ctx.h, self.h, C.c_void_p(d_stops)))

    def update_device(self, d_points=None, d_colors=None):
        """Replace the control points (device f64 x,y per point) and/or solid
        colours (device f32 RGBA per path) of the batched form."""
        self.ctx.check(lib().rz_scene_update_device(
            self.ctx.h, self.h, C.c_void_p(d_points) if d_points else None,
            C.c_void_p(d_colors) if d_colors else None))

    def backward_device(self, tr, view, d_grad_rgba, grad_stride, d_grad_points, d_grad_colors,
                        d_grad_tr=None, bg=None, flatness=DEFAULT_FLATNESS,
                        edge_mode=_abi.RZ_EDGE_ITER, d_grad_stops=None):
        """Adjoint of render_device (batched): gradients are added to the
        buffers (d_grad_stops: f32 RGBA per gradient stop, or None)."""
        v = (C.c_double * 4)(*view)
        b = (C.c_float * 4)(*bg) if bg is not None else None
        self.ctx.check(lib().rz_render_backward_stops_device(
            self.ctx.h, self.h, tr.abi(), flatness, edge_mode, v, b, C.c_void_p(d_grad_rgba),
            grad_stride, C.c_void_p(d_grad_points), C.c_void_p(d_grad_colors),
            C.c_void_p(d_grad_stops) if d_grad_stops else None,
            C.c_void_p(d_grad_tr) if d_grad_tr else None))

    def close(self):
        if self.h:
            lib().rz_scene_free(self.ctx.h, self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def render_torch(dscene, points, colors, tr, view, bg=None, flatness=DEFAULT_FLATNESS, stops=None):
    """Differentiable Scene::render (batched form) as a torch.autograd.Function:
    `points` (N, 2) float tensor of every path's stored control points in scene
    order, `colors` (n_paths, 4) solid premultiplied linear RGBA, `tr` a
    Transform or a 6-element tensor [m00, m01, m02, m10, m11, m12], `stops`
    None or a (n_stops, 4) tensor of the gradient paints' stop colours
    (premultiplied linear RGBA, DeviceScene.n_stops order); returns the
    (H, W, 4) float32 layer of `view`. Gradients flow to points, colors, (a
    tensor) tr and stops through rz_render_backward_stops_device.

    The library's context runs on its own HIP stream: before it reads the
    inputs (or writes a buffer torch just recycled) it waits for everything
    already enqueued on torch's current stream (rz_wait_stream), and each
    call returns only after the context's work is done. The backward pass
    re-uploads the points and colours its forward used, so other renders of
    the same scene in between do not change the gradients."""
    import torch

    tr_is_tensor = isinstance(tr, torch.Tensor)

    def _tr_of(t):
        if tr_is_tensor:
            return Transform(*[float(v) for v in t.detach().to("cpu", torch.float64).reshape(6)])
        return tr

    def _wait(dev):
        s = torch.cuda.current_stream(dev)
        dscene.ctx.check(lib().rz_wait_stream(dscene.ctx.h, C.c_void_p(s.cuda_stream)))

    class _Render(torch.autograd.Function):
        @staticmethod
        def forward(fctx, pts, cols, trt, sts):
            x0, y0 = int(np.floor(view[0])), int(np.floor(view[1]))
            w, h = int(np.ceil(view[2])) - x0, int(np.ceil(view[3])) - y0
            T = _tr_of(trt)
            p64 = pts.detach().to(torch.float64).contiguous()
            c32 = cols.detach().to(torch.float32).contiguous()
            s32 = sts.detach().to(torch.float32).contiguous() if sts is not None else None
            out = torch.empty((h, w, 4), dtype=torch.float32, device=pts.device)
            _wait(pts.device)
            dscene.update_device(p64.data_ptr(), c32.data_ptr())
            if s32 is not None:
                dscene.update_stops_device(s32.data_ptr())
            L = lib()
            for _ in range(4):
                dscene.render_device(T, view, out.data_ptr(), w, bg=bg, batched=True,
                                     flatness=flatness)
                s = L.rz_sync(dscene.ctx.h)
                if s != _abi.RZ_E_CAPACITY:
                    break
            dscene.ctx.check(s)
            fctx.keep = (p64, c32, s32)
            fctx.T = T
            fctx.shape = (h, w)
            fctx.dtypes = (pts.dtype, cols.dtype, trt.dtype if tr_is_tensor else None,
                           sts.dtype if sts is not None else None)
            return out

        @staticmethod
        def backward(fctx, grad):
            h, w = fctx.shape
            p64, c32, s32 = fctx.keep
            g = grad.detach().to(torch.float32).contiguous()
            gp = torch.zeros((dscene.n_points(), 2), dtype=torch.float32, device=g.device)
            gc = torch.zeros((dscene.n, 4), dtype=torch.float32, device=g.device)
            gt = torch.zeros(6, dtype=torch.float64, device=g.device) if tr_is_tensor else None
            gs = torch.zeros_like(s32) if s32 is not None else None
            _wait(g.device)
            # the state the forward rendered (the scene may have been updated since)
            dscene.update_device(p64.data_ptr(), c32.data_ptr())
            if s32 is not None:
                dscene.update_stops_device(s32.data_ptr())
            dscene.backward_device(fctx.T, view, g.data_ptr(), w, gp.data_ptr(), gc.data_ptr(),
                                   gt.data_ptr() if gt is not None else None, bg=bg,
                                   flatness=flatness,
                                   d_grad_stops=gs.data_ptr() if gs is not None else None)
            dscene.ctx.check(lib().rz_sync(dscene.ctx.h))
            gtr = gt.to(fctx.dtypes[2]) if gt is not None else None
            gst = gs.to(fctx.dtypes[3]) if gs is not None else None
            return gp.to(fctx.dtypes[0]), gc.to(fctx.dtypes[1]), gtr, gst

    return _Render.apply(points, colors, tr if tr_is_tensor else None, stops)


class DiffScene:
    """A scene of closed cubic paths parameterised for optimisation: `points`
    (n_paths * 3 * n_cubics, 2) unique control points (consecutive cubics
    share an endpoint), `colors` (n_paths, 4) solid premultiplied RGBA, and
    `index` mapping each stored path point (4 per cubic, scene order) to its
    unique point."""

    def __init__(self, points, colors, n_cubics):
        self.points = np.asarray(points, np.float64).reshape(-1, 2)
        self.colors = np.asarray(colors, np.float32).reshape(-1, 4)
        self.n_paths = self.colors.shape[0]
        self.n_cubics = n_cubics
        m = 3 * n_cubics
        idx = []
        for k in range(self.n_paths):
            for i in range(n_cubics):
                idx += [k * m + 3 * i, k * m + 3 * i + 1, k * m + 3 * i + 2, k * m + (3 * i + 3) % m]
        self.index = np.array(idx, np.int64)

    def paths(self, points=None):
        u = self.points if points is None else np.asarray(points, np.float64).reshape(-1, 2)
        stored = u[self.index].reshape(self.n_paths, -1)
        nc = self.n_cubics
        kinds = np.full(nc, 2, np.uint8)
        offs = np.arange(nc + 1, dtype=np.uint32) * 4
        subs = np.array([0, nc], np.uint32)
        closed = np.array([1], np.uint8)
        return [Path(kinds, offs, stored[k], subs, closed) for k in range(self.n_paths)]

    def scene(self, points=None, colors=None, rule=FillRule.NonZero):
        cols = self.colors if colors is None else np.asarray(colors, np.float32).reshape(-1, 4)
        return Scene.group([Scene.fill(p, LinColor(*c), rule) for p, c in zip(self.paths(points), cols)])


def workload_diff(seed=5, n_paths=256, n_cubics=32, size=1024.0):
    """Differentiable render workload: closed paths of cubics with control
    points uniform in [0, size)^2 and random translucent solid colours, plus a
    target made by moving the points by N(0, 2 px) and the colours by
    N(0, 0.05). Returns (scene, target_points, target_colors)."""
    rng = np.random.default_rng(seed)
    pts = rng.uniform(0.0, size, (n_paths * 3 * n_cubics, 2))
    a = rng.uniform(0.3, 0.9, (n_paths, 1))
    cols = np.concatenate([rng.uniform(0.0, 1.0, (n_paths, 3)) * a, a], 1)
    tp = pts + rng.normal(0.0, 2.0, pts.shape)
    tc = np.clip(cols + rng.normal(0.0, 0.05, cols.shape), 0.0, 1.0)
    tc[:, :3] = np.minimum(tc[:, :3], tc[:, 3:4])
    return DiffScene(pts, cols, n_cubics), tp, tc


class WorkloadSet:
    """Owned set of generated paths (include/rz_workloads.h); scene sets also
    carry a paint and a fill rule per path."""

    def __init__(self, handle, with_paints=False):
        if not handle:
            raise RuntimeError("workload generation failed")
        self.h = handle
        L = lib()
        self.paths = []
        self.paints = []
        self.rules = []
        for i in range(L.rzw_count(handle)):
            p = _abi.rz_path()
            L.rzw_path(handle, i, C.byref(p))
            self.paths.append(Path.from_abi(p))
            if with_paints:
                q = _abi.rz_paint()
                L.rzw_paint(handle, i, C.byref(q))
                self.paints.append(RawPaint(q))
                self.rules.append(int(L.rzw_rule(handle, i)))
        t = _abi.rz_transform()
        L.rzw_transform(handle, C.byref(t))
        self.transform = Transform(*t.m)

    def scene(self):
        return Scene.group([Scene.fill(p, q, r) for p, q, r in zip(self.paths, self.paints, self.rules)])

    def __del__(self):
        try:
            lib().rzw_free(self.h)
        except Exception:
            pass


def workload_cubics64(seed=1, size=512.0):
    return WorkloadSet(lib().rzw_cubics64(seed, size))


def workload_random_walk(seed=2, n_segments=200_000, canvas=4096.0):
    return WorkloadSet(lib().rzw_random_walk(seed, n_segments, canvas))


class GlyphSet(WorkloadSet):
    """Generated glyphs with their atlas placement (rz_atlas_slot per path)."""

    def __init__(self, handle):
        super().__init__(handle)
        L = lib()
        self.slots = []
        for i in range(len(self.paths)):
            sl = _abi.rz_atlas_slot()
            L.rzw_slot(handle, i, C.byref(sl))
            self.slots.append(sl)
        w, h = C.c_uint32(), C.c_uint32()
        L.rzw_atlas_size(handle, C.byref(w), C.byref(h))
        self.atlas_size = (w.value, h.value)


def atlas_slot(x, y, width, height, tr=None):
    """Placement of one glyph in an atlas (rz_atlas_slot)."""
    sl = _abi.rz_atlas_slot()
    sl.tr = (tr or Transform.identity()).abi()
    sl.x, sl.y, sl.width, sl.height = x, y, width, height
    return sl


def mask_atlas(paths, slots, atlas, fill_rule=FillRule.NonZero, flatness=DEFAULT_FLATNESS,
               edge_mode=_abi.RZ_EDGE_MASK, ctx=None):
    """Rasterizer::mask of each path into its view of a 2D float32 atlas, all in
    one pipeline (rz_mask_atlas). Views must not overlap."""
    ctx = ctx or default_context()
    if atlas.dtype != np.float32 or atlas.ndim != 2 or not atlas.flags.c_contiguous:
        raise TypeError("atlas must be a contiguous 2D float32 array")
    n = len(paths)
    keep = [p.abi() for p in paths]
    parr = (C.POINTER(_abi.rz_path) * max(n, 1))(*[C.pointer(p) for p in keep])
    sarr = (_abi.rz_atlas_slot * max(n, 1))(*slots)
    ctx.check(lib().rz_mask_atlas(ctx.h, parr, sarr, n, flatness, int(fill_rule), edge_mode,
                                  atlas.ctypes.data_as(C.POINTER(C.c_float)), atlas.shape[1],
                                  atlas.shape[0]))
    return atlas


class DeviceGlyphs:
    """A glyph set uploaded once with its placement, masked into a device atlas
    by rz_mask_atlas_device."""

    def __init__(self, paths, slots, ctx=None):
        self.ctx = ctx or default_context()
        n = len(paths)
        keep = [p.abi() for p in paths]
        parr = (C.POINTER(_abi.rz_path) * max(n, 1))(*[C.pointer(p) for p in keep])
        h = C.c_void_p()
        self.ctx.check(lib().rz_scene_upload(self.ctx.h, parr, None, None, n, C.byref(h)))
        self.h = h
        sarr = (_abi.rz_atlas_slot * max(n, 1))(*slots)
        self.ctx.check(lib().rz_atlas_set(self.ctx.h, self.h, sarr))

    def mask_device(self, d_atlas, row_stride, fill_rule=FillRule.NonZero,
                    flatness=DEFAULT_FLATNESS, edge_mode=_abi.RZ_EDGE_MASK):
        self.ctx.check(lib().rz_mask_atlas_device(self.ctx.h, self.h, flatness, int(fill_rule),
                                                  edge_mode, C.c_void_p(d_atlas), row_stride))

    def tune(self, d_atlas, row_stride, fill_rule=FillRule.NonZero, flatness=DEFAULT_FLATNESS,
             edge_mode=_abi.RZ_EDGE_MASK, budget_ms=2000.0):
        """rz_tune_atlas_device: search the launch choices for this glyph set."""
        r = _abi.rz_tune_result()
        self.ctx.check(lib().rz_tune_atlas_device(self.ctx.h, self.h, flatness, int(fill_rule),
                                                  edge_mode, C.c_void_p(d_atlas), row_stride,
                                                  float(budget_ms), C.byref(r)))
        return _tune_dict(r)

    def close(self):
        if self.h:
            lib().rz_scene_free(self.ctx.h, self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def workload_glyphs(seed=4, n=65_536, atlas_width=16384):
    """Glyph atlas: n star/blob glyphs of 8-200 quads, 1-3 contours, masks of
    16-128 px shelf-packed into one atlas."""
    return GlyphSet(lib().rzw_glyphs(seed, n, atlas_width))


def workload_scene(seed=3, n_paths=10_000, canvas=2048.0):
    """Batched scene: star/blob paths with solid / linear / radial paints."""
    return WorkloadSet(lib().rzw_scene(seed, n_paths, canvas), with_paints=True)


def workload_circles(count=1024, size=1024.0):
    """The reference's many-circles scene benchmark (benches/scene_bench.rs)."""
    return WorkloadSet(lib().rzw_circles(count, size), with_paints=True)


def rnd_uniform(seed, n):
    out = np.zeros(n, np.float64)
    lib().rzw_rnd_uniform(seed, n, out.ctypes.data_as(C.POINTER(C.c_double)))
    return out

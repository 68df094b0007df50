"""Command line renderer of an SVG path, after the reference's
examples/rasterize.rs (argument handling :47-160, rendering :232-326):

    python -m rasterize_amd [-h <height>] [-w <width>] [-v <view_box>]
        [-t <transform>] [-s <stroke>] [-fg <color>] [-bg <color>]
        [-f bmp|png|rgba] [-rf <flatness>] [-ro] [-ra] <input_file> <output_file>

The path (SVG path data; '-' reads stdin) is transformed by -t, stroked by
-s (width, round joins and caps, on the GPU), fitted into the image
(Transform::fit_size, centred), filled with -fg (default opaque black, or
60% black with -ro) over -bg, or over a 16 px #d0d0d0 checkerboard on
#f0f0f0, and written as BMP (default), PNG or raw RGBA (u32 LE height,
width, then RGBA bytes; '-' writes stdout). -ro adds the outline with its
control points (examples/rasterize.rs:167-228); -ra uses
ActiveEdgeRasterizer."""
import struct
import sys

USAGE = """Very simple tool that accepts SVG path as an input and produces rasterized image

USAGE:
    python -m rasterize_amd [-h <height>] [-w <width>] [-v <view_box>] [-t <transform>] [-s <stroke>]
              [-fg <color>] [-bg <color>] [-f <format>] [-rf <flatness>] [-ro] [-ra]
              <input_file> <output_file>

ARGS:
    -h <height>        height in pixels of the output image
    -w <width>         width in pixels of the output image
    -v <view_box>      view box "x y width height"
    -t <transform>     apply transform (SVG transform list)
    -s <stroke_width>  stroke path before rendering
    -f <format>        output file format (bmp, png, rgba)
    -fg <color>        foreground color
    -bg <color>        background color
    -rf <flatness>     flatness used by rasterizer (default: 0.05)
    -ro                rasterize outline with control points instead of filling
    -ra                rasterize active-edge instead of signed-difference rasterizer
    <input_file>       file containing SVG path ('-' means stdin)
    <output_file>      image rendered in the BMP format ('-' means stdout)
"""


class ArgError(ValueError):
    pass


def parse_view_box(text):
    """FromStr for BBox (src/geometry.rs:690-701): "minx miny width height",
    separated by spaces or commas."""
    vals = []
    for part in text.replace(",", " ").split(" "):
        if len(vals) == 4:
            break
        try:
            vals.append(float(part.strip()))
        except ValueError:
            raise ArgError(f"invalid view box: {text!r}") from None
    if len(vals) < 4:
        raise ArgError(f"invalid view box: {text!r}")
    x, y, w, h = vals
    return (min(x, x + w), min(y, y + h), max(x, x + w), max(y, y + h))


def parse_args(argv):
    a = {"input": None, "output": None, "format": "bmp", "width": 0, "height": 0, "view_box": None,
         "tr": None, "fg": None, "bg": None, "flatness": 0.05, "active_edge": False,
         "stroke": None, "outline": False}
    it = iter(argv)
    positional = []

    def value(flag):
        try:
            return next(it)
        except StopIteration:
            raise ArgError(f"{flag} requires argument") from None
    for arg in it:
        if arg == "-h":
            a["height"] = int(value(arg))
        elif arg == "-w":
            a["width"] = int(value(arg))
        elif arg == "-v":
            a["view_box"] = parse_view_box(value(arg))
        elif arg == "-t":
            a["tr"] = value(arg)
        elif arg == "-f":
            f = value(arg)
            if f not in ("bmp", "png", "rgba"):
                raise ArgError(f"Invalid output format: {f}")
            a["format"] = f
        elif arg == "-fg":
            a["fg"] = value(arg)
        elif arg == "-bg":
            a["bg"] = value(arg)
        elif arg == "-rf":
            fl = float(value(arg))
            if not fl >= 2.220446049250313e-16:
                raise ArgError("flatness is too small")
            a["flatness"] = fl
        elif arg == "-ra":
            a["active_edge"] = True
        elif arg == "-s":
            a["stroke"] = float(value(arg))
        elif arg == "-ro":
            a["outline"] = True
        else:
            positional.append(arg)
    if len(positional) > 2:
        raise ArgError("unexpected positional argment")
    if len(positional) < 2:
        return None
    a["input"], a["output"] = positional
    return a


def _gpu_stroke(path, style):
    return path.stroke(style)


def outline(path, tr, stroke=_gpu_stroke):
    """The path under tr stroked 2 px wide, with 3 px discs at the segment ends
    and its control polygons stroked 1 px wide, control points circled
    (examples/rasterize.rs:167-228), as one black non-zero fill."""
    import math
    import rasterize_amd as rz
    p = path.transform_points(tr)
    round_ = (rz.LineJoin.Round, rz.LineCap.Round)
    out = stroke(p, rz.StrokeStyle(2.0, *round_))
    control_style = rz.StrokeStyle(1.0, *round_)
    r = 3.0
    pts = p.points.reshape(-1, 2)
    for i in range(len(p)):
        b, e = int(p.subpaths[i]), int(p.subpaths[i + 1])
        for k in range(b, e):
            q = [tuple(v) for v in pts[p.point_offsets[k]:p.point_offsets[k + 1]]]
            control = rz.PathBuilder()
            if p.kinds[k] == rz._abi.RZ_QUAD:
                control.move_to(q[0]).line_to(q[1]).circle(r).line_to(q[2])
            elif p.kinds[k] == rz._abi.RZ_CUBIC:
                control.move_to(q[0]).line_to(q[1]).circle(r).move_to(q[3]).line_to(q[2]).circle(r)
            out.extend(rz.PathBuilder().move_to(q[0]).circle(r).build())
            out.extend(stroke(control.build(), control_style))
        start = pts[p.point_offsets[b]]
        end = pts[p.point_offsets[e] - 1]
        if math.hypot(*(start - end)) > r:
            out.extend(rz.PathBuilder().move_to(tuple(end)).circle(r).build())
    return rz.Scene.fill(out, rz.LinColor(0.0, 0.0, 0.0, 1.0))


def build_scene(path, a, stroke=_gpu_stroke):
    """The scene examples/rasterize.rs renders (:248-308): ((width, height),
    scene, background). `stroke(path, style)` strokes (the GPU's by
    default)."""
    import rasterize_amd as rz
    from rasterize_amd.color import parse_rgba
    if a["tr"] is not None:
        path = path.transform_points(rz.Transform.from_svg(a["tr"]))
    if a["stroke"] is not None:
        path = stroke(path, rz.StrokeStyle(a["stroke"], rz.LineJoin.Round, rz.LineCap.Round))
    view_box = a["view_box"] or path.bbox()
    if view_box is None:
        raise ArgError("nothing to render")
    (w, h), tr = rz.Transform.fit_size(view_box, a["width"], a["height"])
    if a["fg"] is not None:
        fg = rz.LinColor.from_rgba8(*parse_rgba(a["fg"]))
    else:
        fg = rz.LinColor(0.0, 0.0, 0.0, 0.6 if a["outline"] else 1.0)
    group = [rz.Scene.fill(path, fg).with_transform(tr)]
    if a["outline"]:
        group.append(outline(path, tr, stroke))
    scene = rz.Scene.group(group)
    if a["bg"] is None:
        board = rz.PathBuilder().checkerboard((0.0, 0.0, float(w), float(h)), 16.0).build()
        scene = rz.Scene.group([
            rz.Scene.fill(board, rz.LinColor.from_rgba8(*parse_rgba("#d0d0d0")), rz.FillRule.EvenOdd),
            scene])
        bg = rz.LinColor.from_rgba8(*parse_rgba("#f0f0f0"))
    else:
        bg = rz.LinColor.from_rgba8(*parse_rgba(a["bg"]))
    return (w, h), scene, bg


def encode(image, fmt):
    import rasterize_amd as rz
    if fmt == "bmp":
        return rz.bmp_bytes(image)
    if fmt == "png":
        return rz.png_bytes(image)
    h, w = image.shape[:2]
    return struct.pack("<II", h, w) + rz.encode_rgba8(image).tobytes()


def main(argv=None):
    argv = sys.argv[1:] if argv is None else argv
    try:
        a = parse_args(argv)
    except (ArgError, ValueError) as e:
        print(f"Error: {e}", file=sys.stderr)
        return 1
    if a is None:
        print(USAGE, file=sys.stderr)
        return 1
    import rasterize_amd as rz
    data = sys.stdin.buffer.read() if a["input"] == "-" else open(a["input"], "rb").read()
    try:
        path = rz.Path.from_svg(data)
        (w, h), scene, bg = build_scene(path, a)
    except (ArgError, ValueError) as e:
        print(f"Error: {e}", file=sys.stderr)
        return 1
    rast = (rz.ActiveEdgeRasterizer(a["flatness"]) if a["active_edge"]
            else rz.SignedDifferenceRasterizer(a["flatness"]))
    layer = scene.render(rast, rz.Transform.identity(), (0.0, 0.0, float(w), float(h)), bg)
    out = encode(layer.image, a["format"])
    if a["output"] == "-":
        sys.stdout.buffer.write(out)
    else:
        with open(a["output"], "wb") as f:
            f.write(out)
    return 0


if __name__ == "__main__":
    sys.exit(main())

"""Paths used by the reference's own tests, rebuilt with PathBuilder.

SQUIRREL: src/path.rs:1107-1114 (the explicit builder form is src/path.rs:1120-1151).
OCTAGON: test_rasterizer, src/rasterize.rs:1075-1085.
FILL_RULE: test_fill_rule, src/rasterize.rs:1125-1131 (SVG commands expanded).
"""
import math

import numpy as np

from rasterize_amd import PathBuilder, Transform


def squirrel():
    return (PathBuilder()
            .move_to((12.0, 1.0))
            .cubic_to((9.79, 1.0), (8.0, 2.31), (8.0, 3.92))
            .cubic_to((8.0, 5.86), (8.5, 6.95), (8.0, 10.0))
            .cubic_to((8.0, 5.5), (5.23, 3.66), (4.0, 3.66))
            .cubic_to((4.05, 3.16), (3.52, 3.0), (3.52, 3.0))
            .cubic_to((3.52, 3.0), (3.3, 3.11), (3.22, 3.34))
            .cubic_to((2.95, 3.03), (2.66, 3.07), (2.66, 3.07))
            .line_to((2.53, 3.65))
            .cubic_to((2.53, 3.65), (0.7, 4.29), (0.68, 6.87))
            .cubic_to((0.88, 7.2), (2.21, 7.47), (3.15, 7.3))
            .cubic_to((4.04, 7.35), (3.82, 8.09), (3.62, 8.29))
            .cubic_to((2.78, 9.13), (2.0, 8.0), (1.0, 8.0))
            .cubic_to((0.0, 8.0), (0.0, 9.0), (1.0, 9.0))
            .cubic_to((2.0, 9.0), (2.0, 10.0), (4.0, 10.0))
            .cubic_to((0.91, 11.2), (4.0, 14.0), (4.0, 14.0))
            .line_to((3.0, 14.0))
            .cubic_to((2.0, 14.0), (2.0, 15.0), (2.0, 15.0))
            .line_to((8.0, 15.0))
            .cubic_to((11.0, 15.0), (13.0, 14.0), (13.0, 11.53))
            .cubic_to((13.0, 10.68), (12.57, 9.74), (12.0, 9.0))
            .cubic_to((10.89, 7.54), (12.23, 6.32), (13.0, 7.0))
            .cubic_to((13.77, 7.68), (16.0, 8.0), (16.0, 5.0))
            .cubic_to((16.0, 2.79), (14.21, 1.0), (12.0, 1.0))
            .close()
            .move_to((2.5, 6.0))
            .cubic_to((2.22, 6.0), (2.0, 5.78), (2.0, 5.5))
            .cubic_to((2.0, 5.22), (2.22, 5.0), (2.5, 5.0))
            .cubic_to((2.78, 5.0), (3.0, 5.22), (3.0, 5.5))
            .cubic_to((3.0, 5.78), (2.78, 6.0), (2.5, 6.0))
            .close()
            .build())


def octagon():
    return (PathBuilder().move_to((1.0, 0.0)).line_to((1.0, 1.0)).line_to((2.0, 2.0))
            .line_to((4.0, 3.0)).line_to((5.0, 3.0)).line_to((7.0, 2.0)).line_to((8.0, 1.0))
            .line_to((8.0, 0.0)).close().build())


OCTAGON_EVENODD = np.array([
    0.0, 1.0, 1.0, 1.0, 1.0, 1.0, 1.0, 1.0, 0.0,
    0.0, 0.5, 1.0, 1.0, 1.0, 1.0, 1.0, 0.5, 0.0,
    0.0, 0.0, 0.25, 0.75, 1.0, 0.75, 0.25, 0.0, 0.0,
    0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0,
]).reshape(4, 9)


def fill_rule_shapes():
    b = PathBuilder()
    b.move_to((50, 0)).line_to((21, 90)).line_to((98, 35)).line_to((2, 35)).line_to((79, 90)).close()
    b.move_to((110, 0)).line_to((200, 0)).line_to((200, 90)).line_to((110, 90)).close()
    b.move_to((130, 20)).line_to((180, 20)).line_to((180, 70)).line_to((130, 70)).close()
    b.move_to((210, 0)).line_to((300, 0)).line_to((300, 90)).line_to((210, 90)).close()
    b.move_to((230, 20)).line_to((230, 70)).line_to((280, 70)).line_to((280, 20)).close()
    return b.build()


def path_size(path, tr, bbox):
    """Path::size (src/path.rs:439-451): (width, height, shifted transform)."""
    min_x, min_y, max_x, max_y = bbox
    mn = (math.floor(min_x) - 1.0, math.floor(min_y) - 1.0)
    mx = (math.ceil(max_x) + 1.0, math.ceil(max_y) + 1.0)
    w = int(round(mx[0] - mn[0]))
    h = int(round(mx[1] - mn[1]))
    shift = Transform.new_translate(1.0 - min_x, 1.0 - min_y)
    return w, h, shift * tr


def rect(x0, y0, x1, y1, reverse=False):
    b = PathBuilder().move_to((x0, y0))
    pts = [(x1, y0), (x1, y1), (x0, y1)]
    if reverse:
        pts = [(x0, y1), (x1, y1), (x1, y0)]
    for p in pts:
        b.line_to(p)
    return b.close().build()

"""ctypes bindings of librz.so (include/rz.h, include/rz_workloads.h).

The library is loaded from this package directory. There is no fallback: if
librz.so is missing the import fails, and GPU entry points fail loudly when no
device is present.
"""
import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "librz.so")

RZ_OK, RZ_E_NAN, RZ_E_ARG, RZ_E_HIP, RZ_E_DEPTH, RZ_E_NOMEM, RZ_E_CAPACITY = range(7)
STATUS_NAMES = {
    RZ_OK: "RZ_OK", RZ_E_NAN: "RZ_E_NAN", RZ_E_ARG: "RZ_E_ARG", RZ_E_HIP: "RZ_E_HIP",
    RZ_E_DEPTH: "RZ_E_DEPTH", RZ_E_NOMEM: "RZ_E_NOMEM", RZ_E_CAPACITY: "RZ_E_CAPACITY",
}
RZ_NONZERO, RZ_EVENODD = 0, 1
RZ_PARAM_SPLIT_QUAD, RZ_PARAM_SPLIT_CUBIC, RZ_PARAM_SPLIT_AUTO, RZ_PARAM_SCAN_DIRECT = 0, 1, 2, 3
RZ_PARAM_GRAPH, RZ_PARAM_GRAPH_REPLAYS = 4, 5
RZ_PARAM_FLAT_LDS, RZ_PARAM_BIN_LDS, RZ_PARAM_FINE_LDS, RZ_PARAM_SCATTER_WAVES, RZ_PARAM_TUNED = 6, 7, 8, 9, 10
RZ_PARAM_SCENE_WAVES, RZ_PARAM_SCENE_QUEUE, RZ_PARAM_GRAPH_GEN = 11, 12, 13
RZ_PARAM_FUSED = 14
RZ_EDGE_MASK, RZ_EDGE_ITER, RZ_EDGE_NONE = 0, 1, 2
RZ_LINE, RZ_QUAD, RZ_CUBIC = 0, 1, 2
RZ_JOIN_MITER, RZ_JOIN_BEVEL, RZ_JOIN_ROUND = 0, 1, 2
RZ_CAP_BUTT, RZ_CAP_SQUARE, RZ_CAP_ROUND = 0, 1, 2
RZ_PAINT_SOLID, RZ_PAINT_LINEAR, RZ_PAINT_RADIAL = 0, 1, 2


class rz_transform(C.Structure):
    _fields_ = [("m", C.c_double * 6)]


class rz_path(C.Structure):
    _fields_ = [
        ("n_segments", C.c_uint32),
        ("n_subpaths", C.c_uint32),
        ("kinds", C.POINTER(C.c_uint8)),
        ("point_offsets", C.POINTER(C.c_uint32)),
        ("points", C.POINTER(C.c_double)),
        ("subpaths", C.POINTER(C.c_uint32)),
        ("closed", C.POINTER(C.c_uint8)),
    ]


class rz_stroke_style(C.Structure):
    _fields_ = [
        ("width", C.c_double),
        ("miter_limit", C.c_double),
        ("line_join", C.c_int32),
        ("line_cap", C.c_int32),
    ]


class rz_shape(C.Structure):
    _fields_ = [
        ("start", C.c_size_t),
        ("width", C.c_size_t),
        ("height", C.c_size_t),
        ("row_stride", C.c_size_t),
        ("col_stride", C.c_size_t),
    ]


class rz_paint(C.Structure):
    _fields_ = [
        ("kind", C.c_int32),
        ("color", C.c_float * 4),
        ("units", C.c_int32),
        ("linear_colors", C.c_int32),
        ("spread", C.c_int32),
        ("tr", C.c_double * 6),
        ("start", C.c_double * 2),
        ("end", C.c_double * 2),
        ("center", C.c_double * 2),
        ("radius", C.c_double),
        ("fcenter", C.c_double * 2),
        ("fradius", C.c_double),
        ("n_stops", C.c_uint32),
        ("stop_positions", C.POINTER(C.c_double)),
        ("stop_colors", C.POINTER(C.c_float)),
    ]


class rz_atlas_slot(C.Structure):
    _fields_ = [("tr", rz_transform), ("x", C.c_uint32), ("y", C.c_uint32), ("width", C.c_uint32),
                ("height", C.c_uint32)]


RZ_NODE_FILL, RZ_NODE_GROUP, RZ_NODE_TRANSFORM, RZ_NODE_OPACITY, RZ_NODE_CLIP = 0, 1, 2, 3, 4
RZ_NODE_STROKE = 5


class rz_scene_node(C.Structure):
    _fields_ = [("kind", C.c_int32), ("fill_rule", C.c_int32), ("units", C.c_int32),
                ("path", C.c_uint32), ("paint", C.c_uint32), ("child", C.c_uint32),
                ("first", C.c_uint32), ("count", C.c_uint32), ("tr", C.c_double * 6),
                ("opacity", C.c_double), ("stroke_width", C.c_double),
                ("miter_limit", C.c_double), ("line_join", C.c_int32), ("line_cap", C.c_int32)]


class rz_tune_result(C.Structure):
    _fields_ = [("split_auto", C.c_int32), ("split_quad", C.c_int32), ("split_cubic", C.c_int32),
                ("flat_lds", C.c_int32), ("bin_lds", C.c_int32), ("fine_lds", C.c_int32),
                ("scatter_log2", C.c_int32), ("scene_waves", C.c_int32), ("scene_queue", C.c_int32),
                ("trials", C.c_uint32),
                ("ms_default", C.c_double),
                ("ms_tuned", C.c_double)]


class rz_kernel_stat(C.Structure):
    _fields_ = [("name", C.c_char * 40), ("calls", C.c_uint64), ("total_ms", C.c_double)]


P = C.POINTER
c_ctx = C.c_void_p
c_dpath = C.c_void_p
c_set = C.c_void_p

# (name, restype, argtypes) of every function include/rz.h and
# include/rz_workloads.h declare
PROTOTYPES = [
    ("rz_ctx_create", C.c_int, [P(c_ctx), C.c_int]),
    ("rz_ctx_destroy", C.c_int, [c_ctx]),
    ("rz_last_error", C.c_char_p, [c_ctx]),
    ("rz_sync", C.c_int, [c_ctx]),
    ("rz_stream", C.c_void_p, [c_ctx]),
    ("rz_wait_stream", C.c_int, [c_ctx, C.c_void_p]),
    ("rz_stream_wait_ctx", C.c_int, [c_ctx, C.c_void_p]),
    ("rz_svg_path_parse", C.c_int, [C.c_char_p, C.c_size_t, P(C.c_void_p), P(C.c_size_t)]),
    ("rz_path_buf_view", C.c_int, [C.c_void_p, P(rz_path)]),
    ("rz_path_stroke", C.c_int, [C.c_void_p, P(rz_path), rz_stroke_style, P(C.c_void_p)]),
    ("rz_svg_transform_parse", C.c_int, [C.c_char_p, C.c_size_t, P(rz_transform), P(C.c_size_t)]),
    ("rz_path_buf_free", C.c_int, [C.c_void_p]),
    ("rz_flatten", C.c_int, [c_ctx, P(rz_path), rz_transform, C.c_double, C.c_int,
                             P(C.c_double), C.c_size_t, P(C.c_size_t)]),
    ("rz_bin", C.c_int, [c_ctx, P(rz_path), rz_transform, C.c_double, C.c_int, C.c_size_t,
                         C.c_size_t, P(C.c_uint32), P(C.c_float), C.c_size_t, P(C.c_int32),
                         P(C.c_float), C.c_size_t, P(C.c_int32), P(C.c_size_t), P(C.c_size_t)]),
    ("rz_mask", C.c_int, [c_ctx, P(rz_path), rz_transform, C.c_double, C.c_int, C.c_int,
                          P(C.c_double), C.c_size_t, rz_shape]),
    ("rz_mask_f32", C.c_int, [c_ctx, P(rz_path), rz_transform, C.c_double, C.c_int, C.c_int,
                              C.c_size_t, C.c_size_t, P(C.c_float), C.c_size_t]),
    ("rz_path_upload", C.c_int, [c_ctx, P(rz_path), P(c_dpath)]),
    ("rz_path_free", C.c_int, [c_ctx, c_dpath]),
    ("rz_mask_device", C.c_int, [c_ctx, c_dpath, rz_transform, C.c_double, C.c_int, C.c_int,
                                 C.c_size_t, C.c_size_t, C.c_void_p, C.c_size_t]),
    ("rz_mask_device_noalloc", C.c_int, [c_ctx, c_dpath, rz_transform, C.c_double, C.c_int,
                                         C.c_int, C.c_size_t, C.c_size_t, C.c_void_p, C.c_size_t]),
    ("rz_fill", C.c_int, [c_ctx, P(rz_path), rz_transform, C.c_double, C.c_int, C.c_int,
                          P(rz_paint), P(C.c_float), C.c_size_t, rz_shape]),
    ("rz_fill_device", C.c_int, [c_ctx, c_dpath, rz_transform, C.c_double, C.c_int, C.c_int,
                                 P(rz_paint), C.c_size_t, C.c_size_t, C.c_void_p, C.c_size_t]),
    ("rz_scene_upload", C.c_int, [c_ctx, P(P(rz_path)), P(rz_paint), P(C.c_int32), C.c_size_t,
                                  P(C.c_void_p)]),
    ("rz_scene_free", C.c_int, [c_ctx, C.c_void_p]),
    ("rz_render_device", C.c_int, [c_ctx, C.c_void_p, rz_transform, C.c_double, C.c_int,
                                   P(C.c_double), P(C.c_float), C.c_int, C.c_void_p, C.c_size_t]),
    ("rz_render", C.c_int, [c_ctx, P(P(rz_path)), P(rz_paint), P(C.c_int32), C.c_size_t,
                            rz_transform, C.c_double, C.c_int, P(C.c_double), P(C.c_float), C.c_int,
                            P(C.c_float), C.c_size_t, P(C.c_int32)]),
    ("rz_render_tree", C.c_int, [c_ctx, P(P(rz_path)), C.c_size_t, P(rz_paint), C.c_size_t,
                                 P(rz_scene_node), C.c_size_t, P(C.c_uint32), C.c_uint32,
                                 rz_transform, C.c_double, C.c_int, P(C.c_double), P(C.c_float),
                                 P(C.c_float), C.c_size_t, P(C.c_int32)]),
    ("rz_scene_update_device", C.c_int, [c_ctx, C.c_void_p, C.c_void_p, C.c_void_p]),
    ("rz_scene_points", C.c_int, [c_ctx, C.c_void_p, P(C.c_size_t)]),
    ("rz_render_backward_device", C.c_int, [c_ctx, C.c_void_p, rz_transform, C.c_double, C.c_int,
                                            P(C.c_double), P(C.c_float), C.c_void_p, C.c_size_t,
                                            C.c_void_p, C.c_void_p, C.c_void_p]),
    ("rz_scene_stops", C.c_int, [c_ctx, C.c_void_p, P(C.c_size_t)]),
    ("rz_scene_update_stops_device", C.c_int, [c_ctx, C.c_void_p, C.c_void_p]),
    ("rz_render_backward_stops_device", C.c_int, [c_ctx, C.c_void_p, rz_transform, C.c_double,
                                                  C.c_int, P(C.c_double), P(C.c_float), C.c_void_p,
                                                  C.c_size_t, C.c_void_p, C.c_void_p, C.c_void_p,
                                                  C.c_void_p]),
    ("rz_encode_rgba8_device", C.c_int, [c_ctx, C.c_void_p, C.c_size_t, C.c_size_t, C.c_size_t,
                                         C.c_void_p, C.c_size_t, C.c_int]),
    ("rz_encode_rgba8", C.c_int, [c_ctx, P(C.c_float), C.c_size_t, C.c_size_t, C.c_size_t,
                                  P(C.c_uint8), C.c_size_t, C.c_int]),
    ("rz_bmp_header", C.c_int, [C.c_size_t, C.c_size_t, P(C.c_uint8)]),
    ("rz_encode_png", C.c_int, [P(C.c_uint8), C.c_size_t, C.c_size_t, C.c_size_t, P(C.c_uint8),
                                C.c_size_t, P(C.c_size_t)]),
    ("rz_atlas_set", C.c_int, [c_ctx, C.c_void_p, P(rz_atlas_slot)]),
    ("rz_mask_atlas_device", C.c_int, [c_ctx, C.c_void_p, C.c_double, C.c_int, C.c_int, C.c_void_p,
                                       C.c_size_t]),
    ("rz_mask_atlas", C.c_int, [c_ctx, P(P(rz_path)), P(rz_atlas_slot), C.c_size_t, C.c_double,
                                C.c_int, C.c_int, P(C.c_float), C.c_size_t, C.c_size_t]),
    ("rz_set_param", C.c_int, [c_ctx, C.c_int, C.c_int]),
    ("rz_get_param", C.c_int, [c_ctx, C.c_int, P(C.c_int)]),
    ("rz_path_bbox", C.c_int, [P(rz_path), rz_transform, P(C.c_int), P(C.c_double)]),
    ("rz_tune_mask_device", C.c_int, [c_ctx, c_dpath, rz_transform, C.c_double, C.c_int, C.c_int,
                                      C.c_size_t, C.c_size_t, C.c_void_p, C.c_size_t, C.c_double,
                                      P(rz_tune_result)]),
    ("rz_tune_render_device", C.c_int, [c_ctx, C.c_void_p, rz_transform, C.c_double, C.c_int,
                                        P(C.c_double), P(C.c_float), C.c_void_p, C.c_size_t,
                                        C.c_double, P(rz_tune_result)]),
    ("rz_tune_atlas_device", C.c_int, [c_ctx, C.c_void_p, C.c_double, C.c_int, C.c_int, C.c_void_p,
                                       C.c_size_t, C.c_double, P(rz_tune_result)]),
    ("rz_tune_save", C.c_int, [c_ctx, C.c_char_p]),
    ("rz_tune_load", C.c_int, [c_ctx, C.c_char_p, P(C.c_size_t)]),
    ("rz_tune_clear", C.c_int, [c_ctx]),
    ("rz_tune_entries", C.c_int, [c_ctx, P(C.c_size_t)]),
    ("rz_profile", C.c_int, [c_ctx, C.c_int]),
    ("rz_profile_stats", C.c_int, [c_ctx, C.c_void_p, C.c_size_t, P(C.c_size_t)]),
    ("rz_last_counts", C.c_int, [c_ctx, P(C.c_size_t), P(C.c_size_t), P(C.c_size_t)]),
    # workloads
    ("rzw_cubics64", c_set, [C.c_uint32, C.c_double]),
    ("rzw_random_walk", c_set, [C.c_uint32, C.c_uint32, C.c_double]),
    ("rzw_scene", c_set, [C.c_uint32, C.c_uint32, C.c_double]),
    ("rzw_circles", c_set, [C.c_uint32, C.c_double]),
    ("rzw_glyphs", c_set, [C.c_uint32, C.c_uint32, C.c_uint32]),
    ("rzw_slot", None, [c_set, C.c_size_t, P(rz_atlas_slot)]),
    ("rzw_atlas_size", None, [c_set, P(C.c_uint32), P(C.c_uint32)]),
    ("rzw_paint", None, [c_set, C.c_size_t, P(rz_paint)]),
    ("rzw_rule", C.c_int32, [c_set, C.c_size_t]),
    ("rzw_count", C.c_size_t, [c_set]),
    ("rzw_path", None, [c_set, C.c_size_t, P(rz_path)]),
    ("rzw_transform", None, [c_set, P(rz_transform)]),
    ("rzw_free", None, [c_set]),
    ("rzw_rnd_uniform", None, [C.c_uint32, C.c_size_t, P(C.c_double)]),
    ("rzw_arc_to", C.c_int, [P(C.c_double), P(C.c_double), C.c_double, C.c_int, C.c_int,
                             P(C.c_double), P(C.c_double)]),
]

_lib = None


def lib():
    """Load librz.so (raises OSError if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise OSError(f"{LIB_PATH} not found: build it with `python -m rasterize_amd.build`")
        # PyTorch ships its own libamdhip64.so.7. A process holds one library of
        # that name, so it must be torch's when torch also runs on the GPU:
        # load torch first (librz then binds to the same HIP runtime).
        try:
            import torch  # noqa: F401
        except ImportError:
            pass
        L = C.CDLL(LIB_PATH)
        for name, res, args in PROTOTYPES:
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def declared_symbols():
    """Names of the functions declared in include/*.h (parsed from the headers)."""
    import re
    inc = os.path.join(os.path.dirname(HERE), "include")
    names = set()
    for fn in ("rz.h", "rz_workloads.h"):
        src = open(os.path.join(inc, fn)).read()
        src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
        for m in re.finditer(r"\b(rzw?_[a-z0-9_]+)\s*\(", src):
            names.add(m.group(1))
    return names

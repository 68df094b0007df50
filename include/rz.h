/*
 * rz.h -- C ABI of the MI355X coverage rasterizer (librz.so).
 *
 * This is the boundary a foreign caller binds (Rust `extern "C"`, ctypes, ...).
 * Plain pointers and sizes only. Each entry point names the reference
 * interface (aslpavel/rasterize v0.6.7) it replaces.
 *
 * Conventions
 *   - Every function returns an rz_status; rz_last_error(ctx) describes the
 *     last failure on that context.
 *   - The caller owns all host buffers. The context owns device scratch, its
 *     HIP stream and its tuned launch parameters. A context is not
 *     thread-safe: use one per host thread (the reference's rasterizers are
 *     `&self` and single-threaded).
 *   - Blocking entry points take host buffers and return when the result is
 *     in them. `_device` entry points take device pointers, are asynchronous
 *     on the context's stream and report errors at the next rz_sync(): the
 *     status of every call queued since the previous sync is merged (first
 *     error, every capacity miss), so RZ_E_CAPACITY there means the queued
 *     calls must all be re-run (the buffers have been grown).
 *   - Coordinates and transforms are f64 as in the reference (Scalar = f64).
 *     Masks are produced in f32 and widened where the reference's image is
 *     f64. Colours are premultiplied linear RGBA f32 (LinColor).
 */
#ifndef RZ_H
#define RZ_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum rz_status {
  RZ_OK = 0,
  RZ_E_NAN = 1,      /* "cannot flatten segment with NaN" (src/path.rs:765-767 panics) */
  RZ_E_ARG = 2,      /* invalid argument (empty subpath, flatness <= 0, bad shape, ...) */
  RZ_E_HIP = 3,      /* HIP runtime error */
  RZ_E_DEPTH = 4,    /* subdivision deeper than 62 levels (non-finite or astronomically large input),
                        or a line starting on the right clip that runs over ~512 px right within one row */
  RZ_E_NOMEM = 5,    /* device allocation failed */
  RZ_E_CAPACITY = 6  /* an async call outgrew the context's buffers; re-run (buffers were grown) */
} rz_status;

/* FillRule (src/path.rs:18-47) */
enum { RZ_NONZERO = 0, RZ_EVENODD = 1 };

/* Right-edge semantics of the signed-difference accumulation.
 *  RZ_EDGE_MASK: SignedDifferenceRasterizer::mask -- clip at x = W-1
 *                (src/rasterize.rs:298-311, 370-387).
 *  RZ_EDGE_ITER: SignedDifferenceRasterizer::mask_iter / fill -- (W+1)-wide
 *                scratch, clip at x = W, column W dropped (src/rasterize.rs:313-355).
 *  RZ_EDGE_NONE: ActiveEdgeRasterizer -- no right clip, columns >= W never
 *                produced (src/rasterize.rs:517-575, 722-727). */
enum { RZ_EDGE_MASK = 0, RZ_EDGE_ITER = 1, RZ_EDGE_NONE = 2 };

/* Segment kinds of rz_path.kinds (src/curve.rs:905-909) */
enum { RZ_LINE = 0, RZ_QUAD = 1, RZ_CUBIC = 2 };

/* Transform [m00, m01, m02, m10, m11, m12] (src/geometry.rs:309-317) */
typedef struct rz_transform {
  double m[6];
} rz_transform;

/* Path {segments, subpaths, closed} (src/path.rs:228-233) as a host SoA view.
 * Segment k has kinds[k]+2 points starting at point index point_offsets[k]
 * (points holds x,y pairs). Subpath i is segments [subpaths[i], subpaths[i+1]).
 * Elliptic arcs are expanded to cubics at build time, as PathBuilder::arc_to
 * does (src/path.rs:945-972). */
typedef struct rz_path {
  uint32_t n_segments;
  uint32_t n_subpaths;
  const uint8_t* kinds;           /* [n_segments] */
  const uint32_t* point_offsets;  /* [n_segments + 1] */
  const double* points;           /* [2 * point_offsets[n_segments]] */
  const uint32_t* subpaths;       /* [n_subpaths + 1] */
  const uint8_t* closed;          /* [n_subpaths] */
} rz_path;

/* Shape (src/image.rs:6-17): element offsets of a strided image view */
typedef struct rz_shape {
  size_t start;
  size_t width;
  size_t height;
  size_t row_stride;
  size_t col_stride;
} rz_shape;

/* Paint (src/rasterize.rs:185-198): LinColor (src/color.rs:271),
 * GradLinear (src/grad.rs:150-211) or GradRadial (src/grad.rs:307-411).
 * Stop colours are given as the user passes them to GradStop (linear
 * premultiplied); the sRGB conversion of non-linear_colors gradients
 * (src/grad.rs:177-179) is applied by the library. */
enum { RZ_PAINT_SOLID = 0, RZ_PAINT_LINEAR = 1, RZ_PAINT_RADIAL = 2 };
enum { RZ_UNITS_USER_SPACE = 0, RZ_UNITS_BBOX = 1 };
enum { RZ_SPREAD_PAD = 0, RZ_SPREAD_REPEAT = 1, RZ_SPREAD_REFLECT = 2 };
typedef struct rz_paint {
  int32_t kind;
  float color[4];        /* solid */
  int32_t units;
  int32_t linear_colors;
  int32_t spread;
  double tr[6];          /* paint transform */
  double start[2];       /* linear */
  double end[2];
  double center[2];      /* radial */
  double radius;
  double fcenter[2];
  double fradius;
  uint32_t n_stops;
  const double* stop_positions; /* [n_stops] */
  const float* stop_colors;     /* [4 * n_stops] */
} rz_paint;

typedef struct rz_ctx rz_ctx;
typedef struct rz_dpath rz_dpath;

/* ---- context ---------------------------------------------------------- */
rz_status rz_ctx_create(rz_ctx** out, int device);
rz_status rz_ctx_destroy(rz_ctx* ctx);
const char* rz_last_error(const rz_ctx* ctx);
/* Wait for the context's stream; returns the first asynchronous error. */
rz_status rz_sync(rz_ctx* ctx);
/* The context's hipStream_t (as void*), for callers that order their own work. */
void* rz_stream(rz_ctx* ctx);
/* Stream ordering with a caller's stream (hipStream_t as void*, NULL = the
 * null stream). rz_wait_stream: work the context enqueues from now on starts
 * after everything already enqueued on `stream` (e.g. the kernels that wrote
 * a device input, or the last use of a recycled output buffer).
 * rz_stream_wait_ctx: work enqueued on `stream` from now on starts after
 * everything the context has enqueued. No reference counterpart (the
 * reference is single-threaded host code). */
rz_status rz_wait_stream(rz_ctx* ctx, void* stream);
rz_status rz_stream_wait_ctx(rz_ctx* ctx, void* stream);

/* ---- SVG path data ------------------------------------------------------ */

/* Path::read_svg_path / FromStr for Path (src/path.rs:528-534, 1058-1064;
 * the parser of src/svg.rs:250-425): SVG path data (M L H V C S Q T A Z and
 * their relative forms, implicit repeats) into an owned path, built with
 * PathBuilder's rules (zero-length lines dropped, arcs expanded to cubics).
 * On a syntax error returns RZ_E_ARG and, if err_offset is not NULL, the
 * byte offset where parsing stopped. Host only. */
typedef struct rz_path_buf rz_path_buf;
rz_status rz_svg_path_parse(const char* text, size_t len, rz_path_buf** out, size_t* err_offset);
/* FromStr for Transform (src/svg.rs:431-600): an SVG transform list
 * (matrix, translate[X|Y], scale[X|Y], rotate with an optional centre, skewX,
 * skewY; deg or rad angles) composed left to right. Host only. */
rz_status rz_svg_transform_parse(const char* text, size_t len, rz_transform* out,
                                 size_t* err_offset);
/* The rz_path view of an owned path (valid until rz_path_buf_free). */
rz_status rz_path_buf_view(const rz_path_buf* path, rz_path* view);
rz_status rz_path_buf_free(rz_path_buf* path);

/* ---- stroking --------------------------------------------------------- */

/* LineJoin and LineCap (src/path.rs:70-116) */
enum { RZ_JOIN_MITER = 0, RZ_JOIN_BEVEL = 1, RZ_JOIN_ROUND = 2 };
enum { RZ_CAP_BUTT = 0, RZ_CAP_SQUARE = 1, RZ_CAP_ROUND = 2 };
/* StrokeStyle (src/path.rs:118-136). miter_limit is LineJoin::Miter's
 * argument (the reference's default style is Miter(4.0), Butt). */
typedef struct rz_stroke_style {
  double width;
  double miter_limit;
  int32_t line_join;
  int32_t line_cap;
} rz_stroke_style;
/* Path::stroke (src/path.rs:370-415): the outline of `path` offset by
 * width / 2 on both sides, joined and capped as `style` says, as a new owned
 * path whose subpaths are all closed (free it with rz_path_buf_free). The
 * offsetting, joins and caps run on the GPU, one work item per segment and
 * direction; the input is copied from and the result to host memory, so the
 * call synchronises the context's stream. RZ_E_ARG for an unknown join or
 * cap. */
rz_status rz_path_stroke(rz_ctx* ctx, const rz_path* path, rz_stroke_style style,
                         rz_path_buf** out);

/* ---- stage hooks (bit-exact parity surfaces) ------------------------- */

/* Path::flatten(tr, flatness, close) (src/path.rs:418-425, 734-796): lines as
 * (x0, y0, x1, y1) f64 quadruples in the reference's order. If cap is too
 * small nothing is written, *n_lines gets the count and RZ_E_CAPACITY is
 * returned. */
rz_status rz_flatten(rz_ctx* ctx, const rz_path* path, rz_transform tr, double flatness,
                     int close, double* lines, size_t cap, size_t* n_lines);

/* Binning stage of the mask pipeline for a width x height image: per 16x16
 * tile its record count (tile_counts[tiles]), the records themselves (4 f32
 * each, tile-major, records of one tile in unspecified order), the band
 * backdrop winding per tile (backdrop[tiles], the integer winding that
 * crossings of the band's bottom edge left of the tile contribute), the
 * free-endpoint events (4 x int32/f32 each) and the right-edge column
 * (right_edge[height], per pixel row the 2^-20 fixed-point sum of the row
 * overlaps of the lines that lie beyond the right clip). Sizes are returned
 * through n_records / n_events; pass NULL buffers to query them.
 * No reference counterpart: this is the GPU decomposition of
 * signed_difference_line (src/rasterize.rs:365-470) exposed for checking. */
rz_status rz_bin(rz_ctx* ctx, const rz_path* path, rz_transform tr, double flatness,
                 int edge_mode, size_t width, size_t height, uint32_t* tile_counts,
                 float* records, size_t records_cap, int32_t* backdrop, float* events,
                 size_t events_cap, int32_t* right_edge, size_t* n_records,
                 size_t* n_events);

/* ---- Rasterizer::mask / mask_iter (drop-in, blocking) --------------------- */

/* Rasterizer::mask (src/rasterize.rs:50-56): SignedDifferenceRasterizer
 * (edge_mode RZ_EDGE_MASK, src/rasterize.rs:298-311) or ActiveEdgeRasterizer
 * (RZ_EDGE_NONE, src/rasterize.rs:537-556) into the caller's f64 image
 * data[data_len] viewed through `shape`. Every pixel of the view is written.
 * flatness: the rasterizer's tolerance (DEFAULT_FLATNESS = 0.05). */
rz_status rz_mask(rz_ctx* ctx, const rz_path* path, rz_transform tr, double flatness,
                  int fill_rule, int edge_mode, double* data, size_t data_len, rz_shape shape);

/* Same computation into a dense row-major f32 buffer out[height * row_stride]. */
rz_status rz_mask_f32(rz_ctx* ctx, const rz_path* path, rz_transform tr, double flatness,
                      int fill_rule, int edge_mode, size_t width, size_t height, float* out,
                      size_t row_stride);

/* ---- device-resident path (inputs already in HBM) ------------------------ */

/* Upload a path once; the handle is reused across calls. */
rz_status rz_path_upload(rz_ctx* ctx, const rz_path* path, rz_dpath** out);
rz_status rz_path_free(rz_ctx* ctx, rz_dpath* dpath);

/* Asynchronous Rasterizer::mask of a resident path into device memory
 * d_out[height * row_stride] (f32). Errors (NaN, capacity) surface at rz_sync. */
rz_status rz_mask_device(rz_ctx* ctx, const rz_dpath* dpath, rz_transform tr, double flatness,
                         int fill_rule, int edge_mode, size_t width, size_t height, float* d_out,
                         size_t row_stride);

/* No-alloc form of rz_mask_device. No counterpart in the reference: it is for
 * callers that cannot allocate device memory (code under HIP graph capture,
 * tight serving loops, harnesses that forbid hipMalloc). It calls no
 * hipMalloc, hipFree or allocating library routine: the pipeline's buffers
 * are carved from one 64 MiB array per GPU (a file-scope __device__ array
 * of librz, held until the process ends) instead of the context's. The path
 * must already be resident (rz_path_upload, done at set-up) and the caller
 * passes the output. A size that cannot fit (the per-tile, per-band and
 * per-flatten-wave parts of the call) is refused with RZ_E_CAPACITY before
 * anything is enqueued; a path that flattens into more lines than the rest
 * of the array holds (about 600k lines for a 4096x4096 mask) is reported at
 * rz_sync as RZ_E_CAPACITY, with the output not written and nothing grown.
 * All no-alloc calls on one GPU share its array: calls that overlap on one
 * GPU (other contexts, other streams) race; calls on one context's stream
 * are ordered. */
rz_status rz_mask_device_noalloc(rz_ctx* ctx, const rz_dpath* dpath, rz_transform tr,
                                 double flatness, int fill_rule, int edge_mode, size_t width,
                                 size_t height, float* d_out, size_t row_stride);

/* ---- Rasterizer::fill ------------------------------------------------------ */

/* Rasterizer::fill (src/rasterize.rs:69-115): the pixels mask_iter yields
 * (alpha >= 1e-6; edge_mode RZ_EDGE_ITER for SignedDifferenceRasterizer,
 * RZ_EDGE_NONE for ActiveEdgeRasterizer) are composited source-over,
 * dst = dst.blend_over(paint.at(pixel centre).with_alpha(alpha)), into the
 * caller's premultiplied linear RGBA f32 image (LinColor, 4 floats per pixel)
 * rgba[rgba_len floats] viewed through `shape` (in pixels). Gradient paints
 * use the pixel-to-paint transform (tr * paint.tr)^-1 (UserSpaceOnUse) or
 * (tr * path.bbox(I).unit_transform() * paint.tr)^-1 (BoundingBox); a
 * non-invertible one draws nothing, as in the reference. */
rz_status rz_fill(rz_ctx* ctx, const rz_path* path, rz_transform tr, double flatness,
                  int fill_rule, int edge_mode, const rz_paint* paint, float* rgba,
                  size_t rgba_len, rz_shape shape);

/* Asynchronous fill of a resident path into a device RGBA f32 image
 * d_rgba[height * row_stride pixels]. */
rz_status rz_fill_device(rz_ctx* ctx, const rz_dpath* dpath, rz_transform tr, double flatness,
                         int fill_rule, int edge_mode, const rz_paint* paint, size_t width,
                         size_t height, float* d_rgba, size_t row_stride);

/* ---- Scene::render ---------------------------------------------------------- */

/* A scene: Scene::group of n Scene::fill(path, paint, fill_rule) nodes
 * (src/scene.rs:64-110), uploaded once and rendered many times. paints and
 * fill_rules may be NULL (opaque black, NonZero), e.g. for glyph sets. */
typedef struct rz_dscene rz_dscene;
rz_status rz_scene_upload(rz_ctx* ctx, const rz_path* const* paths, const rz_paint* paints,
                          const int32_t* fill_rules, size_t n, rz_dscene** out);
rz_status rz_scene_free(rz_ctx* ctx, rz_dscene* scene);

/* Scene::render(rasterizer, tr, Some(view), bg) (src/scene.rs:186-199,
 * 384-459) into a device RGBA f32 layer d_rgba[lh * row_stride pixels], where
 * the layer spans floor(view.min) .. ceil(view.max) (Layer::new,
 * src/scene.rs:483-501); view = {x0, y0, x1, y1}; bg = LinColor or NULL
 * (transparent). Each fill is the rasterizer's fill of its path under
 * translate(-floor(bbox.min)) * tr into the sub-view its bbox covers, in path
 * order. edge_mode selects the rasterizer (RZ_EDGE_ITER: SignedDifference,
 * RZ_EDGE_NONE: ActiveEdge). batched = 1 renders every path in one pipeline
 * (per-tile path lists, one store per pixel); batched = 0 runs the fill
 * pipeline once per path (the plain form). Asynchronous; errors at rz_sync.
 * The batched form has no reference counterpart. */
rz_status rz_render_device(rz_ctx* ctx, const rz_dscene* scene, rz_transform tr, double flatness,
                           int edge_mode, const double* view, const float* bg, int batched,
                           float* d_rgba, size_t row_stride);

/* Blocking Scene::render of n fills into host memory. view = NULL renders the
 * union of the fills' bboxes. layer_xywh receives the layer's origin and
 * size ({0,0,0,0} when no fill is in view: the reference's empty layer); the
 * pixels are written when rgba holds at least w * h pixels (cap_px). */
rz_status rz_render(rz_ctx* ctx, const rz_path* const* paths, const rz_paint* paints,
                    const int32_t* fill_rules, size_t n, rz_transform tr, double flatness,
                    int edge_mode, const double* view, const float* bg, int batched, float* rgba,
                    size_t cap_px, int32_t* layer_xywh);

/* A scene tree (SceneInner, src/scene.rs:19-63) as a node array: Fill (path,
 * paint, fill_rule), Group (children[first .. first + count)), Transform
 * (tr, child), Opacity (opacity, used as given: Scene::opacity's clamp to
 * [0, 1] is the caller's, as in the reference's deserialized scenes; child), Clip (the clip
 * path, fill_rule, units: RZ_UNITS_USER_SPACE or RZ_UNITS_BBOX of the child,
 * child) and Stroke (path, paint, and the StrokeStyle in stroke_width,
 * miter_limit, line_join, line_cap: the path stroked by rz_path_stroke and
 * filled with the non-zero rule, src/scene.rs:297-306). Indices refer to the
 * paths / paints / nodes / children arrays. */
enum { RZ_NODE_FILL = 0, RZ_NODE_GROUP = 1, RZ_NODE_TRANSFORM = 2, RZ_NODE_OPACITY = 3,
       RZ_NODE_CLIP = 4, RZ_NODE_STROKE = 5 };
typedef struct rz_scene_node {
  int32_t kind;
  int32_t fill_rule;
  int32_t units;
  uint32_t path;
  uint32_t paint;
  uint32_t child;
  uint32_t first, count;
  double tr[6];
  double opacity;
  double stroke_width;
  double miter_limit;
  int32_t line_join;
  int32_t line_cap;
} rz_scene_node;

/* Blocking Scene::render(rasterizer, tr, view, bg) (src/scene.rs:186-199) of
 * the tree rooted at nodes[root], with Pipeline::build_rec's bboxes and views
 * (src/scene.rs:268-357) and render_rec's layers (src/scene.rs:397-459): the
 * fills drawn into one layer go through one batched render (each under its
 * own transform), an Opacity or Clip node into a layer of its own, composed
 * into its parent (Layer::compose, src/scene.rs:528-560). Output as rz_render:
 * layer_xywh always, the pixels when rgba holds them (cap_px). edge_mode
 * selects the rasterizer (RZ_EDGE_ITER: SignedDifference, RZ_EDGE_NONE:
 * ActiveEdge); clip masks use its Rasterizer::mask. */
rz_status rz_render_tree(rz_ctx* ctx, const rz_path* const* paths, size_t n_paths,
                         const rz_paint* paints, size_t n_paints, const rz_scene_node* nodes,
                         size_t n_nodes, const uint32_t* children, uint32_t root, rz_transform tr,
                         double flatness, int edge_mode, const double* view, const float* bg,
                         float* rgba, size_t cap_px, int32_t* layer_xywh);

/* ---- differentiable render --------------------------------------------------- */

/* Replace a resident scene's control points (device f64 x,y pairs of every
 * path's points, in path order, rz_scene_points of them) and/or its solid
 * paint colours (device f32 RGBA per path; other paints keep theirs). NULL
 * leaves that part unchanged. Only the batched form sees the update (the
 * plain form keeps the uploaded paths). No reference counterpart. */
rz_status rz_scene_update_device(rz_ctx* ctx, rz_dscene* scene, const double* d_points,
                                 const float* d_colors);
rz_status rz_scene_points(rz_ctx* ctx, const rz_dscene* scene, size_t* n_points);

/* Adjoint of rz_render_device (batched form) with the same arguments: given
 * dL/d(layer) d_grad_rgba (device RGBA f32, grad_stride pixels per row) it
 * adds dL/d(control points) to d_grad_points (f32 x,y per point of the
 * scene), dL/d(solid paint colour) to d_grad_colors (f32 RGBA per path) and,
 * if not NULL, dL/d(transform) to d_grad_tr (f64 [6]). The coverage gradient
 * differentiates the exact box-filtered winding (edge terms of the signed
 * area), the subdivision's fixed-t Bezier evaluation and source-over
 * compositing; subdivision and tile lists are recomputed from the forward
 * pass. No reference counterpart. */
rz_status rz_render_backward_device(rz_ctx* ctx, const rz_dscene* scene, rz_transform tr,
                                    double flatness, int edge_mode, const double* view,
                                    const float* bg, const float* d_grad_rgba, size_t grad_stride,
                                    float* d_grad_points, float* d_grad_colors, double* d_grad_tr);

/* Gradient paints' stops: the scene's given stops are those of its paints, in
 * path order, each paint's stops in the order given (rz_paint.stop_colors;
 * a solid paint has none), rz_scene_stops of them. rz_scene_update_stops_device
 * replaces their colours (device f32 premultiplied linear RGBA per stop; the
 * batched form reads them, converted to sRGB where the gradient interpolates
 * sRGB colours, as GradLinear::new / GradRadial::new do, src/grad.rs:163-194).
 * rz_render_backward_stops_device is rz_render_backward_device that also adds
 * dL/d(stop colour) to d_grad_stops (f32 RGBA per given stop), through
 * GradStops::at's interpolation and the sRGB conversions. No reference
 * counterpart. */
rz_status rz_scene_stops(rz_ctx* ctx, const rz_dscene* scene, size_t* n_stops);
rz_status rz_scene_update_stops_device(rz_ctx* ctx, rz_dscene* scene, const float* d_stop_colors);
rz_status rz_render_backward_stops_device(rz_ctx* ctx, const rz_dscene* scene, rz_transform tr,
                                          double flatness, int edge_mode, const double* view,
                                          const float* bg, const float* d_grad_rgba,
                                          size_t grad_stride, float* d_grad_points,
                                          float* d_grad_colors, float* d_grad_stops,
                                          double* d_grad_tr);

/* ---- output encoders ---------------------------------------------------------- */

/* RGBA::from(LinColor) (src/color.rs:164-175) for every pixel of a device
 * LinColor image d_rgba (width x height, row_stride pixels per row): l2s of
 * the unmultiplied colour, each channel (v * 255 + 0.5) as u8 (saturating),
 * alpha from the premultiplied alpha; 4 bytes per pixel, R, G, B, A, into
 * d_out (out_stride bytes per row). bottom_up = 1 writes the last row first
 * (the BMP row order). Asynchronous on the context's stream. */
rz_status rz_encode_rgba8_device(rz_ctx* ctx, const float* d_rgba, size_t width, size_t height,
                                 size_t row_stride, uint8_t* d_out, size_t out_stride,
                                 int bottom_up);
/* Blocking form over host buffers: rgba (row_stride pixels per row) to out. */
rz_status rz_encode_rgba8(rz_ctx* ctx, const float* rgba, size_t width, size_t height,
                          size_t row_stride, uint8_t* out, size_t out_stride, int bottom_up);
/* ImageRead::write_bmp (src/image.rs:164-213): the 122-byte header (file
 * header + BITMAPV4HEADER, 32-bit BI_BITFIELDS, R G B A byte masks, sRGB) of
 * a width x height image; the file is this header followed by the pixel rows
 * bottom-up (rz_encode_rgba8_device with bottom_up = 1). Host only. */
rz_status rz_bmp_header(size_t width, size_t height, uint8_t* out /* [122] */);
/* ImageRead::write_png (src/image.rs:215-235): a PNG (8-bit RGBA) of the
 * rows rgba8 (top-down, stride bytes apart) into out[cap]; *len receives the
 * file size (pass out = NULL to query it). The image data is stored in
 * uncompressed deflate blocks (the reference's png crate compresses; every
 * decoder reads both). Host only. */
rz_status rz_encode_png(const uint8_t* rgba8, size_t width, size_t height, size_t stride,
                        uint8_t* out, size_t cap, size_t* len);

/* ---- glyph atlas ------------------------------------------------------------ */

/* Where glyph i goes: Rasterizer::mask of its path under `tr` into the atlas
 * view of width x height pixels at (x, y) (src/image.rs:300-311). */
typedef struct rz_atlas_slot {
  rz_transform tr;
  uint32_t x, y, width, height;
} rz_atlas_slot;

/* Glyph sets are uploaded with rz_scene_upload (paints and fill_rules may be
 * NULL); rz_atlas_set places them (one slot per path, resident with the set). */
rz_status rz_atlas_set(rz_ctx* ctx, rz_dscene* glyphs, const rz_atlas_slot* slots);

/* Asynchronous Rasterizer::mask of every glyph into its view of the device
 * atlas d_atlas (f32, row_stride floats per row), all glyphs in one pipeline.
 * Every pixel of every view is written; other pixels are not touched. Views
 * must not overlap. No reference counterpart (the reference masks glyphs one
 * by one). */
rz_status rz_mask_atlas_device(rz_ctx* ctx, const rz_dscene* glyphs, double flatness,
                               int fill_rule, int edge_mode, float* d_atlas, size_t row_stride);

/* Blocking form: n paths and slots, host atlas of atlas_width x atlas_height
 * f32 (dense rows). */
rz_status rz_mask_atlas(rz_ctx* ctx, const rz_path* const* paths, const rz_atlas_slot* slots,
                        size_t n, double flatness, int fill_rule, int edge_mode, float* atlas,
                        size_t atlas_width, size_t atlas_height);

/* ---- launch parameters ----------------------------------------------------
 * No reference counterpart. RZ_PARAM_SPLIT_QUAD / RZ_PARAM_SPLIT_CUBIC: each
 * quad / cubic segment's subdivision tree is processed by 2^value threads
 * (one per subtree at that depth), 0..8; setting either fixes both for the
 * context. RZ_PARAM_SPLIT_AUTO = 1 (the default) chooses them per call from
 * the segments' expected line counts. RZ_PARAM_SCAN_DIRECT (0..16, default
 * 12): a scan of up to 2^value blocks of 2048 counts has each block sum the
 * totals before its own; a longer one scans the totals in a kernel of their
 * own. RZ_PARAM_GRAPH (0 or 1, default 0): rz_mask_device,
 * rz_render_device (batched), rz_mask_atlas_device and
 * rz_render_backward_device replay a frame whose arguments repeat from a
 * HIP graph captured after the same call completed once without growing the
 * context's buffers (the device data they read, e.g. updated scene points,
 * may change between frames). RZ_PARAM_GRAPH_REPLAYS (read only): frames
 * replayed so far. RZ_PARAM_GRAPH_GEN (read only): the context's buffer and
 * parameter generation (a frame captured under another one is not replayed). RZ_PARAM_FLAT_LDS (-1..32, default -1: 4 for one path, 0
 * for scenes and atlases), RZ_PARAM_BIN_LDS, RZ_PARAM_FINE_LDS (0..32,
 * default 0): KB of extra LDS per wave of the flatten, the count pass and the
 * fine stage (fewer waves per SIMD). RZ_PARAM_SCATTER_WAVES (8..18, default
 * 15): the scatter's grid holds at most 2^value waves. RZ_PARAM_SCENE_WAVES
 * (1, 6 or 8, default 6): the batched scene's fine kernel built for that many
 * waves per SIMD (1: registers unbounded). RZ_PARAM_SCENE_QUEUE (0 or 1,
 * default 0): the batched scene's fine stage as a persistent grid taking
 * canvas tiles from a device queue, the tiles covered by the most paths
 * first. RZ_PARAM_FUSED (0 or 1, default 0): flatten and the binning's
 * count pass as one kernel whose waves bin the lines they made from LDS (0:
 * two kernels through an f64 line buffer in HBM, measured 20-35% faster on
 * a single 200k-segment path, a 10k-path scene and a 65k-glyph atlas).
 * RZ_PARAM_TUNED (0 or
 * 1, default 1): a call whose workload has a tuned entry (rz_tune_*) uses that
 * entry's choices instead of the context's. Results do not depend on any of
 * them. */
enum { RZ_PARAM_SPLIT_QUAD = 0, RZ_PARAM_SPLIT_CUBIC = 1, RZ_PARAM_SPLIT_AUTO = 2,
       RZ_PARAM_SCAN_DIRECT = 3, RZ_PARAM_GRAPH = 4, RZ_PARAM_GRAPH_REPLAYS = 5,
       RZ_PARAM_FLAT_LDS = 6, RZ_PARAM_BIN_LDS = 7, RZ_PARAM_FINE_LDS = 8,
       RZ_PARAM_SCATTER_WAVES = 9, RZ_PARAM_TUNED = 10, RZ_PARAM_SCENE_WAVES = 11,
       RZ_PARAM_SCENE_QUEUE = 12, RZ_PARAM_GRAPH_GEN = 13, RZ_PARAM_FUSED = 14 };
rz_status rz_set_param(rz_ctx* ctx, int param, int value);
rz_status rz_get_param(rz_ctx* ctx, int param, int* value);

/* Path::bbox(tr) (src/path.rs:428-431): the exact box of the transformed
 * segments (curve extrema, not control points) as x0, y0, x1, y1 into bbox;
 * *has_bbox = 0 for a path without segments. Host only. */
rz_status rz_path_bbox(const rz_path* path, rz_transform tr, int* has_bbox, double* bbox);

/* ---- launch tuning ---------------------------------------------------------
 * No reference counterpart (the reference's only tunable is the flatness,
 * src/rasterize.rs:279-296, which changes results and is not tuned here).
 * rz_tune_mask_device / rz_tune_render_device (batched form) /
 * rz_tune_atlas_device run the call they are given repeatedly (so they
 * write its output, as the call would), searching the launch parameters
 * above one at a time (virtual-subtree depths 0..6, extra LDS per wave for
 * the flatten, the count pass and the fine stage, scatter grid size), each
 * candidate timed with HIP events over about 3 ms of frames, for at most
 * about budget_ms. The fastest becomes the entry of the call's workload
 * (keyed by entry point, edge mode and log2 buckets of the pixels, the
 * path's slots and its curves' expected line counts), which later calls with
 * that workload use while RZ_PARAM_TUNED is on. Every choice gives the same
 * result bit for bit. rz_tune_save / rz_tune_load keep the entries in a text
 * file, per GPU (arch and CU count; other GPUs' lines are kept by a save and
 * skipped by a load). Not to be called while a capture is open. */
typedef struct rz_tune_result {
  int32_t split_auto;    /* 1: depths from the per-call estimate (the two below) */
  int32_t split_quad, split_cubic;
  int32_t flat_lds;      /* bytes of extra LDS per flatten wave, -1: the built-in choice */
  int32_t bin_lds, fine_lds;
  int32_t scatter_log2;
  int32_t scene_waves, scene_queue;
  uint32_t trials;       /* candidates timed */
  double ms_default;     /* per frame with the context's parameters */
  double ms_tuned;       /* per frame with the chosen ones */
} rz_tune_result;
rz_status rz_tune_mask_device(rz_ctx* ctx, const rz_dpath* path, rz_transform tr, double flatness,
                              int fill_rule, int edge_mode, size_t width, size_t height,
                              float* d_out, size_t row_stride, double budget_ms,
                              rz_tune_result* result);
rz_status rz_tune_render_device(rz_ctx* ctx, const rz_dscene* scene, rz_transform tr,
                                double flatness, int edge_mode, const double* view,
                                const float* bg, float* d_rgba, size_t row_stride,
                                double budget_ms, rz_tune_result* result);
rz_status rz_tune_atlas_device(rz_ctx* ctx, const rz_dscene* glyphs, double flatness,
                               int fill_rule, int edge_mode, float* d_atlas, size_t row_stride,
                               double budget_ms, rz_tune_result* result);
rz_status rz_tune_save(rz_ctx* ctx, const char* path);
rz_status rz_tune_load(rz_ctx* ctx, const char* path, size_t* n_loaded);
rz_status rz_tune_clear(rz_ctx* ctx);
rz_status rz_tune_entries(rz_ctx* ctx, size_t* n_entries);

/* ---- kernel timing -------------------------------------------------------
 * No reference counterpart (the reference only has `tracing` spans in its
 * examples): with profiling on, every kernel the context launches is
 * bracketed by HIP events on the context's stream; rz_profile_stats waits for
 * the stream and returns per-kernel launch counts and total device time. */
typedef struct rz_kernel_stat {
  char name[40];
  uint64_t calls;
  double total_ms;
} rz_kernel_stat;
rz_status rz_profile(rz_ctx* ctx, int enable); /* also clears the statistics */
/* Sizes of the last pipeline run that rz_sync (or a blocking call) collected:
 * flattened lines, cell records and free-endpoint events. */
rz_status rz_last_counts(rz_ctx* ctx, size_t* n_lines, size_t* n_records, size_t* n_events);
rz_status rz_profile_stats(rz_ctx* ctx, rz_kernel_stat* out, size_t cap, size_t* n);

#ifdef __cplusplus
}
#endif

#endif /* RZ_H */

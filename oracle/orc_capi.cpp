// Oracle C API, loaded by tests/ (ctypes), __graft_entry__.smoke() and the
// CPU-baseline leg of bench.py. TEST INFRASTRUCTURE ONLY: the product library
// never links or calls anything here.
//
// The path argument has the same memory layout as the product's rz_path
// (include/rz.h) so the same host arrays feed both.
#include <atomic>
#include <chrono>
#include <cstring>
#include <thread>
#include "orc_dual.hpp"
#include "orc_exact.hpp"
#include "orc_svg.hpp"
#include "orc_paint.hpp"
#include "orc_stroke.hpp"

using namespace orc;

extern "C" {

typedef struct {
  uint32_t n_segments;
  uint32_t n_subpaths;
  const uint8_t* kinds;
  const uint32_t* point_offsets;
  const double* points;
  const uint32_t* subpaths;
  const uint8_t* closed;
} orc_path;

typedef struct {
  int32_t kind;
  float color[4];
  int32_t units;
  int32_t linear_colors;
  int32_t spread;
  double tr[6];
  double start[2];
  double end[2];
  double center[2];
  double radius;
  double fcenter[2];
  double fradius;
  uint32_t n_stops;
  const double* stop_positions;
  const float* stop_colors;
} orc_paint;

enum { ORC_OK = 0, ORC_E_NAN = 1, ORC_E_CAP = 2, ORC_E_ARG = 3 };
}

static Path to_path(const orc_path* p) {
  Path path;
  path.segments.resize(p->n_segments);
  for (uint32_t k = 0; k < p->n_segments; ++k) {
    Segment s;
    s.kind = (SegKind)p->kinds[k];
    uint32_t o = p->point_offsets[k];
    for (int i = 0; i < 4; ++i) s.p[i] = Point{0, 0};
    for (int i = 0; i < s.npts(); ++i) s.p[i] = Point{p->points[2 * (o + i)], p->points[2 * (o + i) + 1]};
    path.segments[k] = s;
  }
  if (p->n_subpaths > 0)
    for (uint32_t i = 0; i <= p->n_subpaths; ++i) path.subpaths.push_back(p->subpaths[i]);
  for (uint32_t i = 0; i < p->n_subpaths; ++i) path.closed.push_back(p->closed[i]);
  return path;
}
static Transform to_tr(const double* m) {
  Transform t;
  for (int i = 0; i < 6; ++i) t.m[i] = m[i];
  return t;
}
static Paint to_paint(const orc_paint* p) {
  if (p->kind == PAINT_SOLID)
    return Paint::solid(LinColor{f4(p->color[0], p->color[1], p->color[2], p->color[3])});
  std::vector<GradStop> stops;
  for (uint32_t i = 0; i < p->n_stops; ++i) {
    const float* c = p->stop_colors + 4 * i;
    stops.push_back(GradStop{p->stop_positions[i], LinColor{f4(c[0], c[1], c[2], c[3])}});
  }
  if (p->kind == PAINT_LINEAR)
    return Paint::linear(stops, (Units)p->units, p->linear_colors != 0, (Spread)p->spread,
                         to_tr(p->tr), P(p->start[0], p->start[1]), P(p->end[0], p->end[1]));
  return Paint::radial(stops, (Units)p->units, p->linear_colors != 0, (Spread)p->spread,
                       to_tr(p->tr), P(p->center[0], p->center[1]), p->radius,
                       P(p->fcenter[0], p->fcenter[1]), p->fradius);
}

// A Path into caller buffers (kinds, point offsets, x/y points, subpath
// offsets, closed flags); sizes always to *n_seg / *n_pts / *n_sub.
static int put_path(const Path& path, uint8_t* kinds, uint32_t* offs, double* pts, uint32_t* subs,
                    uint8_t* closed, size_t cap_seg, size_t cap_pts, size_t cap_sub, size_t* n_seg,
                    size_t* n_pts, size_t* n_sub) {
  size_t np = 0;
  for (const auto& sg : path.segments) np += (size_t)sg.npts();
  *n_seg = path.segments.size();
  *n_pts = np;
  *n_sub = path.nsubpaths();
  if (!kinds) return ORC_OK;
  if (cap_seg < *n_seg || cap_pts < np || cap_sub < path.subpaths.size()) return ORC_E_CAP;
  size_t o = 0;
  offs[0] = 0;
  for (size_t k = 0; k < path.segments.size(); ++k) {
    const Segment& sg = path.segments[k];
    kinds[k] = (uint8_t)sg.kind;
    for (int i = 0; i < sg.npts(); ++i, ++o) {
      pts[2 * o] = sg.p[i].x;
      pts[2 * o + 1] = sg.p[i].y;
    }
    offs[k + 1] = (uint32_t)o;
  }
  for (size_t i = 0; i < path.subpaths.size(); ++i) subs[i] = (uint32_t)path.subpaths[i];
  for (size_t i = 0; i < path.closed.size(); ++i) closed[i] = path.closed[i];
  return ORC_OK;
}

extern "C" {

// Path::flatten(tr, flatness, close): lines as (x0, y0, x1, y1) quadruples.
int orc_flatten(const orc_path* p, const double* tr, double flatness, int close, double* out,
                size_t cap, size_t* n_out) {
  Path path = to_path(p);
  size_t n = 0;
  try {
    flatten(path, to_tr(tr), flatness, close != 0, [&](const Line& l) {
      if (n < cap) {
        out[4 * n] = l.p0.x; out[4 * n + 1] = l.p0.y;
        out[4 * n + 2] = l.p1.x; out[4 * n + 3] = l.p1.y;
      }
      ++n;
    });
  } catch (const std::runtime_error&) {
    *n_out = n;
    return ORC_E_NAN;
  }
  *n_out = n;
  return n > cap ? ORC_E_CAP : ORC_OK;
}

// signed_difference_line on a zeroed-by-caller W x H image (row-major)
int orc_signed_difference_line(double* img, size_t W, size_t H, const double* line) {
  ImageF64 im{img, W * H, Shape::simple(H, W)};
  signed_difference_line(im, Line{P(line[0], line[1]), P(line[2], line[3])});
  return ORC_OK;
}

// SignedDifferenceRasterizer::mask into a strided view; `img` must hold the
// caller's (zeroed) image. rule < 0 keeps the raw accumulation (no fill rule).
int orc_sd_mask(const orc_path* p, const double* tr, double flatness, int rule, double* img,
                size_t len, size_t W, size_t H, size_t start, size_t row_stride, size_t col_stride) {
  Path path = to_path(p);
  ImageF64 im{img, len, Shape{start, W, H, row_stride, col_stride}};
  try {
    if (rule < 0) {
      flatten(path, to_tr(tr), flatness, true, [&](const Line& l) { signed_difference_line(im, l); });
      // prefix along rows only
      for (size_t y = 0; y < H; ++y) {
        double acc = 0.0;
        for (size_t x = 0; x < W; ++x) {
          acc += img[im.shape.offset(y, x)];
          img[im.shape.offset(y, x)] = acc;
        }
      }
    } else {
      sd_mask(path, to_tr(tr), flatness, (FillRule)rule, im);
    }
  } catch (const std::runtime_error&) {
    return ORC_E_NAN;
  }
  return ORC_OK;
}

// SignedDifferenceRasterizer::mask_iter rendered densely: out[y*W+x] = alpha
// for produced pixels, 0 elsewhere.
int orc_sd_mask_iter(const orc_path* p, const double* tr, double flatness, int rule, size_t W,
                     size_t H, double* out) {
  Path path = to_path(p);
  std::memset(out, 0, sizeof(double) * W * H);
  try {
    sd_mask_iter(path, to_tr(tr), flatness, (FillRule)rule, W, H,
                 [&](const Pixel& px) { out[px.y * W + px.x] = px.alpha; });
  } catch (const std::runtime_error&) {
    return ORC_E_NAN;
  }
  return ORC_OK;
}

// ActiveEdgeRasterizer::mask on a zeroed W x H image
int orc_ae_mask(const orc_path* p, const double* tr, double flatness, int rule, size_t W, size_t H,
                double* out) {
  Path path = to_path(p);
  ImageF64 im{out, W * H, Shape::simple(H, W)};
  try {
    ae_mask(path, to_tr(tr), flatness, (FillRule)rule, im);
  } catch (const std::runtime_error&) {
    return ORC_E_NAN;
  }
  return ORC_OK;
}

// ActiveEdgeIter over explicit lines (used by the edge-iterator tests)
int orc_ae_lines(const double* lines, size_t n, size_t W, size_t H, int rule, double* out) {
  std::vector<Line> ls;
  for (size_t i = 0; i < n; ++i)
    ls.push_back(Line{P(lines[4 * i], lines[4 * i + 1]), P(lines[4 * i + 2], lines[4 * i + 3])});
  std::memset(out, 0, sizeof(double) * W * H);
  ae_iter(W, H, (FillRule)rule, ls, [&](const Pixel& px) { out[px.y * W + px.x] = px.alpha; });
  return ORC_OK;
}

// Exact winding-weighted area per pixel of the flattened (closed) path.
int orc_exact_winding_area(const orc_path* p, const double* tr, double flatness, size_t W,
                           size_t H, double* out) {
  Path path = to_path(p);
  std::vector<std::vector<Point>> rings;
  Transform t = to_tr(tr);
  try {
    for (size_t i = 0; i < path.nsubpaths(); ++i) {
      Path sub;
      sub.segments.assign(path.segments.begin() + path.subpaths[i],
                          path.segments.begin() + path.subpaths[i + 1]);
      sub.subpaths = {0, sub.segments.size()};
      sub.closed = {1};
      std::vector<Point> ring;
      flatten(sub, t, flatness, true, [&](const Line& l) { ring.push_back(l.p0); });
      rings.push_back(ring);
    }
  } catch (const std::runtime_error&) {
    return ORC_E_NAN;
  }
  exact_winding_area(rings, W, H, out);
  return ORC_OK;
}

// Exact coverage of {rule(w) holds} per pixel for the flattened path.
int orc_exact_fill_coverage(const orc_path* p, const double* tr, double flatness, int rule,
                            size_t W, size_t H, double* out) {
  Path path = to_path(p);
  std::vector<Line> lines;
  try {
    flatten(path, to_tr(tr), flatness, true, [&](const Line& l) { lines.push_back(l); });
  } catch (const std::runtime_error&) {
    return ORC_E_NAN;
  }
  exact_fill_coverage(lines, W, H, rule, out);
  return ORC_OK;
}

// Rasterizer::fill (signed difference) into a strided LinColor view.
int orc_fill(const orc_path* p, const double* tr, double flatness, int rule, const orc_paint* paint,
             float* rgba, size_t len_px, size_t W, size_t H, size_t start, size_t row_stride) {
  Path path = to_path(p);
  Paint pt = to_paint(paint);
  ImageRGBA im{reinterpret_cast<F4*>(rgba), len_px, Shape{start, W, H, row_stride, 1}};
  try {
    sd_fill(path, to_tr(tr), flatness, (FillRule)rule, pt, im);
  } catch (const std::runtime_error&) {
    return ORC_E_NAN;
  }
  return ORC_OK;
}

// Scene::group(fills).render(SignedDifference, tr, view, bg). Writes the layer
// origin and size to layer_xywh and, if it fits in cap_px, the pixels.
int orc_scene_render(size_t n, const orc_path* paths, const orc_paint* paints, const int32_t* rules,
                     const double* tr, const double* view /*nullable x0,y0,x1,y1*/,
                     const float* bg /*nullable*/, double flatness, float* out, size_t cap_px,
                     int64_t* layer_xywh) {
  std::vector<SceneRef> fills;
  for (size_t i = 0; i < n; ++i) {
    auto node = std::make_shared<SceneNode>();
    node->kind = SceneNode::FILL;
    node->path = std::make_shared<Path>(to_path(&paths[i]));
    node->paint = std::make_shared<Paint>(to_paint(&paints[i]));
    node->rule = (FillRule)rules[i];
    fills.push_back(node);
  }
  SceneRef root;
  if (fills.size() == 1) {
    root = fills[0];
  } else {
    root = std::make_shared<SceneNode>();
    root->kind = SceneNode::GROUP;
    root->children = fills;
  }
  BBox vb;
  if (view) vb = BBox::make(P(view[0], view[1]), P(view[2], view[3]));
  LinColor bgc;
  if (bg) bgc = LinColor{f4(bg[0], bg[1], bg[2], bg[3])};
  Layer layer;
  try {
    layer = scene_render(root.get(), to_tr(tr), view ? &vb : nullptr, bg ? &bgc : nullptr, flatness);
  } catch (const std::runtime_error&) {
    return ORC_E_NAN;
  }
  layer_xywh[0] = layer.x;
  layer_xywh[1] = layer.y;
  layer_xywh[2] = (int64_t)layer.shape.width;
  layer_xywh[3] = (int64_t)layer.shape.height;
  if (layer.image.size() > cap_px) return ORC_E_CAP;
  std::memcpy(out, layer.image.data(), layer.image.size() * sizeof(F4));
  return ORC_OK;
}

// orc_scene_render on `threads` host threads (bands of layer rows; the same
// pixels and values). Used for the full-size scene checks and as the
// multi-core CPU baseline (an upper bound: the reference is single-threaded).
int orc_scene_render_mt(size_t n, const orc_path* paths, const orc_paint* paints,
                        const int32_t* rules, const double* tr, const double* view,
                        const float* bg, double flatness, float* out, size_t cap_px,
                        int64_t* layer_xywh, int threads) {
  std::vector<SceneRef> fills;
  for (size_t i = 0; i < n; ++i) {
    auto node = std::make_shared<SceneNode>();
    node->kind = SceneNode::FILL;
    node->path = std::make_shared<Path>(to_path(&paths[i]));
    node->paint = std::make_shared<Paint>(to_paint(&paints[i]));
    node->rule = (FillRule)rules[i];
    fills.push_back(node);
  }
  auto root = std::make_shared<SceneNode>();
  root->kind = SceneNode::GROUP;
  root->children = fills;
  BBox vb;
  if (view) vb = BBox::make(P(view[0], view[1]), P(view[2], view[3]));
  LinColor bgc;
  if (bg) bgc = LinColor{f4(bg[0], bg[1], bg[2], bg[3])};
  Layer layer;
  try {
    layer = scene_render_mt(root.get(), to_tr(tr), view ? &vb : nullptr, bg ? &bgc : nullptr,
                            flatness, threads);
  } catch (const std::runtime_error&) {
    return ORC_E_NAN;
  }
  layer_xywh[0] = layer.x;
  layer_xywh[1] = layer.y;
  layer_xywh[2] = (int64_t)layer.shape.width;
  layer_xywh[3] = (int64_t)layer.shape.height;
  if (layer.image.size() > cap_px) return ORC_E_CAP;
  std::memcpy(out, layer.image.data(), layer.image.size() * sizeof(F4));
  return ORC_OK;
}

// A glyph atlas: for each glyph i, SignedDifferenceRasterizer::mask of
// paths[i] under tr[6i..6i+6] into a zeroed w x h image, copied into the
// atlas (row-major f32, atlas_w wide) at (x, y) = xywh[4i..4i+4]. Glyphs are
// shared out over `threads` host threads. Pixels outside the views are left.
int orc_mask_atlas_mt(size_t n, const orc_path* paths, const double* tr, const uint32_t* xywh,
                      double flatness, int rule, float* atlas, size_t atlas_w, size_t atlas_h,
                      int threads) {
  std::atomic<size_t> next{0};
  std::atomic<int> rc{ORC_OK};
  auto work = [&]() {
    std::vector<double> img;
    for (size_t i; (i = next.fetch_add(1)) < n;) {
      const uint32_t* q = xywh + 4 * i;
      const size_t x = q[0], y = q[1], w = q[2], h = q[3];
      if (!w || !h) continue;
      if (x + w > atlas_w || y + h > atlas_h) {
        rc = ORC_E_ARG;
        continue;
      }
      img.assign(w * h, 0.0);
      try {
        Path path = to_path(&paths[i]);
        ImageF64 im{img.data(), img.size(), Shape::simple(h, w)};
        sd_mask(path, to_tr(tr + 6 * i), flatness, (FillRule)rule, im);
      } catch (const std::runtime_error&) {
        rc = ORC_E_NAN;
        continue;
      }
      for (size_t r = 0; r < h; ++r)
        for (size_t c = 0; c < w; ++c) atlas[(y + r) * atlas_w + x + c] = (float)img[r * w + c];
    }
  };
  std::vector<std::thread> pool;
  for (int k = 1; k < threads; ++k) pool.emplace_back(work);
  work();
  for (auto& t : pool) t.join();
  return rc;
}

// SignedDifferenceRasterizer::mask_iter as a dense W x H image (pixels it
// does not yield are 0) and its derivative along `tan` (per stored point of
// the path: dx, dy), decisions fixed at the values (see orc_dual.hpp).
int orc_sd_mask_iter_dual(const orc_path* p, const double* tan, const double* tr,
                          double flatness, int rule, size_t W, size_t H, double* alpha,
                          double* dalpha) {
  Path path = to_path(p);
  try {
    sd_mask_iter_dual(path, tan, to_tr(tr), flatness, (FillRule)rule, W, H, alpha, dalpha);
  } catch (const std::runtime_error&) {
    return ORC_E_NAN;
  }
  return ORC_OK;
}

// Path::read_svg_path (src/path.rs:528-534) of `text`: the path's arrays
// (layout of orc_path) into caller buffers of the given capacities; the sizes
// always go to *n_seg / *n_pts / *n_sub (subs holds n_sub + 1 offsets).
// Returns ORC_E_ARG on a syntax error
// (offset in *n_seg), ORC_E_CAP when a buffer is too small.
int orc_svg_parse(const char* text, size_t len, uint8_t* kinds, uint32_t* offs, double* pts,
                  uint32_t* subs, uint8_t* closed, size_t cap_seg, size_t cap_pts,
                  size_t cap_sub, size_t* n_seg, size_t* n_pts, size_t* n_sub) {
  Path path;
  try {
    path = read_svg_path(std::string(text, len));
  } catch (const SvgError& e) {
    *n_seg = e.offset;
    return ORC_E_ARG;
  }
  return put_path(path, kinds, offs, pts, subs, closed, cap_seg, cap_pts, cap_sub, n_seg, n_pts,
                  n_sub);
}

// Path::stroke (src/path.rs:370-415) of the path, written like orc_svg_parse's
// output. join: 0 miter (miter_limit), 1 bevel, 2 round; cap: 0 butt, 1 square,
// 2 round.
int orc_path_stroke(const orc_path* p, double width, int join, double miter_limit, int cap,
                    uint8_t* kinds, uint32_t* offs, double* pts, uint32_t* subs, uint8_t* closed,
                    size_t cap_seg, size_t cap_pts, size_t cap_sub, size_t* n_seg, size_t* n_pts,
                    size_t* n_sub) {
  StrokeStyle st;
  st.width = width;
  st.line_join = (LineJoin)join;
  st.miter_limit = miter_limit;
  st.line_cap = (LineCap)cap;
  const Path out = stroke(to_path(p), st);
  return put_path(out, kinds, offs, pts, subs, closed, cap_seg, cap_pts, cap_sub, n_seg, n_pts,
                  n_sub);
}

// Time Path::stroke: seconds per call over `reps` calls.
double orc_time_stroke(const orc_path* p, double width, int join, double miter_limit, int cap,
                       int reps, size_t* n_seg) {
  StrokeStyle st;
  st.width = width;
  st.line_join = (LineJoin)join;
  st.miter_limit = miter_limit;
  st.line_cap = (LineCap)cap;
  const Path in = to_path(p);
  size_t n = 0;
  const auto t0 = std::chrono::steady_clock::now();
  for (int r = 0; r < reps; ++r) n = stroke(in, st).segments.size();
  const auto t1 = std::chrono::steady_clock::now();
  if (n_seg) *n_seg = n;
  return std::chrono::duration<double>(t1 - t0).count() / (reps > 0 ? reps : 1);
}

// FromStr for Transform (src/svg.rs:592-601); ORC_E_ARG on a syntax error.
int orc_svg_transform(const char* text, size_t len, double* out6) {
  try {
    const Transform t = read_svg_transform(std::string(text, len));
    for (int i = 0; i < 6; ++i) out6[i] = t.m[i];
  } catch (const SvgError&) {
    return ORC_E_ARG;
  }
  return ORC_OK;
}

// Scene::render of a scene tree given as rz_scene_node's layout (kind,
// fill_rule, units, path, paint, child, first, count, tr[6], opacity,
// stroke_width, miter_limit, line_join, line_cap).
typedef struct {
  int32_t kind, fill_rule, units;
  uint32_t path, paint, child, first, count;
  double tr[6];
  double opacity;
  double stroke_width, miter_limit;
  int32_t line_join, line_cap;
} orc_scene_node;

static SceneRef tree_node(const orc_scene_node* nodes, const uint32_t* children, uint32_t id,
                          std::vector<std::shared_ptr<Path>>& paths,
                          std::vector<std::shared_ptr<Paint>>& paints, int depth) {
  if (depth > 256) throw std::runtime_error("scene tree too deep");
  const orc_scene_node& n = nodes[id];
  auto s = std::make_shared<SceneNode>();
  switch (n.kind) {
    case 0:
      s->kind = SceneNode::FILL;
      s->path = paths[n.path];
      s->paint = paints[n.paint];
      s->rule = (FillRule)n.fill_rule;
      break;
    case 1:
      s->kind = SceneNode::GROUP;
      for (uint32_t k = 0; k < n.count; ++k)
        s->children.push_back(tree_node(nodes, children, children[n.first + k], paths, paints, depth + 1));
      break;
    case 2:
      s->kind = SceneNode::TRANSFORM;
      s->tr = to_tr(n.tr);
      s->child = tree_node(nodes, children, n.child, paths, paints, depth + 1);
      break;
    case 3:
      s->kind = SceneNode::OPACITY;
      s->opacity = n.opacity;  // as given (SceneInner::Opacity); Scene::opacity clamps
      s->child = tree_node(nodes, children, n.child, paths, paints, depth + 1);
      break;
    case 5: {
      s->kind = SceneNode::STROKE;
      s->path = paths[n.path];
      StrokeStyle st;
      st.width = n.stroke_width;
      st.miter_limit = n.miter_limit;
      st.line_join = (LineJoin)n.line_join;
      st.line_cap = (LineCap)n.line_cap;
      auto f = std::make_shared<SceneNode>();
      f->kind = SceneNode::FILL;
      f->path = std::make_shared<Path>(stroke(*s->path, st));
      f->paint = paints[n.paint];
      f->rule = NONZERO;
      s->child = f;
      break;
    }
    default:
      s->kind = SceneNode::CLIP;
      s->path = paths[n.path];
      s->rule = (FillRule)n.fill_rule;
      s->units = (Units)n.units;
      s->child = tree_node(nodes, children, n.child, paths, paints, depth + 1);
      break;
  }
  return s;
}

int orc_scene_tree_render(const orc_path* paths, size_t n_paths, const orc_paint* paints,
                          size_t n_paints, const orc_scene_node* nodes, const uint32_t* children,
                          uint32_t root, const double* tr, const double* view, const float* bg,
                          double flatness, int active_edge, float* out, size_t cap_px,
                          int64_t* layer_xywh) {
  std::vector<std::shared_ptr<Path>> ps;
  std::vector<std::shared_ptr<Paint>> pt;
  for (size_t i = 0; i < n_paths; ++i) ps.push_back(std::make_shared<Path>(to_path(&paths[i])));
  for (size_t i = 0; i < n_paints; ++i) pt.push_back(std::make_shared<Paint>(to_paint(&paints[i])));
  BBox vb;
  if (view) vb = BBox::make(P(view[0], view[1]), P(view[2], view[3]));
  LinColor bgc;
  if (bg) bgc = LinColor{f4(bg[0], bg[1], bg[2], bg[3])};
  Layer layer;
  try {
    SceneRef root_node = tree_node(nodes, children, root, ps, pt, 0);
    layer = scene_render(root_node.get(), to_tr(tr), view ? &vb : nullptr, bg ? &bgc : nullptr,
                         flatness, active_edge != 0);
  } catch (const std::runtime_error&) {
    return ORC_E_NAN;
  }
  layer_xywh[0] = layer.x;
  layer_xywh[1] = layer.y;
  layer_xywh[2] = (int64_t)layer.shape.width;
  layer_xywh[3] = (int64_t)layer.shape.height;
  if (layer.image.size() > cap_px) return ORC_E_CAP;
  std::memcpy(out, layer.image.data(), layer.image.size() * sizeof(F4));
  return ORC_OK;
}

// Paint::at for a single point (gradient and colour known-answer tests)
int orc_paint_at(const orc_paint* paint, double x, double y, float* out4) {
  Paint pt = to_paint(paint);
  LinColor c = pt.at(P(x, y));
  for (int i = 0; i < 4; ++i) out4[i] = c.c.v[i];
  return ORC_OK;
}
int orc_radial_offset(const orc_paint* paint, double x, double y, double* out) {
  Paint pt = to_paint(paint);
  return pt.radial_offset(P(x, y), out) ? ORC_OK : ORC_E_ARG;
}
int orc_l2s(const float* in4, float* out4) {
  F4 r = l2s(f4(in4[0], in4[1], in4[2], in4[3]));
  std::memcpy(out4, r.v, sizeof(r.v));
  return ORC_OK;
}
int orc_s2l(const float* in4, float* out4) {
  F4 r = s2l(f4(in4[0], in4[1], in4[2], in4[3]));
  std::memcpy(out4, r.v, sizeof(r.v));
  return ORC_OK;
}
float orc_linear_to_srgb(float v) { return linear_to_srgb(v); }
float orc_srgb_to_linear(float v) { return srgb_to_linear(v); }
int orc_rgba_to_lin(const uint8_t* rgba, float* out4) {
  LinColor c = LinColor::from_rgba(rgba[0], rgba[1], rgba[2], rgba[3]);
  std::memcpy(out4, c.c.v, sizeof(c.c.v));
  return ORC_OK;
}
int orc_lin_to_rgba(const float* in4, uint8_t* out) {
  lin_to_rgba(LinColor{f4(in4[0], in4[1], in4[2], in4[3])}, out);
  return ORC_OK;
}
double orc_spread_at(int spread, double t) { return spread_at((Spread)spread, t); }
int orc_quadratic_solve(double a, double b, double c, double* out2) {
  return quadratic_solve(a, b, c, out2);
}
double orc_alpha_from_winding(int rule, double w) { return alpha_from_winding((FillRule)rule, w); }

// Segment split helpers for the split-equality known-answer test.
// kind: 1 quad, 2 cubic; pts: npts*2 doubles; out: 2 * npts * 2 doubles
int orc_split(int kind, const double* pts, int fast, double* out) {
  Segment s;
  s.kind = (SegKind)kind;
  for (int i = 0; i < 4; ++i) s.p[i] = Point{0, 0};
  for (int i = 0; i < s.npts(); ++i) s.p[i] = P(pts[2 * i], pts[2 * i + 1]);
  Segment a, b;
  if (fast) s.split_fast(&a, &b); else s.split(&a, &b);
  int n = s.npts();
  for (int i = 0; i < n; ++i) {
    out[2 * i] = a.p[i].x; out[2 * i + 1] = a.p[i].y;
    out[2 * (n + i)] = b.p[i].x; out[2 * (n + i) + 1] = b.p[i].y;
  }
  return ORC_OK;
}
double orc_flatness(int kind, const double* pts) {
  Segment s;
  s.kind = (SegKind)kind;
  for (int i = 0; i < 4; ++i) s.p[i] = Point{0, 0};
  for (int i = 0; i < s.npts(); ++i) s.p[i] = P(pts[2 * i], pts[2 * i + 1]);
  return s.flatness();
}

// PathBuilder::arc_to from `src`: writes up to 4 cubics (8 doubles each).
// Returns the number of cubics, or -1 when the arc degenerates into line_to.
int orc_arc_to(const double* src, const double* radii, double x_axis_rot, int large, int sweep,
               const double* dst, double* out) {
  EllipArc arc;
  if (!EllipArc::new_param(P(src[0], src[1]), P(dst[0], dst[1]), radii[0], radii[1], x_axis_rot,
                           large != 0, sweep != 0, &arc))
    return -1;
  std::vector<Segment> cs;
  arc.to_cubics(cs);
  for (size_t i = 0; i < cs.size(); ++i)
    for (int k = 0; k < 4; ++k) {
      out[8 * i + 2 * k] = cs[i].p[k].x;
      out[8 * i + 2 * k + 1] = cs[i].p[k].y;
    }
  return (int)cs.size();
}

// Path::bbox(tr): writes x0, y0, x1, y1; returns ORC_E_ARG for an empty path.
int orc_path_bbox(const orc_path* p, const double* tr, double* out4) {
  Path path = to_path(p);
  BBox b;
  if (!path.bbox(to_tr(tr), &b)) return ORC_E_ARG;
  out4[0] = b.min.x; out4[1] = b.min.y; out4[2] = b.max.x; out4[3] = b.max.y;
  return ORC_OK;
}

// CPU baseline timing: SignedDifferenceRasterizer::mask of the same path
// `reps` times on `threads` host threads (each thread its own image), returns
// the wall time in seconds.
double orc_time_sd_mask(const orc_path* p, const double* tr, double flatness, int rule, size_t W,
                        size_t H, int reps, int threads) {
  Path path = to_path(p);
  Transform t = to_tr(tr);
  if (threads < 1) threads = 1;
  auto t0 = std::chrono::steady_clock::now();
  std::vector<std::thread> pool;
  for (int k = 0; k < threads; ++k) {
    pool.emplace_back([&, k]() {
      std::vector<double> img(W * H);
      for (int r = k; r < reps; r += threads) {
        std::fill(img.begin(), img.end(), 0.0);
        ImageF64 im{img.data(), img.size(), Shape::simple(H, W)};
        sd_mask(path, t, flatness, (FillRule)rule, im);
      }
    });
  }
  for (auto& th : pool) th.join();
  auto t1 = std::chrono::steady_clock::now();
  return std::chrono::duration<double>(t1 - t0).count();
}

// CPU baseline timing of Path::flatten(...).count() (the reference's
// material-big/flatten bench shape).
double orc_time_flatten(const orc_path* p, const double* tr, double flatness, int reps,
                        size_t* n_lines) {
  Path path = to_path(p);
  Transform t = to_tr(tr);
  size_t n = 0;
  auto t0 = std::chrono::steady_clock::now();
  for (int r = 0; r < reps; ++r) {
    n = 0;
    flatten(path, t, flatness, true, [&](const Line&) { ++n; });
  }
  auto t1 = std::chrono::steady_clock::now();
  *n_lines = n;
  return std::chrono::duration<double>(t1 - t0).count();
}

}  // extern "C"

#include "orc_bin_replay.hpp"

extern "C" {
// Sequential replay of the binning stage (see orc_bin_replay.hpp). Records are
// tile-major; within a tile, and for events, the order is the replay's own.
int orc_bin_replay(const orc_path* p, const double* tr, double flatness, int mode, size_t W,
                   size_t H, uint32_t* counts, float* recs, size_t rec_cap, int32_t* bd,
                   float* events, size_t ev_cap, int32_t* redge, size_t* n_recs,
                   size_t* n_events, int phx, int phy) {
  Path path = to_path(p);
  BinReplay r;
  try {
    bin_replay(path, to_tr(tr), flatness, mode, (int)W, (int)H, &r, phx, phy);
  } catch (const std::runtime_error&) {
    return ORC_E_NAN;
  }
  size_t nr = 0;
  for (auto& v : r.recs) nr += v.size();
  *n_recs = nr;
  *n_events = r.events.size();
  if (!counts) return ORC_OK;
  if (nr > rec_cap || r.events.size() > ev_cap) return ORC_E_CAP;
  size_t k = 0;
  for (size_t t = 0; t < r.recs.size(); ++t) {
    counts[t] = (uint32_t)r.recs[t].size();
    for (const auto& q : r.recs[t]) {
      recs[4 * k] = q.x0; recs[4 * k + 1] = q.y0; recs[4 * k + 2] = q.x1; recs[4 * k + 3] = q.y1;
      ++k;
    }
    bd[t] = r.bd[t];
  }
  if (redge)  // int32 out, clamped (as the product's rz_bin reports it)
    for (size_t y = 0; y < H; ++y)
      redge[y] = (int32_t)std::max<int64_t>(INT32_MIN, std::min<int64_t>(INT32_MAX, r.redge[y]));
  for (size_t i = 0; i < r.events.size(); ++i) {
    const auto& e = r.events[i];
    std::memcpy(&events[4 * i], &e.band, 4);
    std::memcpy(&events[4 * i + 1], &e.col, 4);
    events[4 * i + 2] = e.yrel;
    events[4 * i + 3] = e.sign;
  }
  return ORC_OK;
}
}  // extern "C"

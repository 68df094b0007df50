// Oracle: CPU restatement of the reference crate's colour, gradient, fill and
// scene code (src/color.rs, src/grad.rs, src/simd/x86.rs, src/rasterize.rs:69-115,
// src/scene.rs).
//
// TEST INFRASTRUCTURE ONLY (see orc_core.hpp). The f32x4 arithmetic follows the
// default x86+simd build (features ["serde","simd"]): l2s/s2l are the SSE
// polynomials applied to all four lanes, alpha included.
#pragma once
#include <atomic>
#include <thread>
#include <memory>
#include "orc_raster.hpp"

namespace orc {

struct F4 {
  float v[4];
};
inline F4 f4(float a, float b, float c, float d) { return F4{{a, b, c, d}}; }
inline F4 splat(float a) { return F4{{a, a, a, a}}; }
inline F4 operator+(F4 a, F4 b) { return f4(a.v[0] + b.v[0], a.v[1] + b.v[1], a.v[2] + b.v[2], a.v[3] + b.v[3]); }
inline F4 operator-(F4 a, F4 b) { return f4(a.v[0] - b.v[0], a.v[1] - b.v[1], a.v[2] - b.v[2], a.v[3] - b.v[3]); }
inline F4 operator*(F4 a, F4 b) { return f4(a.v[0] * b.v[0], a.v[1] * b.v[1], a.v[2] * b.v[2], a.v[3] * b.v[3]); }
inline F4 operator/(F4 a, F4 b) { return f4(a.v[0] / b.v[0], a.v[1] / b.v[1], a.v[2] / b.v[2], a.v[3] / b.v[3]); }
inline F4 operator*(F4 a, float s) { return a * splat(s); }
inline F4 operator*(float s, F4 a) { return a * splat(s); }
inline F4 f4sqrt(F4 a) {
  return f4(std::sqrt(a.v[0]), std::sqrt(a.v[1]), std::sqrt(a.v[2]), std::sqrt(a.v[3]));
}

// src/simd/x86.rs:196-215
inline F4 l2s(F4 x0) {
  F4 x1 = f4sqrt(x0);
  F4 x2 = f4sqrt(x1);
  F4 x3 = f4sqrt(x2);
  F4 high = -0.01848558f * x0 + 0.6445592f * x1 + 0.70994765f * x2 - 0.33605254f * x3;
  F4 low = x0 * 12.92f;
  F4 r;
  for (int i = 0; i < 4; ++i) r.v[i] = (x0.v[i] <= 0.0031308f) ? low.v[i] : high.v[i];
  return r;
}
// src/simd/x86.rs:217-244
inline F4 s2l(F4 vs) {
  F4 x1 = 2.0843103538116825f * vs - splat(1.0843103538116827f);
  F4 x2 = x1 * x1;
  F4 x3 = x2 * x1;
  F4 high = splat(0.23361048543711943f) + 0.4665843122387033f * x1 + 0.26901741378006355f * x2 +
            0.031661580753065945f * x3;
  F4 low = vs * 0.07739938080495357f;
  F4 r;
  for (int i = 0; i < 4; ++i) r.v[i] = (vs.v[i] <= 0.04045f) ? low.v[i] : high.v[i];
  return r;
}

// src/color.rs:455-486 (scalar, exact powf)
inline float linear_to_srgb(float x0) {
  if (x0 <= 0.0031308f) return x0 * 12.92f;
  float x1 = std::sqrt(x0), x2 = std::sqrt(x1), x3 = std::sqrt(x2);
  return -0.01848558f * x0 + 0.6445592f * x1 + 0.70994765f * x2 - 0.33605254f * x3;
}
inline float srgb_to_linear(float v) {
  if (v <= 0.04045f) return v / 12.92f;
  return std::pow((v + 0.055f) / 1.055f, 2.4f);
}

// LinColor: premultiplied linear RGBA (src/color.rs:271-355)
struct LinColor {
  F4 c;
  float alpha() const { return c.v[3]; }
  // :303-311
  F4 unmultiply() const {
    float a = alpha();
    if (a <= 1e-6f) return splat(0.0f);
    return c / splat(a);
  }
  LinColor into_srgb() const { return LinColor{l2s(unmultiply()) * splat(alpha())}; }
  LinColor into_linear() const { return LinColor{s2l(unmultiply()) * splat(alpha())}; }
  // :342-344: other + self * (1 - other.alpha)
  LinColor blend_over(LinColor other) const {
    return LinColor{other.c + c * (1.0f - other.alpha())};
  }
  LinColor with_alpha(double a) const { return LinColor{c * (float)a}; }
  LinColor lerp(LinColor other, float t) const { return LinColor{other.c * t + c * (1.0f - t)}; }
  // From<RGBA> (:394-406)
  static LinColor from_rgba(uint8_t r, uint8_t g, uint8_t b, uint8_t a8) {
    float a = (float)a8 / 255.0f;
    return LinColor{f4(srgb_to_linear((float)r / 255.0f) * a, srgb_to_linear((float)g / 255.0f) * a,
                       srgb_to_linear((float)b / 255.0f) * a, a)};
  }
};

// Rust `f as u8` (saturating, NaN -> 0)
inline uint8_t sat_u8(float f) {
  if (!(f > 0.0f)) return 0;
  if (f >= 255.0f) return 255;
  return (uint8_t)f;
}
// From<LinColor> for RGBA (src/color.rs:164-175)
inline void lin_to_rgba(LinColor lin, uint8_t out[4]) {
  F4 s = l2s(lin.unmultiply());
  out[0] = sat_u8(s.v[0] * 255.0f + 0.5f);
  out[1] = sat_u8(s.v[1] * 255.0f + 0.5f);
  out[2] = sat_u8(s.v[2] * 255.0f + 0.5f);
  out[3] = sat_u8(lin.alpha() * 255.0f + 0.5f);
}

// ---------------------------------------------------------------------------
// Gradients (src/grad.rs)
// ---------------------------------------------------------------------------
enum Spread { PAD = 0, REPEAT = 1, REFLECT = 2 };
// :24-30
inline double spread_at(Spread s, double t) {
  if (s == PAD) return t;
  if (s == REPEAT) return rem_euclid(t, 1.0);
  return std::fabs(rem_euclid(t + 1.0, 2.0) - 1.0);
}

struct GradStop {
  double position;
  LinColor color;
};

// GradStops (:81-139); `stops` already sorted by position
struct GradStops {
  std::vector<GradStop> stops;
  static GradStops make(std::vector<GradStop> s) {
    std::stable_sort(s.begin(), s.end(), [](const GradStop& a, const GradStop& b) {
      return a.position < b.position;
    });
    if (s.empty()) s.push_back(GradStop{0.0, LinColor{f4(0, 0, 0, 1)}});
    return GradStops{s};
  }
  void convert_to_srgb() {
    for (auto& st : stops) st.color = st.color.into_srgb();
  }
  LinColor at(double t) const {
    // binary_search_by with a comparator that never reports Equal returns the
    // partition point: the number of stops with position < t.
    size_t index = 0;
    while (index < stops.size() && stops[index].position < t) ++index;
    size_t size = stops.size();
    if (index == 0) return stops[0].color;
    if (index == size) return stops[size - 1].color;
    const GradStop& p0 = stops[index - 1];
    const GradStop& p1 = stops[index];
    double ratio = (t - p0.position) / (p1.position - p0.position);
    return p0.color.lerp(p1.color, (float)ratio);
  }
};

enum Units { USER_SPACE_ON_USE = 0, BOUNDING_BOX = 1 };
enum PaintKind { PAINT_SOLID = 0, PAINT_LINEAR = 1, PAINT_RADIAL = 2 };

struct Paint {
  PaintKind kind = PAINT_SOLID;
  LinColor color{};  // solid
  Units units = USER_SPACE_ON_USE;
  bool linear_colors = false;
  Spread spread = PAD;
  Transform tr = Transform::identity();
  Point start{}, end{}, dir{};  // linear
  Point center{}, fcenter{};    // radial
  double radius = 0.0, fradius = 0.0;
  GradStops stops;

  static Paint solid(LinColor c) {
    Paint p;
    p.kind = PAINT_SOLID;
    p.color = c;
    return p;
  }
  // GradLinear::new (:163-194)
  static Paint linear(std::vector<GradStop> stops, Units units, bool linear_colors, Spread spread,
                      Transform tr, Point start, Point end) {
    Paint p;
    p.kind = PAINT_LINEAR;
    p.stops = GradStops::make(stops);
    if (!linear_colors) p.stops.convert_to_srgb();
    p.units = units;
    p.linear_colors = linear_colors;
    p.spread = spread;
    p.tr = tr;
    p.start = start;
    p.end = end;
    Point d = end - start;
    p.dir = d / dot(d, d);
    return p;
  }
  // GradRadial::new (:321-348)
  static Paint radial(std::vector<GradStop> stops, Units units, bool linear_colors, Spread spread,
                      Transform tr, Point center, double radius, Point fcenter, double fradius) {
    Paint p;
    p.kind = PAINT_RADIAL;
    p.stops = GradStops::make(stops);
    if (!linear_colors) p.stops.convert_to_srgb();
    p.units = units;
    p.linear_colors = linear_colors;
    p.spread = spread;
    p.tr = tr;
    p.center = center;
    p.radius = radius;
    p.fcenter = fcenter;
    p.fradius = fradius;
    return p;
  }
  // GradRadial::offset (:361-411)
  bool radial_offset(Point point, double* out) const {
    Point cd = center - fcenter;
    Point pd = point - fcenter;
    double rd = radius - fradius;
    double a = dot(cd, cd) - rd * rd;
    double b = -2.0 * (dot(cd, pd) + fradius * rd);
    double c = dot(pd, pd) - fradius * fradius;
    double r[2];
    int n = quadratic_solve(a, b, c, r);
    if (n == 2) { *out = std::fmax(r[0], r[1]); return true; }
    if (n == 1) { *out = r[0]; return true; }
    return false;
  }
  // Paint::at
  LinColor at(Point point) const {
    if (kind == PAINT_SOLID) return color;
    LinColor c;
    if (kind == PAINT_LINEAR) {
      double t = dot(point - start, dir);
      c = stops.at(spread_at(spread, t));
    } else {
      double off;
      if (!radial_offset(point, &off)) return LinColor{f4(0, 0, 0, 0)};
      c = stops.at(spread_at(spread, off));
    }
    return linear_colors ? c : c.into_linear();
  }
};

struct ImageRGBA {
  F4* data;  // LinColor pixels
  size_t len;
  Shape shape;
};

// Rasterizer::fill default + fill_impl (src/rasterize.rs:69-115) with the
// signed-difference mask_iter, or with ActiveEdgeRasterizer's (active_edge:
// ActiveEdgeIter over the flattened lines, src/rasterize.rs:558-576).
// Rows [r0, r1) of the image only (all of it by default), through
// sd_mask_iter_rows: the same pixels and values (signed difference only).
inline void sd_fill(const Path& path, const Transform& tr, double flatness, FillRule rule,
                    const Paint& paint, ImageRGBA img, size_t r0 = 0, size_t r1 = SIZE_MAX,
                    bool active_edge = false) {
  auto blend = [&](const Pixel& px, LinColor color) {
    F4& dst = img.data[img.shape.offset(px.y, px.x)];
    LinColor d{dst};
    dst = d.blend_over(color.with_alpha(px.alpha)).c;
  };
  const bool whole = r0 == 0 && r1 >= img.shape.height;
  auto mask_iter = [&](auto&& sink) {
    if (active_edge) {
      std::vector<Line> lines;
      flatten(path, tr, flatness, true, [&](const Line& l) { lines.push_back(l); });
      ae_iter(img.shape.width, img.shape.height, rule, lines, sink);
    } else if (whole)
      sd_mask_iter(path, tr, flatness, rule, img.shape.width, img.shape.height, sink);
    else
      sd_mask_iter_rows(path, tr, flatness, rule, img.shape.width, img.shape.height, r0, r1, sink);
  };
  if (paint.kind == PAINT_SOLID) {
    LinColor color = paint.at(P(0.0, 0.0));
    mask_iter([&](const Pixel& px) { blend(px, color); });
    return;
  }
  Transform units_tr;
  if (paint.units == USER_SPACE_ON_USE) {
    units_tr = tr * paint.tr;
  } else {
    BBox bb;
    if (!path.bbox(Transform::identity(), &bb)) return;
    units_tr = tr * bb.unit_transform() * paint.tr;
  }
  Transform pixel_tr;
  if (!units_tr.invert(&pixel_tr)) return;
  mask_iter([&](const Pixel& px) {
    Point pt = P((double)px.x + 0.5, (double)px.y + 0.5);
    blend(px, paint.at(pixel_tr.apply(pt)));
  });
}

// ---------------------------------------------------------------------------
// Scene: Fill / Group / Transform / Opacity / Clip / Stroke (src/scene.rs:19-63,
// 87-200, 256-566). A Stroke node holds, as its child, the Fill of its
// stroked path (Path::stroke, non-zero rule) that Pipeline::build_rec makes
// of it (src/scene.rs:297-306); Scene::bbox takes the unstroked path
// (src/scene.rs:158).
// ---------------------------------------------------------------------------
struct SceneNode;
typedef std::shared_ptr<SceneNode> SceneRef;
struct SceneNode {
  enum Kind { FILL, GROUP, TRANSFORM, OPACITY, CLIP, STROKE } kind;
  std::shared_ptr<Path> path;  // fill; the clip path of a clip; the unstroked path of a stroke
  std::shared_ptr<Paint> paint;
  FillRule rule = NONZERO;
  std::vector<SceneRef> children;
  Transform tr = Transform::identity();
  SceneRef child;
  double opacity = 1.0;
  Units units = USER_SPACE_ON_USE;  // clip
};

// Scene::bbox (src/scene.rs:154-183)
inline bool scene_bbox(const SceneNode* s, const Transform& tr, BBox* out) {
  switch (s->kind) {
    case SceneNode::FILL:
      return s->path->bbox(tr, out);
    case SceneNode::GROUP: {
      bool have = false;
      for (auto& c : s->children) {
        BBox b;
        if (!scene_bbox(c.get(), tr, &b)) continue;
        *out = have ? out->unite(b) : b;
        have = true;
      }
      return have;
    }
    case SceneNode::TRANSFORM:
      return scene_bbox(s->child.get(), tr * s->tr, out);
    case SceneNode::OPACITY:
      return scene_bbox(s->child.get(), tr, out);
    case SceneNode::STROKE:
      return s->path->bbox(tr, out);
    default: {
      Transform ctr = tr;
      if (s->units == BOUNDING_BOX) {
        BBox cb;
        if (!scene_bbox(s->child.get(), Transform::identity(), &cb)) return false;
        ctr = tr * cb.unit_transform();
      }
      BBox a, b;
      if (!scene_bbox(s->child.get(), tr, &a) || !s->path->bbox(ctr, &b)) return false;
      return a.intersect(b, out);
    }
  }
}

struct Layer {
  std::vector<F4> image;
  Shape shape;
  int x = 0, y = 0;
  // Layer::new (src/scene.rs:483-501)
  static Layer make(const BBox& bbox, const LinColor* bg) {
    Layer l;
    int x0 = (int)std::floor(bbox.min.x), x1 = (int)std::ceil(bbox.max.x);
    int y0 = (int)std::floor(bbox.min.y), y1 = (int)std::ceil(bbox.max.y);
    size_t w = (size_t)(x1 - x0), h = (size_t)(y1 - y0);
    l.shape = Shape::simple(h, w);
    l.image.assign(w * h, bg ? bg->c : f4(0, 0, 0, 0));
    l.x = x0;
    l.y = y0;
    return l;
  }
};

struct PipelineNode {
  SceneNode::Kind kind;
  const SceneNode* fill;  // the scene node (fill, opacity, clip)
  std::vector<size_t> children;
  BBox bbox;
  Transform tr;
  Transform clip_tr;
};

// Layer::compose (src/scene.rs:528-560): f(dst, src) over the overlap
template <class F>
inline void layer_compose(Layer& dst, int sx, int sy, size_t sw, size_t sh, F&& f) {
  const long x0 = std::max<long>(dst.x, sx);
  const long x1 = std::min<long>(dst.x + (long)dst.shape.width, sx + (long)sw);
  const long y0 = std::max<long>(dst.y, sy);
  const long y1 = std::min<long>(dst.y + (long)dst.shape.height, sy + (long)sh);
  for (long x = x0; x < x1; ++x)
    for (long y = y0; y < y1; ++y) {
      F4& d = dst.image[dst.shape.offset((size_t)(y - dst.y), (size_t)(x - dst.x))];
      d = f(d, (size_t)(y - sy), (size_t)(x - sx));
    }
}

struct Pipeline {
  std::vector<PipelineNode> nodes;
  // src/scene.rs:268-357
  bool build_rec(const SceneNode* s, const BBox* view, const Transform& tr, size_t* id) {
    auto view_apply = [&](bool have, const BBox& bb, BBox* out) {
      if (!have) return false;
      if (!view) { *out = bb; return true; }
      return view->intersect(bb, out);
    };
    if (s->kind == SceneNode::FILL) {
      BBox pb, vb;
      bool have = s->path->bbox(tr, &pb);
      if (!view_apply(have, pb, &vb)) return false;
      nodes.push_back(PipelineNode{SceneNode::FILL, s, {}, vb, tr, tr});
      *id = nodes.size() - 1;
      return true;
    }
    if (s->kind == SceneNode::STROKE) return build_rec(s->child.get(), view, tr, id);
    if (s->kind == SceneNode::GROUP) {
      std::vector<size_t> ch;
      for (auto& c : s->children) {
        size_t cid;
        if (build_rec(c.get(), view, tr, &cid)) ch.push_back(cid);
      }
      if (ch.empty()) return false;
      bool have = false;
      BBox bb{};
      for (size_t cid : ch) {
        bb = have ? nodes[cid].bbox.unite(bb) : nodes[cid].bbox;
        have = true;
      }
      nodes.push_back(PipelineNode{SceneNode::GROUP, nullptr, ch, bb, tr, tr});
      *id = nodes.size() - 1;
      return true;
    }
    if (s->kind == SceneNode::OPACITY) {
      size_t cid;
      if (!build_rec(s->child.get(), view, tr, &cid)) return false;
      nodes.push_back(PipelineNode{SceneNode::OPACITY, s, {cid}, nodes[cid].bbox, tr, tr});
      *id = nodes.size() - 1;
      return true;
    }
    if (s->kind == SceneNode::CLIP) {
      Transform clip_tr = tr;
      if (s->units == BOUNDING_BOX) {
        BBox cb;
        if (!scene_bbox(s->child.get(), Transform::identity(), &cb)) return false;
        clip_tr = tr * cb.unit_transform();
      }
      BBox pb, clip_bbox;
      if (!s->path->bbox(clip_tr, &pb) || !view_apply(true, pb, &clip_bbox)) return false;
      size_t cid;
      if (!build_rec(s->child.get(), &clip_bbox, tr, &cid)) return false;
      BBox bb;
      if (!clip_bbox.intersect(nodes[cid].bbox, &bb)) return false;
      nodes.push_back(PipelineNode{SceneNode::CLIP, s, {cid}, bb, tr, clip_tr});
      *id = nodes.size() - 1;
      return true;
    }
    return build_rec(s->child.get(), view, tr * s->tr, id);
  }
  // src/scene.rs:397-435
  bool active_edge = false;  // ActiveEdgeRasterizer for fills and clip masks
  void render_rec(size_t id, Layer& layer, double flatness) const {
    const PipelineNode& n = nodes[id];
    if (n.kind == SceneNode::GROUP) {
      for (size_t c : n.children) render_rec(c, layer, flatness);
      return;
    }
    if (n.kind == SceneNode::OPACITY) {
      Layer child = Layer::make(nodes[n.children[0]].bbox, nullptr);
      render_rec(n.children[0], child, flatness);
      const float op = (float)n.fill->opacity;
      layer_compose(layer, child.x, child.y, child.shape.width, child.shape.height,
                    [&](F4 d, size_t r, size_t c) {
                      const LinColor s{child.image[child.shape.offset(r, c)] * op};
                      return LinColor{d}.blend_over(s).c;
                    });
      return;
    }
    if (n.kind == SceneNode::CLIP) {
      Layer mask = Layer::make(n.bbox, nullptr);
      std::vector<double> m(mask.shape.width * mask.shape.height, 0.0);
      Layer child = Layer::make(nodes[n.children[0]].bbox, nullptr);
      render_rec(n.children[0], child, flatness);
      const Transform align = Transform::new_translate(-(double)mask.x, -(double)mask.y);
      ImageF64 mi{m.data(), m.size(), Shape::simple(mask.shape.height, mask.shape.width)};
      if (active_edge)
        ae_mask(*n.fill->path, align * n.clip_tr, flatness, n.fill->rule, mi);
      else
        sd_mask(*n.fill->path, align * n.clip_tr, flatness, n.fill->rule, mi);
      layer_compose(child, mask.x, mask.y, mask.shape.width, mask.shape.height,
                    [&](F4 d, size_t r, size_t c) {
                      return d * (float)m[r * mask.shape.width + c];
                    });
      layer_compose(layer, child.x, child.y, child.shape.width, child.shape.height,
                    [&](F4 d, size_t r, size_t c) {
                      return LinColor{d}.blend_over(LinColor{child.image[child.shape.offset(r, c)]}).c;
                    });
      return;
    }
    long col_min = (long)std::floor(n.bbox.min.x) - layer.x;
    long col_max = (long)std::ceil(n.bbox.max.x) - layer.x + 1;
    long row_min = (long)std::floor(n.bbox.min.y) - layer.y;
    long row_max = (long)std::ceil(n.bbox.max.y) - layer.y + 1;
    Shape vs = view_shape(layer.shape, (size_t)row_min, (size_t)row_max, (size_t)col_min,
                          (size_t)col_max);
    Transform align =
        Transform::new_translate(-std::floor(n.bbox.min.x), -std::floor(n.bbox.min.y));
    ImageRGBA img{layer.image.data(), layer.image.size(), vs};
    sd_fill(*n.fill->path, align * n.tr, flatness, n.fill->rule, *n.fill->paint, img, 0, SIZE_MAX,
            active_edge);
  }
  // render_rec for the layer rows [y0, y1) only: each fill runs over the rows
  // of its sub-view inside the band (sd_fill's row window).
  void render_rec_rows(size_t id, Layer& layer, double flatness, size_t y0, size_t y1) const {
    const PipelineNode& n = nodes[id];
    if (n.kind == SceneNode::GROUP) {
      for (size_t c : n.children) render_rec_rows(c, layer, flatness, y0, y1);
      return;
    }
    long col_min = (long)std::floor(n.bbox.min.x) - layer.x;
    long col_max = (long)std::ceil(n.bbox.max.x) - layer.x + 1;
    long row_min = (long)std::floor(n.bbox.min.y) - layer.y;
    long row_max = (long)std::ceil(n.bbox.max.y) - layer.y + 1;
    Shape vs = view_shape(layer.shape, (size_t)row_min, (size_t)row_max, (size_t)col_min,
                          (size_t)col_max);
    // the sub-view's first layer row (view_shape's clamped row_min)
    const size_t first = std::min((size_t)row_min, layer.shape.height);
    const size_t a = std::max(y0, first), b = std::min(y1, first + vs.height);
    if (a >= b) return;
    Transform align =
        Transform::new_translate(-std::floor(n.bbox.min.x), -std::floor(n.bbox.min.y));
    ImageRGBA img{layer.image.data(), layer.image.size(), vs};
    sd_fill(*n.fill->path, align * n.tr, flatness, n.fill->rule, *n.fill->paint, img, a - first,
            b - first);
  }
};

// Scene::render as scene_render below, the layer split into bands of rows
// rendered by `threads` host threads (each pixel belongs to one band, so the
// threads never write the same pixel). Same pixels, same values.
inline Layer scene_render_mt(const SceneNode* scene, const Transform& tr, const BBox* view,
                             const LinColor* bg, double flatness, int threads, size_t band = 16) {
  Pipeline p;
  size_t root;
  if (!p.build_rec(scene, view, tr, &root)) return Layer{};
  (void)root;
  size_t root_id = p.nodes.size() - 1;
  Layer layer = Layer::make(view ? *view : p.nodes[root_id].bbox, bg);
  const size_t H = layer.shape.height, nb = (H + band - 1) / band;
  std::atomic<size_t> next{0};
  std::atomic<bool> nan{false};
  auto work = [&]() {
    for (size_t b; (b = next.fetch_add(1)) < nb;) {
      try {
        p.render_rec_rows(root_id, layer, flatness, b * band, std::min(H, (b + 1) * band));
      } catch (const std::runtime_error&) {
        nan = true;
      }
    }
  };
  std::vector<std::thread> pool;
  for (int k = 1; k < threads; ++k) pool.emplace_back(work);
  work();
  for (auto& t : pool) t.join();
  if (nan) throw std::runtime_error("cannot flatten segment with NaN");
  return layer;
}

// Scene::render (src/scene.rs:186-199) with the signed-difference rasterizer
// (or the active-edge one)
inline Layer scene_render(const SceneNode* scene, const Transform& tr, const BBox* view,
                          const LinColor* bg, double flatness, bool active_edge = false) {
  Pipeline p;
  p.active_edge = active_edge;
  size_t root;
  if (!p.build_rec(scene, view, tr, &root)) return Layer{};
  (void)root;
  size_t root_id = p.nodes.size() - 1;
  Layer layer = Layer::make(view ? *view : p.nodes[root_id].bbox, bg);
  p.render_rec(root_id, layer, flatness);
  return layer;
}

}  // namespace orc

// f64 helpers of the C ABI: transform algebra and path bounding boxes (host
// and device) and the preparation of a paint for the fine kernels (host). The arithmetic keeps
// the reference's operation order (the library is built with
// -ffp-contract=off), because a gradient's coordinate frame comes from these
// values.
#pragma once
#include <algorithm>
#include <vector>

#include "../../include/rz.h"
#include "rz_geom.hpp"
#include "rz_paint.hpp"

namespace rzh {

using rzg::mkpt;
using rzg::Pt;

struct Tr {
  double m[6];
};
RZ_HD Tr tr_of(const double* m) {
  Tr t;
  for (int i = 0; i < 6; ++i) t.m[i] = m[i];
  return t;
}
// Transform * Transform (src/geometry.rs:519-539)
RZ_HD Tr tr_mul(const Tr& s, const Tr& o) {
  Tr r;
  r.m[0] = s.m[0] * o.m[0] + s.m[1] * o.m[3];
  r.m[1] = s.m[0] * o.m[1] + s.m[1] * o.m[4];
  r.m[2] = s.m[0] * o.m[2] + s.m[1] * o.m[5] + s.m[2];
  r.m[3] = s.m[3] * o.m[0] + s.m[4] * o.m[3];
  r.m[4] = s.m[3] * o.m[1] + s.m[4] * o.m[4];
  r.m[5] = s.m[3] * o.m[2] + s.m[4] * o.m[5] + s.m[5];
  return r;
}
// Transform::invert (src/geometry.rs:370-384)
RZ_HD bool tr_invert(const Tr& t, Tr* out) {
  const double m00 = t.m[0], m01 = t.m[1], m02 = t.m[2], m10 = t.m[3], m11 = t.m[4], m12 = t.m[5];
  const double det = m00 * m11 - m10 * m01;
  if (fabs(det) <= rzg::kEpsilon) return false;
  const double o00 = m11 / det, o01 = -m01 / det, o10 = -m10 / det, o11 = m00 / det;
  out->m[0] = o00;
  out->m[1] = o01;
  out->m[2] = -o00 * m02 - o01 * m12;
  out->m[3] = o10;
  out->m[4] = o11;
  out->m[5] = -o10 * m02 - o11 * m12;
  return true;
}

// BBox (src/geometry.rs:550-673)
struct BBox {
  Pt mn, mx;
};
RZ_HD BBox bbox_new(Pt a, Pt b) {
  BBox r;
  r.mn = mkpt(a.x <= b.x ? a.x : b.x, a.y <= b.y ? a.y : b.y);
  r.mx = mkpt(a.x <= b.x ? b.x : a.x, a.y <= b.y ? b.y : a.y);
  return r;
}
RZ_HD BBox bbox_extend(BBox b, Pt p) {
  if (p.x < b.mn.x) b.mn.x = p.x;
  else if (p.x > b.mx.x) b.mx.x = p.x;
  if (p.y < b.mn.y) b.mn.y = p.y;
  else if (p.y > b.mx.y) b.mx.y = p.y;
  return b;
}
RZ_HD BBox bbox_union_opt(BBox b, const BBox* other) {
  if (!other) return b;
  return bbox_extend(bbox_extend(b, other->mn), other->mx);
}
RZ_HD bool bbox_contains(const BBox& b, Pt p) {
  return b.mn.x <= p.x && p.x <= b.mx.x && b.mn.y <= p.y && p.y <= b.mx.y;
}
// BBox::unit_transform = translate(x, y) . scale(width, height)
RZ_HD Tr bbox_unit_transform(const BBox& b) {
  Tr t, s;
  const double tv[6] = {1.0, 0.0, b.mn.x, 0.0, 1.0, b.mn.y};
  const double sv[6] = {b.mx.x - b.mn.x, 0.0, 0.0, 0.0, b.mx.y - b.mn.y, 0.0};
  for (int i = 0; i < 6; ++i) {
    t.m[i] = tv[i];
    s.m[i] = sv[i];
  }
  return tr_mul(t, s);
}

// Quad::at / Cubic::at (src/curve.rs:432-441, 712-723)
RZ_HD Pt quad_at(const Pt* p, double t) {
  const double t1 = t, t_1 = 1.0 - t;
  const double t2 = t1 * t1, t_2 = t_1 * t_1;
  const double k1 = 2.0 * t1 * t_1;
  return mkpt(t_2 * p[0].x + k1 * p[1].x + t2 * p[2].x, t_2 * p[0].y + k1 * p[1].y + t2 * p[2].y);
}
RZ_HD Pt cubic_at(const Pt* p, double t) {
  const double t1 = t, t_1 = 1.0 - t;
  const double t2 = t1 * t1, t_2 = t_1 * t_1;
  const double t3 = t2 * t1, t_3 = t_2 * t_1;
  const double k1 = 3.0 * t1 * t_2, k2 = 3.0 * t2 * t_1;
  return mkpt(t_3 * p[0].x + k1 * p[1].x + k2 * p[2].x + t3 * p[3].x,
              t_3 * p[0].y + k1 * p[1].y + k2 * p[2].y + t3 * p[3].y);
}

// Segment::bbox (src/curve.rs:270-273, 505-513, 810-818) with the
// extremities of Quad (:535-555) and Cubic (:853-864).
RZ_HD BBox segment_bbox(int kind, const Pt* p, const BBox* init) {
  if (kind == rzg::KIND_LINE) return bbox_union_opt(bbox_new(p[0], p[1]), init);
  if (kind == rzg::KIND_QUAD) {
    BBox b = bbox_union_opt(bbox_new(p[0], p[2]), init);
    if (bbox_contains(b, p[1])) return b;
    const double a0 = (p[2].x - 2.0 * p[1].x) + p[0].x, a1 = (p[2].y - 2.0 * p[1].y) + p[0].y;
    const double b0 = p[1].x - p[0].x, b1 = p[1].y - p[0].y;
    if (fabs(a0) > rzg::kEpsilon) {
      const double t0 = -b0 / a0;
      if (t0 >= 0.0 && t0 <= 1.0) b = bbox_extend(b, quad_at(p, t0));
    }
    if (fabs(a1) > rzg::kEpsilon) {
      const double t1 = -b1 / a1;
      if (t1 >= 0.0 && t1 <= 1.0) b = bbox_extend(b, quad_at(p, t1));
    }
    return b;
  }
  BBox b = bbox_union_opt(bbox_new(p[0], p[3]), init);
  if (bbox_contains(b, p[1]) && bbox_contains(b, p[2])) return b;
  double coef[2][3];
  for (int d = 0; d < 2; ++d) {
    const double v0 = d ? p[0].y : p[0].x, v1 = d ? p[1].y : p[1].x;
    const double v2 = d ? p[2].y : p[2].x, v3 = d ? p[3].y : p[3].x;
    coef[d][0] = ((-1.0 * v0 + 3.0 * v1) - 3.0 * v2) + 1.0 * v3;
    coef[d][1] = (2.0 * v0 - 4.0 * v1) + 2.0 * v2;
    coef[d][2] = -1.0 * v0 + v1;
  }
  for (int d = 0; d < 2; ++d) {
    double r[2];
    const int n = rzp::quadratic_solve(coef[d][0], coef[d][1], coef[d][2], &r[0], &r[1]);
    for (int i = 0; i < n; ++i)
      if (r[i] >= 0.0 && r[i] <= 1.0) b = bbox_extend(b, cubic_at(p, r[i]));
  }
  return b;
}

// Path::bbox(tr) (src/path.rs:428-431, 217-222): false for an empty path.
inline bool path_bbox(const rz_path* path, const Tr& tr, BBox* out) {
  bool have = false;
  BBox acc{};
  for (uint32_t k = 0; k < path->n_segments; ++k) {
    const int kind = path->kinds[k];
    const int np = kind + 2;
    Pt p[4];
    const double* q = path->points + 2 * (size_t)path->point_offsets[k];
    for (int i = 0; i < np; ++i) p[i] = rzg::tr_apply(tr.m, mkpt(q[2 * i], q[2 * i + 1]));
    acc = segment_bbox(kind, p, have ? &acc : nullptr);
    have = true;
  }
  *out = acc;
  return have;
}

// The pixel-to-paint transform of Rasterizer::fill (src/rasterize.rs:82-95):
// (tr * paint.tr)^-1 for UserSpaceOnUse, (tr * bbox.unit_transform() *
// paint.tr)^-1 for BoundingBox with the path's untransformed bbox. False when
// the reference draws nothing (no bbox, or a singular transform).
RZ_HD bool paint_pixel_transform(int units, const double* paint_tr, const Tr& tr,
                                 const BBox* bbox_identity, double* out) {
  Tr units_tr;
  if (units == RZ_UNITS_USER_SPACE) {
    units_tr = tr_mul(tr, tr_of(paint_tr));
  } else {
    if (!bbox_identity) return false;
    units_tr = tr_mul(tr_mul(tr, bbox_unit_transform(*bbox_identity)), tr_of(paint_tr));
  }
  Tr inv;
  if (!tr_invert(units_tr, &inv)) return false;
  for (int i = 0; i < 6; ++i) out[i] = inv.m[i];
  return true;
}

// A paint resolved for fills: DPaint (stop pointers filled in where the stops
// are uploaded) plus the host copies of the stops. `draw` is false when the
// reference's fill draws nothing.
struct PaintPrep {
  rzp::DPaint p;
  std::vector<double> pos;
  std::vector<rzp::C4> col;
  std::vector<rzp::C4> lin;   // the stops' colours as given (before the sRGB conversion)
  std::vector<int32_t> src;   // index of each stop in the paint as given; -1: the default stop
  bool draw = true;
};

// The transform-independent part: the gradient constructors (GradStops::new
// src/grad.rs:86-100, GradLinear::new :163-194, GradRadial::new :321-348).
// Returns false on an invalid paint.
inline bool prepare_paint_static(const rz_paint& in, PaintPrep* out) {
  rzp::DPaint& p = out->p;
  p = rzp::DPaint{};
  p.kind = in.kind;
  p.color = rzp::c4(in.color[0], in.color[1], in.color[2], in.color[3]);
  out->draw = true;
  out->pos.clear();
  out->col.clear();
  out->lin.clear();
  out->src.clear();
  if (in.kind == RZ_PAINT_SOLID) {
    p.n_stops = 0;
    return true;
  }
  if (in.kind != RZ_PAINT_LINEAR && in.kind != RZ_PAINT_RADIAL) return false;
  if (in.spread < RZ_SPREAD_PAD || in.spread > RZ_SPREAD_REFLECT) return false;
  if (in.units != RZ_UNITS_USER_SPACE && in.units != RZ_UNITS_BBOX) return false;
  if (in.n_stops && (!in.stop_positions || !in.stop_colors)) return false;
  p.spread = in.spread;
  p.linear_colors = in.linear_colors != 0;
  // stops: stable sort by position, a default opaque black stop if empty,
  // converted to sRGB unless the gradient interpolates linear colours
  struct Stop {
    double pos;
    rzp::C4 col;
    int32_t src;
  };
  std::vector<Stop> st;
  for (uint32_t i = 0; i < in.n_stops; ++i) {
    const float* c = in.stop_colors + 4 * i;
    st.push_back({in.stop_positions[i], rzp::c4(c[0], c[1], c[2], c[3]), (int32_t)i});
  }
  std::stable_sort(st.begin(), st.end(), [](const Stop& a, const Stop& b) { return a.pos < b.pos; });
  if (st.empty()) st.push_back({0.0, rzp::c4(0.0f, 0.0f, 0.0f, 1.0f), -1});
  for (auto& s : st) {
    out->pos.push_back(s.pos);
    out->col.push_back(p.linear_colors ? s.col : rzp::into_srgb(s.col));
    out->lin.push_back(s.col);
    out->src.push_back(s.src);
  }
  p.n_stops = (int)st.size();
  if (in.kind == RZ_PAINT_LINEAR) {
    p.start = mkpt(in.start[0], in.start[1]);
    const Pt d = mkpt(in.end[0] - in.start[0], in.end[1] - in.start[1]);
    const double dd = d.x * d.x + d.y * d.y;
    p.dir = mkpt(d.x / dd, d.y / dd);
  } else {
    p.center = mkpt(in.center[0], in.center[1]);
    p.fcenter = mkpt(in.fcenter[0], in.fcenter[1]);
    p.radius = in.radius;
    p.fradius = in.fradius;
  }
  return true;
}

// Rasterizer::fill's paint for one call (src/rasterize.rs:76-99).
inline bool prepare_paint(const rz_paint& in, const Tr& tr, const BBox* path_bbox, PaintPrep* out) {
  if (!prepare_paint_static(in, out)) return false;
  if (in.kind == RZ_PAINT_SOLID) return true;
  out->draw = paint_pixel_transform(in.units, in.tr, tr, path_bbox, out->p.ptr);
  return true;
}

}  // namespace rzh

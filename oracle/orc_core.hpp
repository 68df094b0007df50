// Oracle: CPU restatement of the reference crate's geometry, curve, path and
// flatten code (aslpavel/rasterize v0.6.7, Rust).
//
// TEST INFRASTRUCTURE ONLY. This code is the checker that the HIP product path
// is compared against. It is linked by tests/, by __graft_entry__.smoke() and by
// bench.py's CPU-baseline leg, and by nothing else. The product library never
// links, loads or calls it.
//
// Every arithmetic expression keeps the reference's operation order, and the
// file is compiled with -ffp-contract=off (Rust never fuses multiply-add), so
// flattened lines are bit-identical to what the reference produces.
#pragma once
#ifndef _GNU_SOURCE
#define _GNU_SOURCE
#endif
#include <math.h>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <stdexcept>
#include <vector>

namespace orc {

typedef double Scalar;
// src/geometry.rs:14-16
static const double EPSILON = 2.220446049250313e-16;  // f64::EPSILON
static const double EPSILON_SQRT = 1.4901161193847656e-8;
static const double PI = 3.141592653589793;
// src/path.rs:16
static const double DEFAULT_FLATNESS = 0.05;

// ---------------------------------------------------------------------------
// Point (src/geometry.rs:100-294)
// ---------------------------------------------------------------------------
struct Point {
  double x, y;
};
inline Point P(double x, double y) { return Point{x, y}; }
inline Point operator+(Point a, Point b) { return {a.x + b.x, a.y + b.y}; }
inline Point operator-(Point a, Point b) { return {a.x - b.x, a.y - b.y}; }
inline Point operator*(double s, Point p) { return {s * p.x, s * p.y}; }
inline Point operator/(Point p, double s) { return {p.x / s, p.y / s}; }
inline bool operator==(Point a, Point b) { return a.x == b.x && a.y == b.y; }
inline double dot(Point a, Point b) { return a.x * b.x + a.y * b.y; }
inline double cross(Point a, Point b) { return a.x * b.y - a.y * b.x; }
inline double length(Point p) { return std::hypot(p.x, p.y); }
// src/geometry.rs:212-216
inline bool is_close_to(Point a, Point b) {
  return std::fabs(a.x - b.x) < EPSILON && std::fabs(a.y - b.y) < EPSILON;
}
// src/geometry.rs:196-203
inline bool cos_between(Point a, Point b, double* out) {
  double lengths = length(a) * length(b);
  if (lengths < EPSILON) return false;
  *out = dot(a, b) / lengths;
  return true;
}
// src/utils.rs:7-18
inline double clampd(double v, double lo, double hi) {
  if (v < lo) return lo;
  if (v > hi) return hi;
  return v;
}
// src/geometry.rs:186-193
inline bool angle_between(Point a, Point b, double* out) {
  double c;
  if (!cos_between(a, b, &c)) return false;
  double angle = std::acos(clampd(c, -1.0, 1.0));
  *out = cross(a, b) < 0.0 ? -angle : angle;
  return true;
}
// Rust f64::rem_euclid
inline double rem_euclid(double a, double b) {
  double r = std::fmod(a, b);
  return r < 0.0 ? r + std::fabs(b) : r;
}

// ---------------------------------------------------------------------------
// Transform (src/geometry.rs:316-539), stored [m00, m01, m02, m10, m11, m12]
// ---------------------------------------------------------------------------
struct Transform {
  double m[6];
  static Transform identity() { return Transform{{1.0, 0.0, 0.0, 0.0, 1.0, 0.0}}; }
  static Transform new_translate(double tx, double ty) {
    return Transform{{1.0, 0.0, tx, 0.0, 1.0, ty}};
  }
  static Transform new_scale(double sx, double sy) {
    return Transform{{sx, 0.0, 0.0, 0.0, sy, 0.0}};
  }
  // Rust's f64::sin_cos is lowered to one glibc sincos() call on
  // x86_64-linux-gnu; glibc's sincos and separate sin/cos differ in the last
  // bit for ~0.1% of arguments, so the call is made explicitly.
  static Transform new_rotate(double a) {
    double s, c;
    ::sincos(a, &s, &c);
    return Transform{{c, -s, 0.0, s, c, 0.0}};
  }
  // src/geometry.rs:363-367
  Point apply(Point p) const {
    return Point{p.x * m[0] + p.y * m[1] + m[2], p.x * m[3] + p.y * m[4] + m[5]};
  }
  // src/geometry.rs:370-384
  bool invert(Transform* out) const {
    double m00 = m[0], m01 = m[1], m02 = m[2], m10 = m[3], m11 = m[4], m12 = m[5];
    double det = m00 * m11 - m10 * m01;
    if (std::fabs(det) <= EPSILON) return false;
    double o00 = m11 / det;
    double o01 = -m01 / det;
    double o10 = -m10 / det;
    double o11 = m00 / det;
    double o02 = -o00 * m02 - o01 * m12;
    double o12 = -o10 * m02 - o11 * m12;
    *out = Transform{{o00, o01, o02, o10, o11, o12}};
    return true;
  }
  Transform pre_translate(double tx, double ty) const;
  Transform pre_scale(double sx, double sy) const;
  Transform pre_rotate(double a) const;
};
// src/geometry.rs:519-539
inline Transform operator*(const Transform& s, const Transform& o) {
  double s00 = s.m[0], s01 = s.m[1], s02 = s.m[2], s10 = s.m[3], s11 = s.m[4], s12 = s.m[5];
  double o00 = o.m[0], o01 = o.m[1], o02 = o.m[2], o10 = o.m[3], o11 = o.m[4], o12 = o.m[5];
  return Transform{{s00 * o00 + s01 * o10, s00 * o01 + s01 * o11, s00 * o02 + s01 * o12 + s02,
                    s10 * o00 + s11 * o10, s10 * o01 + s11 * o11, s10 * o02 + s11 * o12 + s12}};
}
inline Transform Transform::pre_translate(double tx, double ty) const {
  return *this * new_translate(tx, ty);
}
inline Transform Transform::pre_scale(double sx, double sy) const {
  return *this * new_scale(sx, sy);
}
inline Transform Transform::pre_rotate(double a) const { return *this * new_rotate(a); }

// ---------------------------------------------------------------------------
// BBox (src/geometry.rs:543-688)
// ---------------------------------------------------------------------------
struct BBox {
  Point min, max;
  static BBox make(Point p0, Point p1) {
    double x0 = p0.x, x1 = p1.x, y0 = p0.y, y1 = p1.y;
    if (!(x0 <= x1)) std::swap(x0, x1);
    if (!(y0 <= y1)) std::swap(y0, y1);
    return BBox{{x0, y0}, {x1, y1}};
  }
  double width() const { return max.x - min.x; }
  double height() const { return max.y - min.y; }
  bool contains(Point p) const {
    return min.x <= p.x && p.x <= max.x && min.y <= p.y && p.y <= max.y;
  }
  BBox extend(Point p) const {
    double x0 = min.x, y0 = min.y, x1 = max.x, y1 = max.y;
    if (p.x < x0) x0 = p.x;
    else if (p.x > x1) x1 = p.x;
    if (p.y < y0) y0 = p.y;
    else if (p.y > y1) y1 = p.y;
    return BBox{{x0, y0}, {x1, y1}};
  }
  BBox unite(const BBox& o) const { return extend(o.min).extend(o.max); }
  BBox union_opt(const BBox* o) const { return o ? unite(*o) : *this; }
  // src/geometry.rs:648-657, 677-688
  bool intersect(const BBox& o, BBox* out) const {
    if (min.x > o.max.x || o.min.x > max.x) return false;
    if (min.y > o.max.y || o.min.y > max.y) return false;
    double xmin = std::fmax(min.x, o.min.x), xmax = std::fmin(max.x, o.max.x);
    double ymin = std::fmax(min.y, o.min.y), ymax = std::fmin(max.y, o.max.y);
    *out = make(P(xmin, ymin), P(xmax, ymax));
    return true;
  }
  // src/geometry.rs:662-665
  Transform unit_transform() const {
    return Transform::new_translate(min.x, min.y).pre_scale(width(), height());
  }
};

// src/utils.rs:205-231 -- returns number of roots written to out[0..2]
inline int quadratic_solve(double a, double b, double c, double out[2]) {
  int n = 0;
  if (std::fabs(a) < EPSILON) {
    if (std::fabs(b) > EPSILON) out[n++] = -c / b;
    return n;
  }
  double disc = b * b - 4.0 * a * c;
  if (std::fabs(disc) < EPSILON) {
    out[n++] = -b / (2.0 * a);
  } else if (disc > 0.0) {
    double sq = std::sqrt(disc);
    if (b >= 0.0) {
      double mul = -b - sq;
      out[n++] = mul / (2.0 * a);
      out[n++] = 2.0 * c / mul;
    } else {
      double mul = -b + sq;
      out[n++] = 2.0 * c / mul;
      out[n++] = mul / (2.0 * a);
    }
  }
  return n;
}

// ---------------------------------------------------------------------------
// Segment = Line | Quad | Cubic (src/curve.rs:163-1258)
// ---------------------------------------------------------------------------
enum SegKind : uint8_t { SEG_LINE = 0, SEG_QUAD = 1, SEG_CUBIC = 2 };

struct Segment {
  SegKind kind;
  Point p[4];
  int npts() const { return kind == SEG_LINE ? 2 : (kind == SEG_QUAD ? 3 : 4); }
  Point start() const { return p[0]; }
  Point end() const { return p[npts() - 1]; }

  static Segment line(Point a, Point b) {
    Segment s;
    s.kind = SEG_LINE;
    s.p[0] = a; s.p[1] = b; s.p[2] = s.p[3] = Point{0, 0};
    return s;
  }
  static Segment quad(Point a, Point b, Point c) {
    Segment s;
    s.kind = SEG_QUAD;
    s.p[0] = a; s.p[1] = b; s.p[2] = c; s.p[3] = Point{0, 0};
    return s;
  }
  static Segment cubic(Point a, Point b, Point c, Point d) {
    Segment s;
    s.kind = SEG_CUBIC;
    s.p[0] = a; s.p[1] = b; s.p[2] = c; s.p[3] = d;
    return s;
  }

  // Line: src/curve.rs:233-235; Quad: :413-417; Cubic: :692-697
  double flatness() const {
    switch (kind) {
      case SEG_LINE: return 0.0;
      case SEG_QUAD: {
        Point d = 2.0 * p[1] - p[0] - p[2];
        return d.x * d.x + d.y * d.y;
      }
      default: {
        Point u = 3.0 * p[1] - 2.0 * p[0] - p[3];
        Point v = 3.0 * p[2] - p[0] - 2.0 * p[3];
        return std::fmax(u.x * u.x, v.x * v.x) + std::fmax(u.y * u.y, v.y * v.y);
      }
    }
  }

  // src/curve.rs:236-240, 419-422, 699-702 (Transform::apply per point)
  Segment transform(const Transform& tr) const {
    Segment s = *this;
    for (int i = 0; i < npts(); ++i) s.p[i] = tr.apply(p[i]);
    return s;
  }

  // Segment::split is Curve::split's default, i.e. split_at(0.5)
  // (src/curve.rs:43-45, 1186-1201). Line: :259-263; Quad: :458-468;
  // Cubic: :749-769. Coefficients are computed exactly as the reference does.
  void split(Segment* s0, Segment* s1) const { split_at(0.5, s0, s1); }
  void split_at(double t, Segment* s0, Segment* s1) const {
    switch (kind) {
      case SEG_LINE: {
        Point mid = (1.0 - t) * p[0] + t * p[1];
        *s0 = line(p[0], mid);
        *s1 = line(mid, p[1]);
        return;
      }
      case SEG_QUAD: {
        double t1 = t, t_1 = 1.0 - t;
        double t2 = t1 * t1, t_2 = t_1 * t_1;
        Point mid = t_2 * p[0] + 2.0 * t1 * t_1 * p[1] + t2 * p[2];
        *s0 = quad(p[0], t_1 * p[0] + t * p[1], mid);
        *s1 = quad(mid, t_1 * p[1] + t * p[2], p[2]);
        return;
      }
      default: {
        double t1 = t, t_1 = 1.0 - t;
        double t2 = t1 * t1, t_2 = t_1 * t_1;
        double t3 = t2 * t1, t_3 = t_2 * t_1;
        Point mid = t_3 * p[0] + 3.0 * t1 * t_2 * p[1] + 3.0 * t2 * t_1 * p[2] + t3 * p[3];
        *s0 = cubic(p[0], t_1 * p[0] + t * p[1], t_2 * p[0] + 2.0 * t * t_1 * p[1] + t2 * p[2],
                    mid);
        *s1 = cubic(mid, t_2 * p[1] + 2.0 * t * t_1 * p[2] + t2 * p[3], t_1 * p[2] + t * p[3],
                    p[3]);
        return;
      }
    }
  }

  // Optimised split (Quad::split :449-456, Cubic::split :731-747); used only
  // by the test that pins split() == split_at(0.5).
  void split_fast(Segment* s0, Segment* s1) const {
    if (kind == SEG_QUAD) {
      Point mid = 0.25 * (p[0] + 2.0 * p[1] + p[2]);
      *s0 = quad(p[0], 0.5 * (p[0] + p[1]), mid);
      *s1 = quad(mid, 0.5 * (p[1] + p[2]), p[2]);
    } else if (kind == SEG_CUBIC) {
      Point mid = 0.125 * p[0] + 0.375 * p[1] + 0.375 * p[2] + 0.125 * p[3];
      *s0 = cubic(p[0], 0.5 * p[0] + 0.5 * p[1], 0.25 * p[0] + 0.5 * p[1] + 0.25 * p[2], mid);
      *s1 = cubic(mid, 0.25 * p[1] + 0.5 * p[2] + 0.25 * p[3], 0.5 * p[2] + 0.5 * p[3], p[3]);
    } else {
      split(s0, s1);
    }
  }

  // Quad::at :438-447; Cubic::at :712-723; Line::at :250-253
  Point at(double t) const {
    switch (kind) {
      case SEG_LINE: return (1.0 - t) * p[0] + t * p[1];
      case SEG_QUAD: {
        double t1 = t, t_1 = 1.0 - t;
        double t2 = t1 * t1, t_2 = t_1 * t_1;
        return t_2 * p[0] + 2.0 * t1 * t_1 * p[1] + t2 * p[2];
      }
      default: {
        double t1 = t, t_1 = 1.0 - t;
        double t2 = t1 * t1, t_2 = t_1 * t_1;
        double t3 = t2 * t1, t_3 = t_2 * t_1;
        return t_3 * p[0] + 3.0 * t1 * t_2 * p[1] + 3.0 * t2 * t_1 * p[2] + t3 * p[3];
      }
    }
  }

  // Quad::extremities :575-596; Cubic::extremities :853-864
  int extremities(double out[6]) const {
    int n = 0;
    if (kind == SEG_QUAD) {
      Point a = p[2] - 2.0 * p[1] + p[0];
      Point b = p[1] - p[0];
      if (std::fabs(a.x) > EPSILON) {
        double t0 = -b.x / a.x;
        if (t0 >= 0.0 && t0 <= 1.0) out[n++] = t0;
      }
      if (std::fabs(a.y) > EPSILON) {
        double t1 = -b.y / a.y;
        if (t1 >= 0.0 && t1 <= 1.0) out[n++] = t1;
      }
    } else if (kind == SEG_CUBIC) {
      Point a = -1.0 * p[0] + 3.0 * p[1] - 3.0 * p[2] + 1.0 * p[3];
      Point b = 2.0 * p[0] - 4.0 * p[1] + 2.0 * p[2];
      Point c = -1.0 * p[0] + p[1];
      double r[2];
      int k = quadratic_solve(a.x, b.x, c.x, r);
      for (int i = 0; i < k; ++i)
        if (r[i] >= 0.0 && r[i] <= 1.0) out[n++] = r[i];
      k = quadratic_solve(a.y, b.y, c.y, r);
      for (int i = 0; i < k; ++i)
        if (r[i] >= 0.0 && r[i] <= 1.0) out[n++] = r[i];
    }
    return n;
  }

  // Line::bbox :269-272; Quad::bbox :553-561; Cubic::bbox :810-818
  BBox bbox(const BBox* init) const {
    if (kind == SEG_LINE) return BBox::make(p[0], p[1]).union_opt(init);
    if (kind == SEG_QUAD) {
      BBox b = BBox::make(p[0], p[2]).union_opt(init);
      if (b.contains(p[1])) return b;
      double ts[6];
      int n = extremities(ts);
      for (int i = 0; i < n; ++i) b = b.extend(at(ts[i]));
      return b;
    }
    BBox b = BBox::make(p[0], p[3]).union_opt(init);
    if (b.contains(p[1]) && b.contains(p[2])) return b;
    double ts[6];
    int n = extremities(ts);
    for (int i = 0; i < n; ++i) b = b.extend(at(ts[i]));
    return b;
  }

  // src/curve.rs:1081-1089
  bool has_nans() const {
    for (int i = 0; i < npts(); ++i)
      if (std::isnan(p[i].x) || std::isnan(p[i].y)) return true;
    return false;
  }
};

struct Line {
  Point p0, p1;
};

// ---------------------------------------------------------------------------
// EllipArc (src/ellipse.rs:40-96, 159-214)
// ---------------------------------------------------------------------------
struct EllipArc {
  Point center;
  double rx, ry, phi, eta, eta_delta;

  static bool new_param(Point src, Point dst, double rx, double ry, double x_axis_rot,
                        bool large_flag, bool sweep_flag, EllipArc* out) {
    rx = std::fabs(rx);
    ry = std::fabs(ry);
    double phi = x_axis_rot * PI / 180.0;
    Point q = Transform::new_rotate(-phi).apply(0.5 * (src - dst));
    double x1 = q.x, y1 = q.y;
    double a = x1 / rx, b = y1 / ry;
    double s = a * a + b * b;  // powi(2) is a plain square
    if (s > 1.0) {
      double ss = std::sqrt(s);
      rx = rx * ss;
      ry = ry * ss;
    }
    double rxry = rx * ry, rxy1 = rx * y1, ryx1 = ry * x1;
    double sq = std::sqrt(std::fmax(rxry * rxry / (rxy1 * rxy1 + ryx1 * ryx1) - 1.0, 0.0));
    if (large_flag == sweep_flag) sq = -sq;
    Point center = sq * P(rx * y1 / ry, -ry * x1 / rx);
    double cx = center.x, cy = center.y;
    center = Transform::new_rotate(phi).apply(center) + 0.5 * (dst + src);
    Point v0 = P(1.0, 0.0);
    Point v1 = P((x1 - cx) / rx, (y1 - cy) / ry);
    Point v2 = P((-x1 - cx) / rx, (-y1 - cy) / ry);
    double eta, ed;
    if (!angle_between(v0, v1, &eta)) return false;
    if (!angle_between(v1, v2, &ed)) return false;
    // deviation: NaN angles (a zero or non-finite radius) make the reference's
    // EllipArcCubicIter yield forever; a line, as the SVG notes prescribe
    if (!std::isfinite(eta) || !std::isfinite(ed)) return false;
    ed = rem_euclid(ed, 2.0 * PI);
    if (!sweep_flag && ed > 0.0) ed = ed - 2.0 * PI;
    else if (sweep_flag && ed < 0.0) ed = ed + 2.0 * PI;
    *out = EllipArc{center, rx, ry, phi, eta, ed};
    return true;
  }

  // EllipArcCubicIter
  void to_cubics(std::vector<Segment>& out) const {
    Transform phi_tr = Transform::new_rotate(phi);
    double segment_max_angle = PI / 2.0;
    double segment_count = std::ceil(std::fabs(eta_delta) / segment_max_angle);
    double segment_delta = eta_delta / segment_count;
    double last = segment_count - 1.0;
    for (double index = 0.0; !(index > last); index += 1.0) {
      double eta_1 = eta + segment_delta * index;
      double eta_2 = eta_1 + segment_delta;
      double tn = std::tan((eta_2 - eta_1) / 2.0);
      double sq = std::sqrt(4.0 + 3.0 * (tn * tn));
      double alpha = std::sin(eta_2 - eta_1) * (sq - 1.0) / 3.0;
      Point p0, d0, p3, d3;
      at(phi_tr, eta_1, &p0, &d0);
      at(phi_tr, eta_2, &p3, &d3);
      Point p1 = p0 + alpha * d0;
      Point p2 = p3 - alpha * d3;
      out.push_back(Segment::cubic(p0, p1, p2, p3));
    }
  }
  void at(const Transform& phi_tr, double a, Point* pt, Point* deriv) const {
    double s, c;
    ::sincos(a, &s, &c);  // alpha.sin_cos()
    *pt = phi_tr.apply(P(rx * c, ry * s)) + center;
    *deriv = phi_tr.apply(P(-rx * s, ry * c));
  }
};

// ---------------------------------------------------------------------------
// Path {segments, subpaths, closed} (src/path.rs:228-451) and PathBuilder
// (src/path.rs:799-1056)
// ---------------------------------------------------------------------------
struct Path {
  std::vector<Segment> segments;
  std::vector<size_t> subpaths;  // segments[subpaths[i]..subpaths[i+1]]
  std::vector<uint8_t> closed;
  size_t nsubpaths() const { return subpaths.size() < 2 ? 0 : subpaths.size() - 1; }

  // SubPath::bbox / Path::bbox (src/path.rs:217-222, 428-431)
  bool bbox(const Transform& tr, BBox* out) const {
    bool have = false;
    BBox b{};
    for (size_t i = 0; i < nsubpaths(); ++i) {
      for (size_t k = subpaths[i]; k < subpaths[i + 1]; ++k) {
        b = segments[k].transform(tr).bbox(have ? &b : nullptr);
        have = true;
      }
    }
    if (have) *out = b;
    return have;
  }
};

struct PathBuilder {
  Point position{0.0, 0.0};
  std::vector<Segment> segments;
  std::vector<size_t> subpaths;
  std::vector<uint8_t> closed;

  void subpath_finish(bool close) {
    if (segments.empty() || (!subpaths.empty() && subpaths.back() == segments.size())) return;
    if (subpaths.empty()) subpaths.push_back(0);
    if (close) position = segments[subpaths.back()].start();
    subpaths.push_back(segments.size());
    closed.push_back(close ? 1 : 0);
  }
  PathBuilder& move_to(Point p) {
    subpath_finish(false);
    position = p;
    return *this;
  }
  PathBuilder& close() {
    subpath_finish(true);
    return *this;
  }
  PathBuilder& line_to(Point p) {
    if (!is_close_to(position, p)) {
      segments.push_back(Segment::line(position, p));
      position = p;
    }
    return *this;
  }
  PathBuilder& quad_to(Point p1, Point p2) {
    segments.push_back(Segment::quad(position, p1, p2));
    position = p2;
    return *this;
  }
  PathBuilder& cubic_to(Point p1, Point p2, Point p3) {
    segments.push_back(Segment::cubic(position, p1, p2, p3));
    position = p3;
    return *this;
  }
  // src/path.rs:945-972
  PathBuilder& arc_to(Point radii, double x_axis_rot, bool large, bool sweep, Point p) {
    EllipArc arc;
    if (!EllipArc::new_param(position, p, radii.x, radii.y, x_axis_rot, large, sweep, &arc))
      return line_to(p);
    arc.to_cubics(segments);
    position = p;
    return *this;
  }
  // src/path.rs:977-996
  PathBuilder& circle(double radius) {
    double offset = 0.5522847498307935 * radius;
    Point xo = P(offset, 0.0), yo = P(0.0, offset);
    Point c = position;
    Point p0 = c - P(radius, 0.0), p1 = c - P(0.0, radius);
    Point p2 = c + P(radius, 0.0), p3 = c + P(0.0, radius);
    move_to(p0)
        .cubic_to(p0 - yo, p1 - xo, p1)
        .cubic_to(p1 + xo, p2 - yo, p2)
        .cubic_to(p2 + yo, p3 + xo, p3)
        .cubic_to(p3 - xo, p0 + yo, p0)
        .close()
        .move_to(c);
    return *this;
  }
  Path build() {
    subpath_finish(false);
    Path p;
    p.segments.swap(segments);
    p.subpaths.swap(subpaths);
    p.closed.swap(closed);
    position = Point{0.0, 0.0};
    return p;
  }
};

// ---------------------------------------------------------------------------
// PathFlattenIter (src/path.rs:734-796): in-order DFS over each transformed
// segment's subdivision tree; `close` forces the closing line of every subpath.
// Throws std::runtime_error("cannot flatten segment with NaN") where the
// reference panics.
// ---------------------------------------------------------------------------
template <typename Emit>
inline void flatten(const Path& path, const Transform& tr, double flatness, bool close,
                    Emit&& emit) {
  const double fl = 16.0 * flatness * flatness;
  std::vector<Segment> stack;
  for (size_t i = 0; i < path.nsubpaths(); ++i) {
    size_t b = path.subpaths[i], e = path.subpaths[i + 1];
    for (size_t k = b; k < e; ++k) {
      stack.push_back(path.segments[k].transform(tr));
      while (!stack.empty()) {
        Segment s = stack.back();
        stack.pop_back();
        if (s.has_nans()) throw std::runtime_error("cannot flatten segment with NaN");
        if (s.flatness() < fl) {
          emit(Line{s.start(), s.end()});
          continue;
        }
        Segment s0, s1;
        s.split(&s0, &s1);
        stack.push_back(s1);
        stack.push_back(s0);
      }
    }
    if (path.closed[i] || close) {
      Point a = path.segments[e - 1].end(), bb = path.segments[b].start();
      emit(Line{tr.apply(a), tr.apply(bb)});
    }
  }
}

}  // namespace orc

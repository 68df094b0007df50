// rasterize.hpp -- C++ host API mirroring the reference crate (aslpavel/rasterize
// v0.6.7) over the C ABI in rz.h: Transform, Path/PathBuilder (with elliptic
// arcs expanded to cubics at build time), FillRule, the Rasterizer interface
// with SignedDifferenceRasterizer / ActiveEdgeRasterizer, LinColor and the
// gradient paints. Names, argument meaning and error behaviour follow the
// reference; panics become rz::Error exceptions carrying the rz_status.
#pragma once
#ifndef _GNU_SOURCE
#define _GNU_SOURCE
#endif
#include <math.h>
#include <cmath>
#include <cstdint>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "rz.h"

namespace rz {

using Scalar = double;
constexpr Scalar EPSILON = 2.220446049250313e-16;          // src/geometry.rs:14
constexpr Scalar EPSILON_SQRT = 1.4901161193847656e-8;     // src/geometry.rs:16
constexpr Scalar PI = 3.141592653589793;
constexpr Scalar DEFAULT_FLATNESS = 0.05;                  // src/path.rs:16

struct Error : std::runtime_error {
  rz_status status;
  Error(rz_status s, const std::string& m) : std::runtime_error(m), status(s) {}
};

// Point (src/geometry.rs:100-294)
struct Point {
  Scalar x = 0.0, y = 0.0;
  Point() = default;
  Point(Scalar x_, Scalar y_) : x(x_), y(y_) {}
};
inline Point operator+(Point a, Point b) { return {a.x + b.x, a.y + b.y}; }
inline Point operator-(Point a, Point b) { return {a.x - b.x, a.y - b.y}; }
inline Point operator*(Scalar s, Point p) { return {s * p.x, s * p.y}; }
inline Scalar dot(Point a, Point b) { return a.x * b.x + a.y * b.y; }
inline Scalar cross(Point a, Point b) { return a.x * b.y - a.y * b.x; }
inline bool is_close_to(Point a, Point b) {
  return std::fabs(a.x - b.x) < EPSILON && std::fabs(a.y - b.y) < EPSILON;
}

// Transform (src/geometry.rs:316-539): [m00, m01, m02, m10, m11, m12]
struct Transform {
  Scalar m[6] = {1.0, 0.0, 0.0, 0.0, 1.0, 0.0};
  static Transform identity() { return Transform(); }
  static Transform make(Scalar a, Scalar b, Scalar c, Scalar d, Scalar e, Scalar f) {
    Transform t;
    t.m[0] = a; t.m[1] = b; t.m[2] = c; t.m[3] = d; t.m[4] = e; t.m[5] = f;
    return t;
  }
  static Transform new_translate(Scalar tx, Scalar ty) { return make(1.0, 0.0, tx, 0.0, 1.0, ty); }
  static Transform new_scale(Scalar sx, Scalar sy) { return make(sx, 0.0, 0.0, 0.0, sy, 0.0); }
  // f64::sin_cos is one glibc sincos() call on x86_64-linux-gnu (LLVM merges
  // the pair); sincos and separate sin/cos differ in the last bit for ~0.1%
  // of arguments, so the call is explicit.
  static Transform new_rotate(Scalar a) {
    Scalar s, c;
    ::sincos(a, &s, &c);
    return make(c, -s, 0.0, s, c, 0.0);
  }
  Point apply(Point p) const {
    return {p.x * m[0] + p.y * m[1] + m[2], p.x * m[3] + p.y * m[4] + m[5]};
  }
  Transform operator*(const Transform& o) const {
    const Scalar* s = m;
    const Scalar* q = o.m;
    return make(s[0] * q[0] + s[1] * q[3], s[0] * q[1] + s[1] * q[4], s[0] * q[2] + s[1] * q[5] + s[2],
                s[3] * q[0] + s[4] * q[3], s[3] * q[1] + s[4] * q[4], s[3] * q[2] + s[4] * q[5] + s[5]);
  }
  Transform pre_translate(Scalar tx, Scalar ty) const { return *this * new_translate(tx, ty); }
  Transform pre_scale(Scalar sx, Scalar sy) const { return *this * new_scale(sx, sy); }
  Transform pre_rotate(Scalar a) const { return *this * new_rotate(a); }
  rz_transform abi() const {
    rz_transform t;
    for (int i = 0; i < 6; ++i) t.m[i] = m[i];
    return t;
  }
};

enum class FillRule { NonZero = RZ_NONZERO, EvenOdd = RZ_EVENODD };

// Path {segments, subpaths, closed} stored as the SoA the C ABI consumes.
struct Path {
  std::vector<uint8_t> kinds;
  std::vector<uint32_t> point_offsets{0};
  std::vector<double> points;
  std::vector<uint32_t> subpaths;
  std::vector<uint8_t> closed;

  size_t segments_count() const { return kinds.size(); }
  size_t len() const { return subpaths.size() < 2 ? 0 : subpaths.size() - 1; }
  bool is_empty() const { return subpaths.empty(); }
  rz_path view() const {
    rz_path p;
    p.n_segments = (uint32_t)kinds.size();
    p.n_subpaths = (uint32_t)len();
    p.kinds = kinds.data();
    p.point_offsets = point_offsets.data();
    p.points = points.data();
    p.subpaths = subpaths.data();
    p.closed = closed.data();
    return p;
  }
  void push_segment(uint8_t kind, const Point* pts) {
    kinds.push_back(kind);
    for (int i = 0; i < kind + 2; ++i) {
      points.push_back(pts[i].x);
      points.push_back(pts[i].y);
    }
    point_offsets.push_back(point_offsets.back() + kind + 2u);
  }
  Point seg_start(size_t k) const {
    const double* q = &points[2 * (size_t)point_offsets[k]];
    return {q[0], q[1]};
  }
  Point seg_end(size_t k) const {
    const double* q = &points[2 * (size_t)(point_offsets[k + 1] - 1)];
    return {q[0], q[1]};
  }
};

// Elliptic arc -> cubics (src/ellipse.rs:40-96, 159-214), the expansion
// PathBuilder::arc_to performs at build time (src/path.rs:945-972).
struct EllipArc {
  Point center;
  Scalar rx = 0, ry = 0, phi = 0, eta = 0, eta_delta = 0;

  static bool angle_between(Point a, Point b, Scalar* out) {
    Scalar lengths = std::hypot(a.x, a.y) * std::hypot(b.x, b.y);
    if (lengths < EPSILON) return false;
    Scalar c = dot(a, b) / lengths;
    c = c < -1.0 ? -1.0 : (c > 1.0 ? 1.0 : c);
    Scalar angle = std::acos(c);
    *out = cross(a, b) < 0.0 ? -angle : angle;
    return true;
  }
  static bool new_param(Point src, Point dst, Scalar rx, Scalar ry, Scalar x_axis_rot, bool large,
                        bool sweep, EllipArc* out) {
    rx = std::fabs(rx);
    ry = std::fabs(ry);
    Scalar phi = x_axis_rot * PI / 180.0;
    Point q = Transform::new_rotate(-phi).apply(0.5 * (src - dst));
    Scalar x1 = q.x, y1 = q.y;
    Scalar a = x1 / rx, b = y1 / ry;
    Scalar s = a * a + b * b;
    if (s > 1.0) {
      Scalar ss = std::sqrt(s);
      rx = rx * ss;
      ry = ry * ss;
    }
    Scalar rxry = rx * ry, rxy1 = rx * y1, ryx1 = ry * x1;
    Scalar sq = std::sqrt(std::fmax(rxry * rxry / (rxy1 * rxy1 + ryx1 * ryx1) - 1.0, 0.0));
    if (large == sweep) sq = -sq;
    Point c = sq * Point(rx * y1 / ry, -ry * x1 / rx);
    Scalar cx = c.x, cy = c.y;
    c = Transform::new_rotate(phi).apply(c) + 0.5 * (dst + src);
    Point v0(1.0, 0.0), v1((x1 - cx) / rx, (y1 - cy) / ry), v2((-x1 - cx) / rx, (-y1 - cy) / ry);
    Scalar eta, ed;
    if (!angle_between(v0, v1, &eta) || !angle_between(v1, v2, &ed)) return false;
    // A zero or non-finite radius gives NaN angles, from which the reference's
    // cubic iterator never terminates; such an arc is drawn as the straight
    // line the SVG implementation notes prescribe for a zero radius.
    if (!std::isfinite(eta) || !std::isfinite(ed)) return false;
    ed = std::fmod(ed, 2.0 * PI);
    if (ed < 0.0) ed += 2.0 * PI;  // rem_euclid
    if (!sweep && ed > 0.0) ed -= 2.0 * PI;
    else if (sweep && ed < 0.0) ed += 2.0 * PI;
    out->center = c;
    out->rx = rx;
    out->ry = ry;
    out->phi = phi;
    out->eta = eta;
    out->eta_delta = ed;
    return true;
  }
  template <typename F>
  void to_cubics(F&& emit) const {
    Transform phi_tr = Transform::new_rotate(phi);
    Scalar count = std::ceil(std::fabs(eta_delta) / (PI / 2.0));
    Scalar delta = eta_delta / count;
    auto at = [&](Scalar a, Point* p, Point* d) {
      Scalar s, c;
      ::sincos(a, &s, &c);
      *p = phi_tr.apply(Point(rx * c, ry * s)) + center;
      *d = phi_tr.apply(Point(-rx * s, ry * c));
    };
    for (Scalar i = 0.0; !(i > count - 1.0); i += 1.0) {
      Scalar e1 = eta + delta * i;
      Scalar e2 = e1 + delta;
      Scalar tn = std::tan((e2 - e1) / 2.0);
      Scalar sq = std::sqrt(4.0 + 3.0 * (tn * tn));
      Scalar alpha = std::sin(e2 - e1) * (sq - 1.0) / 3.0;
      Point p0, d0, p3, d3;
      at(e1, &p0, &d0);
      at(e2, &p3, &d3);
      emit(p0, p0 + alpha * d0, p3 - alpha * d3, p3);
    }
  }
};

// PathBuilder (src/path.rs:799-1056)
class PathBuilder {
 public:
  PathBuilder& move_to(Point p) {
    subpath_finish(false);
    position_ = p;
    return *this;
  }
  PathBuilder& close() {
    subpath_finish(true);
    return *this;
  }
  PathBuilder& line_to(Point p) {
    if (!is_close_to(position_, p)) {
      Point q[2] = {position_, p};
      path_.push_segment(RZ_LINE, q);
      position_ = p;
    }
    return *this;
  }
  PathBuilder& quad_to(Point p1, Point p2) {
    Point q[3] = {position_, p1, p2};
    path_.push_segment(RZ_QUAD, q);
    position_ = p2;
    return *this;
  }
  PathBuilder& cubic_to(Point p1, Point p2, Point p3) {
    Point q[4] = {position_, p1, p2, p3};
    path_.push_segment(RZ_CUBIC, q);
    position_ = p3;
    return *this;
  }
  PathBuilder& arc_to(Point radii, Scalar x_axis_rot, bool large, bool sweep, Point p) {
    EllipArc arc;
    if (!EllipArc::new_param(position_, p, radii.x, radii.y, x_axis_rot, large, sweep, &arc))
      return line_to(p);
    arc.to_cubics([&](Point a, Point b, Point c, Point d) {
      Point q[4] = {a, b, c, d};
      path_.push_segment(RZ_CUBIC, q);
    });
    position_ = p;
    return *this;
  }
  PathBuilder& circle(Scalar radius) {
    Scalar off = 0.5522847498307935 * radius;
    Point xo(off, 0.0), yo(0.0, off), c = position_;
    Point p0 = c - Point(radius, 0.0), p1 = c - Point(0.0, radius);
    Point p2 = c + Point(radius, 0.0), p3 = c + Point(0.0, radius);
    return move_to(p0)
        .cubic_to(p0 - yo, p1 - xo, p1)
        .cubic_to(p1 + xo, p2 - yo, p2)
        .cubic_to(p2 + yo, p3 + xo, p3)
        .cubic_to(p3 - xo, p0 + yo, p0)
        .close()
        .move_to(c);
  }
  Point position() const { return position_; }
  Path build() {
    subpath_finish(false);
    Path p = std::move(path_);
    path_ = Path();
    position_ = Point();
    return p;
  }

 private:
  void subpath_finish(bool close) {
    size_t n = path_.kinds.size();
    if (n == 0 || (!path_.subpaths.empty() && path_.subpaths.back() == n)) return;
    if (path_.subpaths.empty()) path_.subpaths.push_back(0);
    if (close) position_ = path_.seg_start(path_.subpaths.back());
    path_.subpaths.push_back((uint32_t)n);
    path_.closed.push_back(close ? 1 : 0);
  }
  Path path_;
  Point position_;
};

// Context: one per host thread (the C ABI context is not thread-safe).
class Context {
 public:
  explicit Context(int device = 0) {
    rz_status s = rz_ctx_create(&ctx_, device);
    if (s != RZ_OK) throw Error(s, "rz_ctx_create failed");
  }
  ~Context() { rz_ctx_destroy(ctx_); }
  Context(const Context&) = delete;
  Context& operator=(const Context&) = delete;
  rz_ctx* get() const { return ctx_; }
  void check(rz_status s) const {
    if (s != RZ_OK) throw Error(s, rz_last_error(ctx_));
  }

 private:
  rz_ctx* ctx_ = nullptr;
};

// Strided f64 image view (src/image.rs:6-56, ImageMut<Pixel = f64>)
struct ImageMut {
  double* data;
  size_t len;
  rz_shape shape;
};

// Rasterizer trait (src/rasterize.rs:44-101)
class Rasterizer {
 public:
  virtual ~Rasterizer() = default;
  virtual const char* name() const = 0;
  // Rasterizer::mask: writes every pixel of `img`
  virtual void mask(const Path& path, Transform tr, ImageMut img, FillRule rule) const = 0;
  // Rasterizer::mask_iter rendered densely: alpha per pixel, 0 where the
  // reference's iterator produces no pixel
  virtual std::vector<float> mask_iter(const Path& path, Transform tr, size_t width, size_t height,
                                       FillRule rule) const = 0;
};

class SignedDifferenceRasterizer : public Rasterizer {
 public:
  explicit SignedDifferenceRasterizer(Context& ctx, Scalar flatness = DEFAULT_FLATNESS)
      : ctx_(ctx), flatness_(flatness) {}
  const char* name() const override { return "signed-difference"; }
  void mask(const Path& path, Transform tr, ImageMut img, FillRule rule) const override {
    rz_path p = path.view();
    ctx_.check(rz_mask(ctx_.get(), &p, tr.abi(), flatness_, (int)rule, RZ_EDGE_MASK, img.data,
                       img.len, img.shape));
  }
  std::vector<float> mask_iter(const Path& path, Transform tr, size_t width, size_t height,
                               FillRule rule) const override {
    std::vector<float> out(width * height);
    rz_path p = path.view();
    ctx_.check(rz_mask_f32(ctx_.get(), &p, tr.abi(), flatness_, (int)rule, RZ_EDGE_ITER, width,
                           height, out.data(), width));
    return out;
  }

 private:
  Context& ctx_;
  Scalar flatness_;
};

class ActiveEdgeRasterizer : public Rasterizer {
 public:
  explicit ActiveEdgeRasterizer(Context& ctx, Scalar flatness = DEFAULT_FLATNESS)
      : ctx_(ctx), flatness_(flatness) {}
  const char* name() const override { return "active-edge"; }
  void mask(const Path& path, Transform tr, ImageMut img, FillRule rule) const override {
    rz_path p = path.view();
    ctx_.check(rz_mask(ctx_.get(), &p, tr.abi(), flatness_, (int)rule, RZ_EDGE_NONE, img.data,
                       img.len, img.shape));
  }
  std::vector<float> mask_iter(const Path& path, Transform tr, size_t width, size_t height,
                               FillRule rule) const override {
    std::vector<float> out(width * height);
    rz_path p = path.view();
    ctx_.check(rz_mask_f32(ctx_.get(), &p, tr.abi(), flatness_, (int)rule, RZ_EDGE_NONE, width,
                           height, out.data(), width));
    return out;
  }

 private:
  Context& ctx_;
  Scalar flatness_;
};

}  // namespace rz

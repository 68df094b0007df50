// SVG path data (https://www.w3.org/TR/SVG11/paths.html#PathData) to a Path:
// the reader behind Path::read_svg_path / FromStr for Path (src/svg.rs:250-425,
// src/path.rs:528-534, 1058-1064), host code behind the C ABI.
//
// The grammar: commands M L H V C S Q T A Z (relative in lower case), numbers
// separated by whitespace or commas or not at all ("M.5-3-11-.11"), a
// command letter repeated implicitly (after M / m as L / l), nothing after Z
// but a new command. Numbers take the reference's value exactly: the digits
// as one integer mantissa (i64, wrapping) times 10 to the decimal exponent,
// the power by binary exponentiation (f64::powi), so parsed points match the
// reference bit for bit.
#include <cmath>
#include <cstdint>
#include <cstring>
#include <string>

#include "../../include/rasterize.hpp"
#include "../../include/rz.h"

namespace {

double pow10i(int e) {  // f64::powi(10, e): repeated squaring, reciprocal last
  const bool recip = e < 0;
  double base = 10.0, r = 1.0;
  for (;;) {
    if (e & 1) r *= base;
    e /= 2;
    if (e == 0) break;
    base *= base;
  }
  return recip ? 1.0 / r : r;
}

class PathData {
 public:
  PathData(const char* s, size_t n) : s_(s), n_(n) {}
  void seek(size_t i) { i_ = i; }
  size_t pos() const { return i_; }

  // Parses the whole text into `b`; false (with the byte offset) on an error.
  bool run(rz::PathBuilder& b, size_t* err) {
    int cmd = 0;  // the command repeated by bare numbers; 0 after Z or at the start
    rz::Point pos(0.0, 0.0), start(0.0, 0.0);
    char last = 0;          // kind of the previous segment: 'Q', 'C' or 0
    rz::Point ctl, end;     // its last control point and end (for T / S)
    for (;;) {
      skip_sep();
      if (i_ >= n_) return true;
      const char c = s_[i_];
      int op;
      if (is_cmd(c)) {
        ++i_;
        op = c;
        cmd = c == 'M' ? 'L' : c == 'm' ? 'l' : (c == 'Z' || c == 'z') ? 0 : c;
      } else {
        if (!cmd) return fail(err);
        op = cmd;
      }
      const bool rel = op >= 'a';
      auto point = [&](rz::Point* out) {
        double x, y;
        if (!number(&x) || !number(&y)) return false;
        *out = rel ? rz::Point(x, y) + pos : rz::Point(x, y);
        return true;
      };
      rz::Point p1, p2, p3;
      char kind = 0;
      switch (op | 0x20) {
        case 'm':
          if (!point(&p3)) return fail(err);
          start = p3;
          b.move_to(p3);
          break;
        case 'l':
          if (!point(&p3)) return fail(err);
          b.line_to(p3);
          break;
        case 'h': {
          double x;
          if (!number(&x)) return fail(err);
          p3 = rz::Point(rel ? pos.x + x : x, pos.y);
          b.line_to(p3);
          break;
        }
        case 'v': {
          double y;
          if (!number(&y)) return fail(err);
          p3 = rz::Point(pos.x, rel ? pos.y + y : y);
          b.line_to(p3);
          break;
        }
        case 'q':
        case 't':
          if ((op | 0x20) == 'q') {
            if (!point(&p1)) return fail(err);
          } else {
            p1 = last == 'Q' ? 2.0 * end - ctl : pos;
          }
          if (!point(&p3)) return fail(err);
          b.quad_to(p1, p3);
          kind = 'Q';
          ctl = p1;
          break;
        case 'c':
        case 's':
          if ((op | 0x20) == 'c') {
            if (!point(&p1)) return fail(err);
          } else {
            p1 = last == 'C' ? 2.0 * end - ctl : pos;
          }
          if (!point(&p2) || !point(&p3)) return fail(err);
          b.cubic_to(p1, p2, p3);
          kind = 'C';
          ctl = p2;
          break;
        case 'a': {
          double rx, ry, rot;
          bool large, sweep;
          if (!number(&rx) || !number(&ry) || !number(&rot) || !flag(&large) || !flag(&sweep) ||
              !point(&p3))
            return fail(err);
          b.arc_to(rz::Point(rx, ry), rot, large, sweep, p3);
          break;
        }
        default:  // 'z'
          p3 = start;
          b.close();
          break;
      }
      last = kind;
      end = p3;
      pos = p3;
    }
  }

 private:
  static bool is_cmd(char c) { return c && std::strchr("MmLlHhVvCcSsQqTtAaZz", c) != nullptr; }
  bool fail(size_t* err) {
    if (err) *err = i_;
    return false;
  }
  void skip_sep() {
    while (i_ < n_ && (s_[i_] == ' ' || s_[i_] == '\t' || s_[i_] == '\r' || s_[i_] == '\n' ||
                       s_[i_] == ','))
      ++i_;
  }
  bool digit() const { return i_ < n_ && s_[i_] >= '0' && s_[i_] <= '9'; }

 public:
  bool number(double* out) {
    skip_sep();
    uint64_t mant = 0;  // wraps as the reference's i64 does
    int64_t exp = 0;
    bool neg = false;
    if (i_ < n_ && (s_[i_] == '-' || s_[i_] == '+')) neg = s_[i_++] == '-';
    size_t nd = 0;
    for (; digit(); ++i_, ++nd) mant = mant * 10u + (uint64_t)(s_[i_] - '0');
    if (i_ < n_ && s_[i_] == '.') {
      ++i_;
      for (; digit(); ++i_, ++nd, --exp) mant = mant * 10u + (uint64_t)(s_[i_] - '0');
    }
    if (nd == 0) return false;
    if (i_ < n_ && (s_[i_] == 'e' || s_[i_] == 'E')) {
      ++i_;
      bool eneg = false;
      if (i_ < n_ && (s_[i_] == '-' || s_[i_] == '+')) eneg = s_[i_++] == '-';
      uint64_t e = 0;
      size_t ne = 0;
      for (; digit(); ++i_, ++ne) e = e * 10u + (uint64_t)(s_[i_] - '0');
      if (ne == 0) return false;
      exp = (int64_t)((uint64_t)exp + (eneg ? 0u - e : e));
    }
    const int64_t m = (int64_t)(neg ? 0u - mant : mant);
    *out = (double)m * pow10i((int32_t)exp);
    return true;
  }

 private:
  bool flag(bool* out) {
    skip_sep();
    if (i_ < n_ && (s_[i_] == '0' || s_[i_] == '1')) {
      *out = s_[i_++] == '1';
      return true;
    }
    return false;
  }

  const char* s_;
  size_t n_;
  size_t i_ = 0;
};

// SVG transform lists (FromStr for Transform, src/svg.rs:431-600): matrix,
// translate, scale / scaleX / scaleY, rotate (with an optional centre), skewX,
// skewY; angles in deg (the default) or rad, length units read and ignored;
// the list composes left to right (tr = tr * next).
class TransformList {
 public:
  TransformList(const char* s, size_t n) : s_(s), n_(n) {}

  bool run(rz::Transform* out, size_t* err) {
    rz::Transform tr;
    for (;;) {
      skip_sep();
      if (i_ >= n_) break;
      const size_t op_at = i_;
      const std::string op = ident();
      skip_sep();
      if (i_ >= n_ || s_[i_] != '(') return fail(err);
      ++i_;
      rz::Transform t;
      double a, b, c, d, e, f;
      if (op == "matrix") {
        if (!number(&a) || !number(&b) || !number(&c) || !number(&d) || !number(&e) || !number(&f))
          return fail(err);
        t = rz::Transform::make(a, c, e, b, d, f);  // columns (m00 m10) (m01 m11) (m02 m12)
      } else if (op == "rotate") {
        if (!angle(&a)) return fail(err);
        t = rz::Transform::new_rotate(a);
        if (length(&b)) {
          if (!length(&c)) return fail(err);
          t = (rz::Transform::new_translate(b, c) * t).pre_translate(-b, -c);
        }
      } else if (op == "translate") {
        if (!length(&a)) return fail(err);
        if (!length(&b)) b = 0.0;
        t = rz::Transform::new_translate(a, b);
      } else if (op == "translateX" || op == "translateY") {
        if (!length(&a)) return fail(err);
        t = op == "translateX" ? rz::Transform::new_translate(a, 0.0)
                               : rz::Transform::new_translate(0.0, a);
      } else if (op == "scale") {
        if (!number(&a)) return fail(err);
        if (!number(&b)) b = a;
        t = rz::Transform::new_scale(a, b);
      } else if (op == "scaleX" || op == "scaleY") {
        if (!number(&a)) return fail(err);
        t = op == "scaleX" ? rz::Transform::new_scale(a, 1.0) : rz::Transform::new_scale(1.0, a);
      } else if (op == "skewX" || op == "skewY") {
        if (!angle(&a)) return fail(err);
        const double ta = std::tan(a);
        t = op == "skewX" ? rz::Transform::make(1.0, ta, 0.0, 0.0, 1.0, 0.0)
                          : rz::Transform::make(1.0, 0.0, 0.0, ta, 1.0, 0.0);
      } else {
        i_ = op_at;
        return fail(err);
      }
      skip_sep();
      if (i_ >= n_ || s_[i_] != ')') return fail(err);
      ++i_;
      tr = tr * t;
    }
    *out = tr;
    return true;
  }

 private:
  bool fail(size_t* err) {
    if (err) *err = i_;
    return false;
  }
  void skip_sep() {
    while (i_ < n_ && (s_[i_] == ' ' || s_[i_] == '\t' || s_[i_] == '\r' || s_[i_] == '\n' ||
                       s_[i_] == ','))
      ++i_;
  }
  // An identifier as the reference keeps it: letters written into a 10-byte
  // ring (src/svg.rs:428, 441-452), so a name of 10 letters or more comes
  // out as its letters past the last multiple of 10; translateX and
  // translateY (10 letters) therefore read as "" and are refused, as there.
  std::string ident() {
    char ring[10];
    size_t len = 0;
    while (i_ < n_ && ((s_[i_] | 0x20) >= 'a' && (s_[i_] | 0x20) <= 'z')) {
      ring[len] = s_[i_++];
      len = (len + 1) % 10;
    }
    return std::string(ring, len);
  }
  bool number(double* out) {
    PathData p(s_, n_);
    p.seek(i_);
    const bool ok = p.number(out);
    i_ = p.pos();
    return ok;
  }
  bool angle(double* out) {  // radians
    if (!number(out)) return false;
    const std::string u = ident();
    if (u.empty() || u == "deg") {
      *out = *out * 3.141592653589793 / 180.0;
      return true;
    }
    return u == "rad";
  }
  bool length(double* out) {
    if (!number(out)) return false;
    (void)ident();
    return true;
  }

  const char* s_;
  size_t n_;
  size_t i_ = 0;
};

}  // namespace

struct rz_path_buf {
  rz::Path path;
};

extern "C" {

rz_status rz_svg_path_parse(const char* text, size_t len, rz_path_buf** out, size_t* err_offset) {
  if (!out || (!text && len)) return RZ_E_ARG;
  *out = nullptr;
  rz::PathBuilder b;
  PathData d(text, len);
  size_t err = 0;
  if (!d.run(b, &err)) {
    if (err_offset) *err_offset = err;
    return RZ_E_ARG;
  }
  *out = new rz_path_buf{b.build()};
  return RZ_OK;
}

rz_status rz_svg_transform_parse(const char* text, size_t len, rz_transform* out,
                                 size_t* err_offset) {
  if (!out || (!text && len)) return RZ_E_ARG;
  rz::Transform tr;
  TransformList t(text, len);
  size_t err = 0;
  if (!t.run(&tr, &err)) {
    if (err_offset) *err_offset = err;
    return RZ_E_ARG;
  }
  *out = tr.abi();
  return RZ_OK;
}

}  // extern "C"

rz_path_buf* rz_path_buf_adopt(rz::Path&& p) { return new rz_path_buf{std::move(p)}; }

extern "C" {

rz_status rz_path_buf_view(const rz_path_buf* p, rz_path* view) {
  if (!p || !view) return RZ_E_ARG;
  *view = p->path.view();
  return RZ_OK;
}

rz_status rz_path_buf_free(rz_path_buf* p) {
  delete p;
  return RZ_OK;
}

}  // extern "C"

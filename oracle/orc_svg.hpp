// Oracle: the SVG path-data parser (src/svg.rs:63-425) and Path::read_svg_path
// (src/path.rs:528-534), restated byte by byte.
//
// TEST INFRASTRUCTURE ONLY (see orc_core.hpp).
#pragma once
#include <string>
#include "orc_core.hpp"

namespace orc {

// f64::powi (LLVM's llvm.powi, lowered to compiler-rt's __powidf2): binary
// exponentiation by repeated squaring on the exponent's bits, the reciprocal
// taken at the end for a negative exponent.
inline double powi(double a, int b) {
  const bool recip = b < 0;
  double r = 1.0;
  for (;;) {
    if (b & 1) r *= a;
    b /= 2;
    if (b == 0) break;
    a *= a;
  }
  return recip ? 1.0 / r : r;
}

struct SvgError {
  size_t offset;
  std::string what;
};

// Parser (src/svg.rs:63-250): one byte of look-back.
struct SvgParser {
  const std::string& in;
  size_t pos = 0;     // next byte of the input
  size_t offset = 0;  // the reference's offset counter
  int buffered = -1;  // the byte put back, or -1

  explicit SvgParser(const std::string& s) : in(s) {}

  int parse_byte() {  // -1 at the end
    if (buffered >= 0) {
      const int b = buffered;
      buffered = -1;
      ++offset;
      return b;
    }
    if (pos < in.size()) {
      ++offset;
      return (unsigned char)in[pos++];
    }
    return -1;
  }
  void unparse_byte(int b) {
    offset = offset ? offset - 1 : 0;
    buffered = b;
  }
  template <class Pred, class Proc>
  size_t parse_while(Pred pred, Proc proc) {
    size_t count = 0;
    for (;;) {
      const int b = parse_byte();
      if (b < 0) break;
      if (!pred(b)) {
        unparse_byte(b);
        break;
      }
      ++count;
      proc(b);
    }
    return count;
  }
  template <class Pred, class Proc>
  bool parse_once(Pred pred, Proc proc) {
    const int b = parse_byte();
    if (b < 0) return false;
    if (pred(b)) {
      proc(b);
      return true;
    }
    unparse_byte(b);
    return false;
  }
  void parse_separators() {
    for (;;) {
      const int b = parse_byte();
      if (b < 0) break;
      if (!(b == ' ' || b == '\t' || b == '\r' || b == '\n' || b == ',')) {
        unparse_byte(b);
        break;
      }
    }
  }
  // src/svg.rs:164-246: integer mantissa (wrapping i64), decimal exponent,
  // value = mantissa as f64 * 10.powi(exponent)
  double parse_scalar() {
    parse_separators();
    int64_t mantissa = 0, exponent = 0;
    int64_t sign = 1;
    auto digit = [](int64_t* v, int b) {
      *v = (int64_t)((uint64_t)*v * 10u + (uint64_t)(b - '0'));
    };
    auto is_digit = [](int b) { return b >= '0' && b <= '9'; };
    parse_once([](int b) { return b == '-' || b == '+'; },
               [&](int b) {
                 if (b == '-') sign = -1;
               });
    const size_t whole = parse_while(is_digit, [&](int b) { digit(&mantissa, b); });
    const bool dot = parse_once([](int b) { return b == '.'; }, [](int) {});
    size_t frac = 0;
    if (dot)
      frac = parse_while(is_digit, [&](int b) {
        digit(&mantissa, b);
        exponent -= 1;
      });
    mantissa = (int64_t)((uint64_t)mantissa * (uint64_t)sign);
    if (whole + frac == 0) throw SvgError{offset, "invalid scalar"};
    const bool e = parse_once([](int b) { return b == 'e' || b == 'E'; }, [](int) {});
    if (e) {
      int64_t sci = 0, sci_sign = 1;
      parse_once([](int b) { return b == '-' || b == '+'; },
                 [&](int b) {
                   if (b == '-') sci_sign = -1;
                 });
      if (parse_while(is_digit, [&](int b) { digit(&sci, b); }) == 0)
        throw SvgError{offset, "invalid scalar"};
      exponent = (int64_t)((uint64_t)exponent + (uint64_t)(sci_sign * sci));
    }
    return (double)mantissa * powi(10.0, (int32_t)exponent);
  }
};

// SvgPathParser (src/svg.rs:250-425) applied to a PathBuilder.
inline Path read_svg_path(const std::string& text) {
  SvgParser p(text);
  PathBuilder b;
  int prev_op = -1;  // Option<u8>
  enum { NONE, MOVE, LINE, QUAD, CUBIC, ARC, CLOSE } prev_kind = NONE;
  Point prev_q1{}, prev_q2{}, prev_c2{}, prev_c3{};
  Point position = P(0.0, 0.0), subpath_start = P(0.0, 0.0);
  auto parse_point = [&]() {
    const double x = p.parse_scalar();
    const double y = p.parse_scalar();
    Point pt = P(x, y);
    if (prev_op >= 0 && prev_op >= 'a' && prev_op <= 'z') pt = pt + position;
    return pt;
  };
  auto parse_flag = [&]() {
    p.parse_separators();
    const int b = p.parse_byte();
    if (b == '0') return false;
    if (b == '1') return true;
    if (b >= 0) p.unparse_byte(b);
    throw SvgError{p.offset, "invalid flag"};
  };
  auto is_op = [](int c) {
    switch (c) {
      case 'M': case 'm': case 'L': case 'l': case 'V': case 'v': case 'H': case 'h':
      case 'C': case 'c': case 'S': case 's': case 'Q': case 'q': case 'T': case 't':
      case 'A': case 'a': case 'Z': case 'z':
        return true;
      default:
        return false;
    }
  };
  for (;;) {
    p.parse_separators();
    int op = p.parse_byte();
    if (op < 0) break;
    if (is_op(op)) {
      prev_op = op == 'm' ? 'l' : op == 'M' ? 'L' : (op == 'Z' || op == 'z') ? -1 : op;
    } else {
      p.unparse_byte(op);
      if (prev_op < 0) throw SvgError{p.offset, "invalid command"};
      op = prev_op;
    }
    Point dst;
    switch (op) {
      case 'M': case 'm': {
        dst = parse_point();
        subpath_start = dst;
        b.move_to(dst);
        prev_kind = MOVE;
        break;
      }
      case 'L': case 'l':
        dst = parse_point();
        b.line_to(dst);
        prev_kind = LINE;
        break;
      case 'V': case 'v': {
        const double y = p.parse_scalar();
        dst = op == 'v' ? P(position.x, position.y + y) : P(position.x, y);
        b.line_to(dst);
        prev_kind = LINE;
        break;
      }
      case 'H': case 'h': {
        const double x = p.parse_scalar();
        dst = op == 'h' ? P(position.x + x, position.y) : P(x, position.y);
        b.line_to(dst);
        prev_kind = LINE;
        break;
      }
      case 'Q': case 'q': {
        const Point p1 = parse_point();
        dst = parse_point();
        b.quad_to(p1, dst);
        prev_q1 = p1;
        prev_q2 = dst;
        prev_kind = QUAD;
        break;
      }
      case 'T': case 't': {
        const Point p1 = prev_kind == QUAD ? 2.0 * prev_q2 - prev_q1 : position;
        dst = parse_point();
        b.quad_to(p1, dst);
        prev_q1 = p1;
        prev_q2 = dst;
        prev_kind = QUAD;
        break;
      }
      case 'C': case 'c': {
        const Point p1 = parse_point();
        const Point p2 = parse_point();
        dst = parse_point();
        b.cubic_to(p1, p2, dst);
        prev_c2 = p2;
        prev_c3 = dst;
        prev_kind = CUBIC;
        break;
      }
      case 'S': case 's': {
        const Point p1 = prev_kind == CUBIC ? 2.0 * prev_c3 - prev_c2 : position;
        const Point p2 = parse_point();
        dst = parse_point();
        b.cubic_to(p1, p2, dst);
        prev_c2 = p2;
        prev_c3 = dst;
        prev_kind = CUBIC;
        break;
      }
      case 'A': case 'a': {
        const double rx = p.parse_scalar();
        const double ry = p.parse_scalar();
        const double rot = p.parse_scalar();
        const bool large = parse_flag();
        const bool sweep = parse_flag();
        dst = parse_point();
        b.arc_to(P(rx, ry), rot, large, sweep, dst);
        prev_kind = ARC;
        break;
      }
      default:  // 'Z' | 'z'
        dst = subpath_start;
        b.close();
        prev_kind = CLOSE;
        break;
    }
    position = dst;
  }
  return b.build();
}

// SvgTransformParser (src/svg.rs:431-590) and FromStr for Transform
// (:592-601): identifiers kept in a 10-byte ring (TRANSFORM_BUF), angles in
// deg (default) or rad, lengths with their units ignored.
inline Transform new_skew(double ax, double ay) {
  Transform t;
  const double v[6] = {1.0, std::tan(ax), 0.0, std::tan(ay), 1.0, 0.0};
  for (int i = 0; i < 6; ++i) t.m[i] = v[i];
  return t;
}

inline Transform read_svg_transform(const std::string& text) {
  SvgParser p(text);
  auto ident = [&]() {
    char buf[10];
    size_t len = 0;
    p.parse_while([](int b) { return (b >= 'a' && b <= 'z') || (b >= 'A' && b <= 'Z'); },
                  [&](int b) {
                    buf[len] = (char)b;
                    len = (len + 1) % 10;
                  });
    return std::string(buf, len);
  };
  auto angle = [&]() {
    const double v = p.parse_scalar();
    const size_t at = p.offset;
    const std::string u = ident();
    if (u.empty() || u == "deg") return v * PI / 180.0;
    if (u == "rad") return v;
    throw SvgError{at, "invalid units"};
  };
  auto length = [&]() {
    const double v = p.parse_scalar();
    (void)ident();
    return v;
  };
  Transform tr = Transform::identity();
  for (;;) {
    p.parse_separators();
    const int b0 = p.parse_byte();
    if (b0 < 0) break;
    p.unparse_byte(b0);
    const size_t at = p.offset;
    const std::string op = ident();
    p.parse_separators();
    if (p.parse_byte() != '(') throw SvgError{p.offset, "bracket expected"};
    Transform t;
    if (op == "matrix") {
      const double m00 = p.parse_scalar(), m10 = p.parse_scalar(), m01 = p.parse_scalar();
      const double m11 = p.parse_scalar(), m02 = p.parse_scalar(), m12 = p.parse_scalar();
      const double v[6] = {m00, m01, m02, m10, m11, m12};
      for (int i = 0; i < 6; ++i) t.m[i] = v[i];
    } else if (op == "rotate") {
      t = Transform::new_rotate(angle());
      double tx = 0.0;
      bool have = true;
      try {
        tx = length();
      } catch (const SvgError&) {
        have = false;
      }
      if (have) {
        const double ty = length();
        t = (Transform::new_translate(tx, ty) * t).pre_translate(-tx, -ty);
      }
    } else if (op == "translate") {
      const double tx = length();
      double ty = 0.0;
      try {
        ty = length();
      } catch (const SvgError&) {
      }
      t = Transform::new_translate(tx, ty);
    } else if (op == "translateX") {
      t = Transform::new_translate(length(), 0.0);
    } else if (op == "translateY") {
      t = Transform::new_translate(0.0, length());
    } else if (op == "scale") {
      const double sx = p.parse_scalar();
      double sy = sx;
      try {
        sy = p.parse_scalar();
      } catch (const SvgError&) {
      }
      t = Transform::new_scale(sx, sy);
    } else if (op == "scaleX") {
      t = Transform::new_scale(p.parse_scalar(), 1.0);
    } else if (op == "scaleY") {
      t = Transform::new_scale(1.0, p.parse_scalar());
    } else if (op == "skewX") {
      t = new_skew(angle(), 0.0);
    } else if (op == "skewY") {
      t = new_skew(0.0, angle());
    } else {
      throw SvgError{at, "invalid transform op"};
    }
    p.parse_separators();
    if (p.parse_byte() != ')') throw SvgError{p.offset, "bracket expected"};
    tr = tr * t;
  }
  return tr;
}

}  // namespace orc

// Path::stroke (src/path.rs:370-415) as independent work items, for the GPU
// (and, built for the host, for bit-exact checks against a sequential
// restatement).
//
// The reference walks each subpath sequentially: every segment's offset
// curves (src/curve.rs:1279-1433) are appended to a list, joined to what the
// list already holds (line_join, or line_cap for the first segment of the way
// back, src/curve.rs:977-1078); at the end of a run the list is closed
// (stroke_close, src/path.rs:708-732) or capped and pushed as one closed
// subpath. Here one item is one (segment, direction) pair, in the order the
// reference visits them:
//
//   open subpath of n segments:   items 0..n-1 forward, n..2n-1 backward
//                                 (item n joins with line_cap): one run
//   closed subpath:               items 0..n-1 forward (run 0),
//                                 n..2n-1 backward (run 1)
//
// An item's offset curves depend on its segment alone. What joins them to the
// rest depends only on the last curve of the nearest earlier item of the run
// that has any (a max-scan over item indices) and, at a run's end, on the
// run's first curve. So every item's output is computed in parallel: a count
// pass, a prefix sum, an emit pass writing the same segments.
//
// Every expression keeps the reference's operation order (the library is
// built with -ffp-contract=off). On the GPU, hypot, sincos, tan, acos and
// fmod are the device library's, which may differ from the host libm in the
// last bit.
#pragma once
#include "rz_geom.hpp"

namespace rzs {

using rzg::mkpt;
using rzg::Pt;
constexpr double kEps = rzg::kEpsilon;
constexpr double kPi = 3.141592653589793;

enum { JOIN_MITER = 0, JOIN_BEVEL = 1, JOIN_ROUND = 2 };
enum { CAP_BUTT = 0, CAP_SQUARE = 1, CAP_ROUND = 2 };

// StrokeStyle (src/path.rs:118-136); miter_limit is LineJoin::Miter's argument
struct Style {
  double width;
  double miter_limit;
  int join;
  int cap;
};

// One outline segment: kind (0 line, 1 quad, 2 cubic) and its kind + 2 points.
struct Seg {
  Pt p[4];
  int kind;
  int pad;
};

RZ_HD Pt add(Pt a, Pt b) { return mkpt(a.x + b.x, a.y + b.y); }
RZ_HD Pt sub(Pt a, Pt b) { return mkpt(a.x - b.x, a.y - b.y); }
RZ_HD Pt scale(double s, Pt a) { return mkpt(s * a.x, s * a.y); }
RZ_HD double dot(Pt a, Pt b) { return a.x * b.x + a.y * b.y; }
RZ_HD double cross(Pt a, Pt b) { return a.x * b.y - a.y * b.x; }

// sqrt(x^2 + y^2) correctly rounded: x^2 + y^2 as an exact double-double
// (products split with fma), the square root of its high part, then one
// rounding decision from the exact residual. The host libm's hypot (what the
// reference calls) is within 0.51 ulp and equals this in all but about 0.1%
// of cases; the device library's is within about 1 ulp. Near a cusp the
// offset construction turns such last-bit differences into different offset
// pieces, so the device takes the correctly rounded value.
RZ_HD double hypot_cr(double x, double y) {
  double a = fabs(x), b = fabs(y);
  if (isinf(a) || isinf(b)) return INFINITY;
  if (isnan(a) || isnan(b)) return a + b;
  if (a < b) {
    const double t = a;
    a = b;
    b = t;
  }
  if (b == 0.0 || b < a * 0x1p-60) return a;  // b^2 / a^2 below half an ulp of 1
  int e = 0;
  (void)frexp(a, &e);
  a = ldexp(a, -e);  // a in [0.5, 1), exact
  b = ldexp(b, -e);
  const double ah = a * a, al = fma(a, a, -ah);
  const double bh = b * b, bl = fma(b, b, -bh);
  const double sh = ah + bh, v = sh - ah;
  const double se = (ah - (sh - v)) + (bh - v);  // two-sum error of ah + bh
  const double lo = se + (al + bl);
  const double s = sh + lo, s_lo = lo - (s - sh);
  double r = sqrt(s);
  const double rr = r * r, rr_lo = fma(r, r, -rr);
  const double d = ((s - rr) - rr_lo) + s_lo;  // (x^2 + y^2) - r^2
  // sqrt(S) > r + u/2  <=>  S - r^2 > r u + u^2 / 4 (u: the ulp above r), and
  // the mirror test with the ulp below
  const double up = nextafter(r, INFINITY), dn = nextafter(r, 0.0);
  const double u = up - r, w = r - dn;
  if (d > r * u + 0.25 * u * u) r = up;
  else if (d < -(r * w) + 0.25 * w * w) r = dn;
  return ldexp(r, e);
}

#if defined(__HIP_DEVICE_COMPILE__) || defined(RZS_HYPOT_CR)
RZ_HD double hyp(double x, double y) { return hypot_cr(x, y); }
#else
RZ_HD double hyp(double x, double y) { return hypot(x, y); }
#endif
RZ_HD double len(Pt a) { return hyp(a.x, a.y); }  // Point::length
// Point::is_close_to (src/geometry.rs:211-216)
RZ_HD bool close_to(Pt a, Pt b) { return fabs(a.x - b.x) < kEps && fabs(a.y - b.y) < kEps; }
// Point::normalize (src/geometry.rs:175-183)
RZ_HD bool unit(Pt a, Pt* out) {
  const double l = len(a);
  if (l < kEps) return false;
  *out = mkpt(a.x / l, a.y / l);
  return true;
}

RZ_HD Seg mk(int kind, Pt a, Pt b, Pt c, Pt d) {
  Seg s;
  s.kind = kind;
  s.pad = 0;
  s.p[0] = a;
  s.p[1] = b;
  s.p[2] = c;
  s.p[3] = d;
  return s;
}
RZ_HD Seg mk_line(Pt a, Pt b) { return mk(0, a, b, mkpt(0.0, 0.0), mkpt(0.0, 0.0)); }
RZ_HD Pt first_pt(const Seg& s) { return s.p[0]; }
RZ_HD Pt last_pt(const Seg& s) { return s.kind == 0 ? s.p[1] : (s.kind == 1 ? s.p[2] : s.p[3]); }
// Curve::reverse
RZ_HD Seg reversed(const Seg& s) {
  if (s.kind == 0) return mk(0, s.p[1], s.p[0], s.p[2], s.p[3]);
  if (s.kind == 1) return mk(1, s.p[2], s.p[1], s.p[0], s.p[3]);
  return mk(2, s.p[3], s.p[2], s.p[1], s.p[0]);
}

// A line as two points
struct Ln {
  Pt a, b;
};
RZ_HD Ln ln(Pt a, Pt b) {
  Ln l;
  l.a = a;
  l.b = b;
  return l;
}
// Line::at (src/curve.rs:250-253)
RZ_HD Pt ln_at(const Ln& l, double t) { return add(scale(1.0 - t, l.a), scale(t, l.b)); }
// Line::intersect (src/curve.rs:204-215): parameters on both lines
RZ_HD bool ln_cross(const Ln& u, const Ln& v, double* t0, double* t1) {
  const double x1 = u.a.x, y1 = u.a.y, x2 = u.b.x, y2 = u.b.y;
  const double x3 = v.a.x, y3 = v.a.y, x4 = v.b.x, y4 = v.b.y;
  const double det = (x4 - x3) * (y1 - y2) - (x1 - x2) * (y4 - y3);
  if (fabs(det) < kEps) return false;
  *t0 = ((y3 - y4) * (x1 - x3) + (x4 - x3) * (y1 - y3)) / det;
  *t1 = ((y1 - y2) * (x1 - x3) + (x2 - x1) * (y1 - y3)) / det;
  return true;
}
RZ_HD bool in01(double t) { return t >= 0.0 && t <= 1.0; }

// line_offset (src/curve.rs:1283-1287): the line moved by dist along its
// normal (y, -x)
RZ_HD bool ln_offset(Pt a, Pt b, double dist, Ln* out) {
  const Pt d = sub(b, a);
  Pt n;
  if (!unit(mkpt(d.y, -d.x), &n)) return false;
  const Pt off = scale(dist, n);
  *out = ln(add(a, off), add(b, off));
  return true;
}

// Segment::ends (Line :195-197, Quad :377-389, Cubic :647-668): tangent lines
// at the start and at the end
RZ_HD void tangents(const Seg& s, Ln* start, Ln* end) {
  if (s.kind == 0) {
    *start = *end = ln(s.p[0], s.p[1]);
    return;
  }
  if (s.kind == 1) {
    const Ln u = ln(s.p[0], s.p[1]), v = ln(s.p[1], s.p[2]);
    if (close_to(s.p[0], s.p[1])) {
      *start = *end = v;
    } else if (close_to(s.p[1], s.p[2])) {
      *start = *end = u;
    } else {
      *start = u;
      *end = v;
    }
    return;
  }
  int i0 = 0, i1 = 0;
  if (!close_to(s.p[0], s.p[1])) i0 = 0;
  else if (!close_to(s.p[1], s.p[2])) i0 = 1;
  else if (!close_to(s.p[2], s.p[3])) i0 = 2;
  if (!close_to(s.p[3], s.p[2])) i1 = 3;
  else if (!close_to(s.p[2], s.p[1])) i1 = 2;
  else if (!close_to(s.p[1], s.p[0])) i1 = 1;
  // four coincident points index before the array in the reference (a
  // panic); offset curves never have them
  if (i1 == 0) i1 = 1;
  *start = ln(s.p[i0], s.p[i0 + 1]);
  *end = ln(s.p[i1 - 1], s.p[i1]);
}

// Point::angle_between (src/geometry.rs:186-205)
RZ_HD bool angle(Pt a, Pt b, double* out) {
  const double l = len(a) * len(b);
  if (l < kEps) return false;
  double c = dot(a, b) / l;
  if (c < -1.0) c = -1.0;
  else if (c > 1.0) c = 1.0;
  const double t = acos(c);
  *out = cross(a, b) < 0.0 ? -t : t;
  return true;
}

// Transform::new_rotate(r).apply(p)
RZ_HD Pt rotate(double r, Pt p) {
  double s, c;
  sincos(r, &s, &c);
  const double m01 = -s;
  return mkpt(p.x * c + p.y * m01 + 0.0, p.x * s + p.y * c + 0.0);
}

// The circular arc of radius r from src to dst (EllipArc::new_param with
// rx = ry = r, no rotation, small arc, src/ellipse.rs:40-96) as its cubics
// (EllipArcCubicIter, src/ellipse.rs:159-214): at most four. False where the
// reference's new_param gives None (or, a documented deviation, where the
// angles are not finite: its cubic iterator would not end).
template <class S>
RZ_HD bool arc(Pt src, Pt dst, double r, bool sweep, S& out) {
  double rx = fabs(r), ry = fabs(r);
  const double phi = 0.0 * kPi / 180.0;
  const Pt q = rotate(-phi, scale(0.5, sub(src, dst)));
  const double x1 = q.x, y1 = q.y;
  const double ax = x1 / rx, by = y1 / ry;
  const double s = ax * ax + by * by;
  if (s > 1.0) {
    const double ss = sqrt(s);
    rx = rx * ss;
    ry = ry * ss;
  }
  const double rxry = rx * ry, rxy1 = rx * y1, ryx1 = ry * x1;
  double sq = sqrt(fmax(rxry * rxry / (rxy1 * rxy1 + ryx1 * ryx1) - 1.0, 0.0));
  if (!sweep) sq = -sq;  // large_flag (false) == sweep_flag
  const Pt c0 = scale(sq, mkpt(rx * y1 / ry, -ry * x1 / rx));
  const double cx = c0.x, cy = c0.y;
  const Pt center = add(rotate(phi, c0), scale(0.5, add(dst, src)));
  const Pt v1 = mkpt((x1 - cx) / rx, (y1 - cy) / ry);
  const Pt v2 = mkpt((-x1 - cx) / rx, (-y1 - cy) / ry);
  double eta, ed;
  if (!angle(mkpt(1.0, 0.0), v1, &eta) || !angle(v1, v2, &ed)) return false;
  if (!isfinite(eta) || !isfinite(ed)) return false;
  double rem = fmod(ed, 2.0 * kPi);  // rem_euclid
  if (rem < 0.0) rem = rem + fabs(2.0 * kPi);
  ed = rem;
  if (!sweep && ed > 0.0) ed = ed - 2.0 * kPi;
  else if (sweep && ed < 0.0) ed = ed + 2.0 * kPi;
  const double count = ceil(fabs(ed) / (kPi / 2.0));
  const double delta = ed / count;
  const double last = count - 1.0;
  double sphi, cphi;
  sincos(phi, &sphi, &cphi);
  const double m01 = -sphi;
  for (double i = 0.0; !(i > last); i += 1.0) {
    const double e1 = eta + delta * i;
    const double e2 = e1 + delta;
    const double tn = tan((e2 - e1) / 2.0);
    const double sqa = sqrt(4.0 + 3.0 * (tn * tn));
    const double alpha = sin(e2 - e1) * (sqa - 1.0) / 3.0;
    Pt p[2], d[2];
    for (int k = 0; k < 2; ++k) {
      double sn, cs;
      sincos(k == 0 ? e1 : e2, &sn, &cs);
      const double px = rx * cs, py = ry * sn;
      p[k] = add(mkpt(px * cphi + py * m01 + 0.0, px * sphi + py * cphi + 0.0), center);
      const double dx = -rx * sn, dy = ry * cs;
      d[k] = mkpt(dx * cphi + dy * m01 + 0.0, dx * sphi + dy * cphi + 0.0);
    }
    out.put(mk(2, p[0], add(p[0], scale(alpha, d[0])), sub(p[1], scale(alpha, d[1])), p[1]));
  }
  return true;
}

// Segment::line_join (src/curve.rs:977-1047) of `a` into `b`: up to four
// segments into `out`
template <class S>
RZ_HD void line_join(const Seg& a, const Seg& b, int join, double width, double miter_limit,
                     S& out) {
  const Pt ea = last_pt(a), sb = first_pt(b);
  if (close_to(ea, sb)) return;
  const Seg bevel = mk_line(ea, sb);
  if (join == JOIN_BEVEL) {
    out.put(bevel);
    return;
  }
  Ln ta0, ta, tb, tb1;  // a's end tangent, b's start tangent
  tangents(a, &ta0, &ta);
  tangents(b, &tb, &tb1);
  double t0, t1;
  const bool hit = ln_cross(ta, tb, &t0, &t1);
  if (join == JOIN_MITER) {
    if (hit && !(in01(t0) && in01(t1))) {
      const Pt u = sub(ta.b, ta.a), v = sub(tb.a, tb.b);
      const double l = len(u) * len(v);
      if (!(l < kEps)) {
        const double c = dot(u, v) / l;  // Point::cos_between
        const double miter = width / sqrt((1.0 - c) / 2.0);
        if (miter < miter_limit) {
          const Pt p = ln_at(ta, t0);
          out.put(mk_line(ta.b, p));
          out.put(mk_line(p, tb.a));
          return;
        }
      }
    }
    out.put(bevel);
    return;
  }
  // round: a bevel where the tangents cross inside both, else an arc
  if (hit && in01(t0) && in01(t1)) {
    out.put(bevel);
    return;
  }
  const bool sweep = cross(sub(ta.b, ta.a), sub(sb, ea)) >= 0.0;
  if (!arc(ta.b, tb.a, width / 2.0, sweep, out)) out.put(bevel);
}

// Segment::line_cap (src/curve.rs:1050-1078) from the end of `a` to the
// start of `b`
template <class S>
RZ_HD void line_cap(const Seg& a, const Seg& b, const Style& st, S& out) {
  const Pt ea = last_pt(a), sb = first_pt(b);
  if (close_to(ea, sb)) return;
  if (st.cap == CAP_BUTT) {
    out.put(mk_line(ea, sb));
  } else if (st.cap == CAP_SQUARE) {
    Ln t0, from;
    tangents(a, &t0, &from);
    Pt tang;
    if (unit(sub(from.b, from.a), &tang)) {
      const Pt e0 = add(ea, scale(st.width / 2.0, tang));
      out.put(mk_line(ea, e0));
      const Pt e1 = add(e0, sub(sb, ea));
      out.put(mk_line(e0, e1));
      out.put(mk_line(e1, sb));
    }
  } else {
    line_join(a, b, JOIN_ROUND, st.width, st.miter_limit, out);
  }
}

// polyline_offset (src/curve.rs:1289-1347): the n control points moved onto
// the offset lines of the control polygon; runs of coincident points go to
// the crossing of the offset lines on either side. False when all points
// coincide.
RZ_HD bool polyline_offset(Pt* ps, int n, double dist) {
  if (n == 0) return true;
  bool have = false;
  Ln prev = ln(mkpt(0.0, 0.0), mkpt(0.0, 0.0));
  int at = 0;
  for (;;) {
    int rep = 1;
    for (int i = at; i < n - 1 && close_to(ps[i], ps[i + 1]); ++i) ++rep;
    if (at + rep >= n) break;
    at += rep;
    Ln next = ln(ps[at - 1], ps[at]);
    ln_offset(ps[at - 1], ps[at], dist, &next);  // the points differ: never fails
    Pt moved = next.a;
    double t0, t1;
    if (have && ln_cross(prev, next, &t0, &t1)) moved = ln_at(prev, t0);
    for (int i = at - rep; i < at; ++i) ps[i] = moved;
    prev = next;
    have = true;
  }
  if (!have) return false;
  for (int i = at; i < n; ++i) ps[i] = prev.b;
  return true;
}

// Quad::at / Cubic::at (src/curve.rs:438-447, 712-723)
RZ_HD Pt quad_at(const Seg& q, double t) {
  const double t1 = t, t_1 = 1.0 - t;
  const double t2 = t1 * t1, t_2 = t_1 * t_1;
  return add(add(scale(t_2, q.p[0]), scale(2.0 * t1 * t_1, q.p[1])), scale(t2, q.p[2]));
}
RZ_HD Pt cubic_at(const Seg& c, double t) {
  const double t1 = t, t_1 = 1.0 - t;
  const double t2 = t1 * t1, t_2 = t_1 * t_1;
  const double t3 = t2 * t1, t_3 = t_2 * t_1;
  return add(add(add(scale(t_3, c.p[0]), scale(3.0 * t1 * t_2, c.p[1])), scale(3.0 * t2 * t_1, c.p[2])),
             scale(t3, c.p[3]));
}

// The diagonal length of Quad::bbox / Cubic::bbox (src/curve.rs:553-561,
// 810-818; BBox::new / contains / extend, src/geometry.rs:552-632)
struct Box {
  double x0, y0, x1, y1;
};
RZ_HD Box box_of(Pt a, Pt b) {
  Box r;
  r.x0 = a.x <= b.x ? a.x : b.x;
  r.x1 = a.x <= b.x ? b.x : a.x;
  r.y0 = a.y <= b.y ? a.y : b.y;
  r.y1 = a.y <= b.y ? b.y : a.y;
  return r;
}
RZ_HD bool box_has(const Box& b, Pt p) { return b.x0 <= p.x && p.x <= b.x1 && b.y0 <= p.y && p.y <= b.y1; }
RZ_HD void box_add(Box& b, Pt p) {
  if (p.x < b.x0) b.x0 = p.x;
  else if (p.x > b.x1) b.x1 = p.x;
  if (p.y < b.y0) b.y0 = p.y;
  else if (p.y > b.y1) b.y1 = p.y;
}
RZ_HD double box_diag(const Box& b) { return len(mkpt(b.x0 - b.x1, b.y0 - b.y1)); }

// quadratic_solve (src/utils.rs:205-231): roots into r, their count returned
RZ_HD int solve2(double a, double b, double c, double r[2]) {
  if (fabs(a) < kEps) {
    if (fabs(b) > kEps) {
      r[0] = -c / b;
      return 1;
    }
    return 0;
  }
  const double disc = b * b - 4.0 * a * c;
  if (fabs(disc) < kEps) {
    r[0] = -b / (2.0 * a);
    return 1;
  }
  if (!(disc > 0.0)) return 0;
  const double sq = sqrt(disc);
  if (b >= 0.0) {
    const double m = -b - sq;
    r[0] = m / (2.0 * a);
    r[1] = 2.0 * c / m;
  } else {
    const double m = -b + sq;
    r[0] = 2.0 * c / m;
    r[1] = m / (2.0 * a);
  }
  return 2;
}

RZ_HD double quad_diag(const Seg& q) {
  Box b = box_of(q.p[0], q.p[2]);
  if (box_has(b, q.p[1])) return box_diag(b);
  // Quad::extremities (src/curve.rs:535-555)
  const Pt a = add(sub(q.p[2], scale(2.0, q.p[1])), q.p[0]);
  const Pt bb = sub(q.p[1], q.p[0]);
  if (fabs(a.x) > kEps) {
    const double t = -bb.x / a.x;
    if (in01(t)) box_add(b, quad_at(q, t));
  }
  if (fabs(a.y) > kEps) {
    const double t = -bb.y / a.y;
    if (in01(t)) box_add(b, quad_at(q, t));
  }
  return box_diag(b);
}

RZ_HD double cubic_diag(const Seg& c) {
  Box b = box_of(c.p[0], c.p[3]);
  if (box_has(b, c.p[1]) && box_has(b, c.p[2])) return box_diag(b);
  // Cubic::extremities (src/curve.rs:853-864)
  const Pt a = add(sub(add(scale(-1.0, c.p[0]), scale(3.0, c.p[1])), scale(3.0, c.p[2])),
                   scale(1.0, c.p[3]));
  const Pt bq = add(sub(scale(2.0, c.p[0]), scale(4.0, c.p[1])), scale(2.0, c.p[2]));
  const Pt cq = add(scale(-1.0, c.p[0]), c.p[1]);
  double r[2];
  int n = solve2(a.x, bq.x, cq.x, r);
  for (int i = 0; i < n; ++i)
    if (in01(r[i])) box_add(b, cubic_at(c, r[i]));
  n = solve2(a.y, bq.y, cq.y, r);
  for (int i = 0; i < n; ++i)
    if (in01(r[i])) box_add(b, cubic_at(c, r[i]));
  return box_diag(b);
}

// quad_offset_should_split (src/curve.rs:1349-1362)
RZ_HD bool quad_split_first(const Seg& q) {
  const Pt p0 = q.p[0], p1 = q.p[1], p2 = q.p[2];
  if (dot(sub(p0, p1), sub(p2, p1)) > 0.0) return true;
  const Pt m = add(add(p0, p1), p2);
  const Pt c_mass = mkpt(m.x / 3.0, m.y / 3.0);
  const double dist = len(sub(c_mass, quad_at(q, 0.5)));
  return quad_diag(q) * 0.1 < dist;
}

// cubic_offset_should_split (src/curve.rs:1414-1433)
RZ_HD bool cubic_split_first(const Seg& c) {
  const Pt p0 = c.p[0], p1 = c.p[1], p2 = c.p[2], p3 = c.p[3];
  if (dot(sub(p3, p0), sub(p2, p1)) < 0.0) return true;
  const double a0 = cross(sub(p3, p0), sub(p1, p0));
  const double a1 = cross(sub(p3, p0), sub(p2, p0));
  if (a0 * a1 < 0.0) return true;
  const Pt m = add(add(add(p0, p1), p2), p3);
  const Pt c_mass = mkpt(m.x / 4.0, m.y / 4.0);
  const double dist = len(sub(c_mass, cubic_at(c, 0.5)));
  return cubic_diag(c) * 0.1 < dist;
}

// Quad::split_at(0.5) (src/curve.rs:458-468) and Cubic::split
// (src/curve.rs:731-747), the halves the offset recursions take
RZ_HD void quad_halves(const Seg& q, Seg* l, Seg* r) {
  const double t = 0.5, t1 = t, t_1 = 1.0 - t;
  const double t2 = t1 * t1, t_2 = t_1 * t_1;
  const Pt mid = add(add(scale(t_2, q.p[0]), scale(2.0 * t1 * t_1, q.p[1])), scale(t2, q.p[2]));
  *l = mk(1, q.p[0], add(scale(t_1, q.p[0]), scale(t, q.p[1])), mid, mkpt(0.0, 0.0));
  *r = mk(1, mid, add(scale(t_1, q.p[1]), scale(t, q.p[2])), q.p[2], mkpt(0.0, 0.0));
}
RZ_HD void cubic_halves(const Seg& c, Seg* l, Seg* r) {
  const Pt p0 = c.p[0], p1 = c.p[1], p2 = c.p[2], p3 = c.p[3];
  const Pt mid = add(add(add(scale(0.125, p0), scale(0.375, p1)), scale(0.375, p2)), scale(0.125, p3));
  *l = mk(2, p0, add(scale(0.5, p0), scale(0.5, p1)),
          add(add(scale(0.25, p0), scale(0.5, p1)), scale(0.25, p2)), mid);
  *r = mk(2, mid, add(add(scale(0.25, p1), scale(0.5, p2)), scale(0.25, p3)),
          add(scale(0.5, p2), scale(0.5, p3)), p3);
}

// Curve::offset (src/curve.rs:275-277, 515-517, 825-827): the offset curves
// of one segment, in the reference's depth-first order (a split happens at
// most three levels deep, so a four-entry stack holds the pending halves).
// Between disconnected offset pieces of a cubic goes a round join of width
// 2 * dist (src/curve.rs:1393-1405).
template <class S>
RZ_HD void seg_offset(const Seg& s, double dist, S& out) {
  if (s.kind == 0) {
    Ln l;
    if (ln_offset(s.p[0], s.p[1], dist, &l)) out.put(mk_line(l.a, l.b));
    return;
  }
  Seg stack[4];
  int depth[4];
  int top = 1;
  stack[0] = s;
  depth[0] = 0;
  bool have_last = false;
  Seg last = s;
  while (top > 0) {
    --top;
    const Seg c = stack[top];
    const int d = depth[top];
    const bool split = s.kind == 1 ? quad_split_first(c) : cubic_split_first(c);
    if (split && d < 3) {
      Seg l, r;
      if (s.kind == 1) quad_halves(c, &l, &r);
      else cubic_halves(c, &l, &r);
      stack[top] = r;
      depth[top] = d + 1;
      stack[top + 1] = l;
      depth[top + 1] = d + 1;
      top += 2;
      continue;
    }
    Pt ps[4] = {c.p[0], c.p[1], c.p[2], c.p[3]};
    if (!polyline_offset(ps, s.kind + 2, dist)) continue;
    const Seg piece = mk(s.kind, ps[0], ps[1], ps[2], ps[3]);
    if (s.kind == 2) {
      if (have_last && !close_to(last_pt(last), first_pt(piece)))
        line_join(last, piece, JOIN_ROUND, dist * 2.0, 0.0, out);
      have_last = true;
      last = piece;
    }
    out.put(piece);
  }
}

// ---------------------------------------------------------------------------
// Work items
// ---------------------------------------------------------------------------

// The input path (rz_path layout) plus, per segment, its subpath.
struct Input {
  const uint8_t* kinds;
  const uint32_t* pt_off;
  const double* pts;
  const uint32_t* subs;    // [n_subpaths + 1]
  const uint8_t* closed;   // [n_subpaths]
  const uint32_t* sub_of;  // [n_segments]
  uint32_t n_items;        // 2 * n_segments
  Style st;
};

// What item g is: its segment (reversed on the way back), whether it joins
// with line_cap, and its run [run_b, run_e) (item indices), run id
// (2 * subpath + 0/1) and the run's closing (0 cap, 1 close forward,
// 2 close backward).
struct Item {
  uint32_t seg;
  bool back, cap;
  uint32_t run_b, run_e, run;
  int closing;
  uint32_t sub;
};

RZ_HD Item item_of(const Input& in, uint32_t g) {
  Item it;
  const uint32_t i = in.sub_of[g >> 1];
  const uint32_t b = in.subs[i], n = in.subs[i + 1] - b;
  const uint32_t j = g - 2 * b;
  const bool closed = in.closed[i] != 0;
  it.sub = i;
  it.back = j >= n;
  it.seg = it.back ? b + (2 * n - 1 - j) : b + j;
  it.cap = !closed && j == n;
  if (closed) {
    it.run_b = 2 * b + (it.back ? n : 0);
    it.run_e = it.run_b + n;
    it.run = 2 * i + (it.back ? 1 : 0);
    it.closing = it.back ? 2 : 1;
  } else {
    it.run_b = 2 * b;
    it.run_e = 2 * b + 2 * n;
    it.run = 2 * i;
    it.closing = 0;
  }
  return it;
}

RZ_HD Seg load_seg(const Input& in, uint32_t k) {
  const int kind = in.kinds[k];
  const double* q = in.pts + 2 * (size_t)in.pt_off[k];
  Seg s = mk(kind, mkpt(q[0], q[1]), mkpt(q[2], q[3]), mkpt(0.0, 0.0), mkpt(0.0, 0.0));
  if (kind >= 1) s.p[2] = mkpt(q[4], q[5]);
  if (kind == 2) s.p[3] = mkpt(q[6], q[7]);
  return s;
}

// The item's segment, in its direction
RZ_HD Seg item_seg(const Input& in, const Item& it) {
  const Seg s = load_seg(in, it.seg);
  return it.back ? reversed(s) : s;
}

// The join in front of an item's curves (item non-empty, an earlier curve in
// its run): `prev` is that curve, `first` the item's first.
template <class S>
RZ_HD void item_join(const Input& in, const Item& it, const Seg& prev, const Seg& first, S& out) {
  if (it.cap) line_cap(prev, first, in.st, out);
  else line_join(prev, first, in.st.join, in.st.width, in.st.miter_limit, out);
}

// The end of a run with any curves: `first` and `last` are its first and
// last curves (stroke_close, src/path.rs:708-732, or the final line_cap of an
// open subpath, src/path.rs:401-408).
template <class S>
RZ_HD void run_end(const Input& in, const Item& it, const Seg& first, const Seg& last, S& out) {
  if (it.closing == 0) {
    line_cap(last, first, in.st, out);
    return;
  }
  const uint32_t b = in.subs[it.sub], e = in.subs[it.sub + 1];
  const Pt s0 = first_pt(load_seg(in, b)), s1 = last_pt(load_seg(in, e - 1));
  Ln off;
  const bool ok = it.closing == 1 ? ln_offset(s1, s0, in.st.width / 2.0, &off)
                                  : ln_offset(s0, s1, in.st.width / 2.0, &off);
  if (ok && len(sub(off.a, off.b)) * 100.0 > in.st.width) {
    const Seg cs = mk_line(off.a, off.b);
    line_join(last, cs, in.st.join, in.st.width, in.st.miter_limit, out);
    out.put(cs);
    line_join(cs, first, in.st.join, in.st.width, in.st.miter_limit, out);
  } else {
    line_join(last, first, in.st.join, in.st.width, in.st.miter_limit, out);
  }
}

// An item's output into `o`: the join in front of its curves, the curves
// themselves when `curves` (a count pass adds their size from the pieces
// pass instead), and, for the last item of a run with curves, the run's end.
// `q` answers from the pieces pass: pieces(g) (curves of item g), prev(g) /
// prev_incl(g) (nearest item < g / <= g with curves, or -1), run_has(r),
// run_first(r) (first item with curves of run r), first(i) / last(i) (first
// and last curve of item i).
template <class Q, class S>
RZ_HD void item_output(const Input& in, const Q& q, uint32_t g, const Item& it, bool curves,
                       S& o) {
  if (q.pieces(g)) {
    const int32_t p = q.prev(g);
    if (p >= (int32_t)it.run_b) item_join(in, it, q.last(p), q.first(g), o);
    if (curves) seg_offset(item_seg(in, it), in.st.width / 2.0, o);
  }
  if (g == it.run_e - 1 && q.run_has(it.run))
    run_end(in, it, q.first(q.run_first(it.run)), q.last(q.prev_incl(g)), o);
}

// Sinks: counting (with the first and last segment) and writing (rz_path
// arrays: kinds, point offsets, x/y pairs).
struct CountSink {
  uint32_t n = 0, pts = 0;
  Seg first, last;
  RZ_HD void put(const Seg& s) {
    if (n == 0) first = s;
    last = s;
    ++n;
    pts += (uint32_t)s.kind + 2u;
  }
};
struct WriteSink {
  uint8_t* kinds;
  uint32_t* pt_off;
  double* pts;
  uint32_t seg, pt;  // next segment / point index
  RZ_HD void put(const Seg& s) {
    kinds[seg] = (uint8_t)s.kind;
    pt_off[seg] = pt;
    for (int i = 0; i < s.kind + 2; ++i) {
      pts[2 * (size_t)(pt + i)] = s.p[i].x;
      pts[2 * (size_t)(pt + i) + 1] = s.p[i].y;
    }
    ++seg;
    pt += (uint32_t)s.kind + 2u;
  }
};

}  // namespace rzs

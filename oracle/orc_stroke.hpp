// Oracle: CPU restatement of Path::stroke (src/path.rs:370-415, 692-732) with
// the segment offsetting, joins and caps it uses (src/curve.rs:977-1078,
// 1279-1433) and the line helpers (src/curve.rs:195-229).
//
// TEST INFRASTRUCTURE ONLY (see orc_core.hpp): the checker the HIP stroke is
// compared against. Sequential, in the reference's order, every expression in
// the reference's operation order (compiled with -ffp-contract=off).
#pragma once
#include "orc_core.hpp"

namespace orc {

// src/path.rs:78-136
enum class LineJoin { Miter = 0, Bevel = 1, Round = 2 };
enum class LineCap { Butt = 0, Square = 1, Round = 2 };
struct StrokeStyle {
  double width = 0.0;
  LineJoin line_join = LineJoin::Miter;
  double miter_limit = 4.0;  // LineJoin::Miter's argument (default Miter(4.0))
  LineCap line_cap = LineCap::Butt;
};

// Point::normal / normalize (src/geometry.rs:169-183)
inline Point normal_of(Point p) { return P(p.y, -p.x); }
inline bool normalize(Point p, Point* out) {
  const double len = length(p);
  if (len < EPSILON) return false;
  *out = P(p.x / len, p.y / len);
  return true;
}

inline Segment seg_of(const Line& l) { return Segment::line(l.p0, l.p1); }
inline Line mk_line(Point a, Point b) { return Line{a, b}; }
// Line::at (src/curve.rs:250-253)
inline Point line_at(const Line& l, double t) { return (1.0 - t) * l.p0 + t * l.p1; }
// Line::length (src/curve.rs:185-188)
inline double line_length(const Line& l) { return length(l.p0 - l.p1); }
// Line::direction (src/curve.rs:227-229)
inline Point direction(const Line& l) { return l.p1 - l.p0; }

// Line::intersect (src/curve.rs:204-215)
inline bool line_intersect(const Line& a, const Line& b, double* t0, double* t1) {
  const double x1 = a.p0.x, y1 = a.p0.y, x2 = a.p1.x, y2 = a.p1.y;
  const double x3 = b.p0.x, y3 = b.p0.y, x4 = b.p1.x, y4 = b.p1.y;
  const double det = (x4 - x3) * (y1 - y2) - (x1 - x2) * (y4 - y3);
  if (std::fabs(det) < EPSILON) return false;
  *t0 = ((y3 - y4) * (x1 - x3) + (x4 - x3) * (y1 - y3)) / det;
  *t1 = ((y1 - y2) * (x1 - x3) + (x2 - x1) * (y1 - y3)) / det;
  return true;
}
// Line::intersect_point (src/curve.rs:217-224)
inline bool line_intersect_point(const Line& a, const Line& b, Point* out) {
  double t0, t1;
  if (!line_intersect(a, b, &t0, &t1)) return false;
  if (!(t0 >= 0.0 && t0 <= 1.0 && t1 >= 0.0 && t1 <= 1.0)) return false;
  *out = line_at(a, t0);
  return true;
}

// Segment::ends: tangent lines at the two ends (Line :195-197, Quad :377-389,
// Cubic :647-668)
inline void seg_ends(const Segment& s, Line* start, Line* end) {
  if (s.kind == SEG_LINE) {
    *start = *end = mk_line(s.p[0], s.p[1]);
  } else if (s.kind == SEG_QUAD) {
    const Line a = mk_line(s.p[0], s.p[1]), b = mk_line(s.p[1], s.p[2]);
    if (is_close_to(s.p[0], s.p[1])) {
      *start = *end = b;
    } else if (is_close_to(s.p[1], s.p[2])) {
      *start = *end = a;
    } else {
      *start = a;
      *end = b;
    }
  } else {
    int i0 = 0;
    for (int i = 0; i < 3; ++i)
      if (!is_close_to(s.p[i], s.p[i + 1])) {
        i0 = i;
        break;
      }
    int i1 = 0;
    for (int i = 3; i >= 1; --i)
      if (!is_close_to(s.p[i], s.p[i - 1])) {
        i1 = i;
        break;
      }
    // all four points close: the reference indexes ps[-1] (a panic); never
    // reached from stroking, whose offset curves have distinct points
    if (i1 == 0) i1 = 1;
    *start = mk_line(s.p[i0], s.p[i0 + 1]);
    *end = mk_line(s.p[i1 - 1], s.p[i1]);
  }
}

// Curve::reverse (Line :279-282, Quad :519-522, Cubic :829-832)
inline Segment seg_reverse(const Segment& s) {
  Segment r = s;
  const int n = s.npts();
  for (int i = 0; i < n; ++i) r.p[i] = s.p[n - 1 - i];
  return r;
}

// Segment::line_join (src/curve.rs:977-1047); at most four segments
inline void line_join(const Segment& self, const Segment& other, const StrokeStyle& style,
                      std::vector<Segment>& out) {
  if (is_close_to(self.end(), other.start())) return;
  const Line bevel = mk_line(self.end(), other.start());
  switch (style.line_join) {
    case LineJoin::Bevel:
      out.push_back(seg_of(bevel));
      return;
    case LineJoin::Miter: {
      Line s0, start, end, e1;
      seg_ends(self, &s0, &start);
      seg_ends(other, &end, &e1);
      double t0, t1;
      if (!line_intersect(start, end, &t0, &t1)) {
        out.push_back(seg_of(bevel));
        return;
      }
      if (t0 >= 0.0 && t0 <= 1.0 && t1 >= 0.0 && t1 <= 1.0) {  // the ends intersect
        out.push_back(seg_of(bevel));
        return;
      }
      const Point p0 = start.p1 - start.p0;
      const Point p1 = end.p0 - end.p1;
      double c;
      if (cos_between(p0, p1, &c)) {
        const double miter_length = style.width / std::sqrt((1.0 - c) / 2.0);
        if (miter_length < style.miter_limit) {
          const Point p = line_at(start, t0);
          out.push_back(Segment::line(start.p1, p));
          out.push_back(Segment::line(p, end.p0));
          return;
        }
      }
      out.push_back(seg_of(bevel));
      return;
    }
    case LineJoin::Round: {
      Line s0, start, end, e1;
      seg_ends(self, &s0, &start);
      seg_ends(other, &end, &e1);
      Point ip;
      if (line_intersect_point(start, end, &ip)) {
        out.push_back(seg_of(bevel));
        return;
      }
      const bool sweep_flag = cross(direction(start), direction(bevel)) >= 0.0;
      const double radius = style.width / 2.0;
      EllipArc arc;
      if (EllipArc::new_param(start.p1, end.p0, radius, radius, 0.0, false, sweep_flag, &arc))
        arc.to_cubics(out);
      else
        out.push_back(seg_of(bevel));
      return;
    }
  }
}

// Segment::line_cap (src/curve.rs:1050-1078)
inline void line_cap(const Segment& self, const Segment& other, const StrokeStyle& style,
                     std::vector<Segment>& out) {
  if (is_close_to(self.end(), other.start())) return;
  const Line butt = mk_line(self.end(), other.start());
  switch (style.line_cap) {
    case LineCap::Butt:
      out.push_back(seg_of(butt));
      return;
    case LineCap::Square: {
      Line s0, from;
      seg_ends(self, &s0, &from);
      Point tang;
      if (normalize(direction(from), &tang)) {
        const Line l0 = mk_line(self.end(), self.end() + style.width / 2.0 * tang);
        out.push_back(seg_of(l0));
        const Line l1 = mk_line(l0.p1, l0.p1 + direction(butt));
        out.push_back(seg_of(l1));
        out.push_back(Segment::line(l1.p1, other.start()));
      }
      return;
    }
    case LineCap::Round: {
      StrokeStyle round = style;
      round.line_join = LineJoin::Round;
      line_join(self, other, round, out);
      return;
    }
  }
}

// line_offset (src/curve.rs:1283-1287)
inline bool line_offset(const Line& l, double dist, Line* out) {
  Point n;
  if (!normalize(normal_of(l.p1 - l.p0), &n)) return false;
  const Point off = dist * n;
  *out = mk_line(l.p0 + off, l.p1 + off);
  return true;
}

// polyline_offset (src/curve.rs:1294-1347): moves the control points of a
// curve onto the offset lines of its control polygon (repeated points onto
// the intersection of the neighbouring offset lines)
inline bool polyline_offset(Point* ps, int n, double dist) {
  if (n == 0) return true;
  bool have_prev = false;
  Line prev{};
  int index = 0;
  for (;;) {
    int repeats = 1;
    for (int i = index; i < n - 1; ++i) {
      if (!is_close_to(ps[i], ps[i + 1])) break;
      ++repeats;
    }
    if (index + repeats >= n) break;
    index += repeats;
    Line next;
    if (!line_offset(mk_line(ps[index - 1], ps[index]), dist, &next))
      throw std::logic_error("polyline implementation error");
    Point point = next.p0;
    if (have_prev) {
      double t0, t1;
      if (line_intersect(prev, next, &t0, &t1)) point = line_at(prev, t0);
    }
    for (int i = index - repeats; i < index; ++i) ps[i] = point;
    prev = next;
    have_prev = true;
  }
  if (!have_prev) return false;
  for (int i = index; i < n; ++i) ps[i] = prev.p1;
  return true;
}

// quad_offset_should_split (src/curve.rs:1349-1362)
inline bool quad_offset_should_split(const Segment& q) {
  const Point p0 = q.p[0], p1 = q.p[1], p2 = q.p[2];
  if (dot(p0 - p1, p2 - p1) > 0.0) return true;
  const Point c_mass = (p0 + p1 + p2) / 3.0;
  const Point c_mid = q.at(0.5);
  const double dist = length(c_mass - c_mid);
  const BBox bb = q.bbox(nullptr);
  const double bbox_diag = length(bb.min - bb.max);
  return bbox_diag * 0.1 < dist;
}

// quad_offset_rec (src/curve.rs:1365-1376)
inline void quad_offset_rec(const Segment& q, double dist, std::vector<Segment>& out, int depth) {
  if (quad_offset_should_split(q) && depth < 3) {
    Segment a, b;
    q.split(&a, &b);  // Quad::split_at(0.5), not the shortcut split()
    quad_offset_rec(a, dist, out, depth + 1);
    quad_offset_rec(b, dist, out, depth + 1);
  } else {
    Point pts[3] = {q.p[0], q.p[1], q.p[2]};
    if (polyline_offset(pts, 3, dist)) out.push_back(Segment::quad(pts[0], pts[1], pts[2]));
  }
}

// cubic_offset_should_split (src/curve.rs:1414-1433)
inline bool cubic_offset_should_split(const Segment& c) {
  const Point p0 = c.p[0], p1 = c.p[1], p2 = c.p[2], p3 = c.p[3];
  if (dot(p3 - p0, p2 - p1) < 0.0) return true;
  const double a0 = cross(p3 - p0, p1 - p0);
  const double a1 = cross(p3 - p0, p2 - p0);
  if (a0 * a1 < 0.0) return true;
  const Point c_mass = (p0 + p1 + p2 + p3) / 4.0;
  const Point c_mid = c.at(0.5);
  const double dist = length(c_mass - c_mid);
  const BBox bb = c.bbox(nullptr);
  const double bbox_diag = length(bb.min - bb.max);
  return bbox_diag * 0.1 < dist;
}

// cubic_offset_rec (src/curve.rs:1379-1411); returns whether *last is set
inline bool cubic_offset_rec(const Segment& c, bool have_last, Segment* last, double dist,
                             std::vector<Segment>& out, int depth) {
  if (cubic_offset_should_split(c) && depth < 3) {
    Segment a, b;
    c.split_fast(&a, &b);
    have_last = cubic_offset_rec(a, have_last, last, dist, out, depth + 1);
    return cubic_offset_rec(b, have_last, last, dist, out, depth + 1);
  }
  Point pts[4] = {c.p[0], c.p[1], c.p[2], c.p[3]};
  if (!polyline_offset(pts, 4, dist)) return have_last;
  const Segment result = Segment::cubic(pts[0], pts[1], pts[2], pts[3]);
  if (have_last && !is_close_to(last->end(), result.start())) {
    StrokeStyle st;
    st.width = dist * 2.0;
    st.line_join = LineJoin::Round;
    st.line_cap = LineCap::Round;
    line_join(*last, result, st, out);
  }
  out.push_back(result);
  *last = result;
  return true;
}

// Curve::offset (Line :275-277, Quad :515-517, Cubic :825-827)
inline void seg_offset(const Segment& s, double dist, std::vector<Segment>& out) {
  if (s.kind == SEG_LINE) {
    Line l;
    if (line_offset(mk_line(s.p[0], s.p[1]), dist, &l)) out.push_back(seg_of(l));
  } else if (s.kind == SEG_QUAD) {
    quad_offset_rec(s, dist, out, 0);
  } else {
    Segment last;
    cubic_offset_rec(s, false, &last, dist, out, 0);
  }
}

// stroke_segment (src/path.rs:692-706): the segment's offset curves, joined to
// what precedes them (with line_join, or line_cap for `cap`)
inline void stroke_segment(std::vector<Segment>& segs, const Segment& s, const StrokeStyle& style,
                           bool cap) {
  const size_t offset = segs.size();
  seg_offset(s, style.width / 2.0, segs);
  if (offset != 0 && offset < segs.size()) {
    const Segment src = segs[offset - 1], dst = segs[offset];
    std::vector<Segment> join;
    if (cap) line_cap(src, dst, style, join);
    else line_join(src, dst, style, join);
    segs.insert(segs.begin() + (ptrdiff_t)offset, join.begin(), join.end());
  }
}

// stroke_close (src/path.rs:708-732)
inline void stroke_close(Point sub_start, Point sub_end, std::vector<Segment>& segs,
                         const StrokeStyle& style, bool forward) {
  if (segs.empty()) return;
  const Segment first = segs.front(), last = segs.back();
  const Line close = forward ? mk_line(sub_end, sub_start) : mk_line(sub_start, sub_end);
  Line off;
  if (line_offset(close, style.width / 2.0, &off) && line_length(off) * 100.0 > style.width) {
    const Segment cs = seg_of(off);
    line_join(last, cs, style, segs);
    segs.push_back(cs);
    line_join(cs, first, style, segs);
  } else {
    line_join(last, first, style, segs);
  }
}

// Path::push (src/path.rs:336-346)
inline void path_push(Path& p, const std::vector<Segment>& segs, bool closed) {
  if (segs.empty()) return;
  if (p.subpaths.empty()) p.subpaths.push_back(0);
  p.segments.insert(p.segments.end(), segs.begin(), segs.end());
  p.subpaths.push_back(p.segments.size());
  p.closed.push_back(closed ? 1 : 0);
}

// Path::stroke (src/path.rs:370-415)
inline Path stroke(const Path& path, const StrokeStyle& style) {
  Path result;
  std::vector<Segment> segs;
  for (size_t i = 0; i < path.nsubpaths(); ++i) {
    const size_t b = path.subpaths[i], e = path.subpaths[i + 1];
    if (b == e) continue;
    const bool closed = path.closed[i] != 0;
    const Point sub_start = path.segments[b].start(), sub_end = path.segments[e - 1].end();
    for (size_t k = b; k < e; ++k) stroke_segment(segs, path.segments[k], style, false);
    size_t k = e;  // the backward walk: segments e-1 .. b, reversed
    if (closed) {
      stroke_close(sub_start, sub_end, segs, style, true);
      path_push(result, segs, true);
      segs.clear();
    } else {
      --k;
      stroke_segment(segs, seg_reverse(path.segments[k]), style, true);
    }
    while (k > b) {
      --k;
      stroke_segment(segs, seg_reverse(path.segments[k]), style, false);
    }
    if (closed) {
      stroke_close(sub_start, sub_end, segs, style, false);
      path_push(result, segs, true);
      segs.clear();
    } else {
      if (!segs.empty()) {
        const Segment last = segs.back(), first = segs.front();
        line_cap(last, first, style, segs);
      }
      path_push(result, segs, true);
      segs.clear();
    }
  }
  return result;
}

}  // namespace orc

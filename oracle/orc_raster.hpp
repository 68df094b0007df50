// Oracle: CPU restatement of the reference crate's rasterizers
// (src/rasterize.rs) -- SignedDifferenceRasterizer and ActiveEdgeRasterizer.
//
// TEST INFRASTRUCTURE ONLY (see orc_core.hpp). Images are row-major f64 with
// an explicit Shape (src/image.rs:6-56) so strided views behave as in the
// reference.
#pragma once
#include <algorithm>
#include "orc_core.hpp"

namespace orc {

enum FillRule { NONZERO = 0, EVENODD = 1 };

// src/path.rs:32-46
inline double alpha_from_winding(FillRule rule, double winding) {
  if (rule == EVENODD) return std::fabs(rem_euclid(winding + 1.0, 2.0) - 1.0);
  double value = std::fabs(winding);
  if (value >= 1.0) return 1.0;
  if (value < 1e-6) return 0.0;
  return value;
}

// src/image.rs:6-56
struct Shape {
  size_t start, width, height, row_stride, col_stride;
  size_t offset(size_t row, size_t col) const { return start + row * row_stride + col * col_stride; }
  static Shape simple(size_t height, size_t width) { return Shape{0, width, height, width, 1}; }
};

// src/image.rs:588-605
inline Shape view_shape(Shape s, size_t row_min, size_t row_max, size_t col_min, size_t col_max) {
  row_min = std::min(row_min, s.height);
  row_max = std::min(std::max(row_max, row_min), s.height);
  col_min = std::min(col_min, s.width);
  col_max = std::min(std::max(col_max, col_min), s.width);
  Shape r = s;
  r.start = s.offset(row_min, col_min);
  r.width = col_max - col_min;
  r.height = row_max - row_min;
  return r;
}

struct ImageF64 {
  double* data;
  size_t len;
  Shape shape;
};

// src/rasterize.rs:922-937
inline void split_at_zero_x(Line line, Line* out, Line* rest, bool* has_rest) {
  Point p0 = line.p0, p1 = line.p1;
  *has_rest = false;
  if (p0.x >= 0.0 && p1.x >= 0.0) {
    *out = line;
    return;
  }
  if (p0.x <= 0.0 && p1.x <= 0.0) {
    *out = Line{P(0.0, p0.y), P(0.0, p1.y)};
    return;
  }
  double t = p0.x / (p0.x - p1.x);
  Point mid = (1.0 - t) * p0 + t * p1;  // Line::at
  *has_rest = true;
  if (p0.x < 0.0) {
    *out = Line{mid, p1};
    *rest = Line{P(0.0, p0.y), mid};
  } else {
    *out = Line{p0, mid};
    *rest = Line{mid, P(0.0, p1.y)};
  }
}

// src/rasterize.rs:365-470: the deposits of one line into an image of
// `shape` over a buffer of `len` elements, each handed to add(offset, value)
// with the offset the reference writes (the one guarded write included only
// when it stays inside the buffer, as `offset < data.len()` does).
template <typename Add>
inline void sd_line_deposits(const Shape& shape, size_t len, Line line, Add&& add) {
  Point p0 = line.p0, p1 = line.p1;
  const double width = (double)shape.width - 1.0;
  if (p0.x > width || p1.x > width) {
    if (p0.x > width && p1.x > width) {
      line = Line{P(width - 0.001, p0.y), P(width - 0.001, p1.y)};
    } else {
      double t = (p0.x - width) / (p0.x - p1.x);
      Point mid = P(width, (1.0 - t) * p0.y + t * p1.y);
      if (p0.x < width) line = Line{p0, mid};
      else line = Line{mid, p1};
    }
  }
  Line rest;
  bool has_rest;
  split_at_zero_x(line, &line, &rest, &has_rest);
  if (has_rest) sd_line_deposits(shape, len, rest, add);

  p0 = line.p0;
  p1 = line.p1;
  const size_t stride = shape.col_stride;

  if (std::fabs(p0.y - p1.y) < EPSILON) return;
  double dir;
  if (p0.y < p1.y) {
    dir = 1.0;
  } else {
    dir = -1.0;
    std::swap(p0, p1);
  }
  double dxdy = (p1.x - p0.x) / (p1.y - p0.y);
  size_t y = (size_t)std::fmax(p0.y, 0.0);  // saturating cast of a non-negative value
  double x = p0.y < 0.0 ? p0.x - p0.y * dxdy : p0.x;
  double x_next = x;
  double yend_f = std::fmax(std::ceil(p1.y), 0.0);
  size_t yend = std::min(shape.height, (size_t)yend_f);
  for (; y < yend; ++y) {
    x = x_next;
    size_t row_offset = shape.offset(y, 0);
    double dy = std::fmin((double)(y + 1), p1.y) - std::fmax((double)y, p0.y);
    double d = dir * dy;
    x_next = x + dxdy * dy;
    double x0, x1;
    if (x < x_next) { x0 = x; x1 = x_next; } else { x0 = x_next; x1 = x; }
    double x0_floor = std::fmax(std::floor(x0), 0.0);
    int x0i = (int)x0_floor;
    double x1_ceil = std::fmin(std::ceil(x1), width);
    int x1i = (int)x1_ceil;
    if (x1i <= x0i + 1) {
      double xmf = 0.5 * (x + x_next) - x0_floor;
      add(row_offset + (size_t)x0i * stride, d * (1.0 - xmf));
      size_t offset = row_offset + (size_t)(x0i + 1) * stride;
      if (offset < len) add(offset, d * xmf);
    } else {
      double s = 1.0 / (x1 - x0);
      double x0f = x0 - x0_floor;
      double x1f = x1 - x1_ceil + 1.0;
      double a0 = 0.5 * s * (1.0 - x0f) * (1.0 - x0f);
      double am = 0.5 * s * x1f * x1f;
      add(row_offset + (size_t)x0i * stride, d * a0);
      if (x1i == x0i + 2) {
        add(row_offset + (size_t)(x0i + 1) * stride, d * (1.0 - a0 - am));
      } else {
        double a1 = s * (1.5 - x0f);
        add(row_offset + (size_t)(x0i + 1) * stride, d * (a1 - a0));
        for (int xi = x0i + 2; xi < x1i - 1; ++xi) add(row_offset + (size_t)xi * stride, d * s);
        double a2 = a1 + (double)(x1i - x0i - 3) * s;
        add(row_offset + (size_t)(x1i - 1) * stride, d * (1.0 - a2 - am));
      }
      add(row_offset + (size_t)x1i * stride, d * am);
    }
  }
}

inline void signed_difference_line(ImageF64 img, Line line) {
  double* data = img.data;
  sd_line_deposits(img.shape, img.len, line, [data](size_t off, double v) { data[off] += v; });
}

// src/rasterize.rs:473-507
inline void signed_difference_to_mask(ImageF64 img, FillRule rule) {
  const Shape s = img.shape;
  for (size_t y = 0; y < s.height; ++y) {
    double acc = 0.0;
    for (size_t x = 0; x < s.width; ++x) {
      size_t off = s.offset(y, x);
      acc += img.data[off];
      if (rule == NONZERO) {
        double value = std::fabs(acc);
        img.data[off] = value > 1.0 ? 1.0 : (value < 1e-6 ? 0.0 : value);
      } else {
        img.data[off] = std::fabs(rem_euclid(acc + 1.0, 2.0) - 1.0);
      }
    }
  }
}

// SignedDifferenceRasterizer::mask (src/rasterize.rs:298-311). `img` is
// assumed zeroed, as Path::mask documents (src/path.rs:509-512).
inline void sd_mask(const Path& path, const Transform& tr, double flatness, FillRule rule,
                    ImageF64 img) {
  flatten(path, tr, flatness, true, [&](const Line& l) { signed_difference_line(img, l); });
  signed_difference_to_mask(img, rule);
}

struct Pixel {
  size_t x, y;
  double alpha;
};

// SignedDifferenceRasterizer::mask_iter (src/rasterize.rs:313-355): a
// (W+1)-wide scratch image, then lazily accumulated rows; pixels with
// |alpha| < 1e-6 are not produced.
template <typename Sink>
inline void sd_mask_iter(const Path& path, const Transform& tr, double flatness, FillRule rule,
                         size_t width, size_t height, Sink&& sink) {
  if (width == 0 || height == 0) return;
  size_t w1 = width + 1;
  std::vector<double> scratch(w1 * height, 0.0);
  ImageF64 img{scratch.data(), scratch.size(), Shape::simple(height, w1)};
  flatten(path, tr, flatness, true, [&](const Line& l) { signed_difference_line(img, l); });
  double winding = 0.0;
  for (size_t index = 0; index < scratch.size(); ++index) {
    size_t y = index / w1;
    size_t x = index - y * w1;
    if (x == 0) winding = 0.0;
    else if (x >= width) continue;
    winding += scratch[index];
    double alpha = alpha_from_winding(rule, winding);
    if (std::fabs(alpha) < 1e-6) continue;
    sink(Pixel{x, y, alpha});
  }
}

// sd_mask_iter restricted to the rows [r0, r1) of the image: the same pixels
// with the same values. Every line still runs its arithmetic over all of its
// rows (x advances row by row), but only deposits that land in the window's
// rows of the (W+1)-wide scratch are kept (a deposit one column past a row's
// end lands in the next row, as in the reference's buffer). Pixel rows are
// reported in image coordinates.
template <typename Sink>
inline void sd_mask_iter_rows(const Path& path, const Transform& tr, double flatness,
                              FillRule rule, size_t width, size_t height, size_t r0, size_t r1,
                              Sink&& sink) {
  r1 = std::min(r1, height);
  if (width == 0 || height == 0 || r0 >= r1) return;
  const size_t w1 = width + 1;
  const Shape full = Shape::simple(height, w1);
  const size_t len = w1 * height, lo = r0 * w1, hi = r1 * w1;
  std::vector<double> scratch(hi - lo, 0.0);
  double* sp = scratch.data();
  flatten(path, tr, flatness, true, [&](const Line& l) {
    const double ylo = std::fmin(l.p0.y, l.p1.y), yhi = std::fmax(l.p0.y, l.p1.y);
    if (yhi + 1.0 < (double)r0 || ylo >= (double)r1) return;  // no deposit in the window
    sd_line_deposits(full, len, l, [&](size_t off, double v) {
      if (off >= lo && off < hi) sp[off - lo] += v;
    });
  });
  double winding = 0.0;
  for (size_t index = 0; index < scratch.size(); ++index) {
    size_t y = index / w1;
    size_t x = index - y * w1;
    if (x == 0) winding = 0.0;
    else if (x >= width) continue;
    winding += scratch[index];
    double alpha = alpha_from_winding(rule, winding);
    if (std::fabs(alpha) < 1e-6) continue;
    sink(Pixel{x, r0 + y, alpha});
  }
}

// ---------------------------------------------------------------------------
// ActiveEdgeRasterizer (src/rasterize.rs:517-920)
// ---------------------------------------------------------------------------
inline double next_ceil(double v) { return std::floor(v) + 1.0; }

struct AEEdge {
  Line line;
  double dxdy;
  bool has_dydx;
  double dydx;
  double dir;
  size_t row;
};

// src/rasterize.rs:760-791
inline bool ae_edge_new(Line line, AEEdge* out) {
  Point p0 = line.p0, p1 = line.p1;
  if (std::fabs(p0.y - p1.y) < EPSILON) return false;
  double dir = 1.0;
  if (!(p0.y <= p1.y)) {
    dir = -1.0;
    std::swap(p0, p1);
  }
  double dx = p1.x - p0.x, dy = p1.y - p0.y;
  double dxdy = dx / dy;
  bool has_dydx = std::fabs(dxdy) > EPSILON_SQRT;
  double dydx = has_dydx ? dy / dx : 0.0;
  if (p0.y < 0.0) p0 = P(p0.x - p0.y * dxdy, 0.0);
  *out = AEEdge{Line{p0, p1}, dxdy, has_dydx, dydx, dir, (size_t)std::floor(p0.y)};
  return true;
}

struct AERowIter {
  bool has_line;
  Line line;
  bool has_rem;
  double reminder;
  bool has_dydx;
  double dydx;
  size_t column;
  double dir;
  // src/rasterize.rs:856-875
  bool next(double* out) {
    if (has_line) {
      has_line = false;
      Point p0 = line.p0, p1 = line.p1;
      double x_ceil = next_ceil(p0.x);
      double x = std::fmin(x_ceil, p1.x);
      double h = has_dydx ? dydx * (x - p0.x) : std::fmin(next_ceil(p0.y), p1.y) - p0.y;
      double y = p0.y + h;
      double s0 = (2.0 * x_ceil - x - p0.x) * h / 2.0;
      double s1 = h - s0;
      if (p1.x > x) {
        line = Line{P(x, y), p1};
        has_line = true;
      }
      double s_prev = has_rem ? reminder : 0.0;
      has_rem = true;
      reminder = dir * s1;
      *out = dir * s0 + s_prev;
      return true;
    }
    if (has_rem) {
      has_rem = false;
      *out = reminder;
      return true;
    }
    return false;
  }
};

// src/rasterize.rs:820-850
inline bool ae_row_iter_new(AEEdge* edge, AERowIter* it) {
  Point p0 = edge->line.p0, p1 = edge->line.p1;
  if (p1.y - p0.y < EPSILON) return false;
  double y_split = std::fmin(next_ceil(p0.y), p1.y);
  double x_split = p0.x + edge->dxdy * (y_split - p0.y);
  Point ps = P(x_split, y_split);
  edge->line = Line{ps, p1};
  double dir;
  Line l;
  if (p0.x <= ps.x) { dir = edge->dir; l = Line{p0, ps}; }
  else { dir = -edge->dir; l = Line{ps, p0}; }
  *it = AERowIter{true, l, false, 0.0, edge->has_dydx, edge->dydx, (size_t)l.p0.x, dir};
  return true;
}

// ActiveEdgeIter (src/rasterize.rs:578-732)
template <typename Sink>
inline void ae_iter(size_t width, size_t height, FillRule rule, const std::vector<Line>& lines,
                    Sink&& sink) {
  std::vector<AEEdge> inactive;
  for (const Line& l0 : lines) {
    Line a, b;
    bool hb;
    split_at_zero_x(l0, &a, &b, &hb);
    AEEdge e;
    if (ae_edge_new(a, &e) && e.row < height) inactive.push_back(e);
    if (hb && ae_edge_new(b, &e) && e.row < height) inactive.push_back(e);
  }
  // sort by row descending (the unstable order of equal rows never matters:
  // equal-row edges are all activated together)
  std::stable_sort(inactive.begin(), inactive.end(),
                   [](const AEEdge& l, const AEEdge& r) { return l.row > r.row; });
  std::vector<AEEdge> active;
  std::vector<AERowIter> iters_inactive, iters_active;
  double winding = 0.0;
  size_t column = 0, row = 0;

  auto next_row = [&]() {
    iters_inactive.clear();
    iters_active.clear();
    std::vector<AEEdge> keep;
    for (AEEdge& e : active) {
      AERowIter it;
      if (ae_row_iter_new(&e, &it)) {
        iters_inactive.push_back(it);
        keep.push_back(e);
      }
    }
    active.swap(keep);
    // partition point: edges with row > current row stay inactive
    size_t idx = 0;
    while (idx < inactive.size() && inactive[idx].row > row) ++idx;
    for (size_t k = idx; k < inactive.size(); ++k) {
      AEEdge e = inactive[k];
      AERowIter it;
      if (ae_row_iter_new(&e, &it)) {
        iters_inactive.push_back(it);
        active.push_back(e);
      }
    }
    inactive.resize(idx);
    std::stable_sort(iters_inactive.begin(), iters_inactive.end(),
                     [](const AERowIter& a, const AERowIter& b) { return b.column < a.column; });
  };
  next_row();
  while (true) {
    if (row >= height) return;
    while (!iters_inactive.empty()) {
      AERowIter it = iters_inactive.back();
      if (it.column <= column) {
        iters_inactive.pop_back();
        iters_active.push_back(it);
      } else {
        break;
      }
    }
    bool any = false;
    double acc = 0.0;
    if (!iters_active.empty()) {
      any = true;
      std::vector<AERowIter> keep;
      for (AERowIter& it : iters_active) {
        double v;
        if (it.next(&v)) {
          acc += v;
          keep.push_back(it);
        }
      }
      iters_active.swap(keep);
    }
    if (any) {
      winding += acc;
    } else if (std::fabs(winding) < EPSILON_SQRT) {
      if (!iters_inactive.empty() && iters_inactive.back().column < width) {
        column = iters_inactive.back().column;
      } else {
        column = 0;
        row += 1;
        winding = 0.0;
        next_row();
      }
      continue;
    }
    sink(Pixel{column, row, alpha_from_winding(rule, winding)});
    column += 1;
    if (column >= width) {
      column = 0;
      row += 1;
      winding = 0.0;
      next_row();
    }
  }
}

// ActiveEdgeRasterizer::mask (src/rasterize.rs:537-556): overwrites only the
// pixels the iterator produces.
inline void ae_mask(const Path& path, const Transform& tr, double flatness, FillRule rule,
                    ImageF64 img) {
  std::vector<Line> lines;
  flatten(path, tr, flatness, true, [&](const Line& l) { lines.push_back(l); });
  ae_iter(img.shape.width, img.shape.height, rule, lines,
          [&](const Pixel& p) { img.data[img.shape.offset(p.y, p.x)] = p.alpha; });
}

}  // namespace orc

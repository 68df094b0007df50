// Seeded synthetic inputs (include/rz_workloads.h) built with the reference's
// LCG (benches/scene_bench.rs:51-112) and PathBuilder semantics.
#include <algorithm>
#include <cmath>
#include <memory>
#include <vector>

#include "../../include/rasterize.hpp"
#include "../../include/rz_workloads.h"

namespace {

// Rnd (benches/scene_bench.rs:51-112)
struct Rnd {
  uint32_t state;
  explicit Rnd(uint32_t seed) : state(seed) {}
  uint32_t step() {
    state = (state * 214013u + 2531011u) & 0x7fffffffu;
    return state >> 16;
  }
  uint32_t uniform_u32() {
    uint32_t hi = step() & 0xffffu;
    uint32_t lo = step() & 0xffffu;
    return (hi << 16) | lo;
  }
  uint64_t uniform_u64() {
    uint64_t hi = uniform_u32();
    uint64_t lo = uniform_u32();
    return (hi << 32) | lo;
  }
  double uniform() {
    const double bpr_recip = std::ldexp(1.0, -53);
    return (double)(uniform_u64() >> 10) * bpr_recip;
  }
  double gauss() {
    const double PI_2 = 2.0 * rz::PI;
    double x = uniform();
    double y = uniform();
    return std::cos(PI_2 * x) * std::sqrt(-2.0 * std::log(1.0 - y));
  }
  rz::Point point(double s) {
    double x = uniform() * s;
    double y = uniform() * s;
    return rz::Point(x, y);
  }
};

}  // namespace

namespace {

// A paint with its own stop storage (rz_paint points into it).
struct PaintStore {
  rz_paint p{};
  std::vector<double> pos;
  std::vector<float> col;
};

// LinColor::from(RGBA) (src/color.rs:394-406, exact powf conversion)
float srgb_to_linear(float v) {
  if (v <= 0.04045f) return v / 12.92f;
  return std::pow((v + 0.055f) / 1.055f, 2.4f);
}
void rgba_to_lin(uint8_t r, uint8_t g, uint8_t b, uint8_t a8, float* out) {
  const float a = (float)a8 / 255.0f;
  out[0] = srgb_to_linear((float)r / 255.0f) * a;
  out[1] = srgb_to_linear((float)g / 255.0f) * a;
  out[2] = srgb_to_linear((float)b / 255.0f) * a;
  out[3] = a;
}

}  // namespace

struct rzw_set {
  std::vector<rz::Path> paths;
  rz::Transform tr;
  std::vector<std::unique_ptr<PaintStore>> paints;  // scenes: one per path
  std::vector<int32_t> rules;
  std::vector<rz_atlas_slot> slots;                 // glyph sets: atlas placement
  uint32_t atlas_w = 0, atlas_h = 0;
};

extern "C" {

rzw_set* rzw_cubics64(uint32_t seed, double size) {
  Rnd rnd(seed);
  rz::PathBuilder pb;
  rz::Point p0 = rnd.point(size);
  pb.move_to(p0);
  for (int k = 0; k < 64; ++k) {
    rz::Point p1 = rnd.point(size);
    rz::Point p2 = rnd.point(size);
    rz::Point p3 = k == 63 ? p0 : rnd.point(size);
    pb.cubic_to(p1, p2, p3);
  }
  pb.close();
  rzw_set* s = new rzw_set();
  s->paths.push_back(pb.build());
  return s;
}

rzw_set* rzw_random_walk(uint32_t seed, uint32_t n_segments, double canvas) {
  Rnd rnd(seed);
  const double step = 40.0;
  const double hi = std::nextafter(canvas, 0.0);
  auto clampc = [&](double v) { return v < 0.0 ? 0.0 : (v > hi ? hi : v); };
  rz::PathBuilder pb;
  rz::Point cur = rnd.point(canvas);
  pb.move_to(cur);
  double sx = 0.0, sy = 0.0;
  for (uint32_t i = 0; i < n_segments; ++i) {
    if (i > 0 && i % 500 == 0) {
      pb.close();
      pb.move_to(cur);
    }
    double gx = rnd.gauss(), gy = rnd.gauss();
    rz::Point next(clampc(cur.x + step * gx), clampc(cur.y + step * gy));
    double u = rnd.uniform();
    if (u < 0.40) {
      pb.line_to(next);
    } else if (u < 0.70) {
      double ax = rnd.gauss(), ay = rnd.gauss();
      pb.quad_to(rz::Point(cur.x + step * ax, cur.y + step * ay), next);
    } else if (u < 0.95) {
      double ax = rnd.gauss(), ay = rnd.gauss(), bx = rnd.gauss(), by = rnd.gauss();
      pb.cubic_to(rz::Point(cur.x + step * ax, cur.y + step * ay),
                  rz::Point(next.x + step * bx, next.y + step * by), next);
    } else {
      double rx = 1.0 + 199.0 * rnd.uniform();
      double ry = 1.0 + 199.0 * rnd.uniform();
      double rot = 360.0 * rnd.uniform();
      bool large = rnd.uniform() < 0.5;
      bool sweep = rnd.uniform() < 0.5;
      pb.arc_to(rz::Point(rx, ry), rot, large, sweep, next);
    }
    cur = next;
    sx += cur.x;
    sy += cur.y;
  }
  pb.close();
  rzw_set* s = new rzw_set();
  s->paths.push_back(pb.build());
  const double cx = sx / (double)(n_segments ? n_segments : 1);
  const double cy = sy / (double)(n_segments ? n_segments : 1);
  const double scale = 0.5 + 1.5 * rnd.uniform();
  const double theta = 2.0 * rz::PI * rnd.uniform();
  s->tr = rz::Transform::new_translate(0.5 * canvas, 0.5 * canvas)
              .pre_rotate(theta)
              .pre_scale(scale, scale)
              .pre_translate(-cx, -cy);
  return s;
}

rzw_set* rzw_scene(uint32_t seed, uint32_t n_paths, double canvas) {
  Rnd rnd(seed);
  rzw_set* s = new rzw_set();
  const double lmin = std::log(8.0), lmax = std::log(1024.0);
  auto rand_color = [&](float* out) {
    const uint8_t r = (uint8_t)(rnd.uniform_u32() % 256), g = (uint8_t)(rnd.uniform_u32() % 256);
    const uint8_t b = (uint8_t)(rnd.uniform_u32() % 256);
    const uint8_t a = (uint8_t)(255.0 * rnd.uniform());
    rgba_to_lin(r, g, b, a, out);
  };
  for (uint32_t i = 0; i < n_paths; ++i) {
    // a star/blob: vertices at increasing angles, jagged radii, every edge a
    // random line, quad or cubic
    const double side = std::exp(lmin + (lmax - lmin) * rnd.uniform());
    const rz::Point c = rnd.point(canvas);
    const double g = rnd.gauss();
    long nseg = std::lround(std::exp(std::log(60.0) + 1.4 * g));
    nseg = nseg < 4 ? 4 : (nseg > 4000 ? 4000 : nseg);
    const double R = 0.5 * side;
    std::vector<double> th(nseg), rr(nseg);
    for (long k = 0; k < nseg; ++k) {
      th[k] = 2.0 * rz::PI * ((double)k + 0.3 * rnd.uniform()) / (double)nseg;
      rr[k] = R * (0.55 + 0.45 * rnd.uniform());
    }
    auto at = [&](double a, double r) { return rz::Point(c.x + r * std::cos(a), c.y + r * std::sin(a)); };
    rz::PathBuilder pb;
    pb.move_to(at(th[0], rr[0]));
    for (long k = 0; k < nseg; ++k) {
      const long k1 = (k + 1) % nseg;
      const double a0 = th[k], a1 = k1 ? th[k1] : th[0] + 2.0 * rz::PI;
      const rz::Point next = at(th[k1], rr[k1]);
      const double u = rnd.uniform();
      if (u < 1.0 / 3.0) {
        pb.line_to(next);
      } else if (u < 2.0 / 3.0) {
        pb.quad_to(at(0.5 * (a0 + a1), R * (0.3 + rnd.uniform())), next);
      } else {
        const rz::Point c1 = at(a0 + (a1 - a0) / 3.0, R * (0.3 + rnd.uniform()));
        const rz::Point c2 = at(a0 + 2.0 * (a1 - a0) / 3.0, R * (0.3 + rnd.uniform()));
        pb.cubic_to(c1, c2, next);
      }
    }
    pb.close();
    s->paths.push_back(pb.build());
    // paint: 60% solid, 25% linear, 15% radial (UserSpaceOnUse, Pad, sRGB stops)
    auto ps = std::make_unique<PaintStore>();
    rz_paint& p = ps->p;
    const double kind = rnd.uniform();
    const double ident[6] = {1, 0, 0, 0, 1, 0};
    for (int k = 0; k < 6; ++k) p.tr[k] = ident[k];
    p.units = RZ_UNITS_USER_SPACE;
    p.spread = RZ_SPREAD_PAD;
    p.linear_colors = 0;
    if (kind < 0.6) {
      p.kind = RZ_PAINT_SOLID;
      rand_color(p.color);
    } else {
      const int nst = 2 + (int)(rnd.uniform_u32() % 5);
      std::vector<double> pos(nst);
      for (int k = 0; k < nst; ++k) pos[k] = rnd.uniform();
      std::sort(pos.begin(), pos.end());
      for (int k = 0; k < nst; ++k) {
        float col[4];
        rand_color(col);
        ps->pos.push_back(pos[k]);
        ps->col.insert(ps->col.end(), col, col + 4);
      }
      p.n_stops = (uint32_t)nst;
      if (kind < 0.85) {
        p.kind = RZ_PAINT_LINEAR;
        p.start[0] = c.x + R * (2.0 * rnd.uniform() - 1.0);
        p.start[1] = c.y + R * (2.0 * rnd.uniform() - 1.0);
        p.end[0] = c.x + R * (2.0 * rnd.uniform() - 1.0);
        p.end[1] = c.y + R * (2.0 * rnd.uniform() - 1.0);
      } else {
        p.kind = RZ_PAINT_RADIAL;
        p.center[0] = p.fcenter[0] = c.x + 0.3 * R * (2.0 * rnd.uniform() - 1.0);
        p.center[1] = p.fcenter[1] = c.y + 0.3 * R * (2.0 * rnd.uniform() - 1.0);
        p.radius = R * (0.2 + 0.8 * rnd.uniform());
        p.fradius = 0.0;
      }
    }
    p.stop_positions = ps->pos.empty() ? nullptr : ps->pos.data();
    p.stop_colors = ps->col.empty() ? nullptr : ps->col.data();
    s->paints.push_back(std::move(ps));
    s->rules.push_back(rnd.uniform() < 0.8 ? RZ_NONZERO : RZ_EVENODD);
  }
  return s;
}

rzw_set* rzw_circles(uint32_t count, double size) {
  Rnd rnd(0);
  rzw_set* s = new rzw_set();
  for (uint32_t i = 0; i < count; ++i) {
    const double x = rnd.uniform() * size, y = rnd.uniform() * size;
    const double r = rnd.uniform() * 10.0 + 30.0;
    rz::PathBuilder pb;
    pb.move_to(rz::Point(x, y)).circle(r);
    s->paths.push_back(pb.build());
    auto ps = std::make_unique<PaintStore>();
    ps->p.kind = RZ_PAINT_SOLID;
    const uint8_t cr = (uint8_t)(rnd.uniform_u32() % 256), cg = (uint8_t)(rnd.uniform_u32() % 256);
    const uint8_t cb = (uint8_t)(rnd.uniform_u32() % 256);
    rgba_to_lin(cr, cg, cb, 255, ps->p.color);
    s->paints.push_back(std::move(ps));
    s->rules.push_back(RZ_NONZERO);
  }
  return s;
}

rzw_set* rzw_glyphs(uint32_t seed, uint32_t n, uint32_t atlas_width) {
  Rnd rnd(seed);
  rzw_set* s = new rzw_set();
  std::vector<uint32_t> side(n);
  for (uint32_t i = 0; i < n; ++i) {
    const uint32_t sz = 16u + rnd.uniform_u32() % 113u;  // U{16..128}
    side[i] = sz;
    const uint32_t q = 8u + rnd.uniform_u32() % 193u;    // U{8..200} quads
    const uint32_t nc = 1u + rnd.uniform_u32() % 3u;     // 1..3 contours
    // glyph space: unit square mapped to [1, sz - 1]^2 (1 px margin)
    const double k = (double)sz - 2.0;
    auto to_px = [&](double u, double v) { return rz::Point(1.0 + k * u, 1.0 + k * v); };
    rz::PathBuilder pb;
    uint32_t left = q;
    for (uint32_t c = 0; c < nc; ++c) {
      const uint32_t m = c + 1 == nc ? left : std::max<uint32_t>(4u, q / nc);
      left = left > m ? left - m : 0u;
      const uint32_t cnt = std::max<uint32_t>(m, 3u);
      // outer contour: a blob round the centre; inner ones: small blobs
      // inside it traced the other way round (holes)
      double cx = 0.5, cy = 0.5, R = 0.5;
      if (c > 0) {
        R = 0.08 + 0.1 * rnd.uniform();
        cx = 0.5 + 0.25 * (2.0 * rnd.uniform() - 1.0);
        cy = 0.5 + 0.25 * (2.0 * rnd.uniform() - 1.0);
      }
      const double dir = c == 0 ? 1.0 : -1.0;
      std::vector<double> th(cnt), rr(cnt);
      for (uint32_t j = 0; j < cnt; ++j) {
        th[j] = dir * 2.0 * rz::PI * ((double)j + 0.4 * rnd.uniform()) / (double)cnt;
        rr[j] = R * (0.7 + 0.3 * rnd.uniform());
      }
      auto at = [&](double a, double r) { return to_px(cx + r * std::cos(a), cy + r * std::sin(a)); };
      pb.move_to(at(th[0], rr[0]));
      for (uint32_t j = 0; j < cnt; ++j) {
        const uint32_t j1 = (j + 1) % cnt;
        const double a0 = th[j], a1 = j1 ? th[j1] : th[0] + dir * 2.0 * rz::PI;
        const double rm = R * (0.8 + 0.45 * rnd.uniform());
        pb.quad_to(at(0.5 * (a0 + a1), rm), at(th[j1], rr[j1]));
      }
      pb.close();
    }
    s->paths.push_back(pb.build());
    s->rules.push_back(RZ_NONZERO);
  }
  // shelf packing, tallest first, rows of atlas_width pixels
  std::vector<uint32_t> order(n);
  for (uint32_t i = 0; i < n; ++i) order[i] = i;
  std::stable_sort(order.begin(), order.end(), [&](uint32_t a, uint32_t b) { return side[a] > side[b]; });
  s->slots.resize(n);
  uint32_t x = 0, y = 0, row_h = 0;
  for (uint32_t i : order) {
    const uint32_t sz = side[i];
    if (x + sz > atlas_width) {
      x = 0;
      y += row_h;
      row_h = 0;
    }
    rz_atlas_slot& sl = s->slots[i];
    sl.tr = rz::Transform::identity().abi();
    sl.x = x;
    sl.y = y;
    sl.width = sz;
    sl.height = sz;
    x += sz;
    row_h = std::max(row_h, sz);
  }
  s->atlas_w = atlas_width;
  s->atlas_h = y + row_h;
  return s;
}

void rzw_slot(const rzw_set* set, size_t i, rz_atlas_slot* out) { *out = set->slots[i]; }

void rzw_atlas_size(const rzw_set* set, uint32_t* width, uint32_t* height) {
  *width = set->atlas_w;
  *height = set->atlas_h;
}

void rzw_paint(const rzw_set* set, size_t i, rz_paint* out) { *out = set->paints[i]->p; }

int32_t rzw_rule(const rzw_set* set, size_t i) { return set->rules[i]; }

size_t rzw_count(const rzw_set* set) { return set ? set->paths.size() : 0; }

void rzw_path(const rzw_set* set, size_t i, rz_path* out) { *out = set->paths[i].view(); }

void rzw_transform(const rzw_set* set, rz_transform* out) { *out = set->tr.abi(); }

void rzw_free(rzw_set* set) { delete set; }

void rzw_rnd_uniform(uint32_t seed, size_t n, double* out) {
  Rnd rnd(seed);
  for (size_t i = 0; i < n; ++i) out[i] = rnd.uniform();
}

int rzw_arc_to(const double* src, const double* radii, double x_axis_rot, int large, int sweep,
               const double* dst, double* out) {
  rz::EllipArc arc;
  if (!rz::EllipArc::new_param(rz::Point(src[0], src[1]), rz::Point(dst[0], dst[1]), radii[0],
                               radii[1], x_axis_rot, large != 0, sweep != 0, &arc))
    return -1;
  int n = 0;
  arc.to_cubics([&](rz::Point a, rz::Point b, rz::Point c, rz::Point d) {
    double* o = out + 8 * n;
    o[0] = a.x; o[1] = a.y; o[2] = b.x; o[3] = b.y;
    o[4] = c.x; o[5] = c.y; o[6] = d.x; o[7] = d.y;
    ++n;
  });
  return n;
}

}  // extern "C"

// Seeded synthetic inputs (include/rz_workloads.h) built with the reference's
// LCG (benches/scene_bench.rs:51-112) and PathBuilder semantics.
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

struct rzw_set {
  std::vector<rz::Path> paths;
  rz::Transform tr;
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

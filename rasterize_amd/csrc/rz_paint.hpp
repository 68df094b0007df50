// Paint evaluation and compositing shared by the fine kernels (device) and the
// host-side paint preparation: premultiplied linear RGBA colours (LinColor),
// the sRGB polynomials of the reference's default x86 SIMD build, gradient
// stops and spreads, linear and two-circle radial gradients, source-over.
//
// Colours are f32x4 as in the reference (every lane, alpha included, goes
// through the same operation), positions and gradient parameters f64. The
// library is built with -ffp-contract=off so no multiply-add is fused.
#pragma once
#include "rz_geom.hpp"

namespace rzp {

using rzg::Pt;

struct C4 {
  float r, g, b, a;
};
RZ_HD C4 c4(float r, float g, float b, float a) {
  C4 c;
  c.r = r;
  c.g = g;
  c.b = b;
  c.a = a;
  return c;
}
RZ_HD C4 splat(float v) { return c4(v, v, v, v); }
RZ_HD C4 add(C4 x, C4 y) { return c4(x.r + y.r, x.g + y.g, x.b + y.b, x.a + y.a); }
RZ_HD C4 sub(C4 x, C4 y) { return c4(x.r - y.r, x.g - y.g, x.b - y.b, x.a - y.a); }
RZ_HD C4 mul(C4 x, C4 y) { return c4(x.r * y.r, x.g * y.g, x.b * y.b, x.a * y.a); }
RZ_HD C4 mul(C4 x, float s) { return c4(x.r * s, x.g * s, x.b * s, x.a * s); }
RZ_HD C4 mul(float s, C4 x) { return c4(s * x.r, s * x.g, s * x.b, s * x.a); }
RZ_HD C4 div(C4 x, C4 y) { return c4(x.r / y.r, x.g / y.g, x.b / y.b, x.a / y.a); }
RZ_HD C4 csqrt(C4 x) { return c4(sqrtf(x.r), sqrtf(x.g), sqrtf(x.b), sqrtf(x.a)); }

// Linear -> sRGB, the default build's f32x4 polynomial applied to every lane
// (src/simd/x86.rs:196-215): 12.92 x for x <= 0.0031308, else
// -0.01848558 x + 0.6445592 x^1/2 + 0.70994765 x^1/4 - 0.33605254 x^1/8.
RZ_HD float l2s1(float x0) {
  const float x1 = sqrtf(x0), x2 = sqrtf(x1), x3 = sqrtf(x2);
  const float high = -0.01848558f * x0 + 0.6445592f * x1 + 0.70994765f * x2 - 0.33605254f * x3;
  return x0 <= 0.0031308f ? x0 * 12.92f : high;
}
// sRGB -> linear (src/simd/x86.rs:217-244): a cubic in u = 2.0843103538116825 v
// - 1.0843103538116827 for v > 0.04045, v / 12.92 (as a product) below.
RZ_HD float s2l1(float vs) {
  const float x1 = 2.0843103538116825f * vs - 1.0843103538116827f;
  const float x2 = x1 * x1;
  const float x3 = x2 * x1;
  const float high = 0.23361048543711943f + 0.4665843122387033f * x1 +
                     0.26901741378006355f * x2 + 0.031661580753065945f * x3;
  return vs <= 0.04045f ? vs * 0.07739938080495357f : high;
}
RZ_HD C4 l2s(C4 x) { return c4(l2s1(x.r), l2s1(x.g), l2s1(x.b), l2s1(x.a)); }
RZ_HD C4 s2l(C4 x) { return c4(s2l1(x.r), s2l1(x.g), s2l1(x.b), s2l1(x.a)); }

// LinColor::unmultiply (src/color.rs:308-316)
RZ_HD C4 unmultiply(C4 c) { return c.a <= 1e-6f ? splat(0.0f) : div(c, splat(c.a)); }
// LinColor::into_srgb / into_linear (src/color.rs:322-332)
RZ_HD C4 into_srgb(C4 c) { return mul(l2s(unmultiply(c)), splat(c.a)); }
// The alpha lane of unmultiply is c.a / c.a: exactly 1 for finite c.a above
// the threshold and NaN (the same NaN as c.a - c.a) for an infinite or NaN
// one, so it needs no division.
RZ_HD C4 into_linear(C4 c) {
  const bool zero = c.a <= 1e-6f;
  const float inv = zero ? 0.0f : c.a;
  const float la = zero ? s2l1(0.0f) : (c.a < INFINITY ? s2l1(1.0f) : s2l1(c.a - c.a));
  const C4 u = zero ? splat(0.0f) : c4(c.r / inv, c.g / inv, c.b / inv, 0.0f);
  return mul(c4(s2l1(u.r), s2l1(u.g), s2l1(u.b), la), splat(c.a));
}
// Derivatives of the two polynomials (where they are differentiable; the
// side of each threshold the evaluation takes).
RZ_HD float l2s1_d(float x0) {
  if (x0 <= 0.0031308f) return 12.92f;
  const float x1 = sqrtf(x0), x2 = sqrtf(x1), x3 = sqrtf(x2);
  return -0.01848558f + 0.6445592f * 0.5f / x1 + 0.70994765f * 0.25f / (x1 * x2) -
         0.33605254f * 0.125f / (x1 * x2 * x3);
}
RZ_HD float s2l1_d(float vs) {
  if (vs <= 0.04045f) return 0.07739938080495357f;
  const float x1 = 2.0843103538116825f * vs - 1.0843103538116827f;
  return 2.0843103538116825f *
         (0.4665843122387033f + 2.0f * 0.26901741378006355f * x1 + 3.0f * 0.031661580753065945f * x1 * x1);
}
// g^T d(f(c / a) * a)/dc for f = s2l1 / l2s1 on r, g, b and f(1) * a on
// alpha (into_linear / into_srgb above a > 1e-6; nothing flows below it).
template <typename F, typename D>
RZ_HD C4 unmultiplied_vjp(C4 c, C4 g, F&& f, D&& df) {
  if (!(c.a > 1e-6f)) return splat(0.0f);
  const float u[3] = {c.r / c.a, c.g / c.a, c.b / c.a};
  const float gc[3] = {g.r, g.g, g.b};
  float out[3];
  float ga = g.a * f(1.0f);
  for (int k = 0; k < 3; ++k) {
    const float d = df(u[k]);
    out[k] = gc[k] * d;
    ga += gc[k] * (f(u[k]) - d * u[k]);
  }
  return c4(out[0], out[1], out[2], ga);
}
RZ_HD C4 into_linear_vjp(C4 c, C4 g) {
  return unmultiplied_vjp(c, g, [](float v) { return s2l1(v); }, [](float v) { return s2l1_d(v); });
}
RZ_HD C4 into_srgb_vjp(C4 c, C4 g) {
  return unmultiplied_vjp(c, g, [](float v) { return l2s1(v); }, [](float v) { return l2s1_d(v); });
}

// Color::lerp for LinColor: other * t + self * (1 - t) (src/color.rs:352-354)
RZ_HD C4 lerp(C4 self, C4 other, float t) { return add(mul(other, t), mul(self, 1.0f - t)); }
// Color::blend_over for LinColor: other + self * (1 - other.alpha)
// (src/color.rs:342-344); `self` is the destination.
RZ_HD C4 blend_over(C4 dst, C4 src) { return add(src, mul(dst, 1.0f - src.a)); }

// f64::rem_euclid
RZ_HD double rem_euclid(double a, double b) {
  const double r = fmod(a, b);
  return r < 0.0 ? r + fabs(b) : r;
}

enum { SPREAD_PAD = 0, SPREAD_REPEAT = 1, SPREAD_REFLECT = 2 };
// GradSpread::at (src/grad.rs:24-30)
RZ_HD double spread_at(int spread, double t) {
  if (spread == SPREAD_PAD) return t;
  if (spread == SPREAD_REPEAT) return rem_euclid(t, 1.0);
  return fabs(rem_euclid(t + 1.0, 2.0) - 1.0);
}

// quadratic_solve (src/utils.rs:204-231): roots of a t^2 + b t + c, the
// cancellation-free pair for two distinct roots. Returns the count.
RZ_HD int quadratic_solve(double a, double b, double c, double* r0, double* r1) {
  if (fabs(a) < rzg::kEpsilon) {
    if (fabs(b) > rzg::kEpsilon) {
      *r0 = -c / b;
      return 1;
    }
    return 0;
  }
  const double disc = b * b - 4.0 * a * c;
  if (fabs(disc) < rzg::kEpsilon) {
    *r0 = -b / (2.0 * a);
    return 1;
  }
  if (disc > 0.0) {
    const double sq = sqrt(disc);
    if (b >= 0.0) {
      const double m = -b - sq;
      *r0 = m / (2.0 * a);
      *r1 = 2.0 * c / m;
    } else {
      const double m = -b + sq;
      *r0 = 2.0 * c / m;
      *r1 = m / (2.0 * a);
    }
    return 2;
  }
  return 0;
}

enum { PAINT_SOLID = 0, PAINT_LINEAR = 1, PAINT_RADIAL = 2 };

// A paint ready for per-pixel evaluation: stops sorted (and in sRGB for
// gradients without linear_colors), the pixel-to-gradient transform resolved.
struct DPaint {
  int kind;
  int spread;
  int linear_colors;
  int n_stops;
  C4 color;            // solid
  double ptr[6];       // pixel centre -> gradient space
  Pt start, dir;       // linear: t = (p - start) . dir
  Pt center, fcenter;  // radial
  double radius, fradius;
  const double* stop_pos;  // [n_stops], ascending
  const C4* stop_col;      // [n_stops]
};

// GradStops::at (src/grad.rs:116-139): the stops with position < t form a
// prefix of the sorted list; its length picks the pair to interpolate.
RZ_HD C4 stops_at(const DPaint& p, double t) {
  int index = 0;
  while (index < p.n_stops && p.stop_pos[index] < t) ++index;
  if (index == 0) return p.stop_col[0];
  if (index == p.n_stops) return p.stop_col[p.n_stops - 1];
  const double p0 = p.stop_pos[index - 1], p1 = p.stop_pos[index];
  const double ratio = (t - p0) / (p1 - p0);
  return lerp(p.stop_col[index - 1], p.stop_col[index], (float)ratio);
}

// GradRadial::offset (src/grad.rs:361-396): the larger root of the two-circle
// equation, false when there is none.
RZ_HD bool radial_offset(const DPaint& p, Pt q, double* out) {
  const Pt cd = rzg::mkpt(p.center.x - p.fcenter.x, p.center.y - p.fcenter.y);
  const Pt pd = rzg::mkpt(q.x - p.fcenter.x, q.y - p.fcenter.y);
  const double rd = p.radius - p.fradius;
  const double a = (cd.x * cd.x + cd.y * cd.y) - rd * rd;
  const double b = -2.0 * ((cd.x * pd.x + cd.y * pd.y) + p.fradius * rd);
  const double c = (pd.x * pd.x + pd.y * pd.y) - p.fradius * p.fradius;
  double r0 = 0.0, r1 = 0.0;
  const int n = quadratic_solve(a, b, c, &r0, &r1);
  if (n == 0) return false;
  *out = n == 2 ? fmax(r0, r1) : r0;
  return true;
}

// The stops a gradient's colour at pixel (x, y) interpolates: indices i0,
// i1 and weights (1 - ratio, ratio) as GradStops::at takes them (one stop
// with weight 1 outside the first and last), and the interpolated colour
// before into_linear. False where the paint is transparent (a radial
// gradient without a root) or solid.
RZ_HD bool gradient_stops_at(const DPaint& p, int x, int y, int* i0, int* i1, float* w1, C4* mixed) {
  if (p.kind == PAINT_SOLID) return false;
  const Pt q = rzg::tr_apply(p.ptr, rzg::mkpt((double)x + 0.5, (double)y + 0.5));
  double t;
  if (p.kind == PAINT_LINEAR) {
    t = (q.x - p.start.x) * p.dir.x + (q.y - p.start.y) * p.dir.y;
  } else if (!radial_offset(p, q, &t)) {
    return false;
  }
  t = spread_at(p.spread, t);
  int index = 0;
  while (index < p.n_stops && p.stop_pos[index] < t) ++index;
  if (index == 0 || index == p.n_stops) {
    *i0 = *i1 = index == 0 ? 0 : p.n_stops - 1;
    *w1 = 0.0f;
    *mixed = p.stop_col[*i0];
    return true;
  }
  const double p0 = p.stop_pos[index - 1], p1 = p.stop_pos[index];
  *i0 = index - 1;
  *i1 = index;
  *w1 = (float)((t - p0) / (p1 - p0));
  *mixed = lerp(p.stop_col[index - 1], p.stop_col[index], *w1);
  return true;
}

// Paint::at for the pixel centre (x + 0.5, y + 0.5) (src/rasterize.rs:90-96,
// src/grad.rs:201-211, 399-411)
RZ_HD C4 paint_at(const DPaint& p, int x, int y) {
  if (p.kind == PAINT_SOLID) return p.color;
  const Pt q = rzg::tr_apply(p.ptr, rzg::mkpt((double)x + 0.5, (double)y + 0.5));
  C4 c;
  if (p.kind == PAINT_LINEAR) {
    const double t = (q.x - p.start.x) * p.dir.x + (q.y - p.start.y) * p.dir.y;
    c = stops_at(p, spread_at(p.spread, t));
  } else {
    double off;
    if (!radial_offset(p, q, &off)) return splat(0.0f);
    c = stops_at(p, spread_at(p.spread, off));
  }
  return p.linear_colors ? c : into_linear(c);
}

}  // namespace rzp

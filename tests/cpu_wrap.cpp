// rzg::wrap_lines (rasterize_amd/csrc/rz_geom.hpp, run by the GPU flatten)
// against the oracle's signed_difference_line deposits (oracle/orc_raster.hpp,
// src/rasterize.rs:365-444) for lines that start exactly on the right clip
// abscissa and end to its right: every vertical line wrap_lines makes, put
// through the same deposit routine, must reproduce the quirk line's deposits
// on every visible pixel (mask: W x H buffer, dense or in a wider image;
// mask_iter: (W + 1) x H). Deterministic cases; prints the worst difference
// and exits 1 above 1e-12.
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <vector>

#include "../oracle/orc_raster.hpp"
#include "../rasterize_amd/csrc/rz_geom.hpp"

namespace {
uint64_t g_state = 0x9e3779b97f4a7c15ull;
double uni(double a, double b) {
  g_state = g_state * 6364136223846793005ull + 1442695040888963407ull;
  return a + (b - a) * (double)(g_state >> 11) * (1.0 / 9007199254740992.0);
}
}  // namespace

int main() {
  const int sizes[][2] = {{5, 4}, {16, 20}, {120, 107}, {1, 3}};
  const double overs[] = {1e-9, 0.35, 0.99, 1.0, 1.5, 3.7, 40.0};
  double worst = 0.0;
  long cases = 0, lines = 0;
  for (auto& sz : sizes) {
    const int W = sz[0], H = sz[1];
    for (int mode = 0; mode < 3; ++mode) {  // 0 mask, 1 mask in a wider image, 2 mask_iter
      const size_t bw = mode == 2 ? (size_t)W + 1 : (size_t)W;
      const size_t rs = mode == 1 ? (size_t)W + 7 : bw;
      const double wc = mode == 2 ? (double)W : (double)W - 1.0;
      const orc::Shape shp{0, bw, (size_t)H, rs, 1};
      const size_t len = (size_t)H * rs;
      const rzg::WrapClip q{wc, (double)rs, W, H};
      for (double over : overs)
        for (int k = 0; k < 40; ++k) {
          const double y0 = k == 0 ? 2.0 : uni(-3.0, H + 3.0);
          const double y1 = k == 1 ? y0 : (k == 2 ? std::floor(y0) + 1.0 : uni(-3.0, H + 3.0));
          rzg::Pt p0 = rzg::mkpt(wc, y0), p1 = rzg::mkpt(wc + over * uni(0.5, 1.0), y1);
          if (k & 1) std::swap(p0.y, p1.y);
          if (!rzg::wraps(p0, p1, q) || rzg::wraps(rzg::mkpt(wc - 1e-12, y0), p1, q)) return 2;
          // rows of slack: the reference's first deposit of a row is not bounds-checked
          std::vector<double> ref(len + 64 * rs, 0.0), got(len + 64 * rs, 0.0);
          orc::sd_line_deposits(shp, ref.size(), orc::Line{orc::P(p0.x, p0.y), orc::P(p1.x, p1.y)},
                                [&](size_t o, double v) { ref[o] += v; });
          rzg::wrap_lines(p0, p1, q, [&](rzg::Pt a, rzg::Pt b) {
            ++lines;
            if (a.x != b.x || a.x < 0.0 || a.x >= W || std::fabs(b.y - a.y) > 1.0) std::exit(3);
            orc::sd_line_deposits(shp, got.size(), orc::Line{orc::P(a.x, a.y), orc::P(b.x, b.y)},
                                  [&](size_t o, double v) { got[o] += v; });
          });
          for (int r = 0; r < H; ++r)
            for (int c = 0; c < W; ++c) {
              const size_t o = (size_t)r * rs + (size_t)c;
              worst = std::fmax(worst, std::fabs(ref[o] - got[o]));
            }
          ++cases;
        }
    }
  }
  std::printf("%ld cases, %ld lines, worst %.3e\n", cases, lines, worst);
  return worst <= 1e-12 ? 0 : 1;
}

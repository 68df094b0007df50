// rzs::hypot_cr (rasterize_amd/csrc/rz_stroke.hpp, the hypot the GPU stroke
// kernels use) built for the host. stdin: pairs of f64; stdout: one f64 per
// pair.
#include <cstdio>

#include "../rasterize_amd/csrc/rz_stroke.hpp"

int main() {
  double xy[2];
  while (fread(xy, 8, 2, stdin) == 2) {
    const double r = rzs::hypot_cr(xy[0], xy[1]);
    fwrite(&r, 8, 1, stdout);
  }
  return 0;
}

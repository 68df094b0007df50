// The product's segment traversal (rasterize_amd/csrc/rz_geom.hpp, the code
// the GPU flatten kernels run), compiled for the host so that a CPU test can
// compare its leaves with the oracle's flatten. stdin: u32 n, f64 tr[6],
// f64 tolerance, then per segment u8 kind (1 quad, 2 cubic) and f64 x,y of
// its 3 or 4 control points; stdout per segment: u32 status, u32 count,
// f64 lines[count][4].
#include <cstdint>
#include <cstdio>
#include <vector>

#include "../rasterize_amd/csrc/rz_geom.hpp"

template <int NP>
static void run(const double* q, const double* m, double fl) {
  rzg::Node<NP> n;
  for (int i = 0; i < NP; ++i) n.p[i] = rzg::tr_apply(m, rzg::mkpt(q[2 * i], q[2 * i + 1]));
  std::vector<double> out;
  const int st = rzg::traverse(n, fl, [&](rzg::Pt a, rzg::Pt b) {
    out.insert(out.end(), {a.x, a.y, b.x, b.y});
  });
  const uint32_t head[2] = {(uint32_t)st, (uint32_t)(out.size() / 4)};
  fwrite(head, 4, 2, stdout);
  fwrite(out.data(), 8, out.size(), stdout);
}

int main() {
  uint32_t n = 0;
  double m[6], tol = 0.0;
  if (fread(&n, 4, 1, stdin) != 1 || fread(m, 8, 6, stdin) != 6 || fread(&tol, 8, 1, stdin) != 1) return 1;
  const double fl = 16.0 * tol * tol;  // src/path.rs:749
  for (uint32_t k = 0; k < n; ++k) {
    uint8_t kind = 0;
    double q[8];
    if (fread(&kind, 1, 1, stdin) != 1) return 1;
    const int np = kind == 1 ? 3 : 4;
    if (fread(q, 8, 2 * np, stdin) != (size_t)(2 * np)) return 1;
    if (np == 3) run<3>(q, m, fl);
    else run<4>(q, m, fl);
  }
  return 0;
}

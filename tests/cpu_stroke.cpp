// The product's stroke (rasterize_amd/csrc/rz_stroke.hpp: the item functions
// the GPU stroke kernels run) compiled for the host, with the kernels' four
// passes (pieces and the max-scan of items with curves, each run's first
// item, counts and their prefix sums, emit) done one item after another, so
// that a CPU test can compare the result with the oracle's Path::stroke bit
// for bit (both use the host libm).
//
// stdin: f64 width, f64 miter_limit, i32 join, i32 cap, u32 n_segments,
// u32 n_subpaths, u32 n_points, then kinds (u8), point offsets (u32,
// n_segments + 1), points (f64 pairs), subpath offsets (u32, n_subpaths + 1),
// closed flags (u8). stdout: u32 n_segments, u32 n_points, u32 n_subpaths,
// then the stroked path's arrays in the same order.
#include <cstdint>
#include <cstdio>
#include <vector>

#include "../rasterize_amd/csrc/rz_stroke.hpp"

struct HostQ {
  const std::vector<uint32_t>& pcs;
  const std::vector<int32_t>& prev_in;  // nearest item <= g with curves
  const std::vector<int32_t>& rfirst;
  const std::vector<uint32_t>& rhas;
  const std::vector<rzs::Seg>& f;
  const std::vector<rzs::Seg>& l;
  uint32_t pieces(uint32_t g) const { return pcs[g]; }
  int32_t prev(uint32_t g) const { return g == 0 ? -1 : prev_in[g - 1]; }
  int32_t prev_incl(uint32_t g) const { return prev_in[g]; }
  bool run_has(uint32_t r) const { return rhas[r] != 0; }
  int32_t run_first(uint32_t r) const { return rfirst[r]; }
  const rzs::Seg& first(int32_t i) const { return f[i]; }
  const rzs::Seg& last(int32_t i) const { return l[i]; }
};

template <typename T>
static bool rd(std::vector<T>& v, size_t n) {
  v.resize(n);
  return n == 0 || fread(v.data(), sizeof(T), n, stdin) == n;
}

int main() {
  double width = 0, limit = 0;
  int32_t join = 0, cap = 0;
  uint32_t ns = 0, nsub = 0, npts = 0;
  if (fread(&width, 8, 1, stdin) != 1 || fread(&limit, 8, 1, stdin) != 1 ||
      fread(&join, 4, 1, stdin) != 1 || fread(&cap, 4, 1, stdin) != 1 ||
      fread(&ns, 4, 1, stdin) != 1 || fread(&nsub, 4, 1, stdin) != 1 ||
      fread(&npts, 4, 1, stdin) != 1)
    return 1;
  std::vector<uint8_t> kinds, closed;
  std::vector<uint32_t> offs, subs;
  std::vector<double> pts;
  if (!rd(kinds, ns) || !rd(offs, ns + 1) || !rd(pts, 2 * (size_t)npts) || !rd(subs, nsub + 1) ||
      !rd(closed, nsub))
    return 1;
  std::vector<uint32_t> sub_of(ns);
  for (uint32_t i = 0; i < nsub; ++i)
    for (uint32_t k = subs[i]; k < subs[i + 1]; ++k) sub_of[k] = i;
  rzs::Input in;
  in.kinds = kinds.data();
  in.pt_off = offs.data();
  in.pts = pts.data();
  in.subs = subs.data();
  in.closed = closed.data();
  in.sub_of = sub_of.data();
  in.n_items = 2 * ns;
  in.st.width = width;
  in.st.miter_limit = limit;
  in.st.join = join;
  in.st.cap = cap;
  const uint32_t n = in.n_items, nr = 2 * nsub;
  std::vector<uint32_t> pcs(n), ppts(n), rhas(nr, 0);
  std::vector<int32_t> prev_in(n), rfirst(nr, -1);
  std::vector<rzs::Seg> f(n), l(n);
  int32_t run = -1;
  for (uint32_t g = 0; g < n; ++g) {  // pieces and the max-scan
    rzs::CountSink c;
    rzs::seg_offset(rzs::item_seg(in, rzs::item_of(in, g)), width / 2.0, c);
    pcs[g] = c.n;
    ppts[g] = c.pts;
    if (c.n) {
      f[g] = c.first;
      l[g] = c.last;
      run = (int32_t)g;
    }
    prev_in[g] = run;
  }
  HostQ q{pcs, prev_in, rfirst, rhas, f, l};
  for (uint32_t g = 0; g < n; ++g) {  // each run's first item with curves
    const rzs::Item it = rzs::item_of(in, g);
    if (pcs[g] && q.prev(g) < (int32_t)it.run_b) {
      rfirst[it.run] = (int32_t)g;
      rhas[it.run] = 1;
    }
  }
  std::vector<uint32_t> seg_off(n + 1, 0), pt_off(n + 1, 0), run_idx(nr + 1, 0);
  for (uint32_t g = 0; g < n; ++g) {  // counts and their prefix sums
    rzs::CountSink c;
    rzs::item_output(in, q, g, rzs::item_of(in, g), false, c);
    seg_off[g + 1] = seg_off[g] + c.n + pcs[g];
    pt_off[g + 1] = pt_off[g] + c.pts + ppts[g];
  }
  for (uint32_t r = 0; r < nr; ++r) run_idx[r + 1] = run_idx[r] + rhas[r];
  const uint32_t o_seg = seg_off[n], o_pts = pt_off[n], o_sub = run_idx[nr];
  std::vector<uint8_t> okinds(o_seg), oclosed(o_sub);
  std::vector<uint32_t> ooffs(o_seg + 1), osubs(o_sub ? o_sub + 1 : 0);
  std::vector<double> opts(2 * (size_t)o_pts);
  for (uint32_t g = 0; g < n; ++g) {  // emit
    const rzs::Item it = rzs::item_of(in, g);
    rzs::WriteSink w;
    w.kinds = okinds.data();
    w.pt_off = ooffs.data();
    w.pts = opts.data();
    w.seg = seg_off[g];
    w.pt = pt_off[g];
    rzs::item_output(in, q, g, it, true, w);
    if (w.seg != seg_off[g + 1] || w.pt != pt_off[g + 1]) return 2;  // count != emit
    if (g == it.run_e - 1 && rhas[it.run]) {
      osubs[run_idx[it.run] + 1] = w.seg;
      oclosed[run_idx[it.run]] = 1;
    }
  }
  ooffs[o_seg] = o_pts;
  if (o_sub) osubs[0] = 0;
  const uint32_t head[3] = {o_seg, o_pts, o_sub};
  fwrite(head, 4, 3, stdout);
  fwrite(okinds.data(), 1, okinds.size(), stdout);
  fwrite(ooffs.data(), 4, ooffs.size(), stdout);
  fwrite(opts.data(), 8, opts.size(), stdout);
  fwrite(osubs.data(), 4, osubs.size(), stdout);
  fwrite(oclosed.data(), 1, oclosed.size(), stdout);
  return 0;
}

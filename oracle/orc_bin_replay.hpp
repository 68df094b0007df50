// Oracle: sequential CPU replay of the product's binning stage, the checker
// for the GPU's tile records, band backdrops and free-endpoint events.
//
// TEST INFRASTRUCTURE ONLY. Flattening is the oracle's own restatement of
// PathFlattenIter (orc_core.hpp) run slot by slot; the per-line cell
// decomposition uses the product's pure geometry functions
// (rasterize_amd/csrc/rz_geom.hpp), because "bit-exact against a replay of the
// same stage" means the same arithmetic executed sequentially: what this
// checks is the GPU's traversal, atomics, scans and memory layout. Whether the
// decomposition itself is right is checked end to end against the oracle's
// reference rasterizer.
#pragma once
#include <algorithm>
#include <cstring>
#include <vector>

#include "../rasterize_amd/csrc/rz_geom.hpp"
#include "orc_core.hpp"

namespace orc {

struct BinReplay {
  int tiles_x = 0, tiles_y = 0;
  std::vector<std::vector<rzg::Rec>> recs;  // per tile
  std::vector<int32_t> bd;                  // per tile, prefix along each band
  std::vector<rzg::Event> events;
};

inline rzg::Pt rp(Point p) { return rzg::mkpt(p.x, p.y); }

inline void bin_replay(const Path& path, const Transform& tr, double flatness, int mode, int W,
                       int H, BinReplay* out) {
  rzg::BinGrid g;
  g.mode = mode;
  g.W = W;
  g.H = H;
  g.tiles_x = (W + rzg::kTile - 1) / rzg::kTile;
  g.tiles_y = (H + rzg::kTile - 1) / rzg::kTile;
  g.wc = mode == rzg::EDGE_MASK ? (double)W - 1.0 : (double)W;
  const size_t tiles = (size_t)g.tiles_x * g.tiles_y;
  out->tiles_x = g.tiles_x;
  out->tiles_y = g.tiles_y;
  out->recs.assign(tiles, {});
  out->bd.assign(tiles, 0);
  out->events.clear();
  const double fl = 16.0 * flatness * flatness;
  auto event = [&](rzg::Pt p, float sign) {
    rzg::Event ev;
    if (rzg::make_event(p, sign, g, &ev)) out->events.push_back(ev);
  };
  auto line = [&](Point a, Point b) {
    rzg::EffLine e = rzg::effective_line(rp(a), rp(b), g);
    for (int k = 0; k < e.n; ++k)
      rzg::bin_piece(
          e.p[k], e.p[k + 1], g, [&](int t, const rzg::Rec& r) { out->recs[t].push_back(r); },
          [&](int band, int col, int dir) { out->bd[(size_t)band * g.tiles_x + col] += dir; });
    if (e.clip0) event(e.p[0], -1.0f);
    if (e.clip1) event(e.p[e.n], 1.0f);
  };
  // one slot: flatten a transformed segment like PathFlattenIter does
  auto slot = [&](const Segment& seg, bool gap_start, bool gap_end) {
    Segment root = seg.transform(tr);
    std::vector<Segment> stack{root};
    while (!stack.empty()) {
      Segment s = stack.back();
      stack.pop_back();
      if (s.has_nans()) throw std::runtime_error("cannot flatten segment with NaN");
      if (s.flatness() < fl) {
        line(s.start(), s.end());
        continue;
      }
      Segment s0, s1;
      s.split(&s0, &s1);
      stack.push_back(s1);
      stack.push_back(s0);
    }
    const bool clip = mode != rzg::EDGE_NONE;
    Point st = root.start(), en = root.end();
    if (gap_start && !(clip && st.x > g.wc)) event(rzg::mkpt(std::fmax(st.x, 0.0), st.y), -1.0f);
    if (gap_end && !(clip && en.x > g.wc)) event(rzg::mkpt(std::fmax(en.x, 0.0), en.y), 1.0f);
  };
  for (size_t i = 0; i < path.nsubpaths(); ++i) {
    size_t b = path.subpaths[i], e = path.subpaths[i + 1];
    for (size_t k = b; k < e; ++k) {
      bool gs = k > b && !(path.segments[k - 1].end() == path.segments[k].start());
      bool ge = k + 1 < e && !(path.segments[k].end() == path.segments[k + 1].start());
      slot(path.segments[k], gs, ge);
    }
    Segment closing = Segment::line(path.segments[e - 1].end(), path.segments[b].start());
    slot(closing, false, false);
  }
  for (int band = 0; band < g.tiles_y; ++band) {
    int32_t acc = 0;
    for (int c = 0; c < g.tiles_x; ++c) {
      acc += out->bd[(size_t)band * g.tiles_x + c];
      out->bd[(size_t)band * g.tiles_x + c] = acc;
    }
  }
}

}  // namespace orc

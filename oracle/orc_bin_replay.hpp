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
  std::vector<int64_t> redge;               // [H]: right-edge column per row
};

inline rzg::Pt rp(Point p) { return rzg::mkpt(p.x, p.y); }
inline rzg::Pt tr_pt(const Transform& tr, Point p) { return rp(tr.apply(p)); }

inline void bin_replay(const Path& path, const Transform& tr, double flatness, int mode, int W,
                       int H, BinReplay* out, int phx = 0, int phy = 0) {
  rzg::BinGrid g;
  g.mode = mode;
  g.W = W;
  g.H = H;
  g.tiles_x = (W + phx + rzg::kTile - 1) / rzg::kTile;
  g.tiles_y = (H + phy + rzg::kTile - 1) / rzg::kTile;
  g.wc = mode == rzg::EDGE_MASK ? (double)W - 1.0 : (double)W;
  g.phx = phx;
  g.phy = phy;
  const size_t tiles = (size_t)g.tiles_x * g.tiles_y;
  out->tiles_x = g.tiles_x;
  out->tiles_y = g.tiles_y;
  out->recs.assign(tiles, {});
  out->bd.assign(tiles, 0);
  out->events.clear();
  std::vector<int64_t> ed((size_t)H + 2, 0);
  const double fl = 16.0 * flatness * flatness;
  auto event = [&](rzg::Pt p, float sign) {
    rzg::Event ev;
    if (rzg::make_event(p, sign, g, &ev)) out->events.push_back(ev);
  };
  auto line = [&](Point a, Point b) {
    rzg::EffLine e = rzg::effective_line(rp(a), rp(b), g);
    if (e.proj) {
      auto add = [&](int k, int v) { ed[(size_t)k] += v; };
      rzg::redge_endpoint(b.y, 1, g, add);
      rzg::redge_endpoint(a.y, -1, g, add);
      return;
    }
    for (int k = 0; k < e.n; ++k)
      rzg::bin_piece(
          e.p[k], e.p[k + 1], g, [&](int t, const rzg::Rec& r) { out->recs[t].push_back(r); },
          [&](int band, int col, int dir) { out->bd[(size_t)band * g.tiles_x + col] += dir; });
    if (e.clip0) event(e.p[0], -1.0f);
    if (e.clip1) event(e.p[e.n], 1.0f);
  };
  // gap between a (previous segment's end) and b (this start): +V(a), -V(b),
  // dropped together when they land on the same band, column and f32 offset;
  // ends beyond the right clip are reported by the clip
  auto gap = [&](Point a0, Point b0) {
    const bool clip = mode != rzg::EDGE_NONE;
    rzg::Pt a = tr_pt(tr, a0), b = tr_pt(tr, b0);
    rzg::Event va, vb;
    bool ha = !(clip && a.x > g.wc) && rzg::make_event(rzg::mkpt(std::fmax(a.x, 0.0), a.y), 1.0f, g, &va);
    bool hb = !(clip && b.x > g.wc) && rzg::make_event(rzg::mkpt(std::fmax(b.x, 0.0), b.y), -1.0f, g, &vb);
    if (ha && hb && va.band == vb.band && va.col == vb.col && va.yrel == vb.yrel) return;
    if (ha) out->events.push_back(va);
    if (hb) out->events.push_back(vb);
  };
  // one slot: flatten a transformed segment like PathFlattenIter does
  auto slot = [&](const Segment& seg) {
    Segment root = seg.transform(tr);
    std::vector<Segment> stack{root};
    while (!stack.empty()) {
      Segment s = stack.back();
      stack.pop_back();
      if (s.has_nans()) throw std::runtime_error("cannot flatten segment with NaN");
      if (s.flatness() < fl) {
        if (mode == rzg::EDGE_NONE) {
          rzg::ae_chain(rp(s.start()), rp(s.end()), (double)H, [&](rzg::Pt a, rzg::Pt b) {
            line(P(a.x, a.y), P(b.x, b.y));
          });
        } else {
          line(s.start(), s.end());
        }
        continue;
      }
      Segment s0, s1;
      s.split(&s0, &s1);
      stack.push_back(s1);
      stack.push_back(s0);
    }
  };
  for (size_t i = 0; i < path.nsubpaths(); ++i) {
    size_t b = path.subpaths[i], e = path.subpaths[i + 1];
    for (size_t k = b; k < e; ++k) {
      if (k > b && !(path.segments[k - 1].end() == path.segments[k].start()))
        gap(path.segments[k - 1].end(), path.segments[k].start());
      slot(path.segments[k]);
    }
    slot(Segment::line(path.segments[e - 1].end(), path.segments[b].start()));
  }
  out->redge.assign((size_t)H, 0);
  int64_t run = 0;
  for (int y = 0; y < H; ++y) out->redge[(size_t)y] = run += ed[(size_t)y];
  for (int band = 0; band < g.tiles_y; ++band) {
    int32_t acc = 0;
    for (int c = 0; c < g.tiles_x; ++c) {
      acc += out->bd[(size_t)band * g.tiles_x + c];
      out->bd[(size_t)band * g.tiles_x + c] = acc;
    }
  }
}

}  // namespace orc

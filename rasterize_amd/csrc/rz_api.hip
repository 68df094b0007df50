// Host side of librz: contexts, device buffers, path upload and the
// orchestration of the mask pipeline behind the C ABI in include/rz.h.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/rz.h"
#include "rz_kernels.hpp"

namespace {

struct DBuf {
  void* p = nullptr;
  size_t cap = 0;
};

}  // namespace

struct rz_ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  std::string err;
  rz_status async_err = RZ_OK;
  // pipeline scratch
  DBuf meta;      // Status | tile_count[tiles] | bd_diff[tiles] | band_ev[tiles_y]
  DBuf tile_off;  // [tiles + 1]
  DBuf ev_off;    // [tiles_y + 1]
  DBuf sums;      // scan block sums
  DBuf recs;      // float4 records
  DBuf events;    // int4 events
  DBuf out;       // dense f32 output for the blocking API
  DBuf lines;     // f64 lines for rz_flatten
  DBuf counts;    // per-slot counts for rz_flatten
  uint32_t rec_cap = 0, ev_cap = 0;
  rzk::Status* host_status = nullptr;  // pinned
  // launch parameters (rz_set_param)
  int split_quad = 3, split_cubic = 4;  // virtual-subtree depth of quad / cubic slots
  uint32_t bin_blocks = 2048;          // grid of the per-line binning kernels
  uint32_t line_cap = 0;
  // kernel timing (rz_profile)
  bool prof = false;
  struct ProfRec {
    const char* name;
    hipEvent_t a, b;
  };
  std::vector<ProfRec> pending;
  std::vector<hipEvent_t> pool;
  std::vector<rz_kernel_stat> stats;
};

struct rz_dpath {
  rzk::DPath view{};
  std::vector<void*> allocs;
  uint32_t n_order = 0;  // padded length of `order`
  uint32_t n_slots = 0;
};

namespace {

rz_status fail(rz_ctx* ctx, rz_status s, const std::string& msg) {
  if (ctx) ctx->err = msg;
  return s;
}

#define RZ_HIP(ctx, expr)                                                              \
  do {                                                                                 \
    hipError_t e_ = (expr);                                                            \
    if (e_ != hipSuccess)                                                              \
      return fail((ctx), RZ_E_HIP, std::string(#expr ": ") + hipGetErrorString(e_));   \
  } while (0)

rz_status ensure(rz_ctx* ctx, DBuf& b, size_t bytes) {
  if (b.cap >= bytes && b.p) return RZ_OK;
  size_t want = std::max<size_t>(bytes, 256);
  if (b.p) {
    // a previous asynchronous call may still read the old buffer
    RZ_HIP(ctx, hipStreamSynchronize(ctx->stream));
    RZ_HIP(ctx, hipFree(b.p));
    b.p = nullptr;
    b.cap = 0;
  }
  hipError_t e = hipMalloc(&b.p, want);
  if (e != hipSuccess) {
    b.p = nullptr;
    return fail(ctx, RZ_E_NOMEM, std::string("hipMalloc: ") + hipGetErrorString(e));
  }
  b.cap = want;
  return RZ_OK;
}

void free_buf(DBuf& b) {
  if (b.p) (void)hipFree(b.p);
  b.p = nullptr;
  b.cap = 0;
}

hipEvent_t take_event(rz_ctx* ctx);
void push_prof(rz_ctx* ctx, const char* name, hipEvent_t a, hipEvent_t b);

// Launch on the context's stream; with profiling on, bracket the kernel with
// events so rz_profile_stats can report its device time.
#define RZ_LAUNCH(ctx, name, kern, grid, block, ...)                              \
  do {                                                                           \
    hipEvent_t a_ = nullptr;                                                     \
    if ((ctx)->prof) {                                                           \
      a_ = take_event(ctx);                                                      \
      (void)hipEventRecord(a_, (ctx)->stream);                                   \
    }                                                                            \
    hipLaunchKernelGGL(kern, grid, block, 0, (ctx)->stream, __VA_ARGS__);        \
    if ((ctx)->prof) {                                                           \
      hipEvent_t b_ = take_event(ctx);                                           \
      (void)hipEventRecord(b_, (ctx)->stream);                                   \
      push_prof(ctx, name, a_, b_);                                              \
    }                                                                            \
  } while (0)

uint32_t blocks_for(size_t n) { return (uint32_t)((n + rzk::kBlock - 1) / rzk::kBlock); }

rz_status check_launch(rz_ctx* ctx, const char* what) {
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return fail(ctx, RZ_E_HIP, std::string(what) + ": " + hipGetErrorString(e));
  return RZ_OK;
}

#define RZ_TRY(expr)                  \
  do {                                \
    rz_status s_ = (expr);            \
    if (s_ != RZ_OK) return s_;       \
  } while (0)

// Exclusive scan of n u32 values: out[0..n] (out[n] = total). Optionally
// copies the offsets to `cursor` and the total to `total`.
rz_status scan_u32(rz_ctx* ctx, const uint32_t* in, uint32_t* out, size_t n, uint32_t* cursor,
                   uint32_t* total) {
  size_t nb = std::max<size_t>(1, (n + rzk::kScanTile - 1) / rzk::kScanTile);
  RZ_TRY(ensure(ctx, ctx->sums, nb * sizeof(uint32_t)));
  uint32_t* sums = (uint32_t*)ctx->sums.p;
  RZ_LAUNCH(ctx, "k_scan_tiles", rzk::k_scan_tiles, dim3((uint32_t)nb), dim3(rzk::kBlock), in,
                     out, n, sums);
  RZ_LAUNCH(ctx, "k_scan_sums", rzk::k_scan_sums, dim3(1), dim3(rzk::kBlock), sums, nb, out, n,
                     total);
  RZ_LAUNCH(ctx, "k_scan_add", rzk::k_scan_add, dim3((uint32_t)nb), dim3(rzk::kBlock), out, n,
                     (const uint32_t*)sums, cursor);
  return check_launch(ctx, "scan");
}

rz_status validate_path(rz_ctx* ctx, const rz_path* p) {
  if (!p) return fail(ctx, RZ_E_ARG, "path is NULL");
  if (p->n_segments && (!p->kinds || !p->point_offsets || !p->points))
    return fail(ctx, RZ_E_ARG, "path arrays are NULL");
  if (p->n_subpaths && (!p->subpaths || !p->closed)) return fail(ctx, RZ_E_ARG, "subpath arrays are NULL");
  for (uint32_t k = 0; k < p->n_segments; ++k) {
    if (p->kinds[k] > RZ_CUBIC) return fail(ctx, RZ_E_ARG, "bad segment kind");
    if (p->point_offsets[k + 1] != p->point_offsets[k] + p->kinds[k] + 2u)
      return fail(ctx, RZ_E_ARG, "point_offsets do not match segment kinds");
  }
  if (p->n_subpaths) {
    if (p->subpaths[0] != 0 || p->subpaths[p->n_subpaths] != p->n_segments)
      return fail(ctx, RZ_E_ARG, "subpath offsets must span all segments");
    for (uint32_t i = 0; i < p->n_subpaths; ++i)
      if (p->subpaths[i + 1] <= p->subpaths[i])
        return fail(ctx, RZ_E_ARG, "empty subpath (SubPath is never empty, src/path.rs:181)");
  } else if (p->n_segments) {
    return fail(ctx, RZ_E_ARG, "segments without subpaths");
  }
  return RZ_OK;
}

template <typename T>
rz_status upload_array(rz_ctx* ctx, rz_dpath* d, const std::vector<T>& v, const T** out) {
  size_t bytes = std::max<size_t>(1, v.size()) * sizeof(T);
  void* p = nullptr;
  hipError_t e = hipMalloc(&p, bytes);
  if (e != hipSuccess) return fail(ctx, RZ_E_NOMEM, std::string("hipMalloc: ") + hipGetErrorString(e));
  d->allocs.push_back(p);
  if (!v.empty()) RZ_HIP(ctx, hipMemcpyAsync(p, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice, ctx->stream));
  *out = (const T*)p;
  return RZ_OK;
}

void dpath_release(rz_dpath* d) {
  for (void* p : d->allocs) (void)hipFree(p);
  d->allocs.clear();
}

}  // namespace

// ---------------------------------------------------------------------------
// Path upload: device copy plus the slot tables the kernels use. A slot is one
// flatten unit in the reference's order (segments of a subpath, then its
// closing line); `order` groups slots by kind so a block never mixes the
// line, quad and cubic code paths.
// ---------------------------------------------------------------------------
static rz_status path_upload_impl(rz_ctx* ctx, const rz_path* p, rz_dpath** out) {
  RZ_TRY(validate_path(ctx, p));
  rz_dpath* d = new rz_dpath();
  const uint32_t ns = p->n_segments, nsub = p->n_subpaths;
  const uint32_t npts = ns ? p->point_offsets[ns] : 0;
  std::vector<double> pts(p->points, p->points + 2 * (size_t)npts);
  std::vector<uint8_t> kinds(p->kinds, p->kinds + ns);
  std::vector<uint32_t> pt_off(p->point_offsets, p->point_offsets + ns);
  const uint32_t nslots = ns + nsub;
  std::vector<uint32_t> slot_item(nslots);
  std::vector<uint8_t> slot_flags(nslots, 0);
  std::vector<double> close_pts(4 * (size_t)nsub);
  std::vector<uint8_t> slot_kind(nslots);
  auto start_of = [&](uint32_t k, double* x, double* y) {
    const double* q = p->points + 2 * (size_t)p->point_offsets[k];
    *x = q[0];
    *y = q[1];
  };
  auto end_of = [&](uint32_t k, double* x, double* y) {
    const double* q = p->points + 2 * (size_t)(p->point_offsets[k + 1] - 1);
    *x = q[0];
    *y = q[1];
  };
  uint32_t s = 0;
  for (uint32_t i = 0; i < nsub; ++i) {
    const uint32_t b = p->subpaths[i], e = p->subpaths[i + 1];
    for (uint32_t k = b; k < e; ++k) {
      slot_item[s] = k;
      slot_kind[s] = p->kinds[k];
      // gap with the previous segment of the subpath (the first segment
      // follows the closing line, which ends exactly at its start)
      if (k > b) {
        double ax, ay, bx, by;
        end_of(k - 1, &ax, &ay);
        start_of(k, &bx, &by);
        if (!(ax == bx && ay == by)) {
          slot_flags[s] |= 1;
          slot_flags[s - 1] |= 2;
        }
      }
      ++s;
    }
    double x0, y0, x1, y1;
    end_of(e - 1, &x0, &y0);
    start_of(b, &x1, &y1);
    close_pts[4 * i + 0] = x0;
    close_pts[4 * i + 1] = y0;
    close_pts[4 * i + 2] = x1;
    close_pts[4 * i + 3] = y1;
    slot_item[s] = 0x80000000u | i;
    slot_kind[s] = RZ_LINE;
    ++s;
  }
  // processing order: kind-grouped, each group padded to whole blocks
  std::vector<uint32_t> order;
  uint32_t nb[3] = {0, 0, 0};
  for (int kind = 0; kind < 3; ++kind) {
    size_t before = order.size();
    for (uint32_t t = 0; t < nslots; ++t)
      if (slot_kind[t] == kind) order.push_back(t);
    size_t cnt = order.size() - before;
    nb[kind] = blocks_for(cnt);
    order.resize(before + (size_t)nb[kind] * rzk::kBlock, ~0u);
  }
  rzk::DPath& v = d->view;
  rz_status st;
  if ((st = upload_array(ctx, d, pts, &v.pts)) != RZ_OK ||
      (st = upload_array(ctx, d, kinds, &v.kinds)) != RZ_OK ||
      (st = upload_array(ctx, d, pt_off, &v.pt_off)) != RZ_OK ||
      (st = upload_array(ctx, d, slot_item, &v.slot_item)) != RZ_OK ||
      (st = upload_array(ctx, d, slot_flags, &v.slot_flags)) != RZ_OK ||
      (st = upload_array(ctx, d, close_pts, &v.close_pts)) != RZ_OK ||
      (st = upload_array(ctx, d, order, &v.order)) != RZ_OK) {
    dpath_release(d);
    delete d;
    return st;
  }
  v.n_slots = nslots;
  v.blocks_line = nb[0];
  v.blocks_quad = nb[1];
  v.blocks_cubic = nb[2];
  d->n_order = (uint32_t)order.size();
  d->n_slots = nslots;
  hipError_t e = hipStreamSynchronize(ctx->stream);
  if (e != hipSuccess) {
    dpath_release(d);
    delete d;
    return fail(ctx, RZ_E_HIP, std::string("upload: ") + hipGetErrorString(e));
  }
  *out = d;
  return RZ_OK;
}

// ---------------------------------------------------------------------------
// Mask pipeline
// ---------------------------------------------------------------------------
namespace {

struct Grid {
  rzg::BinGrid g;
  size_t tiles;
};

Grid make_grid(int mode, size_t W, size_t H) {
  Grid r;
  r.g.mode = mode;
  r.g.W = (int)W;
  r.g.H = (int)H;
  r.g.tiles_x = (int)((W + rzg::kTile - 1) / rzg::kTile);
  r.g.tiles_y = (int)((H + rzg::kTile - 1) / rzg::kTile);
  r.g.wc = mode == RZ_EDGE_MASK ? (double)W - 1.0 : (double)W;
  r.tiles = (size_t)r.g.tiles_x * r.g.tiles_y;
  return r;
}

rz_status check_args(rz_ctx* ctx, double flatness, int rule, int mode, size_t W, size_t H) {
  if (!ctx) return RZ_E_ARG;
  if (!(flatness > 0.0) || !std::isfinite(flatness)) return fail(ctx, RZ_E_ARG, "flatness must be > 0");
  if (rule != RZ_NONZERO && rule != RZ_EVENODD) return fail(ctx, RZ_E_ARG, "bad fill rule");
  if (mode < RZ_EDGE_MASK || mode > RZ_EDGE_NONE) return fail(ctx, RZ_E_ARG, "bad edge mode");
  if (W > (1u << 24) || H > (1u << 24)) return fail(ctx, RZ_E_ARG, "image too large");
  return RZ_OK;
}

struct Pipe {
  Grid grid;
  rzk::Status* st;
  uint32_t* tile_count;
  int32_t* bd;
  uint32_t* band_ev;
  uint32_t* tile_off;
  uint32_t* ev_off;
  int32_t* redge;  // [H + 2] right-edge row differences, then (scanned) per-row sums
};

// Enqueue flatten, count pass, scans and scatter pass. Lines, records and
// events land in ctx->lines / ctx->recs / ctx->events (capacity-checked on the
// device; finish() grows them and asks for a re-run).
rz_status enqueue_bin(rz_ctx* ctx, const rz_dpath* d, const rz_transform& tr, double flatness,
                      int mode, size_t W, size_t H, Pipe* pp) {
  Pipe& P = *pp;
  P.grid = make_grid(mode, W, H);
  const size_t tiles = P.grid.tiles, ty = (size_t)P.grid.g.tiles_y;
  const size_t sw = rzk::kStatusWords;
  const size_t meta_u32 = sw + tiles + tiles + ty + (H + 2);
  RZ_TRY(ensure(ctx, ctx->meta, meta_u32 * 4));
  RZ_TRY(ensure(ctx, ctx->tile_off, (tiles + 1) * 4));
  RZ_TRY(ensure(ctx, ctx->ev_off, (ty + 1) * 4));
  if (ctx->line_cap == 0) {
    ctx->line_cap = (uint32_t)std::min<size_t>(std::max<size_t>((size_t)d->n_slots * 16, 1 << 16), 1u << 28);
    RZ_TRY(ensure(ctx, ctx->lines, (size_t)ctx->line_cap * 32));
  }
  if (ctx->rec_cap == 0) {
    ctx->rec_cap = (uint32_t)std::min<size_t>((size_t)ctx->line_cap + tiles, 1u << 30);
    RZ_TRY(ensure(ctx, ctx->recs, (size_t)ctx->rec_cap * 16));
  }
  if (ctx->ev_cap == 0) {
    ctx->ev_cap = 1 << 14;
    RZ_TRY(ensure(ctx, ctx->events, (size_t)ctx->ev_cap * 16));
  }
  uint32_t* meta = (uint32_t*)ctx->meta.p;
  P.st = (rzk::Status*)meta;
  P.tile_count = meta + sw;
  P.bd = (int32_t*)(meta + sw + tiles);
  P.band_ev = meta + sw + 2 * tiles;
  P.redge = (int32_t*)(meta + sw + 2 * tiles + ty);
  P.tile_off = (uint32_t*)ctx->tile_off.p;
  P.ev_off = (uint32_t*)ctx->ev_off.p;

  RZ_LAUNCH(ctx, "k_fill_u32", rzk::k_fill_u32,
            dim3((uint32_t)std::min<size_t>(blocks_for(meta_u32), 4096)), dim3(rzk::kBlock), meta,
            meta_u32, 0u);
  rzk::Xform x;
  for (int i = 0; i < 6; ++i) x.m[i] = tr.m[i];
  const double fl = 16.0 * flatness * flatness;
  // flatten: quad and cubic slots are split into 2^dq / 2^dc virtual subtrees
  const int dq = ctx->split_quad, dc = ctx->split_cubic;
  const uint32_t bl = d->view.blocks_line, bq = d->view.blocks_quad << dq,
                 bc = d->view.blocks_cubic << dc;
  const uint32_t nblk = bl + bq + bc;
  rzk::LineBufs lb;
  lb.lines = (double4*)ctx->lines.p;
  lb.cap = ctx->line_cap;
  lb.count = &P.st->n_lines;
  lb.st = P.st;
  if (nblk)
    RZ_LAUNCH(ctx, "k_flat_lines", rzk::k_flat_lines, dim3(nblk), dim3(rzk::kBlock), d->view, x,
              fl, lb, bl, bq, dq, dc);
  rzk::BinBufs b;
  b.tile_count = P.tile_count;
  b.bd_diff = P.bd;
  b.band_ev = P.band_ev;
  b.recs = (float4*)ctx->recs.p;
  b.rec_cap = ctx->rec_cap;
  b.events = (int4*)ctx->events.p;
  b.ev_cap = ctx->ev_cap;
  b.redge = P.redge;
  b.st = P.st;
  const uint32_t gb = ctx->bin_blocks;
  const uint32_t gs = std::max<uint32_t>(1, blocks_for(d->n_slots));
  RZ_LAUNCH(ctx, "k_bin_lines<0>", rzk::k_bin_lines<0>, dim3(gb), dim3(rzk::kBlock),
            (const double4*)lb.lines, (const uint32_t*)lb.count, P.grid.g, b);
  RZ_LAUNCH(ctx, "k_gap_events<0>", rzk::k_gap_events<0>, dim3(gs), dim3(rzk::kBlock), d->view, x,
            P.grid.g, b);
  RZ_TRY(check_launch(ctx, "bin pass 0"));
  RZ_TRY(scan_u32(ctx, P.tile_count, P.tile_off, tiles, P.tile_count, &P.st->n_records));
  RZ_TRY(scan_u32(ctx, P.band_ev, P.ev_off, ty, P.band_ev, &P.st->n_events));
  RZ_LAUNCH(ctx, "k_bd_prefix", rzk::k_bd_prefix, dim3((uint32_t)ty + 1), dim3(rzk::kBlock), P.bd,
            P.grid.g.tiles_x, P.redge, (int)(H + 2));
  RZ_LAUNCH(ctx, "k_bin_lines<1>", rzk::k_bin_lines<1>, dim3(gb), dim3(rzk::kBlock),
            (const double4*)lb.lines, (const uint32_t*)lb.count, P.grid.g, b);
  RZ_LAUNCH(ctx, "k_gap_events<1>", rzk::k_gap_events<1>, dim3(gs), dim3(rzk::kBlock), d->view, x,
            P.grid.g, b);
  return check_launch(ctx, "bin pass 1");
}

rz_status enqueue_mask(rz_ctx* ctx, const rz_dpath* d, const rz_transform& tr, double flatness,
                       int rule, int mode, size_t W, size_t H, float* d_out, size_t row_stride) {
  if (W == 0 || H == 0) return RZ_OK;
  Pipe P;
  RZ_TRY(enqueue_bin(ctx, d, tr, flatness, mode, W, H, &P));
  rzk::FineArgs a;
  a.tile_off = P.tile_off;
  a.recs = (const float4*)ctx->recs.p;
  a.bd = P.bd;
  a.band_ev_off = P.ev_off;
  a.events = (const int4*)ctx->events.p;
  a.st = P.st;
  a.out = d_out;
  a.row_stride = row_stride;
  a.W = (int)W;
  a.H = (int)H;
  a.tiles_x = P.grid.g.tiles_x;
  a.tiles_y = P.grid.g.tiles_y;
  a.redge = P.redge;
  if (mode == RZ_EDGE_NONE) {
    a.redge_col = INT32_MAX;
    a.redge_w = 0.0;
  } else {
    const double X = rzg::redge_x(P.grid.g), fx = std::floor(X);
    a.redge_col = (int)fx;
    a.redge_w = 1.0 - (X - fx);
  }
  const uint32_t gf = (uint32_t)P.grid.tiles;
  if (rule == RZ_NONZERO)
    RZ_LAUNCH(ctx, "k_fine_mask<0>", rzk::k_fine_mask<0>, dim3(gf), dim3(rzk::kFineWave), a);
  else
    RZ_LAUNCH(ctx, "k_fine_mask<1>", rzk::k_fine_mask<1>, dim3(gf), dim3(rzk::kFineWave), a);
  RZ_TRY(check_launch(ctx, "k_fine_mask"));
  // the status block is read back at rz_sync
  RZ_HIP(ctx, hipMemcpyAsync(ctx->host_status, P.st, sizeof(rzk::Status), hipMemcpyDeviceToHost,
                             ctx->stream));
  return RZ_OK;
}

// Wait, then interpret the status block: errors, and capacity growth.
rz_status finish(rz_ctx* ctx) {
  RZ_HIP(ctx, hipStreamSynchronize(ctx->stream));
  rzk::Status s = *ctx->host_status;
  std::memset(ctx->host_status, 0, sizeof(s));
  if (s.err == rzg::ST_NAN) return fail(ctx, RZ_E_NAN, "cannot flatten segment with NaN");
  if (s.err == rzg::ST_DEPTH) return fail(ctx, RZ_E_DEPTH, "subdivision depth limit exceeded");
  if (s.overflow) {
    if (s.n_records > ctx->rec_cap) {
      ctx->rec_cap = (uint32_t)std::min<uint64_t>((uint64_t)s.n_records * 5 / 4 + 1024, 0xffffffffu);
      RZ_TRY(ensure(ctx, ctx->recs, (size_t)ctx->rec_cap * 16));
    }
    if (s.n_lines > ctx->line_cap) {
      ctx->line_cap = (uint32_t)std::min<uint64_t>((uint64_t)s.n_lines * 5 / 4 + 1024, 0xffffffffu);
      RZ_TRY(ensure(ctx, ctx->lines, (size_t)ctx->line_cap * 32));
      // binning was skipped, so its record count is unknown: most lines give
      // one record, so size for that rather than re-running twice
      if (ctx->rec_cap < ctx->line_cap) {
        ctx->rec_cap = ctx->line_cap;
        RZ_TRY(ensure(ctx, ctx->recs, (size_t)ctx->rec_cap * 16));
      }
    }
    if (s.n_events > ctx->ev_cap) {
      ctx->ev_cap = s.n_events * 5 / 4 + 1024;
      RZ_TRY(ensure(ctx, ctx->events, (size_t)ctx->ev_cap * 16));
    }
    return fail(ctx, RZ_E_CAPACITY, "pipeline buffers grown; re-run");
  }
  return RZ_OK;
}

// Blocking mask into a device buffer, re-running once after a capacity miss.
rz_status mask_blocking(rz_ctx* ctx, const rz_dpath* d, const rz_transform& tr, double flatness,
                        int rule, int mode, size_t W, size_t H, float* d_out, size_t stride) {
  for (int attempt = 0; attempt < 3; ++attempt) {
    RZ_TRY(enqueue_mask(ctx, d, tr, flatness, rule, mode, W, H, d_out, stride));
    rz_status s = finish(ctx);
    if (s != RZ_E_CAPACITY) return s;
  }
  return fail(ctx, RZ_E_CAPACITY, "pipeline buffers could not be sized");
}

}  // namespace

namespace {
hipEvent_t take_event(rz_ctx* ctx) {
  if (!ctx->pool.empty()) {
    hipEvent_t e = ctx->pool.back();
    ctx->pool.pop_back();
    return e;
  }
  hipEvent_t e = nullptr;
  (void)hipEventCreate(&e);
  return e;
}
void push_prof(rz_ctx* ctx, const char* name, hipEvent_t a, hipEvent_t b) {
  ctx->pending.push_back({name, a, b});
}
// fold finished event pairs into the per-kernel statistics
rz_status collect_prof(rz_ctx* ctx) {
  RZ_HIP(ctx, hipStreamSynchronize(ctx->stream));
  for (auto& r : ctx->pending) {
    float ms = 0.0f;
    (void)hipEventElapsedTime(&ms, r.a, r.b);
    rz_kernel_stat* st = nullptr;
    for (auto& k : ctx->stats)
      if (std::strncmp(k.name, r.name, sizeof(k.name)) == 0) st = &k;
    if (!st) {
      rz_kernel_stat k{};
      std::strncpy(k.name, r.name, sizeof(k.name) - 1);
      ctx->stats.push_back(k);
      st = &ctx->stats.back();
    }
    st->calls += 1;
    st->total_ms += ms;
    ctx->pool.push_back(r.a);
    ctx->pool.push_back(r.b);
  }
  ctx->pending.clear();
  return RZ_OK;
}
}  // namespace

// ---------------------------------------------------------------------------
// C ABI
// ---------------------------------------------------------------------------
extern "C" {

rz_status rz_set_param(rz_ctx* ctx, int param, int value) {
  if (!ctx) return RZ_E_ARG;
  if (value < 0 || value > 8) return fail(ctx, RZ_E_ARG, "parameter out of range 0..8");
  if (param == RZ_PARAM_SPLIT_QUAD) ctx->split_quad = value;
  else if (param == RZ_PARAM_SPLIT_CUBIC) ctx->split_cubic = value;
  else return fail(ctx, RZ_E_ARG, "unknown parameter");
  return RZ_OK;
}

rz_status rz_get_param(rz_ctx* ctx, int param, int* value) {
  if (!ctx || !value) return RZ_E_ARG;
  if (param == RZ_PARAM_SPLIT_QUAD) *value = ctx->split_quad;
  else if (param == RZ_PARAM_SPLIT_CUBIC) *value = ctx->split_cubic;
  else return fail(ctx, RZ_E_ARG, "unknown parameter");
  return RZ_OK;
}

rz_status rz_profile(rz_ctx* ctx, int enable) {
  if (!ctx) return RZ_E_ARG;
  RZ_TRY(collect_prof(ctx));
  ctx->stats.clear();
  ctx->prof = enable != 0;
  return RZ_OK;
}

rz_status rz_profile_stats(rz_ctx* ctx, rz_kernel_stat* out, size_t cap, size_t* n) {
  if (!ctx || !n) return RZ_E_ARG;
  RZ_TRY(collect_prof(ctx));
  *n = ctx->stats.size();
  for (size_t i = 0; i < ctx->stats.size() && i < cap; ++i) out[i] = ctx->stats[i];
  return RZ_OK;
}

rz_status rz_ctx_create(rz_ctx** out, int device) {
  if (!out) return RZ_E_ARG;
  *out = nullptr;
  rz_ctx* ctx = new rz_ctx();
  ctx->device = device;
  hipError_t e = hipSetDevice(device);
  if (e == hipSuccess) e = hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking);
  if (e == hipSuccess) e = hipHostMalloc((void**)&ctx->host_status, sizeof(rzk::Status), hipHostMallocDefault);
  if (e != hipSuccess) {
    delete ctx;
    return RZ_E_HIP;
  }
  std::memset(ctx->host_status, 0, sizeof(rzk::Status));
  *out = ctx;
  return RZ_OK;
}

rz_status rz_ctx_destroy(rz_ctx* ctx) {
  if (!ctx) return RZ_OK;
  (void)hipSetDevice(ctx->device);
  if (ctx->stream) (void)hipStreamSynchronize(ctx->stream);
  DBuf* bufs[] = {&ctx->meta, &ctx->tile_off, &ctx->ev_off, &ctx->sums, &ctx->recs,
                  &ctx->events, &ctx->out, &ctx->lines, &ctx->counts};
  for (DBuf* b : bufs) free_buf(*b);
  for (auto& r : ctx->pending) {
    (void)hipEventDestroy(r.a);
    (void)hipEventDestroy(r.b);
  }
  for (hipEvent_t e : ctx->pool) (void)hipEventDestroy(e);
  if (ctx->host_status) (void)hipHostFree(ctx->host_status);
  if (ctx->stream) (void)hipStreamDestroy(ctx->stream);
  delete ctx;
  return RZ_OK;
}

const char* rz_last_error(const rz_ctx* ctx) { return ctx ? ctx->err.c_str() : "no context"; }

void* rz_stream(rz_ctx* ctx) { return ctx ? (void*)ctx->stream : nullptr; }

rz_status rz_sync(rz_ctx* ctx) {
  if (!ctx) return RZ_E_ARG;
  return finish(ctx);
}

rz_status rz_path_upload(rz_ctx* ctx, const rz_path* path, rz_dpath** out) {
  if (!ctx || !out) return RZ_E_ARG;
  return path_upload_impl(ctx, path, out);
}

rz_status rz_path_free(rz_ctx* ctx, rz_dpath* d) {
  if (!d) return RZ_OK;
  if (ctx) (void)hipStreamSynchronize(ctx->stream);
  dpath_release(d);
  delete d;
  return RZ_OK;
}

rz_status rz_flatten(rz_ctx* ctx, const rz_path* path, rz_transform tr, double flatness, int close,
                     double* lines, size_t cap, size_t* n_lines) {
  if (!ctx || !n_lines) return RZ_E_ARG;
  if (!(flatness > 0.0)) return fail(ctx, RZ_E_ARG, "flatness must be > 0");
  rz_dpath* d = nullptr;
  RZ_TRY(path_upload_impl(ctx, path, &d));
  struct Guard {
    rz_ctx* c;
    rz_dpath* d;
    ~Guard() { rz_path_free(c, d); }
  } guard{ctx, d};
  // Without `close` only explicitly closed subpaths get their closing line;
  // slots of the other closing lines are skipped by zeroing their count.
  const size_t ns = d->n_slots;
  rz_status st = ensure(ctx, ctx->counts, (ns + 1) * 4 * 2 + 64);
  if (st != RZ_OK) return st;
  uint32_t* counts = (uint32_t*)ctx->counts.p;
  uint32_t* offs = counts + ns + 1;
  rzk::Status* dst = (rzk::Status*)(offs + ns + 1);
  RZ_HIP(ctx, hipMemsetAsync(dst, 0, sizeof(rzk::Status), ctx->stream));
  rzk::Xform x;
  for (int i = 0; i < 6; ++i) x.m[i] = tr.m[i];
  const double fl = 16.0 * flatness * flatness;
  const uint32_t nblk = d->n_order / rzk::kBlock;
  if (nblk)
    RZ_LAUNCH(ctx, "k_flatten_count", rzk::k_flatten_count, dim3(nblk), dim3(rzk::kBlock), d->view,
                       x, fl, counts, dst);
  RZ_TRY(check_launch(ctx, "k_flatten_count"));
  if (!close) {
    // zero the counts of closing lines of subpaths that are not closed
    std::vector<uint32_t> zero_slots;
    uint32_t s = 0;
    for (uint32_t i = 0; i < path->n_subpaths; ++i) {
      s += path->subpaths[i + 1] - path->subpaths[i];
      if (!path->closed[i]) zero_slots.push_back(s);
      ++s;
    }
    for (uint32_t zs : zero_slots) RZ_HIP(ctx, hipMemsetAsync(counts + zs, 0, 4, ctx->stream));
  }
  RZ_TRY(scan_u32(ctx, counts, offs, ns, nullptr, nullptr));
  rzk::Status hs;
  uint32_t total = 0;
  RZ_HIP(ctx, hipMemcpyAsync(&hs, dst, sizeof(hs), hipMemcpyDeviceToHost, ctx->stream));
  RZ_HIP(ctx, hipMemcpyAsync(&total, offs + ns, 4, hipMemcpyDeviceToHost, ctx->stream));
  RZ_HIP(ctx, hipStreamSynchronize(ctx->stream));
  if (hs.err == rzg::ST_NAN) return fail(ctx, RZ_E_NAN, "cannot flatten segment with NaN");
  if (hs.err == rzg::ST_DEPTH) return fail(ctx, RZ_E_DEPTH, "subdivision depth limit exceeded");
  *n_lines = total;
  if (!lines) return RZ_OK;
  if (total > cap) return fail(ctx, RZ_E_CAPACITY, "line buffer too small");
  RZ_TRY(ensure(ctx, ctx->lines, (size_t)total * 32 + 32));
  if (nblk)
    RZ_LAUNCH(ctx, "k_flatten_emit", rzk::k_flatten_emit, dim3(nblk), dim3(rzk::kBlock), d->view, x,
                       fl, (const uint32_t*)counts, (const uint32_t*)offs, (double*)ctx->lines.p);
  RZ_TRY(check_launch(ctx, "k_flatten_emit"));
  RZ_HIP(ctx, hipMemcpyAsync(lines, ctx->lines.p, (size_t)total * 32, hipMemcpyDeviceToHost, ctx->stream));
  RZ_HIP(ctx, hipStreamSynchronize(ctx->stream));
  return RZ_OK;
}

rz_status rz_mask_device(rz_ctx* ctx, const rz_dpath* d, rz_transform tr, double flatness,
                         int fill_rule, int edge_mode, size_t width, size_t height, float* d_out,
                         size_t row_stride) {
  RZ_TRY(check_args(ctx, flatness, fill_rule, edge_mode, width, height));
  if (!d || (!d_out && width && height)) return fail(ctx, RZ_E_ARG, "null argument");
  if (row_stride < width) return fail(ctx, RZ_E_ARG, "row_stride < width");
  return enqueue_mask(ctx, d, tr, flatness, fill_rule, edge_mode, width, height, d_out, row_stride);
}

rz_status rz_mask_f32(rz_ctx* ctx, const rz_path* path, rz_transform tr, double flatness,
                      int fill_rule, int edge_mode, size_t width, size_t height, float* out,
                      size_t row_stride) {
  RZ_TRY(check_args(ctx, flatness, fill_rule, edge_mode, width, height));
  if (row_stride < width) return fail(ctx, RZ_E_ARG, "row_stride < width");
  if (width == 0 || height == 0) return RZ_OK;
  if (!out) return fail(ctx, RZ_E_ARG, "out is NULL");
  rz_dpath* d = nullptr;
  RZ_TRY(path_upload_impl(ctx, path, &d));
  const size_t bytes = height * row_stride * sizeof(float);
  rz_status st = ensure(ctx, ctx->out, bytes);
  if (st == RZ_OK)
    st = mask_blocking(ctx, d, tr, flatness, fill_rule, edge_mode, width, height, (float*)ctx->out.p,
                       row_stride);
  if (st == RZ_OK) {
    hipError_t e = hipMemcpy(out, ctx->out.p, bytes, hipMemcpyDeviceToHost);
    if (e != hipSuccess) st = fail(ctx, RZ_E_HIP, std::string("copy back: ") + hipGetErrorString(e));
  }
  rz_path_free(ctx, d);
  return st;
}

rz_status rz_mask(rz_ctx* ctx, const rz_path* path, rz_transform tr, double flatness, int fill_rule,
                  int edge_mode, double* data, size_t data_len, rz_shape shape) {
  if (!ctx) return RZ_E_ARG;
  const size_t W = shape.width, H = shape.height;
  if (W == 0 || H == 0) return RZ_OK;
  if (!data) return fail(ctx, RZ_E_ARG, "data is NULL");
  const size_t last = shape.start + (H - 1) * shape.row_stride + (W - 1) * shape.col_stride;
  if (last >= data_len) return fail(ctx, RZ_E_ARG, "shape exceeds data");
  std::vector<float> tmp(W * H);
  RZ_TRY(rz_mask_f32(ctx, path, tr, flatness, fill_rule, edge_mode, W, H, tmp.data(), W));
  for (size_t y = 0; y < H; ++y)
    for (size_t x = 0; x < W; ++x)
      data[shape.start + y * shape.row_stride + x * shape.col_stride] = (double)tmp[y * W + x];
  return RZ_OK;
}

rz_status rz_bin(rz_ctx* ctx, const rz_path* path, rz_transform tr, double flatness, int edge_mode,
                 size_t width, size_t height, uint32_t* tile_counts, float* records,
                 size_t records_cap, int32_t* backdrop, float* events, size_t events_cap,
                 int32_t* right_edge, size_t* n_records, size_t* n_events) {
  RZ_TRY(check_args(ctx, flatness, RZ_NONZERO, edge_mode, width, height));
  if (width == 0 || height == 0) {
    if (n_records) *n_records = 0;
    if (n_events) *n_events = 0;
    return RZ_OK;
  }
  rz_dpath* d = nullptr;
  RZ_TRY(path_upload_impl(ctx, path, &d));
  struct Guard {
    rz_ctx* c;
    rz_dpath* d;
    ~Guard() { rz_path_free(c, d); }
  } guard{ctx, d};
  Pipe P;
  rz_status st = RZ_OK;
  for (int attempt = 0; attempt < 3; ++attempt) {
    RZ_TRY(enqueue_bin(ctx, d, tr, flatness, edge_mode, width, height, &P));
    RZ_HIP(ctx, hipMemcpyAsync(ctx->host_status, P.st, sizeof(rzk::Status), hipMemcpyDeviceToHost,
                               ctx->stream));
    st = finish(ctx);
    if (st != RZ_E_CAPACITY) break;
  }
  if (st != RZ_OK) return st;
  const size_t tiles = P.grid.tiles, ty = (size_t)P.grid.g.tiles_y;
  std::vector<uint32_t> off(tiles + 1), eoff(ty + 1);
  RZ_HIP(ctx, hipMemcpy(off.data(), P.tile_off, (tiles + 1) * 4, hipMemcpyDeviceToHost));
  RZ_HIP(ctx, hipMemcpy(eoff.data(), P.ev_off, (ty + 1) * 4, hipMemcpyDeviceToHost));
  const size_t nr = off[tiles], ne = eoff[ty];
  if (n_records) *n_records = nr;
  if (n_events) *n_events = ne;
  if (tile_counts)
    for (size_t t = 0; t < tiles; ++t) tile_counts[t] = off[t + 1] - off[t];
  if (backdrop) RZ_HIP(ctx, hipMemcpy(backdrop, P.bd, tiles * 4, hipMemcpyDeviceToHost));
  if (right_edge) RZ_HIP(ctx, hipMemcpy(right_edge, P.redge, height * 4, hipMemcpyDeviceToHost));
  if (records) {
    if (records_cap < nr) return fail(ctx, RZ_E_CAPACITY, "records buffer too small");
    RZ_HIP(ctx, hipMemcpy(records, ctx->recs.p, nr * 16, hipMemcpyDeviceToHost));
  }
  if (events) {
    if (events_cap < ne) return fail(ctx, RZ_E_CAPACITY, "events buffer too small");
    RZ_HIP(ctx, hipMemcpy(events, ctx->events.p, ne * 16, hipMemcpyDeviceToHost));
  }
  return RZ_OK;
}

}  // extern "C"

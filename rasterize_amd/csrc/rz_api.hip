// Host side of librz: contexts, device buffers, path upload and the
// orchestration of the mask pipeline behind the C ABI in include/rz.h.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <atomic>
#include <chrono>
#include <fstream>
#include <mutex>
#include <sstream>
#include <string>
#include <vector>

#include "../../include/rz.h"
#include "rz_host.hpp"
#include "rz_kernels.hpp"
#include "rz_scene.hpp"
#include "rz_backward.hpp"
#include "rz_encode.hpp"
#include "rz_tree.hpp"
#include "rz_stroke_kernels.hpp"
#include "../../include/rasterize.hpp"

namespace {

struct DBuf {
  void* p = nullptr;
  size_t cap = 0;
};

}  // namespace

// ---------------------------------------------------------------------------
// Scratch of the no-alloc entry point (rz_mask_device_noalloc): the pipeline
// buffers of one call, carved from one fixed array. Each GPU the library runs
// on holds its own copy, allocated with the code object and kept until the
// process ends. Every no-alloc call on a GPU shares that GPU's array (calls
// that overlap on one GPU race); each run clears what must start at zero.
// ---------------------------------------------------------------------------
constexpr size_t kNoAllocBytes = (size_t)64 << 20;
static __device__ __align__(256) unsigned char scratch[kNoAllocBytes];

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
char* g_scratch_ptr[kMaxDevices];

// The device address of `scratch` on `device`, which must be current: looked
// up once per device with hipGetSymbolAddress. Host code uses only this
// address, never the array's name (a host-side stand-in).
char* noalloc_scratch(int device) {
  if (device < 0 || device >= kMaxDevices) return nullptr;
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (!g_scratch_ptr[device]) {
    void* p = nullptr;
    if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) == hipSuccess) g_scratch_ptr[device] = (char*)p;
  }
  return g_scratch_ptr[device];
}
}  // namespace

// Launch choices that change only the speed of a run, never its result: the
// flatten's virtual-subtree depths, extra LDS per wave (so fewer waves per
// SIMD) for the flatten, the count pass and the fine stage, and the size of
// the scatter's grid. Set per context (rz_set_param) or per workload by the
// tuner (rz_tune_*).
struct LaunchParams {
  int32_t split_auto = 1;     // depths from the path's line estimate (split_depths)
  int32_t dq = 3, dc = 4;     // else these
  int32_t flat_lds = -1;      // bytes; -1: 4096 for one path (3 waves per SIMD), 0 for batches
  int32_t bin_lds = 0;        // bytes
  int32_t fine_lds = 0;       // bytes
  int32_t scatter_log2 = 15;  // the scatter grid holds at most 2^v waves
  int32_t scene_waves = 6;    // k_scene_fine built for 1 (no bound), 6 or 8 waves per SIMD
  int32_t scene_queue = 0;    // 1: k_scene_fine persistent over a queue of tiles, costliest first
  int32_t fused = 0;          // 1: flatten and count pass as one kernel (k_flat_bin); not tuned
                              // (also keeps the size a multiple of 8: GraphKey has no padding)
};
static_assert(sizeof(LaunchParams) == 10 * sizeof(int32_t), "LaunchParams has padding");
bool same_params(const LaunchParams& a, const LaunchParams& b) {
  return std::memcmp(&a, &b, sizeof(a)) == 0;
}

// What the tuner keys its choices by: the entry point, the edge mode, and
// log2 buckets of the output's pixels, the path's slots and its quads' and
// cubics' expected line counts under the call's transform and flatness.
struct TuneKey {
  int32_t kind;  // 0 mask / fill, 1 batched scene, 2 glyph atlas
  int32_t mode, px_log2, slots_log2, quad_log2, cubic_log2;
};
bool same_tune_key(const TuneKey& a, const TuneKey& b) { return std::memcmp(&a, &b, sizeof(a)) == 0; }
struct TuneEntry {
  TuneKey key;
  LaunchParams p;
  double ms_default, ms_tuned;
};

// Arguments that fix the work a device entry point enqueues, for replaying a
// captured frame (all bytes compared; zero-initialised, padding included).
struct GraphKey {
  int kind;                 // 0 mask, 1 scene render, 2 glyph atlas, 3 scene backward
  int rule, mode, has_bg;
  uint64_t obj;             // serial of the resident path / scene
  uint64_t gen;             // the context's buffer and parameter generation
  double tr[6], flatness, view[4];
  float bg[4];
  size_t w, h, stride, stride2;
  const void* ptr[5];       // device outputs / inputs
  LaunchParams lp;          // the launch choices the frame was enqueued with
};
// keys are compared with memcmp and copied member by member: no padding bytes
static_assert(sizeof(GraphKey) == 4 * sizeof(int) + 2 * sizeof(uint64_t) + 11 * sizeof(double) +
                                      4 * sizeof(float) + 4 * sizeof(size_t) + 5 * sizeof(void*) +
                                      sizeof(LaunchParams),
              "GraphKey has padding");
struct GraphEntry {
  GraphKey key;
  hipGraphExec_t exec;
  uint64_t last_use;
};
namespace {
std::atomic<uint64_t> g_serial{1};  // resident paths and scenes
bool same_key(const GraphKey& a, const GraphKey& b) { return std::memcmp(&a, &b, sizeof(a)) == 0; }
constexpr size_t kMaxGraphs = 8;  // captured frames kept, and settled keys remembered
}

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
  DBuf stops;     // gradient stops of the current fill: f64 positions, then C4 colours
  // batched scene render
  DBuf line_path;  // path of each line
  DBuf bw_w, bw_p, bw_tr;  // scene backward: per-(path, tile) winding / dL/dW, colour below, transform partials
  DBuf bw_stops;           // scene backward: dL/d(stop as read) per stop of the scene
  DBuf sc_order;           // queued scene fine: tile order, queue head, per-tile path counts
  int n_cus = 256;         // compute units of the device
  DBuf sc_pi, sc_counts, sc_rects, sc_rows, sc_rowcnt;
  uint32_t sc_cap_t = 0, sc_cap_b = 0, sc_cap_r = 0;  // per-tile / per-band / per-row slices
  DBuf list_buf;  // binning: record list (records, tiles) and event list
  DBuf wave_lines;  // per flatten wave: its lines' first index and count
  DBuf atlas_tiles;  // per atlas tile: what the fine stage needs (AtlasTile)
  DBuf sk_work, sk_out;  // rz_path_stroke: input path and work arrays, output path
  uint32_t rec_cap = 0, ev_cap = 0;
  uint32_t part_cap = 0;  // overflow record-list entries per part
  rzk::Status* host_status = nullptr;  // pinned, two blocks (the copies of acc[0], acc[1])
  // device, two blocks: the merged status of the runs since the last sync,
  // [0] of runs on the context's own buffers, [1] of no-alloc runs
  rzk::Status* acc = nullptr;
  rzk::Status* acc_run = nullptr;      // where the runs being enqueued merge (acc or acc + 1)
  bool noalloc = false;                // buffers are carved from `scratch`: ensure() never allocates
  // hipGraph replay of repeated frames (RZ_PARAM_GRAPH)
  bool use_graphs = false;
  bool capturing = false;              // a stream capture is open: ensure() never allocates
  uint64_t gen = 1;                    // bumped when a buffer or a launch parameter changes
  uint64_t tick = 0;
  uint64_t replays = 0;                // frames replayed from a captured graph
  std::vector<GraphEntry> graphs;      // captured frames, least recently used evicted
  GraphKey graph_pending{};            // the last eager run
  bool pending_valid = false;
  std::vector<GraphKey> graph_settled;  // eager runs that completed, oldest first
  char* scratch_ptr = nullptr;         // device address of this device's `scratch`
  hipEvent_t xev = nullptr;            // orders the context's stream with a caller's stream
  rzk::Status last{};                  // status of the last run collected by finish()
  // launch parameters: the context's (rz_set_param), and those of the call
  // being enqueued (the context's, or the tuned entry of the call's workload)
  LaunchParams lp, cur;
  bool use_tuned = true;               // apply tuned entries (RZ_PARAM_TUNED)
  bool tune_forced = false;            // the tuner has set `cur` itself
  std::vector<TuneEntry> tuned;
  std::string gpu_id;                  // arch and CU count: tuning cache entries are per GPU
  uint32_t scan_direct = rzk::kScanDirect;  // block totals summed by the add pass itself
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

// Typical expected line count of a path's quads and cubics: the geometric
// mean over segments of (flatness / (16 tol^2))^(1/4) (a segment flattens into
// about 1.4 times that many lines), at tol = 0.05 under the transform the
// estimate was made with. The geometric mean keeps a few large curves of a
// scene from splitting its many small ones.
struct SplitEst {
  double quad = 1.0, cubic = 1.0;
};

struct rz_dpath {
  uint64_t serial = g_serial.fetch_add(1);
  rzk::DPath view{};
  std::vector<void*> allocs;
  uint32_t n_order = 0;  // padded length of `order`
  uint32_t n_slots = 0;
  bool has_bbox = false;  // Path::bbox(identity), for bounding-box paints
  SplitEst est;           // expected line counts of its quads / cubics
  rzh::BBox bbox{};
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
  if (ctx->capturing)
    return fail(ctx, RZ_E_CAPACITY, "a buffer would grow while a frame is being captured");
  ++ctx->gen;
  if (ctx->noalloc)
    return fail(ctx, RZ_E_CAPACITY, "no-alloc call: a buffer of " + std::to_string(bytes) +
                                        " bytes does not fit its part of the scratch (" +
                                        std::to_string(b.cap) + " bytes)");
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
// Launch on the context's stream (timed with events when profiling). `lds`
// is dynamic LDS reserved per workgroup beyond the kernel's static LDS.
#define RZ_LAUNCH_LDS(ctx, name, kern, grid, block, lds, ...)                    \
  do {                                                                           \
    hipEvent_t a_ = nullptr;                                                     \
    if ((ctx)->prof) {                                                           \
      a_ = take_event(ctx);                                                      \
      (void)hipEventRecord(a_, (ctx)->stream);                                   \
    }                                                                            \
    hipLaunchKernelGGL(kern, grid, block, lds, (ctx)->stream, __VA_ARGS__);      \
    if ((ctx)->prof) {                                                           \
      hipEvent_t b_ = take_event(ctx);                                           \
      (void)hipEventRecord(b_, (ctx)->stream);                                   \
      push_prof(ctx, name, a_, b_);                                              \
    }                                                                            \
  } while (0)
#define RZ_LAUNCH(ctx, name, kern, grid, block, ...) \
  RZ_LAUNCH_LDS(ctx, name, kern, grid, block, 0, __VA_ARGS__)

uint32_t blocks_for(size_t n) { return (uint32_t)((n + rzk::kBlock - 1) / rzk::kBlock); }

// Sums behind a SplitEst.
struct SplitSum {
  double q = 0, c = 0;  // sums of log counts
  size_t nq = 0, nc = 0;
  SplitEst est() const {
    SplitEst e;
    e.quad = nq ? std::exp(q / (double)nq) : 1.0;
    e.cubic = nc ? std::exp(c / (double)nc) : 1.0;
    return e;
  }
};

void split_sum(const rz_path* p, const double* m, SplitSum* s) {
  const double fl = 16.0 * 0.05 * 0.05;
  for (uint32_t k = 0; k < p->n_segments; ++k) {
    const double* q = p->points + 2 * (size_t)p->point_offsets[k];
    double f;
    if (p->kinds[k] == RZ_QUAD) {
      rzg::Node<3> n;
      for (int i = 0; i < 3; ++i) n.p[i] = rzg::tr_apply(m, rzg::mkpt(q[2 * i], q[2 * i + 1]));
      f = rzg::flatness(n);
    } else if (p->kinds[k] == RZ_CUBIC) {
      rzg::Node<4> n;
      for (int i = 0; i < 4; ++i) n.p[i] = rzg::tr_apply(m, rzg::mkpt(q[2 * i], q[2 * i + 1]));
      f = rzg::flatness(n);
    } else {
      continue;
    }
    const double e = f > fl && std::isfinite(f) ? 0.25 * std::log(f / fl) : 0.0;  // log of the count
    if (p->kinds[k] == RZ_QUAD) {
      s->q += e;
      ++s->nq;
    } else {
      s->c += e;
      ++s->nc;
    }
  }
}

// Virtual-subtree depths for a call: about four lines per thread for quads,
// two for cubics (measured best on the 200k-segment path and the glyph atlas).
// `scale` is the transform's area scale (|det|), applied on top of the
// estimate's own transform.
// Expected lines per quad / cubic of a path under a transform of determinant
// `scale` at `flatness` (the estimate is made at tol 0.05 under identity).
double line_factor(double scale, double flatness) {
  return 1.41 * std::sqrt(std::sqrt(scale > 0.0 ? scale : 1.0) * 0.05 / flatness);
}

void auto_depths(const SplitEst& e, double scale, double flatness, int* dq, int* dc) {
  const double f = line_factor(scale, flatness);
  auto depth = [](double lines, double per_thread) {
    const double d = std::round(std::log2(std::max(lines, 1.0) / per_thread));
    return (int)std::min(6.0, std::max(0.0, d));
  };
  *dq = depth(e.quad * f, 4.0);
  *dc = depth(e.cubic * f, 2.0);
}

void split_depths(rz_ctx* ctx, const SplitEst& e, double scale, double flatness, int* dq,
                  int* dc) {
  if (!ctx->cur.split_auto) {
    *dq = ctx->cur.dq;
    *dc = ctx->cur.dc;
    return;
  }
  auto_depths(e, scale, flatness, dq, dc);
}

int log2_bucket(double v) { return v >= 1.0 ? (int)std::lround(std::log2(v)) : 0; }

TuneKey tune_key(int kind, int mode, size_t px, uint32_t slots, const SplitEst& e, double scale,
                 double flatness) {
  const double f = line_factor(scale, flatness);
  TuneKey k;
  std::memset(&k, 0, sizeof(k));
  k.kind = kind;
  k.mode = mode;
  k.px_log2 = log2_bucket((double)px);
  k.slots_log2 = log2_bucket((double)slots);
  k.quad_log2 = log2_bucket(e.quad * f);
  k.cubic_log2 = log2_bucket(e.cubic * f);
  return k;
}

// The launch choices of the call about to be enqueued: its workload's tuned
// entry if there is one (and RZ_PARAM_TUNED is on), else the context's.
void select_params(rz_ctx* ctx, const TuneKey& key) {
  if (ctx->tune_forced) return;
  ctx->cur = ctx->lp;
  if (!ctx->use_tuned) return;
  for (const TuneEntry& t : ctx->tuned)
    if (same_tune_key(t.key, key)) {
      ctx->cur = t.p;
      ctx->cur.fused = ctx->lp.fused;  // the context's choice, not the tuner's
      return;
    }
}

double det_of(const rz_transform& tr) { return std::fabs(tr.m[0] * tr.m[4] - tr.m[1] * tr.m[3]); }

TuneKey mask_tune_key(const rz_dpath* d, const rz_transform& tr, double flatness, int mode,
                      size_t W, size_t H) {
  return tune_key(0, mode, W * H, d->n_slots, d->est, det_of(tr), flatness);
}

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
// copies the offsets to `cursor` and the total to `total`; the short array
// `x` and the rows of `r` are done by extra blocks of the second kernel.
rz_status scan_u32(rz_ctx* ctx, const uint32_t* in, uint32_t* out, size_t n, uint32_t* cursor,
                   uint32_t* total, const rzk::SmallScan& x = rzk::SmallScan{},
                   const rzk::RowPrefix& r = rzk::RowPrefix{}) {
  size_t nb = std::max<size_t>(1, (n + rzk::kScanTile - 1) / rzk::kScanTile);
  RZ_TRY(ensure(ctx, ctx->sums, nb * sizeof(uint32_t)));
  uint32_t* sums = (uint32_t*)ctx->sums.p;
  RZ_LAUNCH(ctx, "k_scan_tiles", rzk::k_scan_tiles, dim3((uint32_t)nb), dim3(rzk::kBlock), in,
                     out, n, sums);
  const bool scanned = nb > ctx->scan_direct;
  if (scanned)
    RZ_LAUNCH(ctx, "k_scan_sums", rzk::k_scan_sums, dim3(1), dim3(rzk::kBlock), sums, nb, out, n,
                       total);
  const uint32_t extra = (x.n ? 1u : 0u) + r.rows + (r.n_tail > 0 ? 1u : 0u);
  RZ_LAUNCH(ctx, "k_scan_add", rzk::k_scan_add, dim3((uint32_t)nb + extra), dim3(rzk::kBlock), out,
                     n, (const uint32_t*)sums, (uint32_t)nb, scanned, cursor, total, x, r);
  return check_launch(ctx, "scan");
}

// The tile scan and the band scan: a short band array rides along with the
// tile scan's second kernel, as do the rows of `r`.
rz_status scan_tiles_bands(rz_ctx* ctx, const uint32_t* tiles_in, uint32_t* tiles_out, size_t nt,
                           uint32_t* n_records, uint32_t* bands, uint32_t* bands_out, size_t nb,
                           uint32_t* n_events, const rzk::RowPrefix& r = rzk::RowPrefix{}) {
  if (nb <= 4096) {
    const rzk::SmallScan x{bands, bands_out, bands, n_events, (uint32_t)nb};
    return scan_u32(ctx, tiles_in, tiles_out, nt, nullptr, n_records, x, r);
  }
  RZ_TRY(scan_u32(ctx, tiles_in, tiles_out, nt, nullptr, n_records, rzk::SmallScan{}, r));
  return scan_u32(ctx, bands, bands_out, nb, bands, n_events);
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
  std::vector<uint32_t> gap_slots;
  for (uint32_t t = 0; t < nslots; ++t)
    if (slot_flags[t] & 1) gap_slots.push_back(t);
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
      (st = upload_array(ctx, d, order, &v.order)) != RZ_OK ||
      (st = upload_array(ctx, d, gap_slots, &v.gap_slots)) != RZ_OK) {
    dpath_release(d);
    delete d;
    return st;
  }
  v.n_slots = nslots;
  v.n_gaps = (uint32_t)gap_slots.size();
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
  const rzh::Tr ident = {{1.0, 0.0, 0.0, 0.0, 1.0, 0.0}};
  d->has_bbox = rzh::path_bbox(p, ident, &d->bbox);
  SplitSum sum;
  split_sum(p, ident.m, &sum);
  d->est = sum.est();
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
  r.g.phx = r.g.phy = 0;
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

// The meta block holds the status, per-tile counts and backdrops, per-band
// event counts, the right-edge rows (int64, so at an even word: the block
// itself is 256-byte aligned) and the counters of the binning stage.
size_t redge_words(size_t before) { return (before + 1) & ~(size_t)1; }
size_t mask_meta_words(size_t tiles, size_t ty, size_t H) {
  return redge_words(rzk::kStatusWords + 2 * tiles + ty) + 2 * (H + 2) + rzk::kParts + 1 +
         rzk::kLineParts;
}

struct Pipe {
  Grid grid;
  rzk::Status* st;
  uint32_t* tile_count;
  int32_t* bd;
  uint32_t* band_ev;
  uint32_t* tile_off;
  uint32_t* ev_off;
  long long* redge;  // [H + 2] right-edge row differences, then (scanned) per-row sums
};

// Flatten every slot of `d` into the line parts of ctx->lines (and, for a
// batched scene, the path of each line into ctx->line_path). `line_fill`
// (kLineParts words) must be zero; the number of lines goes to st->n_lines.
rz_status enqueue_flatten(rz_ctx* ctx, const rz_dpath* d, const rzk::Xform& x, double fl,
                          const rzk::PathInfo* pi, const uint32_t* slot_path, rzk::Status* st,
                          uint32_t* line_fill, int dq, int dc, int mode, double H,
                          const rzg::WrapClip& wrap, rzk::LineBufs* lb, bool launch = true) {
  const uint32_t bl = d->view.blocks_line, bq = d->view.blocks_quad << dq,
                 bc = d->view.blocks_cubic << dc;
  const uint32_t nblk = bl + bq + bc;
  lb->lines = (double4*)ctx->lines.p;
  lb->line_path = pi ? (uint32_t*)ctx->line_path.p : nullptr;
  lb->part_cap = ctx->line_cap / rzk::kLineParts;
  lb->fill = line_fill;
  lb->st = st;
  lb->n_waves = nblk * (rzk::kBlock / 64);
  lb->ae_h = H;
  lb->wrap = wrap;
  RZ_TRY(ensure(ctx, ctx->wave_lines, (size_t)std::max<uint32_t>(lb->n_waves, 1) * sizeof(uint2)));
  lb->wave_lines = (uint2*)ctx->wave_lines.p;
  const uint32_t fsub = rzk::kBlock / rzk::kFlatBlock;
  if (nblk && launch) {
    // By default a single path's flatten runs at three waves per SIMD rather
    // than the four its registers allow: 4 KB more LDS per wave. Measured on
    // the 200k-segment path: 169 us at three, 181 us at four (the batched
    // forms are faster at four).
    // ActiveEdgeRasterizer's lines come near x = 0 as chains (rzg::ae_chain)
    const int32_t lds = ctx->cur.flat_lds >= 0 ? ctx->cur.flat_lds : (pi ? 0 : 4096);
    const bool ae = mode == RZ_EDGE_NONE;
    const dim3 grid(nblk * fsub), block(rzk::kFlatBlock);
    // lines kept beyond the right clip: single masks and mask_iter, and
    // batched scenes drawn as mask_iter (not the atlas, not ActiveEdge)
    if (pi && ae) {
      RZ_LAUNCH_LDS(ctx, "k_flat_lines", (rzk::k_flat_lines<true, true, false>), grid, block, lds,
                    d->view, x, fl, bl, bq, dq, dc, pi, slot_path, *lb);
    } else if (pi && mode == RZ_EDGE_ITER) {
      RZ_LAUNCH_LDS(ctx, "k_flat_lines", (rzk::k_flat_lines<true, false, true>), grid, block, lds,
                    d->view, x, fl, bl, bq, dq, dc, pi, slot_path, *lb);
    } else if (pi) {
      RZ_LAUNCH_LDS(ctx, "k_flat_lines", (rzk::k_flat_lines<true, false, false>), grid, block, lds,
                    d->view, x, fl, bl, bq, dq, dc, pi, slot_path, *lb);
    } else if (ae) {
      RZ_LAUNCH_LDS(ctx, "k_flat_lines", (rzk::k_flat_lines<false, true, false>), grid, block, lds,
                    d->view, x, fl, bl, bq, dq, dc, pi, slot_path, *lb);
    } else {
      RZ_LAUNCH_LDS(ctx, "k_flat_lines", (rzk::k_flat_lines<false, false, true>), grid, block, lds,
                    d->view, x, fl, bl, bq, dq, dc, pi, slot_path, *lb);
    }
  }
  return check_launch(ctx, "flatten");
}

// Line capacity: kLineParts parts of whole 64-line chunks.
uint32_t line_cap_for(size_t lines) {
  const size_t unit = (size_t)rzk::kLineParts * 64;
  return (uint32_t)std::min<size_t>((lines + unit - 1) / unit * unit, (size_t)1 << 30);
}

// Binning of the flattened lines around the scans: the count pass, then (after
// the scans) the scatter of the listed records and events. `fill`
// (kParts + 1 words) must be zero when the count pass starts.
template <bool MULTI>
rz_status enqueue_bin_count(rz_ctx* ctx, const rzk::LineBufs& lb, const rzg::BinGrid& g0,
                            const rzk::PathInfo* pi, const rzk::GapArgs& ga, uint32_t* fill,
                            rzk::BinBufs* b, bool launch = true) {
  // one count-pass wave per flatten wave; the record list holds two entries
  // per line slot (records are usually as many as lines or fewer), the
  // overflow list takes the rest
  const uint32_t wpb = rzk::kBinBlock / 64;
  const uint32_t nw = lb.n_waves;  // a multiple of kBlock / 64
  if (ctx->part_cap == 0) ctx->part_cap = 1u << 16;
  const size_t no = (size_t)rzk::kParts * ctx->part_cap;
  const size_t nreg = 2 * (size_t)ctx->line_cap;
  const size_t nl = nreg + no;
  const size_t words = no + (size_t)ctx->ev_cap + nw + (size_t)nw * rzk::kMap;
  RZ_TRY(ensure(ctx, ctx->list_buf, (nl + ctx->ev_cap) * 16 + nl * 8 + words * 4));
  char* lp = (char*)ctx->list_buf.p;
  b->list_rec = (float4*)lp;
  b->ev_list = (int4*)(lp + nl * 16);
  b->list_meta = (uint2*)(lp + (nl + ctx->ev_cap) * 16);
  uint32_t* w = (uint32_t*)(lp + (nl + ctx->ev_cap) * 16 + nl * 8);
  b->ovf_wave = w;
  b->ev_band = b->ovf_wave + no;
  b->wave_used = b->ev_band + ctx->ev_cap;
  b->wave_map = b->wave_used + nw;
  b->n_waves = nw;
  b->wave_lines = lb.wave_lines;
  b->list_regions = nreg;
  b->part_cap = ctx->part_cap;
  b->fill = fill;
  if (nw == 0 || !launch) return RZ_OK;
  // then the gap slots, 64 per wave
  const uint32_t gap_blocks = (ga.p.n_gaps + rzk::kBinBlock - 1) / rzk::kBinBlock;
  RZ_LAUNCH_LDS(ctx, "k_bin_lines", (rzk::k_bin_lines<MULTI>), dim3(nw / wpb + gap_blocks),
            dim3(rzk::kBinBlock), ctx->cur.bin_lds, (const double4*)lb.lines, (const uint32_t*)lb.line_path, g0, pi,
            *b, ga);
  return RZ_OK;
}

// Flatten and count pass as one kernel (k_flat_bin), with `lb` and `b` as
// enqueue_flatten and enqueue_bin_count set them up (launch = false).
template <bool MULTI>
rz_status enqueue_flat_bin(rz_ctx* ctx, const rz_dpath* d, const rzk::Xform& x, double fl,
                           const rzk::PathInfo* pi, const uint32_t* slot_path, int dq, int dc,
                           int mode, const rzk::LineBufs& lb, const rzg::BinGrid& g0,
                           const rzk::BinBufs& b, const rzk::GapArgs& ga) {
  const uint32_t bl = d->view.blocks_line, bq = d->view.blocks_quad << dq;
  const uint32_t gap_waves = (ga.p.n_gaps + 63) / 64;
  const uint32_t nw = lb.n_waves + gap_waves;
  if (nw == 0) return RZ_OK;
  const dim3 grid(nw), block(64);
  if (mode == RZ_EDGE_NONE)
    RZ_LAUNCH_LDS(ctx, "k_flat_bin", (rzk::k_flat_bin<MULTI, true>), grid, block, ctx->cur.bin_lds,
                  d->view, x, fl, bl, bq, dq, dc, pi, slot_path, lb, g0, b, ga);
  else
    RZ_LAUNCH_LDS(ctx, "k_flat_bin", (rzk::k_flat_bin<MULTI, false>), grid, block, ctx->cur.bin_lds,
                  d->view, x, fl, bl, bq, dq, dc, pi, slot_path, lb, g0, b, ga);
  return check_launch(ctx, "flatten + count pass");
}

rz_status enqueue_bin_place(rz_ctx* ctx, const rzk::BinBufs& b, const uint32_t* line_fill) {
  if (b.n_waves == 0) return RZ_OK;
  // many waves, each over regions far apart (neighbouring regions hold
  // records of the same tiles)
  const uint32_t wpb = rzk::kScatterBlock / 64;
  const uint32_t cap = 1u << ctx->cur.scatter_log2;
  const uint32_t grid = (std::min(b.n_waves, cap) + wpb - 1) / wpb;
  // then 64 blocks for the events
  RZ_LAUNCH(ctx, "k_bin_scatter", rzk::k_bin_scatter, dim3(grid + 64), dim3(rzk::kScatterBlock), b,
            line_fill, grid);
  return RZ_OK;
}

// Enqueue flatten, count pass, scans and scatter pass. Lines, records and
// events land in ctx->lines / ctx->recs / ctx->events (capacity-checked on the
// device; finish() grows them and asks for a re-run).
// A mask's lines kept beyond the right clip (rzg::wrap_lines) deposit into an
// image with rows of mask_stride pixels (0: W).
rz_status enqueue_bin(rz_ctx* ctx, const rz_dpath* d, const rz_transform& tr, double flatness,
                      int mode, size_t W, size_t H, Pipe* pp, size_t mask_stride = 0) {
  select_params(ctx, mask_tune_key(d, tr, flatness, mode, W, H));
  Pipe& P = *pp;
  P.grid = make_grid(mode, W, H);
  const size_t tiles = P.grid.tiles, ty = (size_t)P.grid.g.tiles_y;
  const size_t sw = rzk::kStatusWords;
  const size_t ro = redge_words(sw + 2 * tiles + ty);
  const size_t meta_u32 = mask_meta_words(tiles, ty, H);
  RZ_TRY(ensure(ctx, ctx->meta, meta_u32 * 4));
  RZ_TRY(ensure(ctx, ctx->tile_off, (tiles + 1) * 4));
  RZ_TRY(ensure(ctx, ctx->ev_off, (ty + 1) * 4));
  if (ctx->line_cap == 0) {
    ctx->line_cap = line_cap_for(std::max<size_t>((size_t)d->n_slots * 16, 1 << 16));
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
  P.redge = (long long*)(meta + ro);
  P.tile_off = (uint32_t*)ctx->tile_off.p;
  P.ev_off = (uint32_t*)ctx->ev_off.p;

  RZ_LAUNCH(ctx, "k_fill_u32", rzk::k_fill_u32,
            dim3((uint32_t)std::min<size_t>(blocks_for(meta_u32), 4096)), dim3(rzk::kBlock), meta,
            meta_u32, 0u);
  rzk::Xform x;
  for (int i = 0; i < 6; ++i) x.m[i] = tr.m[i];
  const double fl = 16.0 * flatness * flatness;
  // flatten: quad and cubic slots are split into 2^dq / 2^dc virtual subtrees
  rzk::LineBufs lb;
  uint32_t* bin_fill = meta + meta_u32 - (rzk::kParts + 1 + rzk::kLineParts);
  int dq, dc;
  split_depths(ctx, d->est, std::fabs(x.m[0] * x.m[4] - x.m[1] * x.m[3]), flatness, &dq, &dc);
  const bool fused = ctx->cur.fused != 0;
  rzg::WrapClip wrap{-1.0, 1.0, (int)W, (int)H};
  if (mode == RZ_EDGE_MASK)
    wrap = rzg::WrapClip{(double)W - 1.0, (double)(mask_stride ? mask_stride : W), (int)W, (int)H};
  else if (mode == RZ_EDGE_ITER)
    wrap = rzg::WrapClip{(double)W, (double)W + 1.0, (int)W, (int)H};
  RZ_TRY(enqueue_flatten(ctx, d, x, fl, nullptr, nullptr, P.st, bin_fill + rzk::kParts + 1, dq, dc,
                         mode, (double)H, wrap, &lb, !fused));
  rzk::BinBufs b;
  b.tile_count = P.tile_count;
  b.tile_off = P.tile_off;
  b.bd_diff = P.bd;
  b.band_ev = P.band_ev;
  b.recs = (float4*)ctx->recs.p;
  b.rec_cap = ctx->rec_cap;
  b.events = (int4*)ctx->events.p;
  b.ev_cap = ctx->ev_cap;
  b.redge = P.redge;
  b.st = P.st;
  const rzk::GapArgs ga{d->view, x, nullptr};
  RZ_TRY(enqueue_bin_count<false>(ctx, lb, P.grid.g, nullptr, ga, bin_fill, &b, !fused));
  if (fused)
    RZ_TRY(enqueue_flat_bin<false>(ctx, d, x, fl, nullptr, nullptr, dq, dc, mode, lb, P.grid.g, b,
                                   ga));
  RZ_TRY(check_launch(ctx, "bin pass 0"));
  // with the backdrop differences along each tile row and the right-edge
  // row differences
  const rzk::RowPrefix rp{P.bd, P.grid.g.tiles_x, (uint32_t)ty, P.redge, (int)(H + 2)};
  RZ_TRY(scan_tiles_bands(ctx, P.tile_count, P.tile_off, tiles, &P.st->n_records, P.band_ev,
                          P.ev_off, ty, &P.st->n_events, rp));
  RZ_TRY(enqueue_bin_place(ctx, b, lb.fill));
  return check_launch(ctx, "bin pass 1");
}

// Enqueue the whole pipeline with the fine stage writing through `o`
// (MaskOut or FillOut). On capacity overflow the fine kernel writes nothing,
// so a re-run after finish() grew the buffers is exact.
template <class Out>
rz_status enqueue_fine(rz_ctx* ctx, const rz_dpath* d, const rz_transform& tr, double flatness,
                       int rule, int mode, size_t W, size_t H, const Out& o, const char* name0,
                       const char* name1, size_t mask_stride = 0) {
  if (W == 0 || H == 0) return RZ_OK;
  Pipe P;
  RZ_TRY(enqueue_bin(ctx, d, tr, flatness, mode, W, H, &P, mask_stride));
  rzk::FineArgs a;
  a.tile_off = P.tile_off;
  a.recs = (const float4*)ctx->recs.p;
  a.bd = P.bd;
  a.band_ev_off = P.ev_off;
  a.events = (const int4*)ctx->events.p;
  a.st = P.st;
  a.acc = ctx->acc_run;  // the fine kernel merges the run's status (read at rz_sync)
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
    RZ_LAUNCH_LDS(ctx, name0, (rzk::k_fine<0, Out>), dim3(gf), dim3(rzk::kFineWave),
                  ctx->cur.fine_lds, a, o);
  else
    RZ_LAUNCH_LDS(ctx, name1, (rzk::k_fine<1, Out>), dim3(gf), dim3(rzk::kFineWave),
                  ctx->cur.fine_lds, a, o);
  return check_launch(ctx, name0);
}

rz_status enqueue_mask(rz_ctx* ctx, const rz_dpath* d, const rz_transform& tr, double flatness,
                       int rule, int mode, size_t W, size_t H, float* d_out, size_t row_stride) {
  rzk::MaskOut o;
  o.out = d_out;
  o.row_stride = row_stride;
  return enqueue_fine(ctx, d, tr, flatness, rule, mode, W, H, o, "k_fine_mask<0>", "k_fine_mask<1>",
                      row_stride);
}

// Resolve the paint on the host and stage its stops in ctx->stops.
rz_status stage_paint(rz_ctx* ctx, const rz_dpath* d, const rz_transform& tr, const rz_paint* paint,
                      rzh::PaintPrep* pp) {
  if (!paint) return fail(ctx, RZ_E_ARG, "paint is NULL");
  if (!rzh::prepare_paint(*paint, rzh::tr_of(tr.m), d->has_bbox ? &d->bbox : nullptr, pp))
    return fail(ctx, RZ_E_ARG, "invalid paint");
  const size_t n = pp->pos.size();
  if (n) {
    const size_t pos_bytes = (n * sizeof(double) + 15) & ~(size_t)15;
    RZ_TRY(ensure(ctx, ctx->stops, pos_bytes + n * sizeof(rzp::C4)));
    char* base = (char*)ctx->stops.p;
    // pageable sources: the copies are staged before the calls return
    RZ_HIP(ctx, hipMemcpyAsync(base, pp->pos.data(), n * sizeof(double), hipMemcpyHostToDevice,
                               ctx->stream));
    RZ_HIP(ctx, hipMemcpyAsync(base + pos_bytes, pp->col.data(), n * sizeof(rzp::C4),
                               hipMemcpyHostToDevice, ctx->stream));
    pp->p.stop_pos = (const double*)base;
    pp->p.stop_col = (const rzp::C4*)(base + pos_bytes);
  }
  return RZ_OK;
}

rz_status enqueue_fill(rz_ctx* ctx, const rz_dpath* d, const rz_transform& tr, double flatness,
                       int rule, int mode, const rz_paint* paint, size_t W, size_t H, float* d_rgba,
                       size_t row_stride) {
  rzh::PaintPrep pp;
  RZ_TRY(stage_paint(ctx, d, tr, paint, &pp));
  if (!pp.draw || W == 0 || H == 0) return RZ_OK;
  rzk::FillOut o;
  o.img = (rzp::C4*)d_rgba;
  o.row_stride = row_stride;
  o.paint = pp.p;
  return enqueue_fine(ctx, d, tr, flatness, rule, mode, W, H, o, "k_fine_fill<0>", "k_fine_fill<1>");
}

// Wait, then interpret the status merged over every run since the last
// finish: errors, and capacity growth. The accumulated blocks start over.
// No-alloc runs (their own block) never grow anything: a capacity miss there
// means the input does not fit the scratch.
rz_status finish(rz_ctx* ctx) {
  RZ_HIP(ctx, hipMemcpyAsync(ctx->host_status, ctx->acc, 2 * sizeof(rzk::Status),
                             hipMemcpyDeviceToHost, ctx->stream));
  RZ_LAUNCH(ctx, "k_fill_u32", rzk::k_fill_u32, dim3(1), dim3(64), (uint32_t*)ctx->acc,
            (size_t)(2 * rzk::kStatusWords), 0u);
  RZ_HIP(ctx, hipStreamSynchronize(ctx->stream));
  rzk::Status s = ctx->host_status[0];
  const rzk::Status sn = ctx->host_status[1];
  std::memset(ctx->host_status, 0, 2 * sizeof(s));
  ctx->last = s.n_lines || !sn.n_lines ? s : sn;
  const uint32_t err = s.err ? s.err : sn.err;
  if (err == rzg::ST_NAN) return fail(ctx, RZ_E_NAN, "cannot flatten segment with NaN");
  if (err == rzg::ST_DEPTH) return fail(ctx, RZ_E_DEPTH, "subdivision depth limit exceeded");
  if (err == rzg::ST_WRAP)
    return fail(ctx, RZ_E_DEPTH, "a line starting on the right clip runs too far right within a row "
                                 "(its deposits would need more than 256 lines per pixel)");
  // a run that completed with its buffers as they are can be captured next time
  if (ctx->pending_valid && !s.overflow && ctx->graph_pending.gen == ctx->gen) {
    auto& st = ctx->graph_settled;
    bool known = false;
    for (const auto& e : st) known = known || same_key(e, ctx->graph_pending);
    if (!known) {
      if (st.size() >= kMaxGraphs) st.erase(st.begin());
      st.push_back(ctx->graph_pending);
    }
  }
  ctx->pending_valid = false;
  if (sn.overflow && !s.overflow)
    return fail(ctx, RZ_E_CAPACITY,
                "a no-alloc call outgrew the scratch (" + std::to_string(sn.n_lines) +
                    " lines, " + std::to_string(sn.n_records) + " records); its output was not "
                    "written");
  if (s.overflow) {
    ++ctx->gen;
    if (s.n_records > ctx->rec_cap) {
      ctx->rec_cap = (uint32_t)std::min<uint64_t>((uint64_t)s.n_records * 5 / 4 + 1024, 0xffffffffu);
      RZ_TRY(ensure(ctx, ctx->recs, (size_t)ctx->rec_cap * 16));
    }
    if ((s.overflow & 4u) && s.n_line_part > ctx->line_cap / rzk::kLineParts) {
      ctx->line_cap = line_cap_for(((size_t)s.n_line_part * 5 / 4 + 1024) * rzk::kLineParts);
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
    if ((s.overflow & 16u) && s.n_list > ctx->part_cap)
      ctx->part_cap = (uint32_t)std::min<uint64_t>((uint64_t)s.n_list * 5 / 4 + 1024, 1u << 28);
    if (s.overflow & 8u) {
      // a batched scene's per-tile / per-band / per-row slices (totals in pad[])
      auto grow = [](uint32_t need, uint32_t cap) {
        return need > cap ? (uint32_t)std::min<uint64_t>((uint64_t)need * 5 / 4 + 4096, 0xffffffffu) : cap;
      };
      ctx->sc_cap_t = grow(s.pad[0], ctx->sc_cap_t);
      ctx->sc_cap_b = grow(s.pad[1], ctx->sc_cap_b);
      ctx->sc_cap_r = grow(s.pad[2], ctx->sc_cap_r);
    }
    return fail(ctx, RZ_E_CAPACITY, "pipeline buffers grown; re-run");
  }
  return RZ_OK;
}

// Runs of a blocking call at most: each re-run after a capacity miss has grown
// at least one buffer (lines first, as the count pass is skipped then; records,
// events and the record list next; records once more when a full list hid
// some of them from the count).
constexpr int kRuns = 5;

// Blocking mask into a device buffer, re-running after a capacity miss.
rz_status mask_blocking(rz_ctx* ctx, const rz_dpath* d, const rz_transform& tr, double flatness,
                        int rule, int mode, size_t W, size_t H, float* d_out, size_t stride) {
  for (int attempt = 0; attempt < kRuns; ++attempt) {
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
// ---------------------------------------------------------------------------
// Scenes: Scene::group(fills).render(rasterizer, tr, view, bg)
// ---------------------------------------------------------------------------
struct rz_dscene {
  uint64_t serial = g_serial.fetch_add(1);  // renewed by rz_atlas_set
  uint32_t n = 0;
  rz_dpath* all = nullptr;             // every path's segments, for the batched form
  std::vector<rz_dpath*> per;          // one resident path each, for the plain form
  std::vector<void*> allocs;           // ScenePath table, slot -> path map, stops
  const rzk::ScenePath* sp = nullptr;
  rzk::ScenePath* sp_mut = nullptr;
  const uint32_t* slot_path = nullptr;
  const uint32_t* subpaths = nullptr;  // combined path: subpath -> first segment
  uint32_t n_points = 0;               // points of the combined path
  std::vector<rzk::ScenePath> host_sp;  // same table on the host (plain form)
  // gradient stops: as the paints read them (sorted, in the space they are
  // interpolated in), their colours as given, and their index among the
  // scene's given stops (paths in order, each paint's stops as given)
  uint32_t n_stops = 0, n_user_stops = 0;
  rzp::C4* stop_col_dev = nullptr;
  rzp::C4* stop_lin_dev = nullptr;
  int32_t* stop_user_dev = nullptr;   // -1: a gradient's default stop
  uint8_t* stop_flag_dev = nullptr;   // 1: its gradient interpolates linear colours
  // glyph atlas placement (rz_atlas_set): device slots and slice totals
  rzk::AtlasSlot* atlas_slots = nullptr;
  uint32_t* atlas_tile_glyph = nullptr;  // glyph of each atlas tile
  SplitEst atlas_est;                    // line estimate under the slots' transforms
  uint32_t atlas_tiles = 0, atlas_bands = 0, atlas_rows = 0;
  bool atlas_ready = false;
  // host copies of the paths, for Path::bbox(tr) in the plain form
  std::vector<std::vector<uint8_t>> kinds;
  std::vector<std::vector<uint32_t>> offs, subs;
  std::vector<std::vector<double>> pts;
  std::vector<std::vector<uint8_t>> closed;
  rz_path host_path(uint32_t i) const {
    rz_path p;
    p.n_segments = (uint32_t)kinds[i].size();
    p.n_subpaths = subs[i].empty() ? 0 : (uint32_t)subs[i].size() - 1;
    p.kinds = kinds[i].data();
    p.point_offsets = offs[i].data();
    p.points = pts[i].data();
    p.subpaths = subs[i].data();
    p.closed = closed[i].data();
    return p;
  }
};

namespace {

void scene_release(rz_dscene* s) {
  if (s->all) {
    dpath_release(s->all);
    delete s->all;
  }
  for (rz_dpath* d : s->per) {
    dpath_release(d);
    delete d;
  }
  for (void* p : s->allocs) (void)hipFree(p);
  if (s->atlas_slots) (void)hipFree(s->atlas_slots);
  s->atlas_slots = nullptr;
  if (s->atlas_tile_glyph) (void)hipFree(s->atlas_tile_glyph);
  s->atlas_tile_glyph = nullptr;
  s->per.clear();
  s->allocs.clear();
  s->all = nullptr;
}

// The layer of a view (Layer::new, src/scene.rs:483-501).
void layer_of(const double* v, int* lx, int* ly, int* lw, int* lh) {
  const int x0 = (int)std::floor(v[0]), x1 = (int)std::ceil(v[2]);
  const int y0 = (int)std::floor(v[1]), y1 = (int)std::ceil(v[3]);
  *lx = x0;
  *ly = y0;
  *lw = std::max(0, x1 - x0);
  *lh = std::max(0, y1 - y0);
}

// `trs` (6 doubles per path, or NULL): each fill's own transform, drawn
// under render tr * trs[i] (a Transform node above it).
rz_status scene_upload_impl(rz_ctx* ctx, const rz_path* const* paths, const rz_paint* paints,
                            const int32_t* rules, size_t n, rz_dscene** out,
                            const double* trs = nullptr) {
  if (!ctx || !out) return RZ_E_ARG;
  if (n && !paths) return fail(ctx, RZ_E_ARG, "null scene arrays");
  if (n > (1u << 24)) return fail(ctx, RZ_E_ARG, "too many paths");
  rz_dscene* s = new rz_dscene();
  s->n = (uint32_t)n;
  struct Guard {
    rz_dscene* s;
    bool keep = false;
    ~Guard() {
      if (!keep) {
        scene_release(s);
        delete s;
      }
    }
  } guard{s};
  // combined path, slot -> path map, host copies, paints
  std::vector<uint8_t> kinds;
  std::vector<uint32_t> offs{0}, subs{0}, slot_path;
  std::vector<double> pts;
  std::vector<uint8_t> closed;
  std::vector<double> stop_pos;
  std::vector<rzp::C4> stop_col, stop_lin;
  std::vector<int32_t> stop_user;
  std::vector<uint8_t> stop_flag;
  uint32_t user_stops = 0;
  std::vector<size_t> stop_first(n);
  s->host_sp.resize(n);
  for (size_t i = 0; i < n; ++i) {
    const rz_path* p = paths[i];
    RZ_TRY(validate_path(ctx, p));
    const int rule_i = rules ? rules[i] : RZ_NONZERO;
    if (rule_i != RZ_NONZERO && rule_i != RZ_EVENODD) return fail(ctx, RZ_E_ARG, "bad fill rule");
    rz_paint solid{};
    solid.kind = RZ_PAINT_SOLID;
    const rz_paint& paint_i = paints ? paints[i] : solid;
    const uint32_t seg0 = (uint32_t)kinds.size(), pt0 = (uint32_t)(pts.size() / 2);
    const uint32_t ns = p->n_segments, nsub = p->n_subpaths, npts = ns ? p->point_offsets[ns] : 0;
    s->kinds.emplace_back(p->kinds, p->kinds + ns);
    s->offs.emplace_back(p->point_offsets, p->point_offsets + (ns ? ns + 1 : 0));
    s->pts.emplace_back(p->points, p->points + 2 * (size_t)npts);
    s->subs.emplace_back(p->subpaths, p->subpaths + (nsub ? nsub + 1 : 0));
    s->closed.emplace_back(p->closed, p->closed + nsub);
    for (uint32_t k = 0; k < ns; ++k) {
      kinds.push_back(p->kinds[k]);
      offs.push_back(pt0 + p->point_offsets[k + 1]);
    }
    pts.insert(pts.end(), p->points, p->points + 2 * (size_t)npts);
    for (uint32_t j = 0; j < nsub; ++j) {
      subs.push_back(seg0 + p->subpaths[j + 1]);
      closed.push_back(p->closed[j]);
      for (uint32_t k = p->subpaths[j]; k <= p->subpaths[j + 1]; ++k) slot_path.push_back((uint32_t)i);
    }
    rzh::PaintPrep pp;
    if (!rzh::prepare_paint_static(paint_i, &pp)) return fail(ctx, RZ_E_ARG, "invalid paint");
    rzk::ScenePath& S = s->host_sp[i];
    S = rzk::ScenePath{};
    if (trs) {
      S.has_tr = 1;
      for (int k = 0; k < 6; ++k) S.tr[k] = trs[6 * i + k];
    }
    S.seg_begin = seg0;
    S.seg_end = seg0 + ns;
    S.rule = rule_i;
    const rzh::Tr ident = {{1.0, 0.0, 0.0, 0.0, 1.0, 0.0}};
    S.has_bbox_i = rzh::path_bbox(p, ident, &S.bbox_i) ? 1 : 0;
    S.units = paint_i.units;
    for (int k = 0; k < 6; ++k) S.paint_tr[k] = paint_i.tr[k];
    S.paint = pp.p;
    stop_first[i] = stop_pos.size();
    stop_pos.insert(stop_pos.end(), pp.pos.begin(), pp.pos.end());
    stop_col.insert(stop_col.end(), pp.col.begin(), pp.col.end());
    stop_lin.insert(stop_lin.end(), pp.lin.begin(), pp.lin.end());
    for (int32_t k : pp.src) {
      stop_user.push_back(k < 0 ? -1 : (int32_t)user_stops + k);
      stop_flag.push_back(pp.p.linear_colors ? 1 : 0);
    }
    if (paint_i.kind != RZ_PAINT_SOLID) user_stops += paint_i.n_stops;
    rz_dpath* d = nullptr;
    RZ_TRY(path_upload_impl(ctx, p, &d));
    s->per.push_back(d);
  }
  if (kinds.empty()) offs.clear();
  rz_path all;
  all.n_segments = (uint32_t)kinds.size();
  all.n_subpaths = (uint32_t)closed.size();
  all.kinds = kinds.data();
  all.point_offsets = offs.data();
  all.points = pts.data();
  all.subpaths = subs.data();
  all.closed = closed.data();
  RZ_TRY(path_upload_impl(ctx, &all, &s->all));
  // stops, then the ScenePath table pointing into them
  auto up = [&](const void* src, size_t bytes, void** dst) -> rz_status {
    void* p = nullptr;
    hipError_t e = hipMalloc(&p, std::max<size_t>(bytes, 16));
    if (e != hipSuccess) return fail(ctx, RZ_E_NOMEM, std::string("hipMalloc: ") + hipGetErrorString(e));
    s->allocs.push_back(p);
    if (bytes) RZ_HIP(ctx, hipMemcpy(p, src, bytes, hipMemcpyHostToDevice));
    *dst = p;
    return RZ_OK;
  };
  void *dpos = nullptr, *dcol = nullptr, *dsp = nullptr, *dslot = nullptr, *dsubs = nullptr;
  RZ_TRY(up(stop_pos.data(), stop_pos.size() * sizeof(double), &dpos));
  RZ_TRY(up(stop_col.data(), stop_col.size() * sizeof(rzp::C4), &dcol));
  void *dlin = nullptr, *duser = nullptr, *dflag = nullptr;
  RZ_TRY(up(stop_lin.data(), stop_lin.size() * sizeof(rzp::C4), &dlin));
  RZ_TRY(up(stop_user.data(), stop_user.size() * sizeof(int32_t), &duser));
  RZ_TRY(up(stop_flag.data(), stop_flag.size(), &dflag));
  s->n_stops = (uint32_t)stop_col.size();
  s->n_user_stops = user_stops;
  s->stop_col_dev = (rzp::C4*)dcol;
  s->stop_lin_dev = (rzp::C4*)dlin;
  s->stop_user_dev = (int32_t*)duser;
  s->stop_flag_dev = (uint8_t*)dflag;
  for (size_t i = 0; i < n; ++i) {
    rzk::ScenePath& S = s->host_sp[i];
    S.paint.stop_pos = (const double*)dpos + stop_first[i];
    S.paint.stop_col = (const rzp::C4*)dcol + stop_first[i];
  }
  RZ_TRY(up(s->host_sp.data(), n * sizeof(rzk::ScenePath), &dsp));
  RZ_TRY(up(slot_path.data(), slot_path.size() * sizeof(uint32_t), &dslot));
  RZ_TRY(up(subs.data(), subs.size() * sizeof(uint32_t), &dsubs));
  s->sp = (const rzk::ScenePath*)dsp;
  s->sp_mut = (rzk::ScenePath*)dsp;
  s->slot_path = (const uint32_t*)dslot;
  s->subpaths = (const uint32_t*)dsubs;
  s->n_points = (uint32_t)(pts.size() / 2);
  guard.keep = true;
  *out = s;
  return RZ_OK;
}

// The batched form: one pipeline for every path of the scene. With d_out
// null it stops before the fine stage (the adjoint needs only the binned
// state).
// `layer` (x, y, w, h; NULL: the layer of the view, Layer::new) and `load`
// (draw over the layer's pixels rather than start from bg) serve the scene
// tree, whose fills between two Opacity / Clip nodes are drawn in batches
// into layers that are not the view's.
TuneKey scene_tune_key(const rz_dscene* s, const rz_transform& tr, double flatness, int mode,
                       const double* view) {
  int lx, ly, lw, lh;
  layer_of(view, &lx, &ly, &lw, &lh);
  return tune_key(1, mode, (size_t)lw * lh, s->all ? s->all->n_slots : 0,
                  s->all ? s->all->est : SplitEst{}, det_of(tr), flatness);
}

TuneKey atlas_tune_key(const rz_dscene* s, double flatness, int mode) {
  return tune_key(2, mode, (size_t)s->atlas_tiles * 256, s->all ? s->all->n_slots : 0,
                  s->atlas_est, 1.0, flatness);
}

rz_status enqueue_scene_batched(rz_ctx* ctx, const rz_dscene* s, const rz_transform& tr,
                                const double* view, rzp::C4 bg, int mode, double flatness,
                                rzp::C4* d_out, size_t row_stride, const int* layer = nullptr,
                                int load = 0) {
  select_params(ctx, scene_tune_key(s, tr, flatness, mode, view));
  int lx, ly, lw, lh;
  if (layer) {
    lx = layer[0];
    ly = layer[1];
    lw = layer[2];
    lh = layer[3];
  } else {
    layer_of(view, &lx, &ly, &lw, &lh);
  }
  const uint32_t n = s->n;
  if (lw == 0 || lh == 0) return RZ_OK;
  const int ctx_x = (lw + 15) / 16, cty = (lh + 15) / 16;
  if (n == 0) {
    if (load) return RZ_OK;
    RZ_LAUNCH(ctx, "k_fill_rgba", rzk::k_fill_rgba, dim3(1024), dim3(256), d_out, lw, lh, row_stride, bg);
    return check_launch(ctx, "k_fill_rgba");
  }
  const rz_dpath* d = s->all;
  // buffers
  if (ctx->sc_cap_t == 0) {
    ctx->sc_cap_t = (uint32_t)std::max<size_t>((size_t)ctx_x * cty * 16, 1 << 16);
    ctx->sc_cap_b = (uint32_t)std::max<size_t>((size_t)cty * 4 + n * 4, 1 << 12);
    ctx->sc_cap_r = (uint32_t)std::max<size_t>((size_t)lh * 4 + n * 66, 1 << 14);
  }
  if (ctx->line_cap == 0) {
    ctx->line_cap = line_cap_for(std::max<size_t>((size_t)d->n_slots * 16, 1 << 16));
    RZ_TRY(ensure(ctx, ctx->lines, (size_t)ctx->line_cap * 32));
  }
  if (ctx->rec_cap == 0) {
    ctx->rec_cap = ctx->line_cap;
    RZ_TRY(ensure(ctx, ctx->recs, (size_t)ctx->rec_cap * 16));
  }
  if (ctx->ev_cap == 0) {
    ctx->ev_cap = 1 << 14;
    RZ_TRY(ensure(ctx, ctx->events, (size_t)ctx->ev_cap * 16));
  }
  const uint32_t ct = ctx->sc_cap_t, cb = ctx->sc_cap_b, cr = ctx->sc_cap_r;
  const size_t sw = rzk::kStatusWords;
  // per path four 64-bit box keys (k_scene_bbox) after the right-edge rows
  const size_t ro = redge_words(sw + 2 * (size_t)ct + cb);
  const size_t kb = ro + 2 * (size_t)cr;
  const size_t meta_u32 = kb + 8 * (size_t)n + rzk::kParts + 1 + rzk::kLineParts;
  RZ_TRY(ensure(ctx, ctx->meta, meta_u32 * 4));
  RZ_TRY(ensure(ctx, ctx->tile_off, ((size_t)ct + 1) * 4));
  RZ_TRY(ensure(ctx, ctx->ev_off, ((size_t)cb + 1) * 4));
  RZ_TRY(ensure(ctx, ctx->line_path, (size_t)ctx->line_cap * 4));
  RZ_TRY(ensure(ctx, ctx->sc_pi, (size_t)n * sizeof(rzk::PathInfo)));
  RZ_TRY(ensure(ctx, ctx->sc_counts, (size_t)n * 12));
  RZ_TRY(ensure(ctx, ctx->sc_rects, (size_t)n * sizeof(int4)));
  RZ_TRY(ensure(ctx, ctx->sc_rows, (size_t)cty * n * 4));
  RZ_TRY(ensure(ctx, ctx->sc_rowcnt, (size_t)cty * 4));
  uint32_t* meta = (uint32_t*)ctx->meta.p;
  rzk::Status* st = (rzk::Status*)meta;
  uint32_t* tile_count = meta + sw;
  int32_t* bd = (int32_t*)(meta + sw + ct);
  uint32_t* band_ev = meta + sw + 2 * (size_t)ct;
  long long* redge = (long long*)(meta + ro);
  uint32_t* tile_off = (uint32_t*)ctx->tile_off.p;
  uint32_t* ev_off = (uint32_t*)ctx->ev_off.p;
  rzk::PathInfo* pi = (rzk::PathInfo*)ctx->sc_pi.p;
  RZ_LAUNCH(ctx, "k_fill_u32", rzk::k_fill_u32,
            dim3((uint32_t)std::min<size_t>(blocks_for(meta_u32), 4096)), dim3(rzk::kBlock), meta,
            meta_u32, 0u);
  rzk::SceneView v;
  v.vx0 = view[0];
  v.vy0 = view[1];
  v.vx1 = view[2];
  v.vy1 = view[3];
  v.lx = lx;
  v.ly = ly;
  v.lw = lw;
  v.lh = lh;
  for (int k = 0; k < 6; ++k) v.tr[k] = tr.m[k];
  v.mode = mode;
  unsigned long long* bkeys = (unsigned long long*)(meta + kb);
  if (d->n_slots)
    RZ_LAUNCH(ctx, "k_scene_bbox", rzk::k_scene_bbox, dim3((d->n_slots + 63) / 64), dim3(64),
              d->view, s->slot_path, v, s->sp, bkeys);
  RZ_LAUNCH(ctx, "k_scene_setup", rzk::k_scene_setup, dim3(blocks_for(n)), dim3(rzk::kBlock), s->sp,
            n, v, (const unsigned long long*)bkeys, pi, (uint32_t*)ctx->sc_counts.p,
            (int4*)ctx->sc_rects.p);
  RZ_LAUNCH(ctx, "k_scene_offsets", rzk::k_scene_offsets, dim3(1), dim3(rzk::kBlock), n,
            (const uint32_t*)ctx->sc_counts.p, pi, st, ct, cb, cr);
  // flatten every path into its sub-view
  rzk::Xform x{};
  const double fl = 16.0 * flatness * flatness;
  rzk::LineBufs lb;
  uint32_t* bin_fill = meta + meta_u32 - (rzk::kParts + 1 + rzk::kLineParts);
  int dq, dc;
  split_depths(ctx, d->est, std::fabs(tr.m[0] * tr.m[4] - tr.m[1] * tr.m[3]), flatness, &dq, &dc);
  const bool fused = ctx->cur.fused != 0;
  RZ_TRY(enqueue_flatten(ctx, d, x, fl, pi, s->slot_path, st, bin_fill + rzk::kParts + 1, dq, dc,
                         mode, 0.0, rzg::WrapClip{-1.0, 1.0, 0, 0}, &lb, !fused));
  rzk::BinBufs b;
  b.tile_count = tile_count;
  b.tile_off = tile_off;
  b.bd_diff = bd;
  b.band_ev = band_ev;
  b.recs = (float4*)ctx->recs.p;
  b.rec_cap = ctx->rec_cap;
  b.events = (int4*)ctx->events.p;
  b.ev_cap = ctx->ev_cap;
  b.redge = redge;
  b.st = st;
  rzg::BinGrid g0{};
  const rzk::GapArgs ga{d->view, x, s->slot_path};
  RZ_TRY(enqueue_bin_count<true>(ctx, lb, g0, pi, ga, bin_fill, &b, !fused));
  if (fused)
    RZ_TRY(enqueue_flat_bin<true>(ctx, d, x, fl, pi, s->slot_path, dq, dc, mode, lb, g0, b, ga));
  RZ_TRY(check_launch(ctx, "scene pass 0"));
  RZ_TRY(scan_tiles_bands(ctx, tile_count, tile_off, ct, &st->n_records, band_ev, ev_off, cb,
                          &st->n_events));
  RZ_LAUNCH(ctx, "k_scene_bd", rzk::k_scene_bd, dim3(n), dim3(64), (const rzk::PathInfo*)pi, bd,
            redge, (const rzk::Status*)st);
  RZ_TRY(enqueue_bin_place(ctx, b, lb.fill));
  RZ_LAUNCH(ctx, "k_scene_rows", rzk::k_scene_rows, dim3((uint32_t)cty), dim3(64),
            (const int4*)ctx->sc_rects.p, n, (uint32_t*)ctx->sc_rows.p, (uint32_t*)ctx->sc_rowcnt.p,
            (const rzk::Status*)st);
  rzk::SceneFineArgs a;
  a.pi = pi;
  a.rects = (const int4*)ctx->sc_rects.p;
  a.rows = (const uint32_t*)ctx->sc_rows.p;
  a.row_count = (const uint32_t*)ctx->sc_rowcnt.p;
  a.n_paths = n;
  a.tile_off = tile_off;
  a.recs = (const float4*)ctx->recs.p;
  a.bd = bd;
  a.band_ev_off = ev_off;
  a.events = (const int4*)ctx->events.p;
  a.redge = redge;
  a.st = st;
  a.acc = ctx->acc_run;
  a.bg = bg;
  a.load_dst = load;
  a.out = d_out;
  a.row_stride = row_stride;
  a.lw = lw;
  a.lh = lh;
  a.tiles_x = ctx_x;
  uint32_t* queue = nullptr;
  a.order = nullptr;
  a.n_tiles = (uint32_t)(ctx_x * cty);
  // no layer (the adjoint's own tile pass recomposites): the binned state only
  if (d_out) {
    dim3 fgrid((uint32_t)(ctx_x * cty));
    if (ctx->cur.scene_queue) {
      const size_t nt = (size_t)ctx_x * cty, nd = (size_t)(ctx_x + 1) * (cty + 1);
      RZ_TRY(ensure(ctx, ctx->sc_order, (nt + nd + 64) * 4));
      uint32_t* ord = (uint32_t*)ctx->sc_order.p;
      uint32_t* q = ord + nt;
      int* cnt = (int*)(q + 64);
      RZ_LAUNCH(ctx, "k_scene_order", rzk::k_scene_order, dim3(1), dim3(rzk::kOrderBlock),
                (const int4*)ctx->sc_rects.p, s->n, ctx_x, cty, cnt, ord, q, (const rzk::Status*)st);
      queue = q;
      a.order = ord;
      // persistent: as many waves as the chip holds of this kernel
      const uint32_t resident = (uint32_t)ctx->n_cus * 4u * (uint32_t)std::max(ctx->cur.scene_waves, 5);
      fgrid = dim3(std::min<uint32_t>(a.n_tiles, resident));
    }
    const int waves = ctx->cur.scene_waves;
#define RZ_SCENE_FINE(W, Q)                                                                \
  RZ_LAUNCH_LDS(ctx, "k_scene_fine", (rzk::k_scene_fine<W, Q>), fgrid, dim3(64), ctx->cur.fine_lds, \
                a, queue)
    if (queue) {
      if (waves == 8) RZ_SCENE_FINE(8, true);
      else if (waves == 1) RZ_SCENE_FINE(1, true);
      else RZ_SCENE_FINE(6, true);
    } else {
      if (waves == 8) RZ_SCENE_FINE(8, false);
      else if (waves == 1) RZ_SCENE_FINE(1, false);
      else RZ_SCENE_FINE(6, false);
    }
#undef RZ_SCENE_FINE
  } else {
    RZ_LAUNCH(ctx, "k_status_merge", rzk::k_status_merge, dim3(1), dim3(64), (const rzk::Status*)st,
              ctx->acc_run);
  }
  return check_launch(ctx, "scene pass 1");
}

// The plain form: Pipeline::render_rec's Fill arm (src/scene.rs:410-430) run
// path by path, each through the single-path pipeline into its sub-view.
rz_status enqueue_scene_plain(rz_ctx* ctx, const rz_dscene* s, const rz_transform& tr,
                              const double* view, rzp::C4 bg, int mode, double flatness,
                              rzp::C4* d_out, size_t row_stride) {
  int lx, ly, lw, lh;
  layer_of(view, &lx, &ly, &lw, &lh);
  if (lw == 0 || lh == 0) return RZ_OK;
  RZ_LAUNCH(ctx, "k_fill_rgba", rzk::k_fill_rgba, dim3(1024), dim3(256), d_out, lw, lh, row_stride, bg);
  // each path's run merges its status into the context's (enqueue_fine)
  const rzh::Tr T = rzh::tr_of(tr.m);
  for (uint32_t i = 0; i < s->n; ++i) {
    const rz_path hp = s->host_path(i);
    rzh::BBox b;
    if (!rzh::path_bbox(&hp, T, &b)) continue;
    if (view[0] > b.mx.x || b.mn.x > view[2] || view[1] > b.mx.y || b.mn.y > view[3]) continue;
    const double bx0 = std::fmax(view[0], b.mn.x), bx1 = std::fmin(view[2], b.mx.x);
    const double by0 = std::fmax(view[1], b.mn.y), by1 = std::fmin(view[3], b.mx.y);
    const double fx = std::floor(bx0), fy = std::floor(by0);
    int col_min = (int)fx - lx, col_max = (int)std::ceil(bx1) - lx + 1;
    int row_min = (int)fy - ly, row_max = (int)std::ceil(by1) - ly + 1;
    col_min = std::min(std::max(col_min, 0), lw);
    col_max = std::min(std::max(col_max, col_min), lw);
    row_min = std::min(std::max(row_min, 0), lh);
    row_max = std::min(std::max(row_max, row_min), lh);
    const int W = col_max - col_min, H = row_max - row_min;
    if (W <= 0 || H <= 0) continue;
    const rzh::Tr align = {{1.0, 0.0, -fx, 0.0, 1.0, -fy}};
    const rzh::Tr m = rzh::tr_mul(align, T);
    const rzk::ScenePath& S = s->host_sp[i];
    rzk::FillOut o;
    o.img = d_out + (size_t)row_min * row_stride + col_min;
    o.row_stride = row_stride;
    o.paint = S.paint;
    if (S.paint.kind != rzp::PAINT_SOLID &&
        !rzh::paint_pixel_transform(S.units, S.paint_tr, m, S.has_bbox_i ? &S.bbox_i : nullptr,
                                    o.paint.ptr))
      continue;
    rz_transform tm;
    for (int k = 0; k < 6; ++k) tm.m[k] = m.m[k];
    RZ_TRY(enqueue_fine(ctx, s->per[i], tm, flatness, S.rule, mode, (size_t)W, (size_t)H, o,
                        "k_fine_fill<0>", "k_fine_fill<1>"));
  }
  return RZ_OK;
}

// Glyph atlas: every path of the set masked into its own atlas view, in one
// pipeline (flatten and binning in their batched form, one wave per glyph tile).
rz_status enqueue_atlas(rz_ctx* ctx, const rz_dscene* s, int rule, int mode, double flatness,
                        float* d_atlas, size_t row_stride) {
  const uint32_t n = s->n;
  if (n == 0 || s->atlas_tiles == 0) return RZ_OK;
  select_params(ctx, atlas_tune_key(s, flatness, mode));
  const rz_dpath* d = s->all;
  if (ctx->line_cap == 0) {
    ctx->line_cap = line_cap_for(std::max<size_t>((size_t)d->n_slots * 16, 1 << 16));
    RZ_TRY(ensure(ctx, ctx->lines, (size_t)ctx->line_cap * 32));
  }
  if (ctx->rec_cap == 0) {
    ctx->rec_cap = ctx->line_cap;
    RZ_TRY(ensure(ctx, ctx->recs, (size_t)ctx->rec_cap * 16));
  }
  if (ctx->ev_cap == 0) {
    ctx->ev_cap = 1 << 14;
    RZ_TRY(ensure(ctx, ctx->events, (size_t)ctx->ev_cap * 16));
  }
  const uint32_t ct = s->atlas_tiles, cb = s->atlas_bands, cr = s->atlas_rows;
  const size_t sw = rzk::kStatusWords;
  const size_t ro = redge_words(sw + 2 * (size_t)ct + cb);
  const size_t meta_u32 = ro + 2 * (size_t)cr + rzk::kParts + 1 + rzk::kLineParts;
  RZ_TRY(ensure(ctx, ctx->meta, meta_u32 * 4));
  RZ_TRY(ensure(ctx, ctx->tile_off, ((size_t)ct + 1) * 4));
  RZ_TRY(ensure(ctx, ctx->ev_off, ((size_t)cb + 1) * 4));
  RZ_TRY(ensure(ctx, ctx->line_path, (size_t)ctx->line_cap * 4));
  RZ_TRY(ensure(ctx, ctx->sc_pi, (size_t)n * sizeof(rzk::PathInfo)));
  uint32_t* meta = (uint32_t*)ctx->meta.p;
  rzk::Status* st = (rzk::Status*)meta;
  uint32_t* tile_count = meta + sw;
  int32_t* bd = (int32_t*)(meta + sw + ct);
  uint32_t* band_ev = meta + sw + 2 * (size_t)ct;
  long long* redge = (long long*)(meta + ro);
  uint32_t* tile_off = (uint32_t*)ctx->tile_off.p;
  uint32_t* ev_off = (uint32_t*)ctx->ev_off.p;
  rzk::PathInfo* pi = (rzk::PathInfo*)ctx->sc_pi.p;
  RZ_LAUNCH(ctx, "k_fill_u32", rzk::k_fill_u32,
            dim3((uint32_t)std::min<size_t>(blocks_for(meta_u32), 4096)), dim3(rzk::kBlock), meta,
            meta_u32, 0u);
  RZ_LAUNCH(ctx, "k_atlas_setup", rzk::k_atlas_setup, dim3(blocks_for(n)), dim3(rzk::kBlock),
            (const rzk::AtlasSlot*)s->atlas_slots, n, mode, pi);
  rzk::Xform x{};
  const double fl = 16.0 * flatness * flatness;
  rzk::LineBufs lb;
  uint32_t* bin_fill = meta + meta_u32 - (rzk::kParts + 1 + rzk::kLineParts);
  int dq, dc;
  split_depths(ctx, s->atlas_est, 1.0, flatness, &dq, &dc);
  const bool fused = ctx->cur.fused != 0;
  RZ_TRY(enqueue_flatten(ctx, d, x, fl, pi, s->slot_path, st, bin_fill + rzk::kParts + 1, dq, dc,
                         mode, 0.0, rzg::WrapClip{-1.0, 1.0, 0, 0}, &lb, !fused));
  rzk::BinBufs b;
  b.tile_count = tile_count;
  b.tile_off = tile_off;
  b.bd_diff = bd;
  b.band_ev = band_ev;
  b.recs = (float4*)ctx->recs.p;
  b.rec_cap = ctx->rec_cap;
  b.events = (int4*)ctx->events.p;
  b.ev_cap = ctx->ev_cap;
  b.redge = redge;
  b.st = st;
  rzg::BinGrid g0{};
  const rzk::GapArgs ga{d->view, x, s->slot_path};
  RZ_TRY(enqueue_bin_count<true>(ctx, lb, g0, pi, ga, bin_fill, &b, !fused));
  if (fused)
    RZ_TRY(enqueue_flat_bin<true>(ctx, d, x, fl, pi, s->slot_path, dq, dc, mode, lb, g0, b, ga));
  RZ_TRY(check_launch(ctx, "atlas pass 0"));
  RZ_TRY(scan_tiles_bands(ctx, tile_count, tile_off, ct, &st->n_records, band_ev, ev_off, cb,
                          &st->n_events));
  RZ_LAUNCH(ctx, "k_scene_bd", rzk::k_scene_bd, dim3(n), dim3(64), (const rzk::PathInfo*)pi, bd,
            redge, (const rzk::Status*)st);
  RZ_TRY(enqueue_bin_place(ctx, b, lb.fill));
  rzk::AtlasFineArgs a;
  RZ_TRY(ensure(ctx, ctx->atlas_tiles, (size_t)ct * sizeof(rzk::AtlasTile)));
  RZ_LAUNCH(ctx, "k_atlas_tiles", rzk::k_atlas_tiles, dim3(blocks_for(ct)), dim3(rzk::kBlock),
            (const rzk::PathInfo*)pi, (const uint32_t*)s->atlas_tile_glyph, ct,
            (const uint32_t*)tile_off, (const uint32_t*)ev_off, (const int32_t*)bd,
            (rzk::AtlasTile*)ctx->atlas_tiles.p);
  a.tiles = (const rzk::AtlasTile*)ctx->atlas_tiles.p;
  a.n = n;
  a.recs = (const float4*)ctx->recs.p;
  a.events = (const int4*)ctx->events.p;
  a.redge = redge;
  a.st = st;
  a.acc = ctx->acc_run;
  a.out = d_atlas;
  a.row_stride = row_stride;
  if (rule == RZ_NONZERO)
    RZ_LAUNCH_LDS(ctx, "k_atlas_fine<0>", rzk::k_atlas_fine<0>, dim3(ct), dim3(64), ctx->cur.fine_lds, a);
  else
    RZ_LAUNCH_LDS(ctx, "k_atlas_fine<1>", rzk::k_atlas_fine<1>, dim3(ct), dim3(64), ctx->cur.fine_lds, a);
  return check_launch(ctx, "atlas pass 1");
}

// New colours for a scene's given gradient stops (f32 RGBA each): the stops
// the paints read, converted to sRGB unless their gradient interpolates
// linear colours (GradLinear::new / GradRadial::new, src/grad.rs:163-194).
__global__ void k_set_stops(const float* colors, const int32_t* user, const uint8_t* flag,
                            rzp::C4* lin, rzp::C4* col, uint32_t n) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n || user[i] < 0) return;
  const float* c = colors + 4 * (size_t)user[i];
  const rzp::C4 v = rzp::c4(c[0], c[1], c[2], c[3]);
  lin[i] = v;
  col[i] = flag[i] ? v : rzp::into_srgb(v);
}

// dL/d(given stop colour) += J^T dL/d(stop as read), through into_srgb where
// the gradient interpolates sRGB colours.
__global__ void k_stop_grads(const float* acc, const int32_t* user, const uint8_t* flag,
                             const rzp::C4* lin, float* grad, uint32_t n, const rzk::Status* st) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (st->err || st->overflow) return;  // the run was not completed: nothing to add
  if (i >= n || user[i] < 0) return;
  rzp::C4 g = rzp::c4(acc[4 * i], acc[4 * i + 1], acc[4 * i + 2], acc[4 * i + 3]);
  if (!flag[i]) g = rzp::into_srgb_vjp(lin[i], g);
  float* o = grad + 4 * (size_t)user[i];
  o[0] += g.r;
  o[1] += g.g;
  o[2] += g.b;
  o[3] += g.a;
}

// Adjoint of the batched render (see rz_backward.hpp): the forward pipeline
// again (its tile lists and records are recomputed, not kept), then the tile
// pass (compositing adjoint) and the line pass (coverage and subdivision
// adjoint). Gradients are accumulated into the caller's buffers.
rz_status enqueue_scene_backward(rz_ctx* ctx, const rz_dscene* s, const rz_transform& tr,
                                 const double* view, rzp::C4 bg, int mode, double flatness,
                                 const float* grad_out, size_t grad_stride, float* grad_points,
                                 float* grad_colors, float* grad_stops, double* grad_tr) {
  int lx, ly, lw, lh;
  layer_of(view, &lx, &ly, &lw, &lh);
  if (lw == 0 || lh == 0 || s->n == 0) return RZ_OK;
  // the forward pipeline up to the fine stage (k_scene_bwd_tiles composites
  // the paths again itself)
  RZ_TRY(enqueue_scene_batched(ctx, s, tr, view, bg, mode, flatness, nullptr, 0));
  const size_t entries = ctx->sc_cap_t;
  RZ_TRY(ensure(ctx, ctx->bw_w, entries * 256 * sizeof(float)));
  RZ_TRY(ensure(ctx, ctx->bw_p, entries * 256 * sizeof(rzp::C4)));
  const uint32_t ct = ctx->sc_cap_t, cb = ctx->sc_cap_b;
  const size_t sw = rzk::kStatusWords;
  uint32_t* meta = (uint32_t*)ctx->meta.p;
  rzk::Status* st = (rzk::Status*)meta;
  rzk::SceneFineArgs a;
  a.pi = (const rzk::PathInfo*)ctx->sc_pi.p;
  a.rects = (const int4*)ctx->sc_rects.p;
  a.rows = (const uint32_t*)ctx->sc_rows.p;
  a.row_count = (const uint32_t*)ctx->sc_rowcnt.p;
  a.n_paths = s->n;
  a.tile_off = (const uint32_t*)ctx->tile_off.p;
  a.recs = (const float4*)ctx->recs.p;
  a.bd = (const int32_t*)(meta + sw + ct);
  a.band_ev_off = (const uint32_t*)ctx->ev_off.p;
  a.events = (const int4*)ctx->events.p;
  a.redge = (const long long*)(meta + redge_words(sw + 2 * (size_t)ct + cb));
  a.st = st;
  a.acc = nullptr;
  a.bg = bg;
  a.load_dst = 0;
  a.out = nullptr;
  a.row_stride = 0;
  a.lw = lw;
  a.lh = lh;
  a.tiles_x = (lw + 15) / 16;
  rzk::BwdTileArgs bt;
  bt.f = a;
  bt.grad_out = (const rzp::C4*)grad_out;
  bt.grad_stride = grad_stride;
  bt.wbuf = (float*)ctx->bw_w.p;
  bt.pbuf = (rzp::C4*)ctx->bw_p.p;
  bt.grad_colors = grad_colors;
  bt.stop_acc = nullptr;
  bt.stop_base = s->stop_col_dev;
  const bool stops = grad_stops && s->n_user_stops;
  if (stops) {
    RZ_TRY(ensure(ctx, ctx->bw_stops, (size_t)s->n_stops * 4 * sizeof(float)));
    bt.stop_acc = (float*)ctx->bw_stops.p;
    RZ_LAUNCH(ctx, "k_fill_u32", rzk::k_fill_u32, dim3(blocks_for((size_t)s->n_stops * 4)),
              dim3(rzk::kBlock), (uint32_t*)ctx->bw_stops.p, (size_t)s->n_stops * 4, 0u);
  }
  const uint32_t ntiles = (uint32_t)(((lw + 15) / 16) * ((lh + 15) / 16));
  RZ_LAUNCH(ctx, "k_scene_bwd_tiles", rzk::k_scene_bwd_tiles, dim3(ntiles), dim3(64), bt);
  if (stops)
    RZ_LAUNCH(ctx, "k_stop_grads", k_stop_grads, dim3((s->n_stops + 255) / 256), dim3(256),
              (const float*)ctx->bw_stops.p, (const int32_t*)s->stop_user_dev,
              (const uint8_t*)s->stop_flag_dev, (const rzp::C4*)s->stop_lin_dev, grad_stops,
              s->n_stops, (const rzk::Status*)st);
  // line pass over the same flatten items as the forward
  const rz_dpath* d = s->all;
  int dq, dc;
  split_depths(ctx, d->est, std::fabs(tr.m[0] * tr.m[4] - tr.m[1] * tr.m[3]), flatness, &dq, &dc);
  const uint32_t bl = d->view.blocks_line, bq = d->view.blocks_quad << dq,
                 bc = d->view.blocks_cubic << dc;
  const uint32_t nblk = bl + bq + bc;
  if (nblk) {
    const size_t nwaves = (size_t)nblk * rzk::kBlock / 64;
    RZ_TRY(ensure(ctx, ctx->bw_tr, nwaves * 6 * sizeof(double)));
    rzk::BwdLineArgs bl_args;
    bl_args.pi = (const rzk::PathInfo*)ctx->sc_pi.p;
    bl_args.slot_path = s->slot_path;
    bl_args.subpaths = s->subpaths;
    bl_args.wbuf = (const float*)ctx->bw_w.p;
    bl_args.grad_points = grad_points;
    bl_args.grad_tr_part = (double*)ctx->bw_tr.p;
    rzk::Xform x{};
    const double fl = 16.0 * flatness * flatness;
    RZ_LAUNCH(ctx, "k_scene_bwd_lines", rzk::k_scene_bwd_lines, dim3(nblk), dim3(rzk::kBlock), d->view,
              x, fl, bl, bq, dq, dc, (const rzk::PathInfo*)ctx->sc_pi.p, s->slot_path, bl_args,
              (const rzk::Status*)st);
    if (grad_tr)
      RZ_LAUNCH(ctx, "k_sum_tr", rzk::k_sum_tr, dim3(1), dim3(rzk::kBlock),
                (const double*)ctx->bw_tr.p, nwaves, grad_tr, (const rzk::Status*)st);
  }
  return check_launch(ctx, "scene backward");
}

__global__ void k_set_colors(rzk::ScenePath* sp, const float* colors, uint32_t n) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n && sp[i].paint.kind == rzp::PAINT_SOLID)
    sp[i].paint.color = rzp::c4(colors[4 * i], colors[4 * i + 1], colors[4 * i + 2], colors[4 * i + 3]);
}

rz_status atlas_set_impl(rz_ctx* ctx, rz_dscene* s, const rz_atlas_slot* slots) {
  if (!ctx || !s) return RZ_E_ARG;
  if (s->n && !slots) return fail(ctx, RZ_E_ARG, "slots is NULL");
  std::vector<rzk::AtlasSlot> v(s->n);
  uint64_t t = 0, b = 0, r = 0;
  for (uint32_t i = 0; i < s->n; ++i) {
    rzk::AtlasSlot& a = v[i];
    const rz_atlas_slot& q = slots[i];
    if (q.width > (1u << 16) || q.height > (1u << 16) || q.x > (1u << 30) || q.y > (1u << 30))
      return fail(ctx, RZ_E_ARG, "atlas slot too large");
    for (int k = 0; k < 6; ++k) a.m[k] = q.tr.m[k];
    a.x = (int)q.x;
    a.y = (int)q.y;
    a.w = (int)q.width;
    a.h = (int)q.height;
    a.tile_base = (uint32_t)t;
    a.band_base = (uint32_t)b;
    a.redge_base = (uint32_t)r;
    a.pad = 0;
    const uint64_t tx = (q.width + 15) / 16, ty = (q.height + 15) / 16;
    if (q.width && q.height) {
      t += tx * ty;
      b += ty;
      r += q.height + 2;
    }
  }
  if (t > 0xffffffffull || r > 0xffffffffull) return fail(ctx, RZ_E_ARG, "atlas too large");
  SplitSum sum;
  for (uint32_t i = 0; i < s->n; ++i) {
    const rz_path hp = s->host_path(i);
    split_sum(&hp, slots[i].tr.m, &sum);
  }
  s->atlas_est = sum.est();
  std::vector<uint32_t> tile_glyph(t);
  for (uint32_t i = 0; i < s->n; ++i) {
    const uint64_t end = i + 1 < s->n ? v[i + 1].tile_base : t;
    for (uint64_t k = v[i].tile_base; k < end; ++k) tile_glyph[k] = i;
  }
  if (s->atlas_slots || s->atlas_tile_glyph) {
    RZ_HIP(ctx, hipStreamSynchronize(ctx->stream));
    if (s->atlas_slots) RZ_HIP(ctx, hipFree(s->atlas_slots));
    if (s->atlas_tile_glyph) RZ_HIP(ctx, hipFree(s->atlas_tile_glyph));
    s->atlas_slots = nullptr;
    s->atlas_tile_glyph = nullptr;
  }
  void* p = nullptr;
  hipError_t e = hipMalloc(&p, std::max<size_t>(v.size(), 1) * sizeof(rzk::AtlasSlot));
  if (e != hipSuccess) return fail(ctx, RZ_E_NOMEM, std::string("hipMalloc: ") + hipGetErrorString(e));
  s->atlas_slots = (rzk::AtlasSlot*)p;
  if (!v.empty()) RZ_HIP(ctx, hipMemcpy(p, v.data(), v.size() * sizeof(rzk::AtlasSlot), hipMemcpyHostToDevice));
  void* q = nullptr;
  e = hipMalloc(&q, std::max<size_t>(tile_glyph.size(), 1) * sizeof(uint32_t));
  if (e != hipSuccess) return fail(ctx, RZ_E_NOMEM, std::string("hipMalloc: ") + hipGetErrorString(e));
  s->atlas_tile_glyph = (uint32_t*)q;
  if (!tile_glyph.empty())
    RZ_HIP(ctx, hipMemcpy(q, tile_glyph.data(), tile_glyph.size() * sizeof(uint32_t), hipMemcpyHostToDevice));
  s->atlas_tiles = (uint32_t)t;
  s->atlas_bands = (uint32_t)b;
  s->atlas_rows = (uint32_t)r;
  s->atlas_ready = true;
  s->serial = g_serial.fetch_add(1);  // frames captured with the old placement do not match
  return RZ_OK;
}

// ---------------------------------------------------------------------------
// hipGraph replay of repeated frames. With RZ_PARAM_GRAPH on, a device entry
// point whose arguments (GraphKey) match a frame that already completed once
// with the context's buffers as they are now is captured (one stream capture
// of exactly the kernels an eager call enqueues: fills instead of memsets, no
// host copies) and replayed from then on with one hipGraphLaunch. Anything
// that changes a buffer or a launch parameter renews the generation, so older
// captures no longer match. Data the kernels read (resident points and
// colours, rz_scene_update_device) may change between replays.
// ---------------------------------------------------------------------------
GraphKey graph_key(rz_ctx* ctx, int kind, uint64_t obj) {
  GraphKey k;
  std::memset(&k, 0, sizeof(k));
  k.kind = kind;
  k.obj = obj;
  k.gen = ctx->gen;
  k.lp = ctx->cur;
  return k;
}

void drop_graphs(rz_ctx* ctx) {
  for (auto& e : ctx->graphs) (void)hipGraphExecDestroy(e.exec);
  ctx->graphs.clear();
}

template <class F>
rz_status run_graphed(rz_ctx* ctx, const GraphKey& k, F&& enqueue) {
  if (!ctx->use_graphs || ctx->prof) return enqueue();
  for (auto& e : ctx->graphs)
    if (same_key(e.key, k)) {
      e.last_use = ++ctx->tick;
      RZ_HIP(ctx, hipGraphLaunch(e.exec, ctx->stream));
      ++ctx->replays;
      return RZ_OK;
    }
  auto settled = std::find_if(ctx->graph_settled.begin(), ctx->graph_settled.end(),
                              [&](const GraphKey& e) { return same_key(e, k); });
  if (settled != ctx->graph_settled.end()) {
    ctx->graph_settled.erase(settled);
    hipGraph_t g = nullptr;
    RZ_HIP(ctx, hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeRelaxed));
    ctx->capturing = true;
    const rz_status st = enqueue();
    ctx->capturing = false;
    const hipError_t ee = hipStreamEndCapture(ctx->stream, &g);
    hipGraphExec_t ex = nullptr;
    hipError_t ei = hipErrorUnknown;
    if (st == RZ_OK && ee == hipSuccess && g)
      ei = hipGraphInstantiate(&ex, g, nullptr, nullptr, 0);
    if (g) (void)hipGraphDestroy(g);
    (void)hipGetLastError();
    if (ei == hipSuccess) {
      if (ctx->graphs.size() >= kMaxGraphs) {
        auto lru = std::min_element(ctx->graphs.begin(), ctx->graphs.end(),
                                    [](const GraphEntry& a, const GraphEntry& b) {
                                      return a.last_use < b.last_use;
                                    });
        (void)hipGraphExecDestroy(lru->exec);
        ctx->graphs.erase(lru);
      }
      ctx->graphs.push_back(GraphEntry{k, ex, ++ctx->tick});
      RZ_HIP(ctx, hipGraphLaunch(ex, ctx->stream));
      ++ctx->replays;
      return RZ_OK;
    }
    // the capture did not work out: run this frame eagerly
  }
  const rz_status st = enqueue();
  // a run that grew nothing is a candidate for capture once it has completed
  if (st == RZ_OK && ctx->gen == k.gen) {
    ctx->graph_pending = k;
    ctx->pending_valid = true;
  }
  return st;
}

// ---------------------------------------------------------------------------
// Scene trees: Scene::render of Fill / Group / Transform / Opacity / Clip
// nodes (src/scene.rs:186-199, 268-459). Pipeline::build_rec runs on the host
// (bboxes, views, transforms); render_rec draws every run of fills that goes
// into one layer as one batched render (each fill under its own transform),
// and an Opacity or Clip node into a layer of its own, composed into its
// parent by k_compose_over / k_compose_mask.
// ---------------------------------------------------------------------------
struct TreeNode {
  int kind;            // RZ_NODE_*
  rzh::BBox bbox;
  rzh::Tr tr;
  uint32_t path = 0, paint = 0;
  int rule = 0;
  std::vector<int> children;
  int child = -1;
  float opacity = 1.0f;
  rzh::Tr clip_tr;
  bool has_view = false;  // the view the node was built under (fills: their batch's)
  rzh::BBox view;
};

struct TreeBuild {
  const rz_scene_node* nodes;
  size_t n_nodes;
  const uint32_t* children;
  const rz_path* const* paths;
  size_t n_paths;
  size_t n_paints;
  std::vector<TreeNode> out;
  int depth = 0;
  bool bad = false;
  const std::vector<uint32_t>* stroked = nullptr;  // Stroke node -> its stroked path's index

  const rz_scene_node* node(uint32_t id) {
    if (id >= n_nodes || depth > 256) {
      bad = true;
      return nullptr;
    }
    return &nodes[id];
  }
  static rzh::BBox unite(const rzh::BBox& a, const rzh::BBox& b) { return rzh::bbox_union_opt(a, &b); }
  // range_intersect of both axes (src/geometry.rs:648-657, 677-688)
  static bool intersect(const rzh::BBox& a, const rzh::BBox& b, rzh::BBox* out) {
    if (a.mn.x > b.mx.x || b.mn.x > a.mx.x || a.mn.y > b.mx.y || b.mn.y > a.mx.y) return false;
    *out = rzh::bbox_new(rzg::mkpt(std::fmax(a.mn.x, b.mn.x), std::fmax(a.mn.y, b.mn.y)),
                         rzg::mkpt(std::fmin(a.mx.x, b.mx.x), std::fmin(a.mx.y, b.mx.y)));
    return true;
  }
  // Scene::bbox (src/scene.rs:154-183)
  bool scene_bbox(uint32_t id, const rzh::Tr& tr, rzh::BBox* res) {
    const rz_scene_node* n = node(id);
    if (!n) return false;
    ++depth;
    bool have = false;
    switch (n->kind) {
      case RZ_NODE_FILL:
        have = n->path < n_paths && rzh::path_bbox(paths[n->path], tr, res);
        break;
      case RZ_NODE_GROUP:
        for (uint32_t k = 0; k < n->count; ++k) {
          rzh::BBox b;
          if (!scene_bbox(children[n->first + k], tr, &b)) continue;
          *res = have ? unite(*res, b) : b;
          have = true;
        }
        break;
      case RZ_NODE_TRANSFORM:
        have = scene_bbox(n->child, rzh::tr_mul(tr, rzh::tr_of(n->tr)), res);
        break;
      case RZ_NODE_OPACITY:
        have = scene_bbox(n->child, tr, res);
        break;
      case RZ_NODE_STROKE:  // the unstroked path (src/scene.rs:158)
        have = n->path < n_paths && rzh::path_bbox(paths[n->path], tr, res);
        break;
      case RZ_NODE_CLIP: {
        rzh::Tr ctr = tr;
        if (n->units == RZ_UNITS_BBOX) {
          rzh::BBox cb;
          const rzh::Tr ident = {{1.0, 0.0, 0.0, 0.0, 1.0, 0.0}};
          if (!scene_bbox(n->child, ident, &cb)) break;
          ctr = rzh::tr_mul(tr, rzh::bbox_unit_transform(cb));
        }
        rzh::BBox a, b;
        have = scene_bbox(n->child, tr, &a) && n->path < n_paths &&
               rzh::path_bbox(paths[n->path], ctr, &b) && intersect(a, b, res);
        break;
      }
      default:
        bad = true;
    }
    --depth;
    return have;
  }
  // Pipeline::build_rec (src/scene.rs:268-357): the node's index in `out`, or -1
  int build(uint32_t id, const rzh::BBox* view, const rzh::Tr& tr) {
    const rz_scene_node* n = node(id);
    if (!n) return -1;
    ++depth;
    int r = -1;
    auto view_apply = [&](const rzh::BBox& b, rzh::BBox* o) {
      if (!view) {
        *o = b;
        return true;
      }
      return intersect(*view, b, o);
    };
    TreeNode t;
    t.kind = n->kind;
    t.tr = tr;
    t.has_view = view != nullptr;
    if (view) t.view = *view;
    switch (n->kind) {
      case RZ_NODE_FILL:
      case RZ_NODE_STROKE: {
        // a Stroke is the Fill of its stroked path, non-zero rule
        // (src/scene.rs:297-306)
        const bool stroke = n->kind == RZ_NODE_STROKE;
        const uint32_t path = stroke ? (stroked ? (*stroked)[id] : ~0u) : n->path;
        rzh::BBox pb;
        if (path >= n_paths || n->paint >= n_paints) {
          bad = true;
          break;
        }
        if (!rzh::path_bbox(paths[path], tr, &pb) || !view_apply(pb, &t.bbox)) break;
        t.kind = RZ_NODE_FILL;
        t.path = path;
        t.paint = n->paint;
        t.rule = stroke ? RZ_NONZERO : n->fill_rule;
        out.push_back(t);
        r = (int)out.size() - 1;
        break;
      }
      case RZ_NODE_GROUP: {
        bool have = false;
        for (uint32_t k = 0; k < n->count; ++k) {
          const int c = build(children[n->first + k], view, tr);
          if (c < 0) continue;
          t.children.push_back(c);
          t.bbox = have ? unite(out[c].bbox, t.bbox) : out[c].bbox;
          have = true;
        }
        if (!have) break;
        out.push_back(t);
        r = (int)out.size() - 1;
        break;
      }
      case RZ_NODE_TRANSFORM:
        r = build(n->child, view, rzh::tr_mul(tr, rzh::tr_of(n->tr)));
        break;
      case RZ_NODE_OPACITY: {
        const int c = build(n->child, view, tr);
        if (c < 0) break;
        t.child = c;
        t.bbox = out[c].bbox;
        t.opacity = (float)n->opacity;  // as given: Scene::opacity clamps, the tree does not
        out.push_back(t);
        r = (int)out.size() - 1;
        break;
      }
      case RZ_NODE_CLIP: {
        if (n->path >= n_paths) {
          bad = true;
          break;
        }
        t.clip_tr = tr;
        if (n->units == RZ_UNITS_BBOX) {
          rzh::BBox cb;
          const rzh::Tr ident = {{1.0, 0.0, 0.0, 0.0, 1.0, 0.0}};
          if (!scene_bbox(n->child, ident, &cb)) break;
          t.clip_tr = rzh::tr_mul(tr, rzh::bbox_unit_transform(cb));
        }
        rzh::BBox pb, clip_bbox;
        if (!rzh::path_bbox(paths[n->path], t.clip_tr, &pb) || !view_apply(pb, &clip_bbox)) break;
        const int c = build(n->child, &clip_bbox, tr);
        if (c < 0 || !intersect(clip_bbox, out[c].bbox, &t.bbox)) break;
        t.child = c;
        t.path = n->path;
        t.rule = n->fill_rule;
        out.push_back(t);
        r = (int)out.size() - 1;
        break;
      }
      default:
        bad = true;
    }
    --depth;
    return r;
  }
};

// A layer on the device (Layer::new: the pixel box around a bbox).
struct DLayer {
  int x = 0, y = 0, w = 0, h = 0;
  rzp::C4* d = nullptr;
};

struct TreeRender {
  rz_ctx* ctx;
  TreeBuild* tb;
  const rz_paint* paints;
  int mode;
  double flatness;
  std::vector<void*> allocs;            // layers and masks, freed at the end
  std::vector<rz_dscene*> scenes;       // batches
  std::vector<rz_dpath*> dpaths;        // clip paths
  std::vector<int> batch;               // fills waiting for their layer

  ~TreeRender() {
    (void)hipStreamSynchronize(ctx->stream);
    for (void* p : allocs) (void)hipFree(p);
    for (rz_dscene* s : scenes) {
      scene_release(s);
      delete s;
    }
    for (rz_dpath* d : dpaths) {
      dpath_release(d);
      delete d;
    }
  }
  rz_status alloc(size_t bytes, void** p) {
    hipError_t e = hipMalloc(p, std::max<size_t>(bytes, 16));
    if (e != hipSuccess) return fail(ctx, RZ_E_NOMEM, std::string("hipMalloc: ") + hipGetErrorString(e));
    allocs.push_back(*p);
    return RZ_OK;
  }
  rz_status new_layer(const rzh::BBox& b, rzp::C4 fill, DLayer* L) {
    const double v[4] = {b.mn.x, b.mn.y, b.mx.x, b.mx.y};
    layer_of(v, &L->x, &L->y, &L->w, &L->h);
    void* p = nullptr;
    RZ_TRY(alloc((size_t)L->w * L->h * sizeof(rzp::C4), &p));
    L->d = (rzp::C4*)p;
    if (L->w && L->h)
      RZ_LAUNCH(ctx, "k_fill_rgba", rzk::k_fill_rgba, dim3(1024), dim3(256), L->d, L->w, L->h,
                (size_t)L->w, fill);
    return check_launch(ctx, "k_fill_rgba");
  }
  // the fills collected for `L`, as one batched render over its pixels
  rz_status flush(const DLayer& L) {
    if (batch.empty()) return RZ_OK;
    std::vector<int> b;
    b.swap(batch);
    if (!L.w || !L.h) return RZ_OK;
    const size_t n = b.size();
    std::vector<const rz_path*> ps(n);
    std::vector<rz_paint> pt(n);
    std::vector<int32_t> rules(n);
    std::vector<double> trs(6 * n);
    for (size_t i = 0; i < n; ++i) {
      const TreeNode& f = tb->out[b[i]];
      ps[i] = tb->paths[f.path];
      pt[i] = paints[f.paint];
      rules[i] = f.rule;
      for (int k = 0; k < 6; ++k) trs[6 * i + k] = f.tr.m[k];
    }
    rz_dscene* s = nullptr;
    RZ_TRY(scene_upload_impl(ctx, ps.data(), pt.data(), rules.data(), n, &s, trs.data()));
    scenes.push_back(s);
    const TreeNode& f0 = tb->out[b[0]];
    const double inf = INFINITY;
    double view[4] = {-inf, -inf, inf, inf};
    if (f0.has_view) {
      view[0] = f0.view.mn.x;
      view[1] = f0.view.mn.y;
      view[2] = f0.view.mx.x;
      view[3] = f0.view.mx.y;
    }
    const int layer[4] = {L.x, L.y, L.w, L.h};
    rz_transform ident = {{1.0, 0.0, 0.0, 0.0, 1.0, 0.0}};
    return enqueue_scene_batched(ctx, s, ident, view, rzp::splat(0.0f), mode, flatness, L.d,
                                 (size_t)L.w, layer, 1);
  }
  // Pipeline::render_rec (src/scene.rs:397-459)
  rz_status render(int id, const DLayer& L) {
    const TreeNode& n = tb->out[id];
    switch (n.kind) {
      case RZ_NODE_FILL: {
        // fills of one batch were built under one view (a new view only
        // comes with a Clip, whose child draws into its own layer)
        if (!batch.empty()) {
          const TreeNode& f0 = tb->out[batch[0]];
          const bool same = f0.has_view == n.has_view &&
                            (!n.has_view || std::memcmp(&f0.view, &n.view, sizeof(n.view)) == 0);
          if (!same) RZ_TRY(flush(L));
        }
        batch.push_back(id);
        return RZ_OK;
      }
      case RZ_NODE_GROUP:
        for (int c : n.children) RZ_TRY(render(c, L));
        return RZ_OK;
      case RZ_NODE_OPACITY: {
        RZ_TRY(flush(L));
        DLayer C;
        RZ_TRY(new_layer(tb->out[n.child].bbox, rzp::splat(0.0f), &C));
        RZ_TRY(render(n.child, C));
        RZ_TRY(flush(C));
        return compose_over(L, C, n.opacity);
      }
      case RZ_NODE_CLIP: {
        RZ_TRY(flush(L));
        DLayer C;
        RZ_TRY(new_layer(tb->out[n.child].bbox, rzp::splat(0.0f), &C));
        RZ_TRY(render(n.child, C));
        RZ_TRY(flush(C));
        // the clip's mask over the node's box: Rasterizer::mask of the clip
        // path under translate(-mask.x, -mask.y) * clip_tr
        int mx, my, mw, mh;
        const double v[4] = {n.bbox.mn.x, n.bbox.mn.y, n.bbox.mx.x, n.bbox.mx.y};
        layer_of(v, &mx, &my, &mw, &mh);
        if (mw && mh) {
          void* m = nullptr;
          RZ_TRY(alloc((size_t)mw * mh * sizeof(float), &m));
          rz_dpath* dp = nullptr;
          RZ_TRY(path_upload_impl(ctx, tb->paths[n.path], &dp));
          dpaths.push_back(dp);
          const rzh::Tr align = {{1.0, 0.0, -(double)mx, 0.0, 1.0, -(double)my}};
          const rzh::Tr mt = rzh::tr_mul(align, n.clip_tr);
          rz_transform t;
          for (int k = 0; k < 6; ++k) t.m[k] = mt.m[k];
          const int mmode = mode == RZ_EDGE_NONE ? RZ_EDGE_NONE : RZ_EDGE_MASK;
          RZ_TRY(enqueue_mask(ctx, dp, t, flatness, n.rule, mmode, (size_t)mw, (size_t)mh, (float*)m,
                              (size_t)mw));
          // child * mask over their overlap
          rzk::ComposeRect r;
          if (overlap(C.x, C.y, C.w, C.h, mx, my, mw, mh, &r)) {
            const uint32_t grid = (uint32_t)std::min<size_t>(((size_t)r.w * r.h + 255) / 256, 4096);
            RZ_LAUNCH(ctx, "k_compose_mask", rzk::k_compose_mask, dim3(grid), dim3(256), C.d,
                      (size_t)C.w, (const float*)m, (size_t)mw, r);
          }
        }
        return compose_over(L, C, 1.0f);
      }
      default:
        return fail(ctx, RZ_E_ARG, "bad scene node");
    }
  }
  // the overlap of layer a (dst) and layer b (src), as a ComposeRect
  static bool overlap(int ax, int ay, int aw, int ah, int bx, int by, int bw, int bh,
                      rzk::ComposeRect* r) {
    const int x0 = std::max(ax, bx), x1 = std::min(ax + aw, bx + bw);
    const int y0 = std::max(ay, by), y1 = std::min(ay + ah, by + bh);
    if (x1 <= x0 || y1 <= y0) return false;
    *r = rzk::ComposeRect{x1 - x0, y1 - y0, x0 - ax, y0 - ay, x0 - bx, y0 - by};
    return true;
  }
  rz_status compose_over(const DLayer& dst, const DLayer& src, float opacity) {
    rzk::ComposeRect r;
    if (!overlap(dst.x, dst.y, dst.w, dst.h, src.x, src.y, src.w, src.h, &r)) return RZ_OK;
    const uint32_t grid = (uint32_t)std::min<size_t>(((size_t)r.w * r.h + 255) / 256, 4096);
    RZ_LAUNCH(ctx, "k_compose_over", rzk::k_compose_over, dim3(grid), dim3(256), dst.d,
              (size_t)dst.w, (const rzp::C4*)src.d, (size_t)src.w, r, opacity);
    return check_launch(ctx, "k_compose_over");
  }
};

rz_status check_scene_args(rz_ctx* ctx, double flatness, int mode, const double* view) {
  if (!ctx) return RZ_E_ARG;
  if (!(flatness > 0.0) || !std::isfinite(flatness)) return fail(ctx, RZ_E_ARG, "flatness must be > 0");
  if (mode < RZ_EDGE_MASK || mode > RZ_EDGE_NONE) return fail(ctx, RZ_E_ARG, "bad edge mode");
  if (!view) return fail(ctx, RZ_E_ARG, "view is NULL");
  for (int k = 0; k < 4; ++k)
    if (!std::isfinite(view[k])) return fail(ctx, RZ_E_ARG, "view must be finite");
  if (view[2] - view[0] > (double)(1 << 24) || view[3] - view[1] > (double)(1 << 24))
    return fail(ctx, RZ_E_ARG, "view too large");
  return RZ_OK;
}

// ---------------------------------------------------------------------------
// Tuning: a coordinate search over the launch choices of one workload. Each
// candidate is timed with HIP events over repeated frames on the context's
// stream (kernel by kernel, no graph replay); the fastest becomes the
// workload's entry, which select_params applies to later calls. Every choice
// gives the same result bit for bit, so the search only ever trades speed.
// ---------------------------------------------------------------------------
template <typename F>
rz_status tune_search(rz_ctx* ctx, const TuneKey& key, int auto_dq, int auto_dc, F&& frame,
                      double budget_ms, rz_tune_result* res) {
  if (ctx->capturing) return fail(ctx, RZ_E_ARG, "tuning inside a stream capture");
  hipEvent_t ea = nullptr, eb = nullptr;
  RZ_HIP(ctx, hipEventCreate(&ea));
  if (hipEventCreate(&eb) != hipSuccess) {
    (void)hipEventDestroy(ea);
    return fail(ctx, RZ_E_HIP, "hipEventCreate failed");
  }
  const bool graphs = ctx->use_graphs;
  ctx->use_graphs = false;
  ctx->tune_forced = true;
  struct Trial {
    LaunchParams p;
    double ms;
  };
  std::vector<Trial> seen;
  const auto t0 = std::chrono::steady_clock::now();
  auto elapsed_ms = [&]() {
    return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
  };
  auto timed = [&](int reps, double* ms) -> rz_status {
    RZ_HIP(ctx, hipEventRecord(ea, ctx->stream));
    for (int r = 0; r < reps; ++r) RZ_TRY(frame());
    RZ_HIP(ctx, hipEventRecord(eb, ctx->stream));
    RZ_TRY(finish(ctx));
    float t = 0.0f;
    RZ_HIP(ctx, hipEventElapsedTime(&t, ea, eb));
    *ms = (double)t / reps;
    return RZ_OK;
  };
  auto measure = [&](const LaunchParams& p, double* ms) -> rz_status {
    for (const Trial& t : seen)
      if (same_params(t.p, p)) {
        *ms = t.ms;
        return RZ_OK;
      }
    ctx->cur = p;
    // a frame that grew a buffer is run again (as a caller would)
    rz_status st = RZ_E_CAPACITY;
    for (int a = 0; a < 4 && st == RZ_E_CAPACITY; ++a) {
      st = frame();
      if (st == RZ_OK) st = finish(ctx);
    }
    RZ_TRY(st);
    double one = 0.0;
    RZ_TRY(timed(1, &one));
    // about 3 ms of frames, 3 to 50 of them
    const int reps = (int)std::min(50.0, std::max(3.0, 3.0 / std::max(one, 0.01)));
    RZ_TRY(timed(reps, ms));
    seen.push_back(Trial{p, *ms});
    return RZ_OK;
  };
  auto explicit_depths = [&](LaunchParams p) {
    if (p.split_auto) {
      p.split_auto = 0;
      p.dq = auto_dq;
      p.dc = auto_dc;
    }
    return p;
  };
  const LaunchParams base = ctx->lp;
  LaunchParams best = base;
  double ms_default = 0.0, best_ms = 0.0;
  rz_status st = measure(base, &ms_default);
  best_ms = ms_default;
  auto consider = [&](const LaunchParams& p) -> rz_status {
    if (elapsed_ms() > budget_ms) return RZ_OK;
    double ms = 0.0;
    RZ_TRY(measure(p, &ms));
    if (ms < best_ms) {
      best_ms = ms;
      best = p;
    }
    return RZ_OK;
  };
  auto sweep_depths = [&]() -> rz_status {
    for (int v = 0; v <= 6; ++v) {
      LaunchParams p = explicit_depths(best);
      p.dq = v;
      RZ_TRY(consider(p));
    }
    for (int v = 0; v <= 6; ++v) {
      LaunchParams p = explicit_depths(best);
      p.dc = v;
      RZ_TRY(consider(p));
    }
    return RZ_OK;
  };
  if (st == RZ_OK) st = sweep_depths();
  for (int32_t v : {-1, 0, 4096, 8192, 16384}) {
    if (st != RZ_OK) break;
    LaunchParams p = best;
    p.flat_lds = v;
    st = consider(p);
  }
  for (int32_t v : {0, 4096, 8192, 16384}) {
    if (st != RZ_OK) break;
    LaunchParams p = best;
    p.bin_lds = v;
    st = consider(p);
  }
  for (int32_t v : {0, 8192, 16384, 24576}) {
    if (st != RZ_OK) break;
    LaunchParams p = best;
    p.fine_lds = v;
    st = consider(p);
  }
  for (int32_t v = 12; v <= 17 && st == RZ_OK; ++v) {
    LaunchParams p = best;
    p.scatter_log2 = v;
    st = consider(p);
  }
  if (key.kind == 1) {
    for (int32_t v : {6, 8, 1}) {
      if (st != RZ_OK) break;
      LaunchParams p = best;
      p.scene_waves = v;
      st = consider(p);
    }
    for (int32_t v : {0, 1}) {
      if (st != RZ_OK) break;
      LaunchParams p = best;
      p.scene_queue = v;
      st = consider(p);
    }
  }
  if (st == RZ_OK) st = sweep_depths();
  ctx->tune_forced = false;
  ctx->use_graphs = graphs;
  ++ctx->gen;
  (void)hipEventDestroy(ea);
  (void)hipEventDestroy(eb);
  RZ_TRY(st);
  bool found = false;
  for (TuneEntry& t : ctx->tuned)
    if (same_tune_key(t.key, key)) {
      t = TuneEntry{key, best, ms_default, best_ms};
      found = true;
    }
  if (!found) ctx->tuned.push_back(TuneEntry{key, best, ms_default, best_ms});
  if (res) {
    res->split_auto = best.split_auto;
    res->split_quad = best.split_auto ? auto_dq : best.dq;
    res->split_cubic = best.split_auto ? auto_dc : best.dc;
    res->flat_lds = best.flat_lds;
    res->bin_lds = best.bin_lds;
    res->fine_lds = best.fine_lds;
    res->scatter_log2 = best.scatter_log2;
    res->scene_waves = best.scene_waves;
    res->scene_queue = best.scene_queue;
    res->trials = (uint32_t)seen.size();
    res->ms_default = ms_default;
    res->ms_tuned = best_ms;
  }
  return RZ_OK;
}

// The tuning cache file: lines "<gpu> <kind> <mode> <px> <slots> <quad>
// <cubic> <split_auto> <dq> <dc> <flat_lds> <bin_lds> <fine_lds> <scatter>
// <scene_waves> <scene_queue> <ms_default> <ms_tuned>"; '#' starts a comment. Entries of other GPUs are
// kept by a save and skipped by a load.
std::string tune_line(const std::string& gpu, const TuneEntry& t) {
  std::ostringstream o;
  o << gpu << ' ' << t.key.kind << ' ' << t.key.mode << ' ' << t.key.px_log2 << ' '
    << t.key.slots_log2 << ' ' << t.key.quad_log2 << ' ' << t.key.cubic_log2 << ' '
    << t.p.split_auto << ' ' << t.p.dq << ' ' << t.p.dc << ' ' << t.p.flat_lds << ' '
    << t.p.bin_lds << ' ' << t.p.fine_lds << ' ' << t.p.scatter_log2 << ' ' << t.p.scene_waves << ' '
    << t.p.scene_queue << ' ' << t.ms_default << ' ' << t.ms_tuned;
  return o.str();
}

bool parse_tune_line(const std::string& line, std::string* gpu, TuneEntry* t) {
  std::istringstream in(line);
  std::memset(&t->key, 0, sizeof(t->key));
  LaunchParams p;
  if (!(in >> *gpu >> t->key.kind >> t->key.mode >> t->key.px_log2 >> t->key.slots_log2 >>
        t->key.quad_log2 >> t->key.cubic_log2 >> p.split_auto >> p.dq >> p.dc >> p.flat_lds >>
        p.bin_lds >> p.fine_lds >> p.scatter_log2 >> p.scene_waves >> p.scene_queue >>
        t->ms_default >> t->ms_tuned))
    return false;
  // only choices that rz_set_param could make
  if (p.split_auto < 0 || p.split_auto > 1 || p.dq < 0 || p.dq > 8 || p.dc < 0 || p.dc > 8 ||
      p.flat_lds < -1 || p.flat_lds > 32768 || p.bin_lds < 0 || p.bin_lds > 32768 ||
      p.fine_lds < 0 || p.fine_lds > 32768 || p.scatter_log2 < 8 || p.scatter_log2 > 18 ||
      (p.scene_waves != 1 && p.scene_waves != 6 && p.scene_waves != 8) || p.scene_queue < 0 ||
      p.scene_queue > 1)
    return false;
  t->p = p;
  return true;
}

}  // namespace

extern "C" {

rz_status rz_set_param(rz_ctx* ctx, int param, int value) {
  if (!ctx) return RZ_E_ARG;
  if (param == RZ_PARAM_GRAPH_REPLAYS || param == RZ_PARAM_GRAPH_GEN)
    return fail(ctx, RZ_E_ARG, "read-only parameter");
  ++ctx->gen;  // frames captured with the old value no longer match
  if (param == RZ_PARAM_GRAPH) {
    if (value != 0 && value != 1) return fail(ctx, RZ_E_ARG, "parameter out of range 0..1");
    ctx->use_graphs = value != 0;
    if (!ctx->use_graphs) drop_graphs(ctx);
    return RZ_OK;
  }
  if (param == RZ_PARAM_SCAN_DIRECT) {
    if (value < 0 || value > 16) return fail(ctx, RZ_E_ARG, "parameter out of range 0..16");
    ctx->scan_direct = 1u << value;
    return RZ_OK;
  }
  auto range = [&](int lo, int hi) {
    if (value >= lo && value <= hi) return RZ_OK;
    return fail(ctx, RZ_E_ARG, "parameter out of range " + std::to_string(lo) + ".." + std::to_string(hi));
  };
  LaunchParams& lp = ctx->lp;
  if (param == RZ_PARAM_SPLIT_QUAD || param == RZ_PARAM_SPLIT_CUBIC) {
    RZ_TRY(range(0, 8));
    (param == RZ_PARAM_SPLIT_QUAD ? lp.dq : lp.dc) = value;
    lp.split_auto = 0;
  } else if (param == RZ_PARAM_SPLIT_AUTO) {
    RZ_TRY(range(0, 1));
    lp.split_auto = value;
  } else if (param == RZ_PARAM_FLAT_LDS) {
    RZ_TRY(range(-1, 32));
    lp.flat_lds = value < 0 ? -1 : value * 1024;
  } else if (param == RZ_PARAM_BIN_LDS) {
    RZ_TRY(range(0, 32));
    lp.bin_lds = value * 1024;
  } else if (param == RZ_PARAM_FINE_LDS) {
    RZ_TRY(range(0, 32));
    lp.fine_lds = value * 1024;
  } else if (param == RZ_PARAM_SCATTER_WAVES) {
    RZ_TRY(range(8, 18));
    lp.scatter_log2 = value;
  } else if (param == RZ_PARAM_SCENE_WAVES) {
    if (value != 1 && value != 6 && value != 8)
      return fail(ctx, RZ_E_ARG, "parameter must be 1, 6 or 8");
    lp.scene_waves = value;
  } else if (param == RZ_PARAM_SCENE_QUEUE) {
    RZ_TRY(range(0, 1));
    lp.scene_queue = value;
  } else if (param == RZ_PARAM_FUSED) {
    RZ_TRY(range(0, 1));
    lp.fused = value;
  } else if (param == RZ_PARAM_TUNED) {
    RZ_TRY(range(0, 1));
    ctx->use_tuned = value != 0;
  } else {
    return fail(ctx, RZ_E_ARG, "unknown parameter");
  }
  return RZ_OK;
}

rz_status rz_get_param(rz_ctx* ctx, int param, int* value) {
  if (!ctx || !value) return RZ_E_ARG;
  const LaunchParams& lp = ctx->lp;
  if (param == RZ_PARAM_SPLIT_QUAD) *value = lp.dq;
  else if (param == RZ_PARAM_SPLIT_CUBIC) *value = lp.dc;
  else if (param == RZ_PARAM_SPLIT_AUTO) *value = lp.split_auto;
  else if (param == RZ_PARAM_FLAT_LDS) *value = lp.flat_lds < 0 ? -1 : lp.flat_lds / 1024;
  else if (param == RZ_PARAM_BIN_LDS) *value = lp.bin_lds / 1024;
  else if (param == RZ_PARAM_FINE_LDS) *value = lp.fine_lds / 1024;
  else if (param == RZ_PARAM_SCATTER_WAVES) *value = lp.scatter_log2;
  else if (param == RZ_PARAM_SCENE_WAVES) *value = lp.scene_waves;
  else if (param == RZ_PARAM_SCENE_QUEUE) *value = lp.scene_queue;
  else if (param == RZ_PARAM_FUSED) *value = lp.fused;
  else if (param == RZ_PARAM_TUNED) *value = ctx->use_tuned ? 1 : 0;
  else if (param == RZ_PARAM_SCAN_DIRECT) *value = 31 - __builtin_clz(ctx->scan_direct);
  else if (param == RZ_PARAM_GRAPH) *value = ctx->use_graphs ? 1 : 0;
  else if (param == RZ_PARAM_GRAPH_REPLAYS) *value = (int)std::min<uint64_t>(ctx->replays, INT32_MAX);
  else if (param == RZ_PARAM_GRAPH_GEN) *value = (int)(ctx->gen & 0x7fffffff);
  else return fail(ctx, RZ_E_ARG, "unknown parameter");
  return RZ_OK;
}

rz_status rz_path_bbox(const rz_path* path, rz_transform tr, int* has_bbox, double* bbox) {
  if (!path || !has_bbox || !bbox) return RZ_E_ARG;
  if (path->n_segments && (!path->kinds || !path->point_offsets || !path->points)) return RZ_E_ARG;
  rzh::BBox b;
  *has_bbox = rzh::path_bbox(path, rzh::tr_of(tr.m), &b) ? 1 : 0;
  if (*has_bbox) {
    bbox[0] = b.mn.x;
    bbox[1] = b.mn.y;
    bbox[2] = b.mx.x;
    bbox[3] = b.mx.y;
  }
  return RZ_OK;
}

rz_status rz_tune_mask_device(rz_ctx* ctx, const rz_dpath* d, rz_transform tr, double flatness,
                              int fill_rule, int edge_mode, size_t width, size_t height,
                              float* d_out, size_t row_stride, double budget_ms,
                              rz_tune_result* res) {
  RZ_TRY(check_args(ctx, flatness, fill_rule, edge_mode, width, height));
  if (!d || !d_out || !width || !height) return fail(ctx, RZ_E_ARG, "null argument or empty mask");
  if (row_stride < width) return fail(ctx, RZ_E_ARG, "row_stride < width");
  int adq, adc;
  auto_depths(d->est, det_of(tr), flatness, &adq, &adc);
  return tune_search(ctx, mask_tune_key(d, tr, flatness, edge_mode, width, height), adq, adc,
                     [&]() {
                       return enqueue_mask(ctx, d, tr, flatness, fill_rule, edge_mode, width,
                                           height, d_out, row_stride);
                     },
                     budget_ms, res);
}

rz_status rz_tune_render_device(rz_ctx* ctx, const rz_dscene* scene, rz_transform tr,
                                double flatness, int edge_mode, const double* view,
                                const float* bg, float* d_rgba, size_t row_stride,
                                double budget_ms, rz_tune_result* res) {
  RZ_TRY(check_scene_args(ctx, flatness, edge_mode, view));
  if (!scene || !d_rgba) return fail(ctx, RZ_E_ARG, "null argument");
  int lx, ly, lw, lh;
  layer_of(view, &lx, &ly, &lw, &lh);
  if (row_stride < (size_t)lw) return fail(ctx, RZ_E_ARG, "row_stride < layer width");
  const rzp::C4 c = bg ? rzp::c4(bg[0], bg[1], bg[2], bg[3]) : rzp::splat(0.0f);
  int adq = 0, adc = 0;
  if (scene->all) auto_depths(scene->all->est, det_of(tr), flatness, &adq, &adc);
  return tune_search(ctx, scene_tune_key(scene, tr, flatness, edge_mode, view), adq, adc,
                     [&]() {
                       return enqueue_scene_batched(ctx, scene, tr, view, c, edge_mode, flatness,
                                                    (rzp::C4*)d_rgba, row_stride);
                     },
                     budget_ms, res);
}

rz_status rz_tune_atlas_device(rz_ctx* ctx, const rz_dscene* glyphs, double flatness,
                               int fill_rule, int edge_mode, float* d_atlas, size_t row_stride,
                               double budget_ms, rz_tune_result* res) {
  RZ_TRY(check_args(ctx, flatness, fill_rule, edge_mode, 0, 0));
  if (!glyphs || !glyphs->atlas_ready) return fail(ctx, RZ_E_ARG, "glyph placement not set");
  if (!d_atlas) return fail(ctx, RZ_E_ARG, "d_atlas is NULL");
  int adq, adc;
  auto_depths(glyphs->atlas_est, 1.0, flatness, &adq, &adc);
  return tune_search(ctx, atlas_tune_key(glyphs, flatness, edge_mode), adq, adc,
                     [&]() {
                       return enqueue_atlas(ctx, glyphs, fill_rule, edge_mode, flatness, d_atlas,
                                            row_stride);
                     },
                     budget_ms, res);
}

rz_status rz_tune_save(rz_ctx* ctx, const char* path) {
  if (!ctx || !path) return RZ_E_ARG;
  std::vector<std::string> keep;
  {
    std::ifstream in(path);
    std::string line, gpu;
    TuneEntry t;
    while (std::getline(in, line))
      if (!line.empty() && line[0] != '#' && parse_tune_line(line, &gpu, &t) && gpu != ctx->gpu_id)
        keep.push_back(line);
  }
  const std::string tmp = std::string(path) + ".tmp";
  {
    std::ofstream out(tmp, std::ios::trunc);
    if (!out) return fail(ctx, RZ_E_ARG, std::string("cannot write ") + tmp);
    out << "# rasterize_amd launch tuning: gpu kind mode px slots quad cubic split_auto dq dc "
           "flat_lds bin_lds fine_lds scatter scene_waves scene_queue ms_default ms_tuned\n";
    for (const std::string& l : keep) out << l << '\n';
    out.precision(6);
    for (const TuneEntry& t : ctx->tuned) out << tune_line(ctx->gpu_id, t) << '\n';
    if (!out) return fail(ctx, RZ_E_ARG, std::string("cannot write ") + tmp);
  }
  if (std::rename(tmp.c_str(), path) != 0)
    return fail(ctx, RZ_E_ARG, std::string("cannot replace ") + path);
  return RZ_OK;
}

rz_status rz_tune_load(rz_ctx* ctx, const char* path, size_t* n_loaded) {
  if (!ctx || !path) return RZ_E_ARG;
  std::ifstream in(path);
  if (!in) return fail(ctx, RZ_E_ARG, std::string("cannot read ") + path);
  std::string line, gpu;
  size_t n = 0;
  while (std::getline(in, line)) {
    TuneEntry t;
    if (line.empty() || line[0] == '#' || !parse_tune_line(line, &gpu, &t) || gpu != ctx->gpu_id)
      continue;
    bool found = false;
    for (TuneEntry& e : ctx->tuned)
      if (same_tune_key(e.key, t.key)) {
        e = t;
        found = true;
      }
    if (!found) ctx->tuned.push_back(t);
    ++n;
  }
  ++ctx->gen;
  if (n_loaded) *n_loaded = n;
  return RZ_OK;
}

rz_status rz_tune_clear(rz_ctx* ctx) {
  if (!ctx) return RZ_E_ARG;
  ctx->tuned.clear();
  ++ctx->gen;
  return RZ_OK;
}

rz_status rz_tune_entries(rz_ctx* ctx, size_t* n) {
  if (!ctx || !n) return RZ_E_ARG;
  *n = ctx->tuned.size();
  return RZ_OK;
}

rz_status rz_last_counts(rz_ctx* ctx, size_t* n_lines, size_t* n_records, size_t* n_events) {
  if (!ctx) return RZ_E_ARG;
  if (n_lines) *n_lines = ctx->last.n_lines;
  if (n_records) *n_records = ctx->last.n_records;
  if (n_events) *n_events = ctx->last.n_events;
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
  if (e == hipSuccess)
    e = hipHostMalloc((void**)&ctx->host_status, 2 * sizeof(rzk::Status), hipHostMallocDefault);
  if (e == hipSuccess) e = hipMalloc((void**)&ctx->acc, 2 * sizeof(rzk::Status));
  if (e == hipSuccess) e = hipMemset(ctx->acc, 0, 2 * sizeof(rzk::Status));
  if (e == hipSuccess) e = hipEventCreateWithFlags(&ctx->xev, hipEventDisableTiming);
  if (e == hipSuccess) ctx->scratch_ptr = noalloc_scratch(device);
  if (e == hipSuccess) {
    hipDeviceProp_t prop;
    e = hipGetDeviceProperties(&prop, device);
    if (e == hipSuccess) {
      std::string arch = prop.gcnArchName;
      arch = arch.substr(0, arch.find(':'));
      ctx->gpu_id = arch + "/" + std::to_string(prop.multiProcessorCount) + "cu";
      ctx->n_cus = std::max(1, prop.multiProcessorCount);
    }
  }
  if (e != hipSuccess) {
    if (ctx->acc) (void)hipFree(ctx->acc);
    if (ctx->host_status) (void)hipHostFree(ctx->host_status);
    if (ctx->stream) (void)hipStreamDestroy(ctx->stream);
    delete ctx;
    return RZ_E_HIP;
  }
  std::memset(ctx->host_status, 0, 2 * sizeof(rzk::Status));
  ctx->acc_run = ctx->acc;
  *out = ctx;
  return RZ_OK;
}

rz_status rz_ctx_destroy(rz_ctx* ctx) {
  if (!ctx) return RZ_OK;
  (void)hipSetDevice(ctx->device);
  if (ctx->stream) (void)hipStreamSynchronize(ctx->stream);
  DBuf* bufs[] = {&ctx->meta,      &ctx->tile_off,  &ctx->ev_off,    &ctx->sums,
                  &ctx->recs,      &ctx->events,    &ctx->out,       &ctx->lines,
                  &ctx->counts,    &ctx->stops,     &ctx->line_path,
                  &ctx->bw_w,      &ctx->bw_p,      &ctx->bw_tr,     &ctx->bw_stops,
                  &ctx->sc_order,
                  &ctx->sc_pi,     &ctx->sc_counts,
                  &ctx->sc_rects,  &ctx->sc_rows,   &ctx->sc_rowcnt, &ctx->list_buf,
                  &ctx->wave_lines, &ctx->atlas_tiles};
  drop_graphs(ctx);
  for (DBuf* b : bufs) free_buf(*b);
  for (auto& r : ctx->pending) {
    (void)hipEventDestroy(r.a);
    (void)hipEventDestroy(r.b);
  }
  for (hipEvent_t e : ctx->pool) (void)hipEventDestroy(e);
  if (ctx->host_status) (void)hipHostFree(ctx->host_status);
  if (ctx->acc) (void)hipFree(ctx->acc);
  if (ctx->xev) (void)hipEventDestroy(ctx->xev);
  if (ctx->stream) (void)hipStreamDestroy(ctx->stream);
  delete ctx;
  return RZ_OK;
}

const char* rz_last_error(const rz_ctx* ctx) { return ctx ? ctx->err.c_str() : "no context"; }

void* rz_stream(rz_ctx* ctx) { return ctx ? (void*)ctx->stream : nullptr; }

rz_status rz_wait_stream(rz_ctx* ctx, void* stream) {
  if (!ctx) return RZ_E_ARG;
  RZ_HIP(ctx, hipEventRecord(ctx->xev, (hipStream_t)stream));
  RZ_HIP(ctx, hipStreamWaitEvent(ctx->stream, ctx->xev, 0));
  return RZ_OK;
}

rz_status rz_stream_wait_ctx(rz_ctx* ctx, void* stream) {
  if (!ctx) return RZ_E_ARG;
  RZ_HIP(ctx, hipEventRecord(ctx->xev, ctx->stream));
  RZ_HIP(ctx, hipStreamWaitEvent((hipStream_t)stream, ctx->xev, 0));
  return RZ_OK;
}

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

}  // extern "C"

// Owned path buffers (rz_svg.cpp)
rz_path_buf* rz_path_buf_adopt(rz::Path&& p);

namespace {
// Consecutive sub-buffers of one device buffer, 256-byte aligned.
struct Carve {
  char* base;
  size_t at = 0;
  template <typename T>
  T* take(size_t n) {
    T* p = (T*)(base + at);
    at += (std::max<size_t>(n, 1) * sizeof(T) + 255) & ~(size_t)255;
    return p;
  }
};
}  // namespace

extern "C" {

rz_status rz_path_stroke(rz_ctx* ctx, const rz_path* path, rz_stroke_style style,
                         rz_path_buf** out) {
  if (!ctx || !out) return RZ_E_ARG;
  *out = nullptr;
  RZ_TRY(validate_path(ctx, path));
  if (style.line_join < RZ_JOIN_MITER || style.line_join > RZ_JOIN_ROUND ||
      style.line_cap < RZ_CAP_BUTT || style.line_cap > RZ_CAP_ROUND)
    return fail(ctx, RZ_E_ARG, "unknown line join or line cap");
  const uint32_t ns = path->n_segments, nsub = path->n_subpaths;
  if ((uint64_t)ns * 2 >= 0x7fffffffu) return fail(ctx, RZ_E_ARG, "too many segments to stroke");
  rz::Path res;
  if (ns == 0) {
    *out = rz_path_buf_adopt(std::move(res));
    return RZ_OK;
  }
  const uint32_t npts = path->point_offsets[ns];
  const uint32_t n_items = 2 * ns, n_runs = 2 * nsub;
  const uint32_t nblk = (n_items + rzk::kStrokeBlock - 1) / rzk::kStrokeBlock;
  std::vector<uint32_t> sub_of(ns);
  for (uint32_t i = 0; i < nsub; ++i)
    for (uint32_t k = path->subpaths[i]; k < path->subpaths[i + 1]; ++k) sub_of[k] = i;
  // device layout: the input path, then the work arrays
  size_t need = 0;
  {
    Carve c{nullptr};
    c.take<uint8_t>(ns);
    c.take<uint32_t>(ns + 1);
    c.take<double>(2 * (size_t)npts);
    c.take<uint32_t>(nsub + 1);
    c.take<uint8_t>(nsub);
    c.take<uint32_t>(ns);
    c.take<rzs::Seg>(n_items);
    c.take<rzs::Seg>(n_items);
    for (int k = 0; k < 3; ++k) c.take<uint32_t>(n_items);  // pieces, piece_pts, incl
    c.take<int32_t>(nblk);
    c.take<int32_t>(n_runs);
    c.take<uint32_t>(n_runs);
    for (int k = 0; k < 4; ++k) c.take<uint32_t>(n_items + 1);  // counts and their sums
    c.take<uint32_t>(n_runs + 1);
    c.take<uint32_t>(4);
    need = c.at;
  }
  RZ_TRY(ensure(ctx, ctx->sk_work, need));
  Carve c{(char*)ctx->sk_work.p};
  rzs::Input in;
  uint8_t* d_kinds = c.take<uint8_t>(ns);
  uint32_t* d_off = c.take<uint32_t>(ns + 1);
  double* d_pts = c.take<double>(2 * (size_t)npts);
  uint32_t* d_subs = c.take<uint32_t>(nsub + 1);
  uint8_t* d_closed = c.take<uint8_t>(nsub);
  uint32_t* d_sub_of = c.take<uint32_t>(ns);
  hipStream_t s = ctx->stream;
  RZ_HIP(ctx, hipMemcpyAsync(d_kinds, path->kinds, ns, hipMemcpyHostToDevice, s));
  RZ_HIP(ctx, hipMemcpyAsync(d_off, path->point_offsets, (ns + 1) * 4, hipMemcpyHostToDevice, s));
  RZ_HIP(ctx, hipMemcpyAsync(d_pts, path->points, 16 * (size_t)npts, hipMemcpyHostToDevice, s));
  RZ_HIP(ctx, hipMemcpyAsync(d_subs, path->subpaths, (nsub + 1) * 4, hipMemcpyHostToDevice, s));
  RZ_HIP(ctx, hipMemcpyAsync(d_closed, path->closed, nsub, hipMemcpyHostToDevice, s));
  RZ_HIP(ctx, hipMemcpyAsync(d_sub_of, sub_of.data(), ns * 4, hipMemcpyHostToDevice, s));
  in.kinds = d_kinds;
  in.pt_off = d_off;
  in.pts = d_pts;
  in.subs = d_subs;
  in.closed = d_closed;
  in.sub_of = d_sub_of;
  in.n_items = n_items;
  in.st.width = style.width;
  in.st.miter_limit = style.miter_limit;
  in.st.join = style.line_join;
  in.st.cap = style.line_cap;
  rzk::StrokeBufs b;
  b.first = c.take<rzs::Seg>(n_items);
  b.last = c.take<rzs::Seg>(n_items);
  b.pieces = c.take<uint32_t>(n_items);
  b.piece_pts = c.take<uint32_t>(n_items);
  b.incl = c.take<int32_t>(n_items);
  b.blk = c.take<int32_t>(nblk);
  b.run_first = c.take<int32_t>(n_runs);
  b.run_has = c.take<uint32_t>(n_runs);
  b.cnt_seg = c.take<uint32_t>(n_items + 1);
  b.cnt_pts = c.take<uint32_t>(n_items + 1);
  uint32_t* seg_off = c.take<uint32_t>(n_items + 1);
  uint32_t* pt_off = c.take<uint32_t>(n_items + 1);
  uint32_t* run_idx = c.take<uint32_t>(n_runs + 1);
  uint32_t* tot = c.take<uint32_t>(4);
  const dim3 grid(nblk), blk(rzk::kStrokeBlock);
  RZ_LAUNCH(ctx, "k_stroke_pieces", rzk::k_stroke_pieces, grid, blk, in, b);
  RZ_LAUNCH(ctx, "k_stroke_blocks", rzk::k_stroke_blocks, dim3(1), dim3(rzk::kBlock), b.blk, nblk);
  RZ_LAUNCH(ctx, "k_fill_u32", rzk::k_fill_u32, dim3(blocks_for(n_runs)), dim3(rzk::kBlock),
            (uint32_t*)b.run_first, (size_t)n_runs, ~0u);
  RZ_LAUNCH(ctx, "k_fill_u32", rzk::k_fill_u32, dim3(blocks_for(n_runs)), dim3(rzk::kBlock),
            b.run_has, (size_t)n_runs, 0u);
  RZ_LAUNCH(ctx, "k_stroke_runs", rzk::k_stroke_runs, grid, blk, in, b);
  RZ_LAUNCH(ctx, "k_stroke_count", rzk::k_stroke_count, grid, blk, in, b);
  RZ_TRY(check_launch(ctx, "stroke count"));
  RZ_TRY(scan_u32(ctx, b.cnt_seg, seg_off, n_items, nullptr, tot + 0));
  RZ_TRY(scan_u32(ctx, b.cnt_pts, pt_off, n_items, nullptr, tot + 1));
  RZ_TRY(scan_u32(ctx, b.run_has, run_idx, n_runs, nullptr, tot + 2));
  uint32_t h_tot[4] = {0, 0, 0, 0};
  RZ_HIP(ctx, hipMemcpyAsync(h_tot, tot, 12, hipMemcpyDeviceToHost, s));
  RZ_HIP(ctx, hipStreamSynchronize(s));
  const uint32_t o_seg = h_tot[0], o_pts = h_tot[1], o_sub = h_tot[2];
  size_t need_out = 0;
  {
    Carve co{nullptr};
    co.take<uint8_t>(o_seg);
    co.take<uint32_t>(o_seg + 1);
    co.take<double>(2 * (size_t)o_pts);
    co.take<uint32_t>(o_sub + 1);
    co.take<uint8_t>(o_sub);
    need_out = co.at;
  }
  RZ_TRY(ensure(ctx, ctx->sk_out, need_out));
  Carve co{(char*)ctx->sk_out.p};
  rzk::StrokeOut o;
  o.seg_off = seg_off;
  o.pt_off = pt_off;
  o.run_idx = run_idx;
  o.kinds = co.take<uint8_t>(o_seg);
  o.seg_pts = co.take<uint32_t>(o_seg + 1);
  o.pts = co.take<double>(2 * (size_t)o_pts);
  o.subs = co.take<uint32_t>(o_sub + 1);
  o.closed = co.take<uint8_t>(o_sub);
  RZ_LAUNCH(ctx, "k_stroke_emit", rzk::k_stroke_emit, grid, blk, in, b, o);
  RZ_TRY(check_launch(ctx, "stroke emit"));
  res.kinds.resize(o_seg);
  res.point_offsets.resize((size_t)o_seg + 1);
  res.points.resize(2 * (size_t)o_pts);
  res.subpaths.resize(o_sub ? (size_t)o_sub + 1 : 0);
  res.closed.resize(o_sub);
  if (o_seg) {
    RZ_HIP(ctx, hipMemcpyAsync(res.kinds.data(), o.kinds, o_seg, hipMemcpyDeviceToHost, s));
    RZ_HIP(ctx, hipMemcpyAsync(res.point_offsets.data(), o.seg_pts, (size_t)o_seg * 4,
                               hipMemcpyDeviceToHost, s));
    RZ_HIP(ctx, hipMemcpyAsync(res.points.data(), o.pts, 16 * (size_t)o_pts, hipMemcpyDeviceToHost,
                               s));
  }
  if (o_sub) {
    RZ_HIP(ctx, hipMemcpyAsync(res.subpaths.data() + 1, o.subs + 1, (size_t)o_sub * 4,
                               hipMemcpyDeviceToHost, s));
    RZ_HIP(ctx, hipMemcpyAsync(res.closed.data(), o.closed, o_sub, hipMemcpyDeviceToHost, s));
  }
  RZ_HIP(ctx, hipStreamSynchronize(s));
  res.point_offsets[o_seg] = o_pts;
  if (o_sub) res.subpaths[0] = 0;
  *out = rz_path_buf_adopt(std::move(res));
  return RZ_OK;
}

rz_status rz_mask_device(rz_ctx* ctx, const rz_dpath* d, rz_transform tr, double flatness,
                         int fill_rule, int edge_mode, size_t width, size_t height, float* d_out,
                         size_t row_stride) {
  RZ_TRY(check_args(ctx, flatness, fill_rule, edge_mode, width, height));
  if (!d || (!d_out && width && height)) return fail(ctx, RZ_E_ARG, "null argument");
  if (row_stride < width) return fail(ctx, RZ_E_ARG, "row_stride < width");
  select_params(ctx, mask_tune_key(d, tr, flatness, edge_mode, width, height));
  GraphKey k = graph_key(ctx, 0, d->serial);
  std::memcpy(k.tr, tr.m, sizeof(k.tr));
  k.flatness = flatness;
  k.rule = fill_rule;
  k.mode = edge_mode;
  k.w = width;
  k.h = height;
  k.stride = row_stride;
  k.ptr[0] = d_out;
  return run_graphed(ctx, k, [&]() {
    return enqueue_mask(ctx, d, tr, flatness, fill_rule, edge_mode, width, height, d_out, row_stride);
  });
}

// The no-alloc form of rz_mask_device: the same pipeline on buffers carved
// from this GPU's `scratch` instead of the context's own (which are left as
// they are). The fixed parts (per tile, per band, per row, per flatten wave,
// events, overflow list) are sized from the call; the rest is shared out
// per line (f64 line, record, two list entries). Inputs that cannot fit are
// refused here; a path that flattens into more lines than fit is reported at
// rz_sync (RZ_E_CAPACITY, nothing written, nothing grown).
rz_status rz_mask_device_noalloc(rz_ctx* ctx, const rz_dpath* d, rz_transform tr, double flatness,
                                 int fill_rule, int edge_mode, size_t width, size_t height,
                                 float* d_out, size_t row_stride) {
  RZ_TRY(check_args(ctx, flatness, fill_rule, edge_mode, width, height));
  if (!d || (!d_out && width && height)) return fail(ctx, RZ_E_ARG, "null argument");
  if (row_stride < width) return fail(ctx, RZ_E_ARG, "row_stride < width");
  if (width == 0 || height == 0) return RZ_OK;
  if (!ctx->scratch_ptr)
    return fail(ctx, RZ_E_HIP, "no-alloc scratch: hipGetSymbolAddress failed at context creation");
  // the sizes enqueue_mask will ask for, as its helpers compute them
  const Grid gr = make_grid(edge_mode, width, height);
  const size_t tiles = gr.tiles, ty = (size_t)gr.g.tiles_y;
  const size_t meta_u32 = mask_meta_words(tiles, ty, height);
  int dq, dc;
  select_params(ctx, mask_tune_key(d, tr, flatness, edge_mode, width, height));
  split_depths(ctx, d->est, std::fabs(tr.m[0] * tr.m[4] - tr.m[1] * tr.m[3]), flatness, &dq, &dc);
  const size_t nblk = (size_t)d->view.blocks_line + ((size_t)d->view.blocks_quad << dq) +
                      ((size_t)d->view.blocks_cubic << dc);
  const size_t nw = nblk * (rzk::kBlock / 64);
  const uint32_t ev_cap = 1u << 14, part_cap = 1u << 12;
  auto scan_blocks = [](size_t n) {
    return std::max<size_t>(1, (n + rzk::kScanTile - 1) / rzk::kScanTile);
  };
  const size_t sums = std::max(scan_blocks(tiles), scan_blocks(ty)) * 4;
  const size_t no = (size_t)rzk::kParts * part_cap;
  const size_t list_fixed = no * 24 + (size_t)ev_cap * 16 +
                            (no + ev_cap + nw + nw * rzk::kMap) * 4;  // + 48 per line
  auto al = [](size_t b) { return (b + 255) & ~(size_t)255; };
  const size_t fixed = al(meta_u32 * 4) + al((tiles + 1) * 4) + al((ty + 1) * 4) + al(sums) +
                       al(std::max<size_t>(nw, 1) * 8) + al((size_t)ev_cap * 16) +
                       al(list_fixed) + al(tiles * 16) + 4 * 256;
  const size_t unit = (size_t)rzk::kLineParts * 64;  // line capacity comes in whole parts
  const size_t per_line = 32 + 16 + 48;
  if (fixed + unit * per_line > kNoAllocBytes || tiles + unit > 0xffffffffull)
    return fail(ctx, RZ_E_CAPACITY,
                "no-alloc call: a " + std::to_string(width) + "x" + std::to_string(height) +
                    " mask of this path needs more than the " +
                    std::to_string(kNoAllocBytes >> 20) + " MiB scratch");
  const size_t line_cap = std::min<size_t>((kNoAllocBytes - fixed) / per_line / unit * unit,
                                           (size_t)1 << 30);
  const size_t rec_cap = line_cap + tiles;
  // carve, in order; every part starts 256-byte aligned
  char* p = ctx->scratch_ptr;
  auto take = [&](size_t bytes) {
    DBuf b;
    b.p = p;
    b.cap = al(bytes);
    p += b.cap;
    return b;
  };
  struct Saved {
    rz_ctx* c;
    DBuf meta, tile_off, ev_off, sums, lines, recs, events, list_buf, wave_lines;
    uint32_t line_cap, rec_cap, ev_cap, part_cap;
    ~Saved() {
      c->meta = meta; c->tile_off = tile_off; c->ev_off = ev_off; c->sums = sums;
      c->lines = lines; c->recs = recs; c->events = events; c->list_buf = list_buf;
      c->wave_lines = wave_lines;
      c->line_cap = line_cap; c->rec_cap = rec_cap; c->ev_cap = ev_cap; c->part_cap = part_cap;
      c->noalloc = false;
      c->acc_run = c->acc;
    }
  } saved{ctx,          ctx->meta,   ctx->tile_off, ctx->ev_off,   ctx->sums,
          ctx->lines,   ctx->recs,   ctx->events,   ctx->list_buf, ctx->wave_lines,
          ctx->line_cap, ctx->rec_cap, ctx->ev_cap, ctx->part_cap};
  ctx->noalloc = true;
  ctx->acc_run = ctx->acc + 1;
  ctx->meta = take(meta_u32 * 4);
  ctx->tile_off = take((tiles + 1) * 4);
  ctx->ev_off = take((ty + 1) * 4);
  ctx->sums = take(sums);
  ctx->wave_lines = take(std::max<size_t>(nw, 1) * 8);
  ctx->events = take((size_t)ev_cap * 16);
  ctx->list_buf = take(list_fixed + 48 * line_cap);
  ctx->recs = take(rec_cap * 16);
  ctx->lines = take(line_cap * 32);
  ctx->line_cap = (uint32_t)line_cap;
  ctx->rec_cap = (uint32_t)rec_cap;
  ctx->ev_cap = ev_cap;
  ctx->part_cap = part_cap;
  if ((size_t)(p - ctx->scratch_ptr) > kNoAllocBytes)
    return fail(ctx, RZ_E_CAPACITY, "no-alloc call: scratch layout overflow");
  return enqueue_mask(ctx, d, tr, flatness, fill_rule, edge_mode, width, height, d_out, row_stride);
}

rz_status rz_fill_device(rz_ctx* ctx, const rz_dpath* d, rz_transform tr, double flatness,
                         int fill_rule, int edge_mode, const rz_paint* paint, size_t width,
                         size_t height, float* d_rgba, size_t row_stride) {
  RZ_TRY(check_args(ctx, flatness, fill_rule, edge_mode, width, height));
  if (!d || (!d_rgba && width && height)) return fail(ctx, RZ_E_ARG, "null argument");
  if (row_stride < width) return fail(ctx, RZ_E_ARG, "row_stride < width");
  return enqueue_fill(ctx, d, tr, flatness, fill_rule, edge_mode, paint, width, height, d_rgba,
                      row_stride);
}

rz_status rz_fill(rz_ctx* ctx, const rz_path* path, rz_transform tr, double flatness,
                  int fill_rule, int edge_mode, const rz_paint* paint, float* rgba,
                  size_t rgba_len, rz_shape shape) {
  RZ_TRY(check_args(ctx, flatness, fill_rule, edge_mode, shape.width, shape.height));
  const size_t W = shape.width, H = shape.height;
  if (W == 0 || H == 0) return RZ_OK;
  if (!rgba) return fail(ctx, RZ_E_ARG, "rgba is NULL");
  if (!paint) return fail(ctx, RZ_E_ARG, "paint is NULL");
  const size_t last = shape.start + (H - 1) * shape.row_stride + (W - 1) * shape.col_stride;
  if (4 * last + 4 > rgba_len) return fail(ctx, RZ_E_ARG, "shape exceeds data");
  rz_dpath* d = nullptr;
  RZ_TRY(path_upload_impl(ctx, path, &d));
  struct Guard {
    rz_ctx* c;
    rz_dpath* d;
    ~Guard() { rz_path_free(c, d); }
  } guard{ctx, d};
  // the destination view goes to a dense device image and back
  const size_t bytes = W * H * 4 * sizeof(float);
  std::vector<float> dense(W * H * 4);
  for (size_t y = 0; y < H; ++y)
    for (size_t x = 0; x < W; ++x) {
      const size_t o = 4 * (shape.start + y * shape.row_stride + x * shape.col_stride);
      std::memcpy(&dense[4 * (y * W + x)], rgba + o, 4 * sizeof(float));
    }
  RZ_TRY(ensure(ctx, ctx->out, bytes));
  RZ_HIP(ctx, hipMemcpy(ctx->out.p, dense.data(), bytes, hipMemcpyHostToDevice));
  rz_status st = RZ_OK;
  for (int attempt = 0; attempt < kRuns; ++attempt) {
    st = enqueue_fill(ctx, d, tr, flatness, fill_rule, edge_mode, paint, W, H, (float*)ctx->out.p, W);
    if (st != RZ_OK) return st;
    st = finish(ctx);
    if (st != RZ_E_CAPACITY) break;
  }
  if (st != RZ_OK) return st;
  RZ_HIP(ctx, hipMemcpy(dense.data(), ctx->out.p, bytes, hipMemcpyDeviceToHost));
  for (size_t y = 0; y < H; ++y)
    for (size_t x = 0; x < W; ++x) {
      const size_t o = 4 * (shape.start + y * shape.row_stride + x * shape.col_stride);
      std::memcpy(rgba + o, &dense[4 * (y * W + x)], 4 * sizeof(float));
    }
  return RZ_OK;
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
  for (int attempt = 0; attempt < kRuns; ++attempt) {
    RZ_TRY(enqueue_bin(ctx, d, tr, flatness, edge_mode, width, height, &P));
    RZ_LAUNCH(ctx, "k_status_merge", rzk::k_status_merge, dim3(1), dim3(64),
              (const rzk::Status*)P.st, ctx->acc_run);
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
  if (right_edge) {
    // kept in int64 on the device; the hook reports int32 (values beyond it
    // are clamped)
    std::vector<long long> re64(height);
    RZ_HIP(ctx, hipMemcpy(re64.data(), P.redge, height * 8, hipMemcpyDeviceToHost));
    for (size_t y = 0; y < height; ++y)
      right_edge[y] = (int32_t)std::max<long long>(INT32_MIN, std::min<long long>(INT32_MAX, re64[y]));
  }
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

rz_status rz_scene_upload(rz_ctx* ctx, const rz_path* const* paths, const rz_paint* paints,
                          const int32_t* fill_rules, size_t n, rz_dscene** out) {
  return scene_upload_impl(ctx, paths, paints, fill_rules, n, out);
}

rz_status rz_scene_free(rz_ctx* ctx, rz_dscene* scene) {
  if (!scene) return RZ_OK;
  if (ctx) (void)hipStreamSynchronize(ctx->stream);
  scene_release(scene);
  delete scene;
  return RZ_OK;
}

rz_status rz_render_device(rz_ctx* ctx, const rz_dscene* scene, rz_transform tr, double flatness,
                           int edge_mode, const double* view, const float* bg, int batched,
                           float* d_rgba, size_t row_stride) {
  RZ_TRY(check_scene_args(ctx, flatness, edge_mode, view));
  if (!scene) return fail(ctx, RZ_E_ARG, "scene is NULL");
  int lx, ly, lw, lh;
  layer_of(view, &lx, &ly, &lw, &lh);
  if (lw && lh && !d_rgba) return fail(ctx, RZ_E_ARG, "d_rgba is NULL");
  if (row_stride < (size_t)lw) return fail(ctx, RZ_E_ARG, "row_stride < layer width");
  const rzp::C4 c = bg ? rzp::c4(bg[0], bg[1], bg[2], bg[3]) : rzp::splat(0.0f);
  if (batched) {
    select_params(ctx, scene_tune_key(scene, tr, flatness, edge_mode, view));
    GraphKey k = graph_key(ctx, 1, scene->serial);
    std::memcpy(k.tr, tr.m, sizeof(k.tr));
    std::memcpy(k.view, view, sizeof(k.view));
    if (bg) std::memcpy(k.bg, bg, sizeof(k.bg));
    k.has_bg = bg != nullptr;
    k.flatness = flatness;
    k.mode = edge_mode;
    k.stride = row_stride;
    k.ptr[0] = d_rgba;
    return run_graphed(ctx, k, [&]() {
      return enqueue_scene_batched(ctx, scene, tr, view, c, edge_mode, flatness, (rzp::C4*)d_rgba,
                                   row_stride);
    });
  }
  return enqueue_scene_plain(ctx, scene, tr, view, c, edge_mode, flatness, (rzp::C4*)d_rgba,
                             row_stride);
}

rz_status rz_render(rz_ctx* ctx, const rz_path* const* paths, const rz_paint* paints,
                    const int32_t* fill_rules, size_t n, rz_transform tr, double flatness,
                    int edge_mode, const double* view, const float* bg, int batched, float* rgba,
                    size_t cap_px, int32_t* layer_xywh) {
  if (!ctx || !layer_xywh) return RZ_E_ARG;
  // Pipeline::build (src/scene.rs:268-357): without a view the layer is the
  // union of the fills' bboxes; no fill in view means an empty layer
  double v[4];
  bool any = false;
  const rzh::Tr T = rzh::tr_of(tr.m);
  for (size_t i = 0; i < n; ++i) {
    RZ_TRY(validate_path(ctx, paths ? paths[i] : nullptr));
    rzh::BBox b;
    if (!rzh::path_bbox(paths[i], T, &b)) continue;
    if (view) {
      if (view[0] > b.mx.x || b.mn.x > view[2] || view[1] > b.mx.y || b.mn.y > view[3]) continue;
      any = true;
      continue;
    }
    if (!any) {
      v[0] = b.mn.x;
      v[1] = b.mn.y;
      v[2] = b.mx.x;
      v[3] = b.mx.y;
    } else {
      v[0] = std::fmin(v[0], b.mn.x);
      v[1] = std::fmin(v[1], b.mn.y);
      v[2] = std::fmax(v[2], b.mx.x);
      v[3] = std::fmax(v[3], b.mx.y);
    }
    any = true;
  }
  for (int k = 0; k < 4; ++k) layer_xywh[k] = 0;
  if (!any) return RZ_OK;
  if (view)
    for (int k = 0; k < 4; ++k) v[k] = view[k];
  RZ_TRY(check_scene_args(ctx, flatness, edge_mode, v));
  int lx, ly, lw, lh;
  layer_of(v, &lx, &ly, &lw, &lh);
  layer_xywh[0] = lx;
  layer_xywh[1] = ly;
  layer_xywh[2] = lw;
  layer_xywh[3] = lh;
  const size_t px = (size_t)lw * lh;
  if (!rgba || cap_px < px) return RZ_OK;
  rz_dscene* s = nullptr;
  RZ_TRY(scene_upload_impl(ctx, paths, paints, fill_rules, n, &s));
  struct Guard {
    rz_ctx* c;
    rz_dscene* s;
    ~Guard() { rz_scene_free(c, s); }
  } guard{ctx, s};
  RZ_TRY(ensure(ctx, ctx->out, std::max<size_t>(px, 1) * 16));
  rz_status st = RZ_OK;
  for (int attempt = 0; attempt < kRuns; ++attempt) {
    st = rz_render_device(ctx, s, tr, flatness, edge_mode, v, bg, batched, (float*)ctx->out.p,
                          (size_t)lw);
    if (st != RZ_OK) return st;
    st = finish(ctx);
    if (st != RZ_E_CAPACITY) break;
  }
  if (st != RZ_OK) return st;
  RZ_HIP(ctx, hipMemcpy(rgba, ctx->out.p, px * 16, hipMemcpyDeviceToHost));
  return RZ_OK;
}

rz_status rz_render_tree(rz_ctx* ctx, const rz_path* const* paths, size_t n_paths,
                         const rz_paint* paints, size_t n_paints, const rz_scene_node* nodes,
                         size_t n_nodes, const uint32_t* children, uint32_t root, rz_transform tr,
                         double flatness, int edge_mode, const double* view, const float* bg,
                         float* rgba, size_t cap_px, int32_t* layer_xywh) {
  if (!ctx || !layer_xywh || (n_nodes && !nodes)) return RZ_E_ARG;
  if (!(flatness > 0.0) || !std::isfinite(flatness)) return fail(ctx, RZ_E_ARG, "flatness must be > 0");
  if (edge_mode < RZ_EDGE_MASK || edge_mode > RZ_EDGE_NONE) return fail(ctx, RZ_E_ARG, "bad edge mode");
  for (size_t i = 0; i < n_paths; ++i) RZ_TRY(validate_path(ctx, paths ? paths[i] : nullptr));
  for (int k = 0; k < 4; ++k) layer_xywh[k] = 0;
  // Stroke nodes: their paths stroked on the GPU, appended to the path list
  std::vector<const rz_path*> all(paths, paths + n_paths);
  std::vector<uint32_t> stroked(n_nodes, ~0u);
  struct Owned {
    std::vector<rz_path_buf*> bufs;
    ~Owned() {
      for (rz_path_buf* b : bufs) rz_path_buf_free(b);
    }
  } owned;
  std::vector<rz_path> views;
  size_t n_strokes = 0;
  for (size_t i = 0; i < n_nodes; ++i) n_strokes += nodes[i].kind == RZ_NODE_STROKE;
  views.reserve(n_strokes);  // `all` points into it
  for (size_t i = 0; i < n_nodes; ++i) {
    if (nodes[i].kind != RZ_NODE_STROKE) continue;
    if (nodes[i].path >= n_paths) return fail(ctx, RZ_E_ARG, "malformed scene tree (stroke path index)");
    const rz_stroke_style st{nodes[i].stroke_width, nodes[i].miter_limit, nodes[i].line_join,
                             nodes[i].line_cap};
    rz_path_buf* b = nullptr;
    RZ_TRY(rz_path_stroke(ctx, paths[nodes[i].path], st, &b));
    owned.bufs.push_back(b);
    rz_path v;
    rz_path_buf_view(b, &v);
    views.push_back(v);
    all.push_back(&views.back());
    stroked[i] = (uint32_t)all.size() - 1;
  }
  TreeBuild tb{nodes, n_nodes, children, all.data(), all.size(), n_paints, {}};
  tb.stroked = &stroked;
  rzh::BBox vb;
  if (view) {
    for (int k = 0; k < 4; ++k)
      if (!std::isfinite(view[k])) return fail(ctx, RZ_E_ARG, "view must be finite");
    vb = rzh::bbox_new(rzg::mkpt(view[0], view[1]), rzg::mkpt(view[2], view[3]));
  }
  const int root_id = tb.build(root, view ? &vb : nullptr, rzh::tr_of(tr.m));
  if (tb.bad) return fail(ctx, RZ_E_ARG, "malformed scene tree (index out of range, or deeper than 256)");
  if (root_id < 0) return RZ_OK;  // nothing in view: the reference's empty layer
  const rzh::BBox lb = view ? vb : tb.out[root_id].bbox;
  double v[4] = {lb.mn.x, lb.mn.y, lb.mx.x, lb.mx.y};
  RZ_TRY(check_scene_args(ctx, flatness, edge_mode, v));
  int lx, ly, lw, lh;
  layer_of(v, &lx, &ly, &lw, &lh);
  layer_xywh[0] = lx;
  layer_xywh[1] = ly;
  layer_xywh[2] = lw;
  layer_xywh[3] = lh;
  if (!rgba || cap_px < (size_t)lw * lh || !lw || !lh) return RZ_OK;
  const rzp::C4 c = bg ? rzp::c4(bg[0], bg[1], bg[2], bg[3]) : rzp::splat(0.0f);
  rz_status st = RZ_OK;
  for (int attempt = 0; attempt < kRuns; ++attempt) {
    TreeRender R{ctx, &tb, paints, edge_mode, flatness, {}, {}, {}, {}};
    DLayer L;
    st = R.new_layer(lb, c, &L);
    if (st == RZ_OK) st = R.render(root_id, L);
    if (st == RZ_OK) st = R.flush(L);
    if (st != RZ_OK) return st;
    st = finish(ctx);
    if (st == RZ_OK) {
      RZ_HIP(ctx, hipMemcpy(rgba, L.d, (size_t)lw * lh * 16, hipMemcpyDeviceToHost));
      return RZ_OK;
    }
    if (st != RZ_E_CAPACITY) return st;
  }
  return st;
}

rz_status rz_atlas_set(rz_ctx* ctx, rz_dscene* glyphs, const rz_atlas_slot* slots) {
  return atlas_set_impl(ctx, glyphs, slots);
}

rz_status rz_mask_atlas_device(rz_ctx* ctx, const rz_dscene* glyphs, double flatness, int fill_rule,
                               int edge_mode, float* d_atlas, size_t row_stride) {
  RZ_TRY(check_args(ctx, flatness, fill_rule, edge_mode, 0, 0));
  if (!glyphs || !glyphs->atlas_ready) return fail(ctx, RZ_E_ARG, "glyph placement not set");
  if (glyphs->n && !d_atlas) return fail(ctx, RZ_E_ARG, "d_atlas is NULL");
  select_params(ctx, atlas_tune_key(glyphs, flatness, edge_mode));
  GraphKey k = graph_key(ctx, 2, glyphs->serial);
  k.flatness = flatness;
  k.rule = fill_rule;
  k.mode = edge_mode;
  k.stride = row_stride;
  k.ptr[0] = d_atlas;
  return run_graphed(ctx, k, [&]() {
    return enqueue_atlas(ctx, glyphs, fill_rule, edge_mode, flatness, d_atlas, row_stride);
  });
}

rz_status rz_mask_atlas(rz_ctx* ctx, const rz_path* const* paths, const rz_atlas_slot* slots,
                        size_t n, double flatness, int fill_rule, int edge_mode, float* atlas,
                        size_t atlas_width, size_t atlas_height) {
  RZ_TRY(check_args(ctx, flatness, fill_rule, edge_mode, 0, 0));
  if (n && (!slots || !atlas)) return fail(ctx, RZ_E_ARG, "null argument");
  for (size_t i = 0; i < n; ++i)
    if ((size_t)slots[i].x + slots[i].width > atlas_width ||
        (size_t)slots[i].y + slots[i].height > atlas_height)
      return fail(ctx, RZ_E_ARG, "glyph slot outside the atlas");
  rz_dscene* s = nullptr;
  RZ_TRY(scene_upload_impl(ctx, paths, nullptr, nullptr, n, &s));
  struct Guard {
    rz_ctx* c;
    rz_dscene* s;
    ~Guard() { rz_scene_free(c, s); }
  } guard{ctx, s};
  RZ_TRY(atlas_set_impl(ctx, s, slots));
  const size_t bytes = atlas_width * atlas_height * sizeof(float);
  RZ_TRY(ensure(ctx, ctx->out, std::max<size_t>(bytes, 4)));
  // pixels outside every glyph view keep the caller's values
  RZ_HIP(ctx, hipMemcpy(ctx->out.p, atlas, bytes, hipMemcpyHostToDevice));
  rz_status st = RZ_OK;
  for (int attempt = 0; attempt < kRuns; ++attempt) {
    st = enqueue_atlas(ctx, s, fill_rule, edge_mode, flatness, (float*)ctx->out.p, atlas_width);
    if (st != RZ_OK) return st;
    st = finish(ctx);
    if (st != RZ_E_CAPACITY) break;
  }
  if (st != RZ_OK) return st;
  RZ_HIP(ctx, hipMemcpy(atlas, ctx->out.p, bytes, hipMemcpyDeviceToHost));
  return RZ_OK;
}

rz_status rz_scene_update_device(rz_ctx* ctx, rz_dscene* scene, const double* d_points,
                                 const float* d_colors) {
  if (!ctx || !scene) return RZ_E_ARG;
  if (d_points && scene->n_points)
    RZ_HIP(ctx, hipMemcpyAsync((void*)scene->all->view.pts, d_points,
                               (size_t)scene->n_points * 2 * sizeof(double),
                               hipMemcpyDeviceToDevice, ctx->stream));
  if (d_colors && scene->n)
    RZ_LAUNCH(ctx, "k_set_colors", k_set_colors, dim3((scene->n + 255) / 256), dim3(256),
              scene->sp_mut, d_colors, scene->n);
  return check_launch(ctx, "scene update");
}

rz_status rz_scene_stops(rz_ctx* ctx, const rz_dscene* scene, size_t* n_stops) {
  if (!ctx || !scene || !n_stops) return RZ_E_ARG;
  *n_stops = scene->n_user_stops;
  return RZ_OK;
}

rz_status rz_scene_update_stops_device(rz_ctx* ctx, rz_dscene* scene, const float* d_stop_colors) {
  if (!ctx || !scene) return RZ_E_ARG;
  if (!d_stop_colors || !scene->n_user_stops) return RZ_OK;
  RZ_LAUNCH(ctx, "k_set_stops", k_set_stops, dim3((scene->n_stops + 255) / 256), dim3(256),
            d_stop_colors, (const int32_t*)scene->stop_user_dev, (const uint8_t*)scene->stop_flag_dev,
            scene->stop_lin_dev, scene->stop_col_dev, scene->n_stops);
  return check_launch(ctx, "stop update");
}

rz_status rz_encode_rgba8_device(rz_ctx* ctx, const float* d_rgba, size_t width, size_t height,
                                 size_t row_stride, uint8_t* d_out, size_t out_stride,
                                 int bottom_up) {
  if (!ctx) return RZ_E_ARG;
  if (width == 0 || height == 0) return RZ_OK;
  if (!d_rgba || !d_out) return fail(ctx, RZ_E_ARG, "null argument");
  if (row_stride < width || out_stride < 4 * width) return fail(ctx, RZ_E_ARG, "stride too small");
  if (width > (1u << 24) || height > (1u << 24)) return fail(ctx, RZ_E_ARG, "image too large");
  const size_t n = width * height;
  const uint32_t grid = (uint32_t)std::min<size_t>((n + 255) / 256, 16384);
  RZ_LAUNCH(ctx, "k_encode_rgba8", rze::k_encode_rgba8, dim3(grid), dim3(256),
            (const rzp::C4*)d_rgba, row_stride, (int)width, (int)height, d_out, out_stride,
            bottom_up ? 1 : 0);
  return check_launch(ctx, "k_encode_rgba8");
}

rz_status rz_encode_rgba8(rz_ctx* ctx, const float* rgba, size_t width, size_t height,
                          size_t row_stride, uint8_t* out, size_t out_stride, int bottom_up) {
  if (!ctx) return RZ_E_ARG;
  if (width == 0 || height == 0) return RZ_OK;
  if (!rgba || !out) return fail(ctx, RZ_E_ARG, "null argument");
  if (row_stride < width || out_stride < 4 * width) return fail(ctx, RZ_E_ARG, "stride too small");
  if (width > (1u << 24) || height > (1u << 24)) return fail(ctx, RZ_E_ARG, "image too large");
  // the image goes to the device densely, the bytes come back densely
  const size_t in_bytes = width * height * 16, out_bytes = width * height * 4;
  RZ_TRY(ensure(ctx, ctx->out, in_bytes + out_bytes));
  char* d = (char*)ctx->out.p;
  RZ_HIP(ctx, hipMemcpy2DAsync(d, width * 16, rgba, row_stride * 16, width * 16, height,
                               hipMemcpyHostToDevice, ctx->stream));
  RZ_TRY(rz_encode_rgba8_device(ctx, (const float*)d, width, height, width,
                                (uint8_t*)(d + in_bytes), width * 4, bottom_up));
  RZ_HIP(ctx, hipMemcpy2DAsync(out, out_stride, d + in_bytes, width * 4, width * 4, height,
                               hipMemcpyDeviceToHost, ctx->stream));
  RZ_HIP(ctx, hipStreamSynchronize(ctx->stream));
  return RZ_OK;
}

rz_status rz_bmp_header(size_t width, size_t height, uint8_t* out) {
  if (!out || width > 0xffffu * 0xffffu || height > 0x7fffffffu ||
      width * height > (0xffffffffu - rze::kBmpHeader) / 4)
    return RZ_E_ARG;
  rze::bmp_header((uint32_t)width, (uint32_t)height, out);
  return RZ_OK;
}

rz_status rz_encode_png(const uint8_t* rgba8, size_t width, size_t height, size_t stride,
                        uint8_t* out, size_t cap, size_t* len) {
  if (!len || width > 0x7fffffffu || height > 0x7fffffffu) return RZ_E_ARG;
  if (height && (width * 4 + 1) * height > 0x7fffffffu) return RZ_E_ARG;  // one IDAT chunk
  *len = rze::png_size(width, height);
  if (!out) return RZ_OK;
  if (cap < *len) return RZ_E_CAPACITY;
  if (width && height && (!rgba8 || stride < 4 * width)) return RZ_E_ARG;
  rze::png_encode(rgba8, width, height, stride, out);
  return RZ_OK;
}

rz_status rz_scene_points(rz_ctx* ctx, const rz_dscene* scene, size_t* n_points) {
  if (!ctx || !scene || !n_points) return RZ_E_ARG;
  *n_points = scene->n_points;
  return RZ_OK;
}

rz_status rz_render_backward_stops_device(rz_ctx* ctx, const rz_dscene* scene, rz_transform tr,
                                          double flatness, int edge_mode, const double* view,
                                          const float* bg, const float* d_grad_rgba,
                                          size_t grad_stride, float* d_grad_points,
                                          float* d_grad_colors, float* d_grad_stops,
                                          double* d_grad_tr) {
  RZ_TRY(check_scene_args(ctx, flatness, edge_mode, view));
  if (!scene || !d_grad_rgba || !d_grad_points || !d_grad_colors)
    return fail(ctx, RZ_E_ARG, "null argument");
  const rzp::C4 c = bg ? rzp::c4(bg[0], bg[1], bg[2], bg[3]) : rzp::splat(0.0f);
  select_params(ctx, scene_tune_key(scene, tr, flatness, edge_mode, view));
  GraphKey k = graph_key(ctx, 3, scene->serial);
  std::memcpy(k.tr, tr.m, sizeof(k.tr));
  std::memcpy(k.view, view, sizeof(k.view));
  if (bg) std::memcpy(k.bg, bg, sizeof(k.bg));
  k.has_bg = bg != nullptr;
  k.flatness = flatness;
  k.mode = edge_mode;
  k.stride = grad_stride;
  k.ptr[0] = d_grad_rgba;
  k.ptr[1] = d_grad_points;
  k.ptr[2] = d_grad_colors;
  k.ptr[3] = d_grad_tr;
  k.ptr[4] = d_grad_stops;
  return run_graphed(ctx, k, [&]() {
    return enqueue_scene_backward(ctx, scene, tr, view, c, edge_mode, flatness, d_grad_rgba,
                                  grad_stride, d_grad_points, d_grad_colors, d_grad_stops,
                                  d_grad_tr);
  });
}

rz_status rz_render_backward_device(rz_ctx* ctx, const rz_dscene* scene, rz_transform tr,
                                    double flatness, int edge_mode, const double* view,
                                    const float* bg, const float* d_grad_rgba, size_t grad_stride,
                                    float* d_grad_points, float* d_grad_colors, double* d_grad_tr) {
  return rz_render_backward_stops_device(ctx, scene, tr, flatness, edge_mode, view, bg, d_grad_rgba,
                                         grad_stride, d_grad_points, d_grad_colors, nullptr,
                                         d_grad_tr);
}

}  // extern "C"

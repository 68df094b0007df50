/*
 * rz_workloads.h -- seeded synthetic inputs for the benchmark configurations.
 *
 * Not part of the reference's API: these build the inputs the benchmark and
 * the parity tests share, from the reference's own LCG `Rnd`
 * (benches/scene_bench.rs:51-112), so every consumer sees identical paths.
 * Paths are built with PathBuilder semantics (src/path.rs:799-1056), arcs
 * expanded to cubics as PathBuilder::arc_to does.
 */
#ifndef RZ_WORKLOADS_H
#define RZ_WORKLOADS_H

#include "rz.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct rzw_set rzw_set;

/* One closed subpath of 64 cubic Beziers with control points uniform in
 * [0, size)^2, the last endpoint back at the first point (self-intersecting). */
rzw_set* rzw_cubics64(uint32_t seed, double size);

/* One path of n_segments commands: 40% lines, 30% quads, 25% cubics, 5%
 * elliptic arcs (radii U(1,200), rotation U(0,360) degrees, random flags);
 * endpoints a random walk with N(0, 40) steps clamped to [0, canvas)^2,
 * a new closed subpath every 500 commands. The set also carries the random
 * affine transform: scale U(0.5, 2), rotation U[0, 2pi), centroid moved to the
 * canvas centre. */
rzw_set* rzw_random_walk(uint32_t seed, uint32_t n_segments, double canvas);

/* A scene of n_paths closed star/blob paths (segment count exp(N(ln 60, 1.4))
 * clamped to 4..4000, each edge a line, quad or cubic with equal odds),
 * bbox side log-uniform in 8..1024 px, centre uniform in [0, canvas)^2.
 * Paints: 60% solid, 25% linear, 15% radial (focus at the centre), 2-6
 * sorted random stops, random sRGB colours with random alpha, UserSpaceOnUse,
 * Pad, sRGB interpolation; fill rules 80% NonZero, 20% EvenOdd. */
rzw_set* rzw_scene(uint32_t seed, uint32_t n_paths, double canvas);

/* The reference's scene benchmark (benches/scene_bench.rs:6-47): `count`
 * circles (PathBuilder::circle) of radius U(30, 40) centred uniformly in
 * [0, size)^2, opaque random colours, Rnd seed 0. */
rzw_set* rzw_circles(uint32_t count, double size);

/* A glyph set: n closed paths of U{8..200} quadratic Beziers split over 1-3
 * contours (an outer blob and smaller inner blobs traced the other way round,
 * i.e. holes), each fitted with a 1 px margin into its own mask of side
 * U{16..128} px, shelf-packed (tallest first) into an atlas atlas_width px
 * wide. Placement per glyph through rzw_slot, atlas size through
 * rzw_atlas_size. */
rzw_set* rzw_glyphs(uint32_t seed, uint32_t n, uint32_t atlas_width);
void rzw_slot(const rzw_set* set, size_t i, rz_atlas_slot* out);
void rzw_atlas_size(const rzw_set* set, uint32_t* width, uint32_t* height);

/* Paint and fill rule of path i of a scene set. */
void rzw_paint(const rzw_set* set, size_t i, rz_paint* out);
int32_t rzw_rule(const rzw_set* set, size_t i);

/* Number of paths in the set and a view of path i (valid while the set lives). */
size_t rzw_count(const rzw_set* set);
void rzw_path(const rzw_set* set, size_t i, rz_path* out);
/* The set's transform (identity unless the generator draws one). */
void rzw_transform(const rzw_set* set, rz_transform* out);
void rzw_free(rzw_set* set);

/* The reference's Rnd: first `n` values of uniform() for the given seed
 * (used to pin the generator against the LCG's definition). */
void rzw_rnd_uniform(uint32_t seed, size_t n, double* out);

/* PathBuilder::arc_to from `src` (src/path.rs:945-972): writes the cubics
 * (8 doubles each, at most 4) and returns their count, or -1 when the arc
 * degenerates into line_to. */
int rzw_arc_to(const double* src, const double* radii, double x_axis_rot, int large, int sweep,
               const double* dst, double* out);

#ifdef __cplusplus
}
#endif

#endif /* RZ_WORKLOADS_H */

/*
 * rb_oracle.h -- CPU restatement of the renderbox2 ray-tracing hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker: it is loaded by
 * tests/, by __graft_entry__.smoke() and by the cpu_baseline leg of bench.py,
 * and nowhere else.  The product library (renderbox2_amd/) never links,
 * loads or calls it.
 *
 * Every function restates the reference algorithm named in its comment
 * (paths relative to renderbox2/ in the reference tree).  Arithmetic is plain
 * IEEE-754 binary32; the file is compiled with -ffp-contract=off so that every
 * multiply and add rounds on its own unless an fmaf() is written out.
 *
 * The pins of each function (closed forms, FP64 brute force, invariants,
 * known-answer vectors) live in tests/test_oracle.py and
 * tests/test_oracle_render.py.
 */
#ifndef RB_ORACLE_H
#define RB_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- math ------------------------------------------------------------- */

/* Matrix4x4::invert, src/core/transform.h:205-333 (cofactor expansion). */
void orc_invert4x4(const float* m, float* out);

/* CudaBvh::woopify_tri, src/accelerators/sbvh/cudabvh.cpp:303-322.
 * out = 3 rows of float4 (12 floats). */
void orc_woopify(const float* v0, const float* v1, const float* v2, float* out);

/* ---- SBVH build ------------------------------------------------------- */

typedef struct orc_bvh orc_bvh;

/* SplitBVHBuilder::run/build_node, src/accelerators/sbvh/splitbvhbuilder.cpp:234-627
 * with Platform defaults (src/accelerators/sbvh/platform.h:81-90) except the
 * leaf preferences, and BuildParams::m_split_alpha (src/accelerators/sbvh/bvh.h:110).
 * Returns NULL on allocation failure. */
orc_bvh* orc_bvh_build(const float* verts, uint32_t nv, const uint32_t* idx, uint32_t nt,
                       uint32_t min_leaf, uint32_t max_leaf, float split_alpha,
                       float node_cost, float tri_cost);
void orc_bvh_free(orc_bvh* b);

/* BVH::Stats (src/accelerators/sbvh/bvh.h:78-98) filled as in
 * src/accelerators/sbvh/splitbvhbuilder.cpp:200-213 and bvnode.cpp:64-100.
 * stats: [0]=n_inner [1]=n_leaf [2]=n_tri_refs [3]=n_child_nodes [4]=duplicates
 *        [5]=max_depth  ; sah written to *sah. */
void orc_bvh_stats(const orc_bvh* b, uint64_t stats[6], float* sah);

/* Node dump for invariant tests: per node (pre-order from the root)
 * is_leaf, child0, child1 (node numbers, -1 for leaves), lo, hi (into the
 * tri-index list), bounds[6].  Returns the node count when buffers are NULL. */
uint32_t orc_bvh_nodes(const orc_bvh* b, int32_t* info5, float* bounds6);
uint32_t orc_bvh_tri_indices(const orc_bvh* b, uint32_t* out);

/* CudaBvh::create_compact(bvh, 16) -- the "Compact2" layout,
 * src/accelerators/sbvh/cudabvh.cpp:216-296.  Sizes are in uint4 rows (nodes,
 * woop) and uint32 entries (index).  Call once with NULL buffers for sizes. */
int orc_bvh_compact2(const orc_bvh* b, uint32_t* nodes, size_t* n_node_rows,
                     uint32_t* woop, size_t* n_woop_rows, uint32_t* tri_index,
                     size_t* n_index);

/* ---- traversal -------------------------------------------------------- */

/* kernel_kepler_trace, src/accelerators/sbvh/kepler_dynamic_fetch.h:67-363,
 * restated for one ray at a time (the warp vote of :234-243 reduces to the
 * ray's own predicate).  Ties at equal t go to the lowest triangle id.
 *   rays:   n x 8 floats (ox,oy,oz,tmin, dx,dy,dz,tmax)
 *   out_id: original triangle id or -1; out_t: hit t (tmax on a miss)
 *   counts: optional n x 3 (inner nodes fetched, triangles tested, leaf
 *           terminators reached).
 * nthreads <= 0 uses every core (OpenMP). */
void orc_trace_compact2(const uint32_t* nodes, const uint32_t* woop, const uint32_t* tri_index,
                        const float* rays, uint32_t n, int any_hit,
                        int32_t* out_id, float* out_t, uint32_t* counts, int nthreads);

/* FP64 brute force over every triangle (Moller-Trumbore, the textbook form
 * of intersect_ray_triangle, src/core/primitives.cu:51-70, without its
 * epsilon cull).  out_t64 is the double t; out_margin is the smallest of
 * (u, v, 1-u-v) at the hit, scaled by nothing (used to flag grazing hits);
 * out_gap is the relative distance to the next different triangle's t. */
void orc_bruteforce_f64(const float* verts, const uint32_t* idx, uint32_t nt,
                        const float* rays, uint32_t n, int32_t* out_id, double* out_t64,
                        double* out_margin, double* out_gap, int nthreads);

/* The Woop test of kepler_dynamic_fetch.h:260-282 on a single triangle.
 * orc_woop_hits: for ray i and triangle ids[i] (skip if < 0) write the t and
 * whether it is a hit inside (tmin, tmax) with u, v >= 0, u + v <= 1. */
int orc_woop_hit(const float* v0, const float* v1, const float* v2, const float* ray, float* t_out);
void orc_woop_hits(const float* verts, const uint32_t* idx, const int32_t* ids, const float* rays, uint32_t n,
                   float* out_t, uint8_t* out_hit);

/* ---- RNG -------------------------------------------------------------- */

/* Philox4x32-10 (Salmon et al., SC'11), the counter-based generator that
 * replaces Thrust default_random_engine / cuRAND XORWOW. */
void orc_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]);

/* ---- sampling / camera / ray generation -------------------------------- */

/* sin and cos of 2*pi*u by quarter-turn reduction and Taylor series. */
void orc_sincos_2pi(float u, float* s, float* c);
/* cosine-weighted direction about unit n (Duff et al. basis, Malley map). */
void orc_cosine_dir(const float* n, float u1, float u2, float* out);

/* PerspectiveCamera ctor (src/core/camera.cu:30-57): raster->camera and
 * camera->world matrices, row-major. */
void orc_camera_matrices(const float* pos, const float* at, const float* up, float hfov,
                         uint32_t width, uint32_t height, float* r2c, float* c2w);
/* kernel_generate_perspective_camera_samples (src/renderers/samplerrenderer.cu:48-96)
 * with generate_ray (src/core/camera.cu:91-98). rays: w*h*spp x 8 floats. */
void orc_raygen_camera(const float* r2c, const float* c2w, uint32_t width, uint32_t x0, uint32_t y0,
                       uint32_t w, uint32_t h, uint32_t spp, uint64_t seed, int jitter, float* rays);
/* Diffuse rays from primary hits (trace benchmark ray set ii). */
void orc_raygen_diffuse(const float* verts, const uint32_t* idx, const float* prim, const int32_t* hit_id,
                        const float* hit_t, uint32_t np, uint32_t nout, uint64_t seed, float* out);


/* ---- render passes ------------------------------------------------------ */

/* Film of a tile from per-sample radiance Ls[w*h*spp*3] (sample = pixel*spp
 * + s): filter 0 box / 1 triangle / 2 Gaussian with half-widths fxw, fyw and
 * Gaussian falloff falpha (src/core/filter.h:69-132, film.cpp:45-86,
 * samplerrenderer.cu:103-169).  Sample positions are the camera rays' jitter
 * (Philox, stream 1).  The render passes below end with it. */
void orc_film(const float* Ls, uint32_t width, uint32_t x0, uint32_t y0, uint32_t w, uint32_t h, uint32_t spp,
              uint64_t seed, uint32_t filter, float fxw, float fyw, float falpha, float* out_rgb);
/* AO (src/integrators/integrator_ao.cu:42-85) over the tile, filtered film
 * of spp samples per pixel into out_rgb[w*h*3]. */
void orc_render_ao(const uint32_t* nodes, const uint32_t* woop, const uint32_t* tidx, const float* verts,
                   const uint32_t* idx, const float* r2c, const float* c2w, uint32_t width, uint32_t height,
                   uint32_t x0, uint32_t y0, uint32_t w, uint32_t h, uint32_t spp, uint64_t seed, uint32_t nao,
                   float radius, uint32_t filter, float fxw, float fyw, float falpha, float* out_rgb,
                   const float* tri_n, const uint8_t* has_n);
/* Path tracer (src/integrators/integrator_path.cu:43-261): NEE with MIS,
 * Lambert / mirror / emitter materials (mats4: kind bits + rgb per material). */
void orc_render_pt(const uint32_t* nodes, const uint32_t* woop, const uint32_t* tidx, const float* verts,
                   const uint32_t* idx, const uint32_t* tri_mat, const float* mats4, uint32_t nmats,
                   const uint32_t* lights, uint32_t nl, const float* r2c, const float* c2w, uint32_t width,
                   uint32_t x0, uint32_t y0, uint32_t w, uint32_t h, uint32_t spp, uint64_t seed,
                   uint32_t max_depth, uint32_t rr_start, float rr_prob, uint32_t filter, float fxw, float fyw,
                   float falpha, float* out_rgb, const float* tri_n, const uint8_t* has_n);
/* x^y from log2 / exp2 series (no math library): the renderer's pow. */
float orc_powf(float x, float y);
/* BSDF of one material (m12: kind bits, rgb, param[8] as rb_material) for
 * wo -> wi about the geometric normal ng: out = f.rgb, pdf. */
void orc_bsdf_eval(const float* m12, const float* wo, const float* wi, const float* ng, float* out);
/* One BSDF sample from (u1, u2): out = wi.xyz, weight or f .rgb, pdf. */
void orc_bsdf_sample(const float* m12, const float* wo, const float* ng, float u1, float u2, float* out);
/* Decision log of the SBVH build (checker support).  After
 * orc_build_log_start(), every node that orc_bvh_build on this thread
 * visits is recorded: its level, bounds and references (after degenerate
 * removal: tri id bits + box, 7 floats each, from ref_off), the leaf /
 * object / spatial SAH it computed (FLT_MAX where not computed; the spatial
 * one only when sp_tried), and what it did (kind 0 leaf, 1 object split,
 * 2 spatial split) with the children's reference counts. */
typedef struct {
    int32_t level, kind;
    uint32_t num;
    float bounds[6];
    float leaf_sah, obj_sah, sp_sah;
    int32_t obj_dim, sp_dim, sp_tried;
    uint32_t obj_left;
    float sp_pos;
    uint32_t ln, rn;
    uint64_t ref_off;
} orc_node_log;
void orc_build_log_start(void);
void orc_build_log_stop(void);
size_t orc_build_log_nodes(orc_node_log* out);
size_t orc_build_log_refs(float* out7);
/* Number of BDPT techniques able to build a path z_1..z_n (mask bit i-1 =
 * z_i specular) when camera vertices exist for bounces 0..max_depth and the
 * light-vertex cache holds depths 0..lvc_depth-1; a path's weight is 1/that. */
uint32_t orc_bdpt_techniques(uint32_t mask, uint32_t n, uint32_t max_depth, uint32_t lvc_depth);
/* BDPT with a light-vertex cache (src/integrators/integrator_bdpt_lvc_sk.cu:47-435):
 * light_paths light sub-paths of up to light_max_depth cached vertices, then
 * the camera pass of the path tracer plus `connections` cache connections
 * per diffuse camera vertex. */
void orc_render_bdpt(const uint32_t* nodes, const uint32_t* woop, const uint32_t* tidx, const float* verts,
                     const uint32_t* idx, const uint32_t* tri_mat, const float* mats4, uint32_t nmats,
                     const uint32_t* lights, uint32_t nl, const float* r2c, const float* c2w, uint32_t width,
                     uint32_t x0, uint32_t y0, uint32_t w, uint32_t h, uint32_t spp, uint64_t seed,
                     uint32_t max_depth, uint32_t rr_start, float rr_prob, uint32_t light_paths,
                     uint32_t light_max_depth, uint32_t connections, uint32_t filter, float fxw, float fyw,
                     float falpha, float* out_rgb, const float* tri_n, const uint8_t* has_n);
/* Ray-cast debug integrator (src/integrators/integrator_raycast.cu:37-160):
 * closest hit of the camera rays coloured by shade mode 0 primitive id
 * (15-colour palette, id mod 15), 1 material colour, 2 shading normal (the
 * geometric normal for a triangle without normals), 3 barycentrics (u, v,
 * 0), 4 uniform white; black on a miss; then the film.
 * tri_n / has_n (every render entry point): per-triangle vertex normals (9
 * floats) and whether each triangle has them, or NULL for none. */
void orc_render_raycast(const uint32_t* nodes, const uint32_t* woop, const uint32_t* tidx, const float* verts,
                        const uint32_t* idx, const uint32_t* tri_mat, const float* m12, uint32_t nmats,
                        const float* r2c, const float* c2w, uint32_t width, uint32_t x0, uint32_t y0, uint32_t w,
                        uint32_t h, uint32_t spp, uint64_t seed, uint32_t shade, uint32_t filter, float fxw,
                        float fyw, float falpha, float* out_rgb, const float* tri_n, const uint8_t* has_n);
/* The IntersectionBuffer fields kernel_kepler_trace writes for a hit
 * (kepler_dynamic_fetch.h:312-360) from (id, t) of ray i: position o + t d,
 * geometric normal, shading normal (interpolated, or the geometric one),
 * incoming direction -d, barycentric (u, v) of the position, material id,
 * front (geometric normal against d), epsilon 1e-3 t.  out, per ray:
 * [0..2] position, [3..5] Ng, [6..8] Ns, [9..11] -d, [12..13] uv,
 * [14] epsilon, [15] t; mid_front[2 i] material id, [2 i + 1] front;
 * misses leave their entries untouched. */
void orc_fill(const float* rays, const int32_t* ids, const float* ts, uint32_t n, const float* verts,
              const uint32_t* idx, const uint32_t* tri_mat, const float* tri_n, const uint8_t* has_n, float* out,
              uint32_t* mid_front);

#ifdef __cplusplus
}
#endif
#endif

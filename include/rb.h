/*
 * rb.h -- C-ABI of the MI355X ray-tracing core (librb.so).
 *
 * The drop-in boundary for renderbox2's hot path.  Plain pointers and sizes;
 * device pointers are HIP device allocations; `stream` arguments are
 * hipStream_t values passed as void* (NULL = the library's default stream).
 * Every call returns an rb_status; rb_last_error() gives the message of the
 * last failing call on the calling thread.
 *
 * Reference interfaces replaced (paths relative to renderbox2/ in the
 * reference tree) are cited on each declaration.
 */
#ifndef RB_H
#define RB_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status ------------------------------------------------------------ */

/* Replaces checkCuda's print-and-assert (src/core/defs.h:55-63) and the
 * constant 0.0f of CudaTracer::trace (src/accelerators/sbvh/cudatracer.cu:182). */
typedef enum rb_status {
    RB_OK = 0,
    RB_ERR_INVALID = 1,     /* bad argument / handle / size */
    RB_ERR_OOM = 2,         /* device or host allocation failed, or over budget */
    RB_ERR_DEVICE = 3,      /* HIP runtime error */
    RB_ERR_BUSY = 4,        /* resource in use (e.g. scene with queued requests) */
    RB_ERR_UNSUPPORTED = 5, /* feature not built / not available */
    RB_ERR_NOT_READY = 6    /* asynchronous result not complete yet */
} rb_status;

size_t rb_last_error(char* buf, size_t n);
const char* rb_status_string(rb_status s);
const char* rb_version(void);

/* ---- records ----------------------------------------------------------- */

/* Ray, src/core/primitives.h:145-208: float4(o.xyz, tmin), float4(d.xyz, tmax). */
typedef struct rb_ray {
    float ox, oy, oz, tmin;
    float dx, dy, dz, tmax;
} rb_ray;

/* Lean trace result: the (triangle id, t) pair the reference's
 * STORE_RESULT writes (src/accelerators/sbvh/cudatracerkernels.h:180).
 * id = original triangle index, -1 on a miss (t = tmax then). */
typedef struct rb_hit {
    int32_t id;
    float t;
} rb_hit;

/* Intersection record, the fields of IntersectionBuffer
 * (src/core/intersection.h:63-80) that the tracer writes, plus t.  SoA,
 * device pointers, each of n entries; NULL fields are skipped.  For a hit:
 * position o + t d; geometric normal unit cross(v1 - v0, v2 - v0); shading
 * normal the triangle's vertex normals weighted by the position's
 * barycentrics and normalised (the geometric normal for a triangle without
 * normals); incoming direction -d; uv the barycentric weights of v0 and v1
 * at the position (p = u v0 + v v1 + (1 - u - v) v2); material id; front =
 * geometric normal . d < 0; epsilon 1e-3 t.  Misses write intersected = 0
 * and t = tmax and leave the other fields as they were. */
typedef struct rb_isect_soa {
    float* position[3];
    float* geometric_normal[3];
    float* shading_normal[3];
    float* incoming_direction[3];
    float* uv;                /* 2 floats per entry */
    uint32_t* primitive_id;
    uint32_t* material_id;
    uint8_t* front;
    uint8_t* intersected;
    float* epsilon;
    float* t;
} rb_isect_soa;

/* ---- device / memory / streams ------------------------------------------ */

/* MemoryAllocator (src/memory/memoryallocator.h:82-86) equivalent plus the
 * HBM budget that admission control reads. */
rb_status rb_device_init(int device);
rb_status rb_device_name(char* buf, size_t n, uint32_t* num_cus);
rb_status rb_malloc(void** ptr, size_t bytes);
rb_status rb_free(void* ptr);
rb_status rb_memcpy_htod(void* dst, const void* src, size_t bytes, void* stream);
rb_status rb_memcpy_dtoh(void* dst, const void* src, size_t bytes, void* stream);
rb_status rb_memcpy_dtod(void* dst, const void* src, size_t bytes, void* stream);
rb_status rb_memset(void* dst, int value, size_t bytes, void* stream);
/* Device-to-device copy as a plain streaming kernel (16-byte loads/stores):
 * usable inside captured graphs, and the known-size reference kernel for
 * HBM traffic counters. bytes must be a multiple of 16 and pointers 16-aligned. */
rb_status rb_copy_kernel(void* dst, const void* src, size_t bytes, void* stream);
rb_status rb_mem_info(size_t* free_bytes, size_t* total_bytes, size_t* library_bytes);
/* Self-test, no counterpart in the reference: runs every one of the 2^32
 * single-precision bit patterns through the traversal's reciprocal (the
 * hardware reciprocal refined by one fused Newton step, taken when the
 * refined value is normal, else the division) on the current device.
 * *fast_inputs = inputs that take the refined value (4,227,858,434 on
 * gfx950), *mismatches = inputs whose reciprocal differs from the IEEE
 * division 1.0f / x (must be 0: the traversal's bit-exact results rest on
 * it).  Synchronous; tens of milliseconds of device time. */
rb_status rb_selftest_rcp(unsigned long long* fast_inputs, unsigned long long* mismatches);
rb_status rb_stream_create(void** stream);
rb_status rb_stream_destroy(void* stream);
rb_status rb_stream_synchronize(void* stream);
rb_status rb_event_create(void** ev);
rb_status rb_event_destroy(void* ev);
rb_status rb_event_record(void* ev, void* stream);
rb_status rb_event_elapsed_ms(void* start, void* stop, float* ms);
rb_status rb_event_synchronize(void* ev);

/* ---- host scene (Scene, src/core/scene.h:55-199) --------------------------- */

typedef struct rb_scene_s* rb_scene_t;

/* Material: one BSDF layer of Material + *BsdfParams (src/core/material.h,
 * src/core/params.h:150-240) or a diffuse emitter.  `color` is the albedo /
 * reflectance / transmittance tint, or the emitted radiance.  param[]:
 *   RB_MAT_GLASS      [0] eta outside (etai), [1] eta inside (etat)
 *   RB_MAT_ORENNAYAR  [0] sigma in degrees (clamped to [0, 90]); [1] A and
 *                     [2] B are filled in by rb_scene_add_material
 *   RB_MAT_MICROFACET [0] Blinn exponent, [1] Fresnel: 0 none, 1 conductor,
 *                     2 dielectric; conductor: [2..4] eta rgb, [5..7] k rgb;
 *                     dielectric: [2] etai, [3] etat
 *   RB_MAT_LAYERED    [0] number of layers n (1..4), [1..n] ids of single-layer
 *                     materials (not emitters) already in the scene; `color` unused.
 *                     A Material of several BSDF layers (src/core/material.h:44-52):
 *                     the equal mixture f = (1/n) sum f_i, sampled by picking a
 *                     layer uniformly (material.cu:80-95); see rb_scene_add_layered_material
 * Mirror and glass are specular (delta) BSDFs. */
enum { RB_MAT_LAMBERT = 0, RB_MAT_MIRROR = 1, RB_MAT_EMITTER = 2, RB_MAT_GLASS = 3, RB_MAT_ORENNAYAR = 4,
       RB_MAT_MICROFACET = 5, RB_MAT_LAYERED = 6 };
enum { RB_FRESNEL_NONE = 0, RB_FRESNEL_CONDUCTOR = 1, RB_FRESNEL_DIELECTRIC = 2 };
typedef struct rb_material {
    uint32_t kind;
    float color[3];
    float param[8];
} rb_material;

rb_status rb_scene_create(rb_scene_t* out);
rb_status rb_scene_destroy(rb_scene_t s);
/* Scene::add_material (src/core/scene.h:133) */
rb_status rb_scene_add_material(rb_scene_t s, const rb_material* m, uint32_t* id);
/* A layered material (Material with num_valid_layers = n, src/core/material.h:44-52,
 * read from the scene file's <layers>, customfilereader.cpp:86-225) over n (1..4)
 * single-layer, non-emitting materials of the scene.  It is the equal mixture of
 * its layers: f = (1/n) sum f_i; a sampled direction comes from one layer picked
 * uniformly, with pdf (1/n) sum pdf_i over the non-specular layers (a specular
 * layer's sample keeps its own weight); light samples and connections evaluate
 * the non-specular layers.  The reference's material_sample_f returns sum f_i
 * with the undivided sum of pdfs (the mixture's f / pdf ratio), while its
 * material_f returns sum f_i and material_pdf (1/n) sum pdf_i -- its light
 * samples and BSDF samples estimate different materials; here all three
 * describe the mixture. */
rb_status rb_scene_add_layered_material(rb_scene_t s, const uint32_t* layer_ids, uint32_t n, uint32_t* id);
/* Scene::add_mesh (src/core/scene.cpp:83-123): triangles index the mesh's
 * own vertices; the scene offsets them.  Emissive meshes register each
 * triangle as a DiffuseAreaLight (src/core/scene.cpp:100-118). */
rb_status rb_scene_add_mesh(rb_scene_t s, const float* verts, uint32_t nv, const uint32_t* idx,
                            uint32_t nt, uint32_t material_id);
/* Scene::add_mesh with a Mesh's shading data (src/core/scene.cpp:83-123;
 * Scene h_normals / h_texcoords / h_nor_indices / h_tex_indices,
 * src/core/scene.h:142-147; SceneBuffer, src/core/buffer.h:322-335):
 * nn normals (xyz) and ntc texture coordinates (uv) with their own index
 * triples per triangle (local to the mesh, as idx).  normals (texcoords)
 * may be NULL: no triangle of the mesh has them; nidx (tidx) may be NULL
 * with nn (ntc) == nv: one per vertex, indexed like the vertices.  A
 * triangle whose triple is 0xFFFFFFFF x 3 has none.  Triangles with normals
 * shade with the interpolated normal (kepler_dynamic_fetch.h:327-334),
 * the others with the geometric one; texture coordinates are kept with the
 * scene (no pass of the reference reads them). */
rb_status rb_scene_add_mesh_ex(rb_scene_t s, const float* verts, uint32_t nv, const uint32_t* idx, uint32_t nt,
                               const float* normals, uint32_t nn, const uint32_t* nidx, const float* texcoords,
                               uint32_t ntc, const uint32_t* tidx, uint32_t material_id);
rb_status rb_scene_counts(rb_scene_t s, uint32_t* nv, uint32_t* nt, uint32_t* nlights,
                          uint32_t* nmaterials);
rb_status rb_scene_geometry(rb_scene_t s, float* verts, uint32_t* idx, uint32_t* tri_material,
                            uint32_t* light_tris);
/* The scene's shading data: counts of normals / texcoords and whether any
 * triangle has them (then nidx / tidx hold 3 entries per triangle, scene
 * indices or 0xFFFFFFFF; else those arrays are empty and not written). */
rb_status rb_scene_shading_counts(rb_scene_t s, uint32_t* nnormals, uint32_t* ntexcoords, int* has_nidx,
                                  int* has_tidx);
rb_status rb_scene_shading(rb_scene_t s, float* normals, uint32_t* nidx, float* texcoords, uint32_t* tidx);
rb_status rb_scene_materials(rb_scene_t s, rb_material* out);
rb_status rb_scene_bounds(rb_scene_t s, float lo[3], float hi[3]);

/* Seeded synthetic scenes (the configurations the benchmarks run).  No
 * reference counterpart: the reference reads OBJ/XML scenes
 * (src/io/customfilereader.cpp:47-309). */
enum {
    RB_GEN_ICOSPHERES = 0, /* count spheres, subdivision `subdiv`, radius U[r0,r1], centres U[c0,c1]^3 */
    RB_GEN_THIN_QUADS = 1, /* count quads, aspect 1:50, random orientation, centres U[c0,c1]^3 */
    RB_GEN_ROOM = 2,       /* closed box [c0,c1]^3 of 2048 triangles facing inward */
    RB_GEN_EMITTER_QUADS = 3 /* count axis-aligned emissive quads of area `r0` near the ceiling */
};
typedef struct rb_gen_params {
    uint32_t kind;
    uint32_t count;
    uint32_t subdiv;
    uint64_t seed;
    float r0, r1;   /* radius range, quad length range, or emitter area */
    float c0, c1;   /* placement range */
    int32_t material_id;       /* >= 0: every triangle gets it */
    uint32_t mix_materials;    /* 1: per-object 70% Lambert U[0.2,0.8] / 30% mirror */
    float emit[3];             /* emitter radiance for RB_GEN_EMITTER_QUADS */
    uint32_t subdiv_hi;        /* >subdiv: a seeded share `hi_count` of spheres uses it */
    uint32_t hi_count;
    uint32_t smooth;           /* icospheres: 1 = with their analytic vertex normals */
} rb_gen_params;
rb_status rb_scene_generate(rb_scene_t s, const rb_gen_params* p);

/* ---- SBVH (BVH / SplitBVHBuilder / CudaBvh) --------------------------------- */

typedef struct rb_host_bvh_s* rb_host_bvh_t;
typedef struct rb_bvh_s* rb_bvh_t;

/* Platform + BVH::BuildParams (src/accelerators/sbvh/platform.h:78-135,
 * src/accelerators/sbvh/bvh.h:100-113). Defaults: min_leaf 1,
 * max_leaf 0x7FFFFFF, split_alpha 1e-5, node/tri cost 1/1. threads 0 = all. */
typedef struct rb_build_params {
    uint32_t min_leaf;
    uint32_t max_leaf;
    float split_alpha;
    float node_cost;
    float tri_cost;
    uint32_t threads;
} rb_build_params;
void rb_build_params_default(rb_build_params* p);

/* BVH::Stats (src/accelerators/sbvh/bvh.h:78-98). */
typedef struct rb_bvh_stats {
    float sah;
    uint32_t n_inner, n_leaf, n_tri_refs, n_child_nodes, max_depth;
    uint64_t duplicates;
    double build_ms;
} rb_bvh_stats;

/* BVH::BVH(Scene*, Platform, BuildParams) (src/accelerators/sbvh/bvh.h:115,
 * src/accelerators/sbvh/splitbvhbuilder.cpp:177-214). */
rb_status rb_bvh_build_host(rb_scene_t s, const rb_build_params* p, rb_host_bvh_t* out);
rb_status rb_bvh_build_arrays(const float* verts, uint32_t nv, const uint32_t* idx, uint32_t nt,
                              const rb_build_params* p, rb_host_bvh_t* out);
rb_status rb_bvh_host_stats(rb_host_bvh_t b, rb_bvh_stats* out);
rb_status rb_bvh_host_release(rb_host_bvh_t b);

/* CudaBvh(const BVH&, BVHLayout) + get_*_buffer
 * (src/accelerators/sbvh/cudabvh.h:80-97).  REF_COMPACT2 is byte-compatible
 * with the reference's Compact2 (uint4 rows / uint4 Woop rows with
 * terminators / uint32 index); MI355X is this library's traversal layout.
 * Call with NULL buffers to get the sizes in bytes. */
enum { RB_LAYOUT_REF_COMPACT2 = 1, RB_LAYOUT_MI355X = 2,
       RB_LAYOUT_MI355X_WIDE = 3 /* no counterpart: the MI355X triangles and index with the
                                    wide (4-child, 64-byte) records the upload builds as nodes;
                                    RB_ERR_UNSUPPORTED where the upload would keep the binary
                                    walk only */ };
rb_status rb_bvh_export(rb_host_bvh_t b, uint32_t layout, void* nodes, size_t* nodes_bytes,
                        void* tris, size_t* tris_bytes, void* index, size_t* index_bytes);

/* CudaBvh::serialize / CudaBvh(filename) (src/accelerators/sbvh/cudabvh.cpp:108-214):
 * the reference's text format, read and written for compatibility. */
rb_status rb_bvh_save_text(rb_host_bvh_t b, const char* path);
rb_status rb_bvh_load_text_compact2(const char* path, void* nodes, size_t* nodes_bytes,
                                    void* tris, size_t* tris_bytes, void* index,
                                    size_t* index_bytes);

/* Binary BVH cache (the reference's cache is the decimal text form above,
 * used or rebuilt by a config flag with nothing tying it to the scene,
 * src/core/application.cpp:131-161).  The file holds the host tree (every
 * layout exports from it) and a key hashed from the scene's vertices and
 * triangles and the tree-shaping build parameters (not `threads`): loading
 * it for other geometry or parameters fails with RB_ERR_INVALID, as does a
 * corrupt or truncated file (checksum).  rb_bvh_build_cached loads
 * <cache_dir>/rbvh_<key>.bin if present and valid, else builds and writes it
 * (atomically: temporary file + rename); *from_cache tells which. */
rb_status rb_bvh_save_binary(rb_host_bvh_t b, const char* path);
rb_status rb_bvh_load_binary(const char* path, rb_scene_t scene, const rb_build_params* p, rb_host_bvh_t* out);
rb_status rb_bvh_build_cached(rb_scene_t scene, const rb_build_params* p, const char* cache_dir, rb_host_bvh_t* out,
                              int* from_cache);

/* CudaTracer(const CudaBvh&, MemoryAllocator&) (src/accelerators/sbvh/cudatracer.cu:43-127)
 * plus Scene::gpu_prepare (src/core/scene.cpp:166-301): uploads the MI355X
 * layout and the scene's shading data (vertices, materials, lights).
 * `scene` may be NULL for trace-only use. */
rb_status rb_bvh_upload(rb_host_bvh_t b, rb_scene_t scene, rb_bvh_t* out);
rb_status rb_bvh_release(rb_bvh_t b);
rb_status rb_bvh_device_bytes(rb_bvh_t b, size_t* bytes);
/* Device view of a resident BVH for kernels that trace in their own code
 * through rb_trace_one() (include/rb_trace_device.h); the reference hands
 * its device tracer the same three buffers as texture objects
 * (CudaTracer::get_bvh / BvhStruct, src/accelerators/sbvh/cudatracer.h:44-55,89-101). */
typedef struct rb_bvh_dev {
    const void* nodes;      /* float4[4 * n_nodes]: MI355X node records */
    const void* tris;       /* float4[3 * n_refs]: Woop rows */
    const uint32_t* index;  /* uint32[n_refs]: original triangle id per reference */
    uint32_t n_nodes, n_refs;
    uint32_t* errors;       /* device word for RB_TRACE_ERR_* flags (rb_trace_errors); may be NULL */
    const void* wnodes;     /* float4[4 * n_wnodes]: wide (4-child) records, which rb_trace_one
                               walks for closest hits (the same (id, t)); NULL: the binary
                               records only.  Set when rb_trace_config.wide is on */
    uint32_t n_wnodes;
    uint32_t n_tris;        /* triangles of the scene the BVH was uploaded with (0: none) */
    const void* tri_v;      /* float4[3 * n_tris]: v0.xyz|material id bits, v1.xyz|flags bits
                               (bit 0: the triangle has vertex normals), v2.xyz|0; NULL
                               without a scene.  Read by rb_fill_one */
    const void* tri_n;      /* float4[3 * n_tris]: vertex normals n0, n1, n2 of the triangles
                               that have them; NULL when no triangle has any */
} rb_bvh_dev;
rb_status rb_bvh_device_view(rb_bvh_t b, rb_bvh_dev* out);

/* ---- trace ---------------------------------------------------------------- */

enum {
    RB_CLOSEST = 0,
    RB_ANY = 1,       /* terminate on the first hit (occlusion) */
    RB_MASKED = 2,    /* mask[i] == 0 skips ray i (writes a miss) */
    RB_WIDE_COUNTS = 4 /* rb_trace_counts only: count the wide (4-child) walk a
                          closest-hit trace takes by default instead of the
                          binary one (no counterpart in the reference) */
};

/* CudaTracer::trace(rb, isects, scene, stencil, any_hit)
 * (src/accelerators/sbvh/cudatracer.h:87, cudatracer.cu:134-183), lean
 * output form.  rays/hits/mask are device pointers; asynchronous on
 * `stream`.  n == 0 is a no-op. */
rb_status rb_trace(rb_bvh_t bvh, const rb_ray* rays, uint32_t n, uint32_t flags,
                   const uint8_t* mask, rb_hit* hits, void* stream);

/* The same call writing the reference IntersectionBuffer fields (+ t), the
 * result-write block of kernel_kepler_trace (kepler_dynamic_fetch.h:312-360).
 * Needs a bvh uploaded with its scene. `hits` may be NULL (scratch used). */
rb_status rb_trace_fill(rb_bvh_t bvh, const rb_ray* rays, uint32_t n, uint32_t flags,
                        const uint8_t* mask, rb_hit* hits, const rb_isect_soa* out,
                        void* stream);

/* rb_trace_fill without any device allocation.  No counterpart in the
 * reference: it is for callers that cannot allocate (code under HIP graph
 * capture, serving loops, hosts that forbid hipMalloc).  The (id, t) staging
 * and the trace's work counter live in a static device array of the library
 * (8 MiB + 1,280 bytes per device, held until the process ends), so at most
 * 1,048,576 rays per call (more: RB_ERR_INVALID).  rb_device_init must have
 * run (RB_ERR_INVALID otherwise).  Every call on a device shares that
 * array: calls that overlap on one device (different streams or threads)
 * race and must be serialised by the caller.  Same results as rb_trace_fill. */
rb_status rb_trace_fill_noalloc(rb_bvh_t bvh, const rb_ray* rays, uint32_t n, uint32_t flags,
                                const uint8_t* mask, const rb_isect_soa* out, void* stream);

/* One ray per thread through the device function rb_trace_one(): the
 * reference's kernel_trace_ray form (cuda_tracer_functions.cu:42-279)
 * rather than the persistent kernel.  Same results as rb_trace, bit for
 * bit; kept for in-kernel callers' reference and comparison. */
rb_status rb_trace_simple(rb_bvh_t bvh, const rb_ray* rays, uint32_t n, uint32_t flags, rb_hit* hits, void* stream);

/* One ray per thread through rb_trace_one_isect(), the reference's
 * trace_ray with its Intersection argument (cuda_tracer_functions.cu:236-276):
 * each thread fills its ray's intersection record and writes the fields
 * `out` names.  Needs a bvh uploaded with its scene.  Same records as
 * rb_trace_fill, bit for bit (closest hit; with RB_ANY only `intersected`
 * and `t` are written, as the reference fills nothing else for any hit). */
rb_status rb_trace_simple_fill(rb_bvh_t bvh, const rb_ray* rays, uint32_t n, uint32_t flags,
                               const rb_isect_soa* out, void* stream);

/* Batched trace (no reference counterpart: the reference issues one
 * synchronous launch per call, cudatracer.cu:134-183).  Ray batches of
 * uneven size against several resident BVHs go through one persistent
 * launch per 64 batches; each batch keeps its own rays/hits buffers and
 * closest/any flag (masks are not supported here).  Results equal those of
 * one rb_trace call per batch. */
typedef struct rb_batch_desc {
    rb_bvh_t bvh;
    const rb_ray* rays;
    rb_hit* hits;
    uint32_t n;
    uint32_t flags;
} rb_batch_desc;
rb_status rb_trace_batch(const rb_batch_desc* batches, uint32_t nb, void* stream);

/* Traversal tuning: the knobs the autotuner searches (the reference
 * hard-codes STACK_SIZE 64 / DYNAMIC_FETCH_THRESHOLD 20 / 720 warps of
 * (32,4), kepler_dynamic_fetch.h:62-63 and cudatracer.cu:170-171). */
typedef struct rb_trace_config {
    uint32_t block;          /* threads per block: 64, 128, 256, 512, 1024 */
    uint32_t lds_stack;      /* stack entries per lane kept in LDS: 8..24 (8..16 at
                                block 512, 8..12 at 1024); stack plus cached nodes
                                must fit the device's LDS per block */
    uint32_t chunk;          /* rays claimed per work-counter atomic (multiple of 64) */
    uint32_t fetch_threshold;/* refill when fewer active lanes than this (0..64) */
    uint32_t waves_per_cu;   /* persistent grid size = CUs x this / (block/64) */
    uint32_t speculative;    /* 0 (default): each ray tests its leaves in its own
                                traversal order, results are reproducible and equal
                                the single-ray restatement; 1: the reference's
                                speculative while-while (kepler_dynamic_fetch.h:222-243),
                                whose leaf order depends on the other rays of the
                                wave -- in rare boundary cases ids and t may differ */
    uint32_t leaf_gate;      /* non-speculative order: a wave leaves its inner-node phase
                                once at most this many lanes still search for a leaf
                                (0..63; 0 = when all have one).  Any value gives the
                                same results: only the interleaving of lanes changes */
    uint32_t lds_nodes;      /* top-of-tree node records each block copies into LDS at
                                start and reads from there (0..512, multiple of 64; 0 =
                                none).  The MI355X layout stores the first 512 records
                                breadth-first, so these are the top levels of the tree.
                                Any value gives the same results */
    uint32_t wide;           /* 1 (default): closest-hit traces without a visit-count
                                buffer walk 4-child records (each binary record's
                                children, inner ones replaced by their own two; 64
                                bytes, boxes in 8-bit steps that hold the exact ones),
                                about half the dependent fetches per ray; lds_nodes
                                such records go to LDS.  The same (id, t) as the binary
                                walk, whose order any-hit traces and visit counts keep;
                                0: the binary records for every trace */
} rb_trace_config;
rb_status rb_trace_get_config(rb_trace_config* c);
rb_status rb_trace_set_config(const rb_trace_config* c);
/* Autotuner (no reference counterpart; the reference hard-codes the launch
 * shape): times candidate configurations on the caller's device rays by
 * coordinate descent over block, lds_stack, chunk, fetch_threshold,
 * waves_per_cu and leaf_gate (speculative is kept as set), installs the
 * fastest and returns it with its time per launch. */
rb_status rb_trace_autotune(rb_bvh_t bvh, const rb_ray* rays, uint32_t n, uint32_t flags, rb_hit* hits,
                            uint32_t reps, rb_trace_config* best, float* best_ms, void* stream);
/* Traversal error flags raised on the device since the last call (the
 * reference has no error path: its 64-entry local stack is written past,
 * kepler_dynamic_fetch.h:62-71,199-216).  RB_TRACE_ERR_STACK: a ray needed
 * more stack entries than the traversal holds -- unreachable for a BVH that
 * rb_bvh_upload accepted (it refuses trees that deep), set only if the
 * device data was corrupted or a hand-built rb_bvh_dev was traced.  The
 * pushes past the stack are dropped, so results of the launches since the
 * last call are not trustworthy.  Waits for the device, returns the flags
 * in *flags (0 = none), clears them, and returns RB_ERR_INVALID when any
 * was set. */
enum { RB_TRACE_ERR_STACK = 1 };
rb_status rb_trace_errors(uint32_t* flags);
/* Per-ray traversal counters of the device kernel (inner nodes, triangles,
 * leaves) for traffic accounting; counts is device uint32[3*n].  The
 * counts follow the binary walk (the reference's visit order, whatever
 * rb_trace_config.wide says); with RB_WIDE_COUNTS (closest hit, not
 * speculative) they follow the wide walk, inner = 4-child records visited
 * (RB_ERR_UNSUPPORTED when the BVH has no wide records). */
rb_status rb_trace_counts(rb_bvh_t bvh, const rb_ray* rays, uint32_t n, uint32_t flags,
                          rb_hit* hits, uint32_t* counts, void* stream);

/* ---- ray generation ---------------------------------------------------- */

/* PerspectiveCamera (src/core/camera.h:48-71, camera.cu:30-98). */
typedef struct rb_camera {
    float position[3];
    float lookat[3];
    float up[3];
    float hfov_deg;
    uint32_t width, height;
} rb_camera;

/* The two matrices PerspectiveCamera::generate_ray applies
 * (t_raster_to_camera, t_camera_to_world; src/core/camera.cu:30-57,91-98),
 * row-major 4x4, computed on the host with the reference's transform chain. */
rb_status rb_camera_matrices(const rb_camera* cam, float raster_to_camera[16], float camera_to_world[16]);

/* kernel_generate_perspective_camera_samples (src/renderers/samplerrenderer.cu:48-96):
 * spp rays per pixel of the tile [x0,x0+w) x [y0,y0+h), jittered with the
 * Philox stream (seed, pixel, sample); jitter = 0 puts rays at pixel centres.
 * rays: device rb_ray[w*h*spp], ordered (pixel-major, sample-minor). */
rb_status rb_raygen_camera(const rb_camera* cam, uint32_t x0, uint32_t y0, uint32_t w, uint32_t h,
                           uint32_t spp, uint64_t seed, int jitter, rb_ray* rays, void* stream);

/* ---- scene files ------------------------------------------------------- */

/* Wavefront OBJ (ObjFileReader::read_file, src/io/objfilereader.cpp:117-217):
 * every object of the file is added with `material_id`.  Vertices and
 * triangles keep file order and file-global indexing (Scene::add_mesh,
 * src/core/scene.cpp:83-113); polygons are fanned into triangles, negative
 * indices resolved, `#` starts a comment anywhere on a line.  `vn` normals
 * and `vt` texture coordinates are kept with their corner indices (v, v/vt,
 * v//vn, v/vt/vn); a triangle has normals (texcoords) when all three of its
 * corners name one (rb_scene_add_mesh_ex). */
rb_status rb_scene_load_obj(rb_scene_t s, const char* path, uint32_t material_id);
/* The reference's XML scene file (CustomFileReader::process and get_camera,
 * src/io/customfilereader.cpp:47-340): its mesh file (path relative to the
 * XML file), materials (lambertian, mirror, glass, orennayar and Blinn
 * microfacet <bxdf> layers; a <bsdf> of 2..4 layers becomes an RB_MAT_LAYERED
 * material whose layer materials are numbered after the emitters, so the named
 * materials and emitters keep the reference's numbering; a <layers> count that
 * differs from the <bxdf> count is RB_ERR_INVALID), object -> material
 * assignments, diffuse lights and, when `cam` is not NULL, the camera for a
 * width x height film (fov from <fov>, 90 degrees when absent). */
rb_status rb_scene_load_xml(rb_scene_t s, const char* path, uint32_t width, uint32_t height, rb_camera* cam);

/* Writes a host RGB float image (w*h*3, row-major, top row first) by suffix
 * (ImageWriter::write, src/io/imagewriter.cpp:34-105): .exr (OpenEXR,
 * uncompressed half B/G/R), .pfm (float32), .jpg / .jpeg (baseline JPEG,
 * quality 100 as the reference writes it), .ppm / .png (8-bit: clamp to
 * [0,1], gamma 1/2.2, round, as ImageWriter::to_byte). */
rb_status rb_image_write(const char* path, const float* rgb, uint32_t w, uint32_t h);

/* Incoherent diffuse rays from primary hits: ray i starts at the hit of
 * primary i (or a Philox-redrawn hit primary when i missed), with a
 * cosine-distributed direction about the geometric normal facing the
 * incoming ray and tmin = 1e-3 t.  Needs the bvh's scene. */
rb_status rb_raygen_diffuse(rb_bvh_t bvh, const rb_ray* primaries, const rb_hit* hits,
                            uint32_t n_primary, uint32_t n_out, uint64_t seed, rb_ray* out,
                            void* stream);

/* ---- render passes ------------------------------------------------------- */

/* Parameters of AmbientOcclusionIntegratorParams / PathIntegratorParams /
 * BdptLvcIntegratorParams (src/integrators/integrator_*.h, defaults in the
 * integrator XML files) plus the sampler's spp and the RNG seed. */
typedef struct rb_render_params {
    uint32_t spp;
    uint64_t seed;
    uint32_t ao_samples;   /* AO rays per primary hit */
    float ao_radius;       /* AO ray tmax */
    uint32_t max_depth;    /* last bounce index that is shaded */
    uint32_t rr_start;     /* Russian roulette once bounce > rr_start */
    float rr_prob;         /* roulette continue probability cap */
    uint32_t light_paths;  /* BDPT: light sub-paths traced per frame */
    uint32_t lvc_capacity; /* BDPT: light-vertex cache size */
    uint32_t connections;  /* BDPT: LVC connections per camera vertex */
    uint32_t light_max_depth; /* BDPT: light sub-path depth */
    unsigned long long* ray_counter; /* optional device counter: += rays traced by the call */
    /* Film reconstruction filter (Filter*, src/core/filter.h:69-132; the
     * renderer's <filter> config): RB_FILTER_BOX / _TRIANGLE / _GAUSSIAN with
     * half-widths in pixels and the Gaussian falloff.  A pixel takes every
     * sample of the tile whose splat window (kernel_update_film,
     * src/renderers/samplerrenderer.cu:103-169) covers it, weighted through
     * the 16x16 filter table, and writes max(0, sum(w L) / sum(w)); the box
     * of half-width 0.5 (the default) is the mean of the pixel's own samples. */
    uint32_t filter;
    float filter_width[2];
    float filter_alpha;
    /* Ray-cast debug shading (RayCastIntegratorParams::m_shade,
     * src/core/params.h:139-151): RB_SHADE_* below. */
    uint32_t shade;
    /* Optional device counters (NULL: none), for measurement: camera-path
     * vertices shaded with a BSDF at each bounce are added to
     * path_counter[0..max_depth] (PT, BDPT), light sub-paths traced at each
     * depth to light_path_counter[0..light_max_depth - 1] (BDPT). */
    unsigned long long* path_counter;
    unsigned long long* light_path_counter;
} rb_render_params;
enum { RB_FILTER_BOX = 0, RB_FILTER_TRIANGLE = 1, RB_FILTER_GAUSSIAN = 2 };
enum { RB_SHADE_PRIMITIVE_ID = 0, RB_SHADE_MATERIAL_ID = 1, RB_SHADE_NORMALS = 2, RB_SHADE_UVS = 3,
       RB_SHADE_UNIFORM = 4 };
void rb_render_params_default(rb_render_params* p);

/* One render of the tile [x0,x0+w) x [y0,y0+h): camera rays (spp per pixel),
 * the integrator, and the filtered film (see `filter` above) into out_rgb
 * (device float[w*h*3], row-major).  Replaces SamplerRenderer::render
 * (src/renderers/samplerrenderer.cpp:58-107) + Integrator::render
 * (src/core/integrator.h:59) + kernel_update_film
 * (src/renderers/samplerrenderer.cu:103-169).  Needs a bvh uploaded with its
 * scene.  Asynchronous on `stream`; work buffers are stream-ordered allocations. */
rb_status rb_render_ao(rb_bvh_t bvh, const rb_camera* cam, uint32_t x0, uint32_t y0, uint32_t w, uint32_t h,
                       const rb_render_params* p, float* out_rgb, void* stream);
/* Wavefront path tracer (IntegratorPath, src/integrators/integrator_path.cu:43-283):
 * per bounce shade -> any-hit shadow rays + closest-hit MIS and bounce rays
 * -> connect -> compaction of live paths; queue sizes stay on the device. */
rb_status rb_render_pt(rb_bvh_t bvh, const rb_camera* cam, uint32_t x0, uint32_t y0, uint32_t w, uint32_t h,
                       const rb_render_params* p, float* out_rgb, void* stream);
/* BDPT with a light-vertex cache (IntegratorBdptLvc, src/integrators/integrator_bdpt_lvc.cpp:62-74,
 * integrator_bdpt_lvc_sk.cu:47-514): light pass of light_paths sub-paths
 * (light_paths = 0: lvc_capacity / light_max_depth of them) of up to
 * light_max_depth vertices cached at fixed slots, then the camera pass with
 * direct light and `connections` cache connections per camera vertex. */
rb_status rb_render_bdpt(rb_bvh_t bvh, const rb_camera* cam, uint32_t x0, uint32_t y0, uint32_t w, uint32_t h,
                         const rb_render_params* p, float* out_rgb, void* stream);
/* Ray-cast debug integrator (IntegratorRayCast, src/integrators/integrator_raycast.cu:37-160):
 * one closest-hit trace of the camera rays, each sample coloured by
 * p->shade -- PRIMITIVE_ID: the reference's 15-colour palette indexed by
 * triangle id mod 15; MATERIAL_ID: the hit material's colour; NORMALS: the
 * geometric normal (unit cross(v1 - v0, v2 - v0)); UVS: the barycentric
 * weights (u, v, 0) of v0 and v1; UNIFORM: white -- black on a miss, then
 * the film. */
rb_status rb_render_raycast(rb_bvh_t bvh, const rb_camera* cam, uint32_t x0, uint32_t y0, uint32_t w, uint32_t h,
                            const rb_render_params* p, float* out_rgb, void* stream);

/* A render pass captured into a HIP graph (no reference counterpart; the
 * reference launches every kernel of every iteration from the host,
 * src/core/application.cpp:225-239).  kind: RB_REQ_AO, RB_REQ_PT, RB_REQ_BDPT
 * or RB_REQ_RAYCAST (the serving kinds below).  The graph records the whole pass --
 * camera rays, every bounce up to max_depth (kernels skip work past the
 * live-path count held in device memory), film -- into out_rgb, and owns its
 * scratch.  rb_render_graph_launch(g, seed, stream) replays it for `seed`:
 * the image equals rb_render_* with that seed, bit for bit.  Replays of one
 * graph must not overlap (one stream, or synchronised); ray_counter must be
 * NULL.  Destroy waits for the device. */
typedef struct rb_render_graph_s* rb_render_graph_t;
rb_status rb_render_graph_create(uint32_t kind, rb_bvh_t bvh, const rb_camera* cam, uint32_t x0, uint32_t y0,
                                 uint32_t w, uint32_t h, const rb_render_params* p, float* out_rgb,
                                 rb_render_graph_t* graph);
rb_status rb_render_graph_launch(rb_render_graph_t graph, uint64_t seed, void* stream);
rb_status rb_render_graph_destroy(rb_render_graph_t graph);

/* ---- serving ------------------------------------------------------------ */

/* Continuous-batching request server (no reference counterpart: the
 * reference renders one configured frame per process run,
 * src/core/application.cpp:225-239).  Scenes stay resident; requests are
 * admitted while fewer than max_inflight run and their estimated device
 * memory fits the budget; a scheduler thread inside the library launches,
 * tracks completion and records latency (submit -> result ready).  The
 * running requests either each own a HIP stream so their kernels overlap on
 * the GPU, or advance in lockstep with one batched trace per round (`mode`
 * below).  Results are the same either way, bit for bit. */
typedef struct rb_server_s* rb_server_t;

enum { RB_SCHED_DEADLINE = 0, RB_SCHED_FIFO = 1 };

typedef struct rb_server_config {
    uint32_t max_inflight; /* requests on the device at once (1..32) */
    uint64_t hbm_budget;   /* bytes of request work memory (0: 60% of free HBM) */
    uint32_t max_queue;    /* waiting requests before submit says RB_ERR_BUSY (0: 65536) */
    uint32_t policy;       /* RB_SCHED_DEADLINE (default): earliest of submit + 4 x the
                              estimated GPU time first; RB_SCHED_FIFO: arrival order */
    uint32_t mode;         /* RB_SERVE_STREAMS (default): each request's pass on its own
                              stream, the streams overlapping on the GPU; RB_SERVE_LOCKSTEP:
                              the in-flight requests advance together on one stream, one
                              integrator step each per round and one batched trace launch
                              of all their rays per round */
} rb_server_config;
enum { RB_SERVE_STREAMS = 0, RB_SERVE_LOCKSTEP = 1 };

enum { RB_REQ_AO = 0, RB_REQ_PT = 1, RB_REQ_BDPT = 2, RB_REQ_TRACE = 3, RB_REQ_RAYCAST = 4 };
enum { RB_REQ_QUEUED = 0, RB_REQ_RUNNING = 1, RB_REQ_DONE = 2, RB_REQ_FAILED = 3 };

typedef struct rb_request {
    uint32_t kind;           /* RB_REQ_* */
    uint32_t scene;          /* id from rb_server_add_scene */
    rb_camera camera;        /* render kinds */
    uint32_t x0, y0, w, h;   /* render kinds: tile */
    rb_render_params params; /* render kinds (ray_counter and the path counters are ignored) */
    float* out_host;         /* render kinds: optional host image w*h*3 (valid once done) */
    const rb_ray* rays;      /* RB_REQ_TRACE: device rays / hits, n, RB_CLOSEST or RB_ANY */
    rb_hit* hits;
    uint32_t n;
    uint32_t flags;
} rb_request;

typedef struct rb_request_status {
    int32_t state;           /* RB_REQ_* */
    int32_t result;          /* rb_status of a finished request */
    double queue_ms;         /* submit -> admitted */
    double latency_ms;       /* submit -> result ready */
    uint64_t rays;           /* rays traced for this request */
} rb_request_status;

typedef struct rb_server_stats {
    uint64_t submitted, completed, failed, rejected, deferrals;
    double p50_ms, p99_ms, mean_ms, max_ms;
    uint64_t rays;
    double first_submit_to_last_done_s;
    uint64_t peak_budget_bytes;
    /* device memory the slots' arenas held at most (output + scratch of the
     * largest request each slot ran), and scratch allocations that did not
     * fit their arena and came from the stream-ordered pool instead */
    uint64_t peak_arena_bytes;
    uint64_t arena_misses;
} rb_server_stats;

rb_status rb_server_create(const rb_server_config* cfg, rb_server_t* out);
rb_status rb_server_destroy(rb_server_t s); /* finishes running work, fails queued requests */
rb_status rb_server_add_scene(rb_server_t s, rb_bvh_t bvh, uint32_t* scene_id);
rb_status rb_server_remove_scene(rb_server_t s, uint32_t scene_id); /* RB_ERR_BUSY while referenced */
/* RB_ERR_OOM: the request alone exceeds the budget; RB_ERR_BUSY: queue full. */
rb_status rb_server_submit(rb_server_t s, const rb_request* req, uint64_t* request_id);
rb_status rb_server_poll(rb_server_t s, uint64_t request_id, rb_request_status* st); /* RB_ERR_NOT_READY */
rb_status rb_server_wait(rb_server_t s, uint64_t request_id, rb_request_status* st);
rb_status rb_server_drain(rb_server_t s);
rb_status rb_server_get_stats(rb_server_t s, rb_server_stats* out);
rb_status rb_server_reset_stats(rb_server_t s);
/* Estimated device bytes the server reserves for a request. */
rb_status rb_server_request_bytes(const rb_request* req, uint64_t* bytes);
/* Failure injection for tests: the next k request admissions fail their
 * device allocation (the request ends RB_ERR_OOM, others are unaffected). */
rb_status rb_server_inject_alloc_failures(rb_server_t s, uint32_t k);

#ifdef __cplusplus
}
#endif
#endif /* RB_H */

// rb_internal.h -- host-side structures shared by the librb.so sources.
#pragma once

#include <atomic>
#include <cfloat>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <memory>
#include <string>
#include <vector>

#include "../../include/rb.h"

namespace rb {

// ---------------------------------------------------------------- errors --

void set_error(const char* fmt, ...);
rb_status fail(rb_status s, const char* fmt, ...);

// ------------------------------------------------------------ host math --
// Component-wise float3 operators with the rounding of the reference's
// host build (helper_math.h semantics; this file is compiled with
// -ffp-contract=off so every product and sum rounds separately).

struct F3 {
    float x, y, z;
};
inline F3 f3(float x, float y, float z) { return F3{x, y, z}; }
inline F3 operator-(F3 a, F3 b) { return f3(a.x - b.x, a.y - b.y, a.z - b.z); }
inline F3 operator+(F3 a, F3 b) { return f3(a.x + b.x, a.y + b.y, a.z + b.z); }
inline F3 operator*(F3 a, float s) { return f3(a.x * s, a.y * s, a.z * s); }
inline float dot(F3 a, F3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
inline F3 cross(F3 a, F3 b) { return f3(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x); }
inline F3 normalize(F3 v) { return v * (1.0f / std::sqrt(dot(v, v))); }
// IEEE fmin / fmax written out so they inline (a NaN operand yields the
// other; for equal operands, such as -0 and +0, the second is returned, as
// the x86-64 library routines do).  The library calls cost a fifth of an
// SBVH build.
inline float fmin_(float a, float b) { return (a < b || b != b) ? a : b; }
inline float fmax_(float a, float b) { return (a > b || b != b) ? a : b; }
inline F3 vmin(F3 a, F3 b) { return f3(fmin_(a.x, b.x), fmin_(a.y, b.y), fmin_(a.z, b.z)); }
inline F3 vmax(F3 a, F3 b) { return f3(fmax_(a.x, b.x), fmax_(a.y, b.y), fmax_(a.z, b.z)); }
inline float get(const F3& v, int d) { return d == 0 ? v.x : (d == 1 ? v.y : v.z); }
inline float& getr(F3& v, int d) { return d == 0 ? v.x : (d == 1 ? v.y : v.z); }

inline uint32_t fbits(float f) { uint32_t u; std::memcpy(&u, &f, 4); return u; }

struct Box {
    F3 lo{FLT_MAX, FLT_MAX, FLT_MAX};
    F3 hi{-FLT_MAX, -FLT_MAX, -FLT_MAX};
    bool valid() const { return lo.x <= hi.x && lo.y <= hi.y && lo.z <= hi.z; }
    void insert(F3 p) { lo = vmin(lo, p); hi = vmax(hi, p); }
    void grow(const Box& b) { insert(b.lo); insert(b.hi); }
    void intersect(const Box& b) { lo = vmax(lo, b.lo); hi = vmin(hi, b.hi); }
    float area() const {
        if (!valid()) return 0.0f;
        F3 d = hi - lo;
        return (d.x * d.y + d.y * d.z + d.z * d.x) * 2.0f;
    }
};

// ---------------------------------------------------------------- scene --

// Index of a triangle corner that has no shading normal / texture coordinate.
constexpr uint32_t kNoIndex = 0xFFFFFFFFu;

struct Scene {
    std::vector<float> verts;          // xyz per vertex
    std::vector<uint32_t> idx;         // 3 per triangle (scene-global vertex ids)
    std::vector<uint32_t> tri_mat;     // material per triangle
    std::vector<rb_material> mats;
    std::vector<uint32_t> light_tris;  // triangle ids of emissive triangles
    // Shading data of Scene (src/core/scene.h:142-147): per-corner indices
    // into their own arrays, 3 per triangle, kNoIndex for a triangle without
    // any; an empty index array means no triangle has that data.
    std::vector<float> normals;        // xyz per shading normal
    std::vector<uint32_t> nidx;
    std::vector<float> texcoords;      // uv per texture coordinate
    std::vector<uint32_t> tidx;
    uint32_t num_tris() const { return (uint32_t)(idx.size() / 3); }
    uint32_t num_verts() const { return (uint32_t)(verts.size() / 3); }
    bool tri_has_normals(uint32_t t) const { return !nidx.empty() && nidx[3 * (size_t)t] != kNoIndex; }
};

// Scene::add_mesh (src/core/scene.cpp:83-123) with its checks: vertices and
// triangles (indices local to the mesh), optional normals / texcoords with
// per-corner index triples (NULL indices with normals / texcoords: one per
// vertex, indexed like the vertices; a kNoIndex triple: that triangle has
// none), a material per triangle (tri_mat) or one for all.  Emissive
// triangles become lights.  Nothing is added when a check fails.
rb_status append_mesh(Scene& S, const float* verts, uint32_t nv, const uint32_t* idx, uint32_t nt,
                      const float* normals, uint32_t nn, const uint32_t* nidx, const float* texcoords, uint32_t ntc,
                      const uint32_t* tidx, const uint32_t* tri_mat, uint32_t material_id);

// rb_scene_add_material's checks and derived parameters (Oren-Nayar A, B;
// unused layer slots cleared), in place; a layered material's layer ids are
// looked up in `mats`.
rb_status prepare_material(const std::vector<rb_material>& mats, rb_material* m);

// Copies a vector into a caller array (nothing for an empty vector, whose
// data() may be NULL).
template <class T>
inline void copy_vec(void* dst, const std::vector<T>& v) {
    if (dst && !v.empty()) std::memcpy(dst, v.data(), v.size() * sizeof(T));
}

// Limits every path that grows a scene checks before it reads or copies
// anything: at most 2^32 - 1 vertices and 2^31 - 1 triangles (triangle ids
// travel as int32 in trace results, -1 = miss).
constexpr uint64_t kMaxSceneVerts = 0xFFFFFFFFull;
constexpr uint64_t kMaxSceneTris = 0x7FFFFFFFull;
inline rb_status scene_room(const Scene& S, uint64_t add_verts, uint64_t add_tris) {
    if ((uint64_t)S.num_verts() + add_verts > kMaxSceneVerts || (uint64_t)S.num_tris() + add_tris > kMaxSceneTris)
        return fail(RB_ERR_INVALID, "scene would exceed 2^32 - 1 vertices or 2^31 - 1 triangles (%u + %llu "
                    "vertices, %u + %llu triangles)", S.num_verts(), (unsigned long long)add_verts, S.num_tris(),
                    (unsigned long long)add_tris);
    return RB_OK;
}

// ------------------------------------------------------------ host BVH --

struct BNode {
    Box b;
    BNode* c[2] = {nullptr, nullptr};
    uint32_t* tris = nullptr;  // leaf triangle ids, in Compact2 order
    uint32_t ntris = 0;
    bool leaf = false;
};

struct HostBvh {
    BNode* root = nullptr;
    std::vector<std::unique_ptr<std::vector<BNode>>> node_blocks;
    std::vector<std::unique_ptr<std::vector<uint32_t>>> tri_blocks;
    std::vector<float> verts;     // copy of the geometry the tree was built on
    std::vector<uint32_t> idx;
    rb_build_params params{};
    rb_bvh_stats stats{};
};

// Deepest node (root = depth 0) the SBVH builder makes: a node at this
// depth becomes a leaf whatever its size.
constexpr uint32_t kMaxTreeDepth = 64;

HostBvh* build_sbvh(const float* verts, uint32_t nv, const uint32_t* idx, uint32_t nt,
                    const rb_build_params& p, rb_status* st);
// Binary tree cache (rb_bvh_cache.cpp): key = hash of geometry + the build
// parameters that shape the tree.
uint64_t bvh_cache_key(const float* verts, uint32_t nv, const uint32_t* idx, uint32_t nt, const rb_build_params& p);
rb_status save_bvh_binary(const HostBvh& b, const char* path);
rb_status load_bvh_binary(const char* path, const float* verts, uint32_t nv, const uint32_t* idx, uint32_t nt,
                          const rb_build_params& p, HostBvh** out);
void woopify(const float* v0, const float* v1, const float* v2, float out[12]);

// MI355X layout arrays (host copies).  The first kLdsNodesMax inner
// records are the top of the tree in breadth-first order (the records a
// trace block may keep in LDS); the rest follow in depth-first order.
constexpr uint32_t kLdsNodesMax = 512;
struct Mi355Layout {
    std::vector<uint32_t> nodes;  // 16 words per node
    std::vector<float> tris;      // 12 floats per triangle reference
    std::vector<uint32_t> index;  // 1 per triangle reference
};
rb_status export_mi355x(const HostBvh& b, Mi355Layout& out);
// Wide (4-child) records of an MI355X binary node array, 16 words each:
// words 0..2 the origin p (float), word 3 the biased exponents of the step
// sizes s = 2^e (byte a: e + 127 of axis a), words 4..9 the children's 8-bit
// steps (byte c of each word = child c) for lo.x, hi.x, lo.y, hi.y, lo.z,
// hi.z (plane = fmaf(q, s, p), a box holding the exact one), words 10..13
// their links (>= 0: wide record; < 0: leaf, as in the binary records;
// kWideEmpty: no child), words 14..15 unused.  Record 0 is the root, the
// records are numbered breadth-first.  Closest-hit results do not depend on
// the visit order (ties at equal t go to the lowest triangle id) and every
// leaf the binary walk tests is tested, so the wide traversal returns the
// binary one's (id, t).
constexpr uint32_t kWideWords = 16;
constexpr int32_t kWideEmpty = 0x7FFFFFFF;
rb_status export_wide(const std::vector<uint32_t>& bin, std::vector<uint32_t>& wide);
// Depth of the deepest inner record of an MI355X node array (root = 0);
// UINT32_MAX when a link leaves the array or the links form a cycle.  A
// traversal holds at most depth + 1 stack entries (the sentinel, then one
// postponed child per inner node on the current path).
uint32_t layout_depth(const std::vector<uint32_t>& nodes);
// Entries of a trace's stack (the kernels' and rb_trace_one's).
constexpr uint32_t kTraceStackEntries = 96;
// Whether the wide walk's stack fits a tree of this binary depth: a wide
// record pushes up to three children and the wide tree is half as deep.
inline bool wide_fits_stack(uint32_t depth) {
    return depth != UINT32_MAX && 3 * ((depth + 2) / 2) + 1 < kTraceStackEntries;
}
rb_status export_compact2(const HostBvh& b, std::vector<uint32_t>& nodes, std::vector<uint32_t>& woop,
                          std::vector<uint32_t>& index);

}  // namespace rb

struct rb_scene_s {
    rb::Scene s;
};
struct rb_host_bvh_s {
    rb::HostBvh* b;
};

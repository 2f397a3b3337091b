/*
 * rb_oracle.c -- CPU restatement of the renderbox2 ray-tracing hot path.
 *
 * TEST INFRASTRUCTURE ONLY (the checker).  Loaded by tests/, by
 * __graft_entry__.smoke() and by bench.py's cpu_baseline leg; never by the
 * product library.  Build: oracle/Makefile (gcc -O2 -ffp-contract=off -fopenmp).
 *
 * Reference paths below are relative to renderbox2/ in the reference tree.
 */
#define _GNU_SOURCE
#include "rb_oracle.h"

#include <float.h>
#include <limits.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>

#ifdef _OPENMP
#include <omp.h>
#endif

/* ------------------------------------------------------------------------ */
/* float3 helpers: the component-wise forms of the CUDA-Samples helper_math  */
/* operators the reference uses (normalize = v * (1/sqrt(dot(v,v))) on host). */
/* ------------------------------------------------------------------------ */

typedef struct { float x, y, z; } f3;

static inline f3 mk3(float x, float y, float z) { f3 r = {x, y, z}; return r; }
static inline f3 sub3(f3 a, f3 b) { return mk3(a.x - b.x, a.y - b.y, a.z - b.z); }
static inline f3 add3(f3 a, f3 b) { return mk3(a.x + b.x, a.y + b.y, a.z + b.z); }
static inline f3 mul3s(f3 a, float s) { return mk3(a.x * s, a.y * s, a.z * s); }
static inline float dot3(f3 a, f3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
static inline f3 cross3(f3 a, f3 b)
{
    return mk3(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}
static inline f3 normalize3(f3 v) { return mul3s(v, 1.0f / sqrtf(dot3(v, v))); }
static inline f3 min3v(f3 a, f3 b) { return mk3(fminf(a.x, b.x), fminf(a.y, b.y), fminf(a.z, b.z)); }
static inline f3 max3v(f3 a, f3 b) { return mk3(fmaxf(a.x, b.x), fmaxf(a.y, b.y), fmaxf(a.z, b.z)); }
static inline float get3(f3 v, int d) { return d == 0 ? v.x : (d == 1 ? v.y : v.z); }
static inline void set3(f3* v, int d, float s)
{
    if (d == 0) v->x = s; else if (d == 1) v->y = s; else v->z = s;
}

static inline uint32_t f2u(float f) { uint32_t u; memcpy(&u, &f, 4); return u; }
static inline float u2f(uint32_t u) { float f; memcpy(&f, &u, 4); return f; }

/* (int)f as x86 cvttss2si computes it (INT_MIN for NaN / out of range); the
 * reference's spatial binning relies on that conversion
 * (src/accelerators/sbvh/splitbvhbuilder.cpp:445-446). */
static inline int trunc_int(float f)
{
    if (!(f > -2147483648.0f && f < 2147483648.0f)) return INT_MIN;
    return (int)f;
}

/* ------------------------------------------------------------------------ */
/* AABB, src/core/primitives.h:52-138                                        */
/* ------------------------------------------------------------------------ */

typedef struct { f3 lo, hi; } aabb;

static inline aabb aabb_empty(void)
{
    aabb b; b.lo = mk3(FLT_MAX, FLT_MAX, FLT_MAX); b.hi = mk3(-FLT_MAX, -FLT_MAX, -FLT_MAX);
    return b;
}
static inline int aabb_valid(const aabb* b)
{
    return b->lo.x <= b->hi.x && b->lo.y <= b->hi.y && b->lo.z <= b->hi.z;
}
static inline void aabb_insert(aabb* b, f3 p) { b->lo = min3v(b->lo, p); b->hi = max3v(b->hi, p); }
static inline void aabb_grow(aabb* b, const aabb* o) { aabb_insert(b, o->lo); aabb_insert(b, o->hi); }
static inline void aabb_intersect(aabb* b, const aabb* o) { b->lo = max3v(b->lo, o->lo); b->hi = min3v(b->hi, o->hi); }
static inline float aabb_area(const aabb* b)
{
    if (!aabb_valid(b)) return 0.0f;
    f3 d = sub3(b->hi, b->lo);
    return (d.x * d.y + d.y * d.z + d.z * d.x) * 2.0f;
}

/* ------------------------------------------------------------------------ */
/* Matrix4x4::invert, src/core/transform.h:205-333                           */
/* ------------------------------------------------------------------------ */

void orc_invert4x4(const float* M, float* out)
{
    float inv[16], det;
#define E(i) M[i]
    inv[0] = E(5) * E(10) * E(15) - E(5) * E(11) * E(14) - E(9) * E(6) * E(15) + E(9) * E(7) * E(14) + E(13) * E(6) * E(11) - E(13) * E(7) * E(10);
    inv[4] = -E(4) * E(10) * E(15) + E(4) * E(11) * E(14) + E(8) * E(6) * E(15) - E(8) * E(7) * E(14) - E(12) * E(6) * E(11) + E(12) * E(7) * E(10);
    inv[8] = E(4) * E(9) * E(15) - E(4) * E(11) * E(13) - E(8) * E(5) * E(15) + E(8) * E(7) * E(13) + E(12) * E(5) * E(11) - E(12) * E(7) * E(9);
    inv[12] = -E(4) * E(9) * E(14) + E(4) * E(10) * E(13) + E(8) * E(5) * E(14) - E(8) * E(6) * E(13) - E(12) * E(5) * E(10) + E(12) * E(6) * E(9);
    inv[1] = -E(1) * E(10) * E(15) + E(1) * E(11) * E(14) + E(9) * E(2) * E(15) - E(9) * E(3) * E(14) - E(13) * E(2) * E(11) + E(13) * E(3) * E(10);
    inv[5] = E(0) * E(10) * E(15) - E(0) * E(11) * E(14) - E(8) * E(2) * E(15) + E(8) * E(3) * E(14) + E(12) * E(2) * E(11) - E(12) * E(3) * E(10);
    inv[9] = -E(0) * E(9) * E(15) + E(0) * E(11) * E(13) + E(8) * E(1) * E(15) - E(8) * E(3) * E(13) - E(12) * E(1) * E(11) + E(12) * E(3) * E(9);
    inv[13] = E(0) * E(9) * E(14) - E(0) * E(10) * E(13) - E(8) * E(1) * E(14) + E(8) * E(2) * E(13) + E(12) * E(1) * E(10) - E(12) * E(2) * E(9);
    inv[2] = E(1) * E(6) * E(15) - E(1) * E(7) * E(14) - E(5) * E(2) * E(15) + E(5) * E(3) * E(14) + E(13) * E(2) * E(7) - E(13) * E(3) * E(6);
    inv[6] = -E(0) * E(6) * E(15) + E(0) * E(7) * E(14) + E(4) * E(2) * E(15) - E(4) * E(3) * E(14) - E(12) * E(2) * E(7) + E(12) * E(3) * E(6);
    inv[10] = E(0) * E(5) * E(15) - E(0) * E(7) * E(13) - E(4) * E(1) * E(15) + E(4) * E(3) * E(13) + E(12) * E(1) * E(7) - E(12) * E(3) * E(5);
    inv[14] = -E(0) * E(5) * E(14) + E(0) * E(6) * E(13) + E(4) * E(1) * E(14) - E(4) * E(2) * E(13) - E(12) * E(1) * E(6) + E(12) * E(2) * E(5);
    inv[3] = -E(1) * E(6) * E(11) + E(1) * E(7) * E(10) + E(5) * E(2) * E(11) - E(5) * E(3) * E(10) - E(9) * E(2) * E(7) + E(9) * E(3) * E(6);
    inv[7] = E(0) * E(6) * E(11) - E(0) * E(7) * E(10) - E(4) * E(2) * E(11) + E(4) * E(3) * E(10) + E(8) * E(2) * E(7) - E(8) * E(3) * E(6);
    inv[11] = -E(0) * E(5) * E(11) + E(0) * E(7) * E(9) + E(4) * E(1) * E(11) - E(4) * E(3) * E(9) - E(8) * E(1) * E(7) + E(8) * E(3) * E(5);
    inv[15] = E(0) * E(5) * E(10) - E(0) * E(6) * E(9) - E(4) * E(1) * E(10) + E(4) * E(2) * E(9) + E(8) * E(1) * E(6) - E(8) * E(2) * E(5);
    det = E(0) * inv[0] + E(1) * inv[4] + E(2) * inv[8] + E(3) * inv[12];
#undef E
    det = 1.0f / det;
    for (int i = 0; i < 16; i++) out[i] = inv[i] * det;
}

/* CudaBvh::woopify_tri, src/accelerators/sbvh/cudabvh.cpp:303-322: the
 * columns of M are (v0-v2, v1-v2, normalize(cross(v0-v2, v1-v2)), v2) with
 * w = (0,0,0,1); the rows kept are row 2 of M^-1 (with its w negated), row 0
 * and row 1. */
void orc_woopify(const float* a, const float* b, const float* c, float* out)
{
    f3 v0 = mk3(a[0], a[1], a[2]), v1 = mk3(b[0], b[1], b[2]), v2 = mk3(c[0], c[1], c[2]);
    f3 c0 = sub3(v0, v2), c1 = sub3(v1, v2);
    f3 c2 = normalize3(cross3(sub3(v0, v2), sub3(v1, v2)));
    float m[16] = {
        c0.x, c1.x, c2.x, v2.x,
        c0.y, c1.y, c2.y, v2.y,
        c0.z, c1.z, c2.z, v2.z,
        0.0f, 0.0f, 0.0f, 1.0f,
    };
    float inv[16];
    orc_invert4x4(m, inv);
    out[0] = inv[8]; out[1] = inv[9]; out[2] = inv[10]; out[3] = -inv[11];
    out[4] = inv[0]; out[5] = inv[1]; out[6] = inv[2]; out[7] = inv[3];
    out[8] = inv[4]; out[9] = inv[5]; out[10] = inv[6]; out[11] = inv[7];
}

/* ------------------------------------------------------------------------ */
/* SBVH builder: SplitBVHBuilder, src/accelerators/sbvh/splitbvhbuilder.cpp  */
/* ------------------------------------------------------------------------ */

enum { MAX_DEPTH = 64, MAX_SPATIAL_DEPTH = 48, NUM_BINS = 128 };

typedef struct { uint32_t tri; aabb b; } ref_t;

typedef struct {
    int leaf;
    aabb b;
    int32_t child[2];
    uint32_t lo, hi;
} node_t;

struct orc_bvh {
    node_t* nodes; uint32_t n_nodes, cap_nodes;
    uint32_t* tris; uint32_t n_tris, cap_tris;
    int32_t root;
    uint64_t duplicates;
    uint32_t max_depth;
    float node_cost, tri_cost;
    /* copies of the scene for the layout stage */
    const float* verts; const uint32_t* idx;
};

typedef struct {
    orc_bvh* bvh;
    const float* verts; const uint32_t* idx;
    ref_t* st; size_t n, cap;          /* m_ref_stack */
    aabb* right_bounds;                /* m_right_bounds */
    float min_overlap;
    uint32_t min_leaf, max_leaf;
    int sort_dim;
    int oom;
    aabb bin_b[3][NUM_BINS];
    uint32_t bin_enter[3][NUM_BINS], bin_exit[3][NUM_BINS];
} builder_t;

/* Platform::get_triangle_cost / get_node_cost with batch size 1,
 * src/accelerators/sbvh/platform.h:102-110. */
static inline float tri_cost(const builder_t* B, uint32_t n) { return (float)n * B->bvh->tri_cost; }
static inline float node_cost(const builder_t* B, uint32_t n) { return (float)n * B->bvh->node_cost; }

static f3 vtx(const builder_t* B, uint32_t i) { return mk3(B->verts[3 * i], B->verts[3 * i + 1], B->verts[3 * i + 2]); }

static int push_ref(builder_t* B, ref_t r)
{
    if (B->n == B->cap) {
        size_t nc = B->cap ? B->cap * 2 : 1024;
        ref_t* p = (ref_t*)realloc(B->st, nc * sizeof(ref_t));
        if (!p) { B->oom = 1; return 0; }
        B->st = p; B->cap = nc;
    }
    B->st[B->n++] = r;
    return 1;
}

static int32_t new_node(builder_t* B)
{
    orc_bvh* T = B->bvh;
    if (T->n_nodes == T->cap_nodes) {
        uint32_t nc = T->cap_nodes ? T->cap_nodes * 2 : 1024;
        node_t* p = (node_t*)realloc(T->nodes, (size_t)nc * sizeof(node_t));
        if (!p) { B->oom = 1; return -1; }
        T->nodes = p; T->cap_nodes = nc;
    }
    return (int32_t)T->n_nodes++;
}

/* sort_compare, splitbvhbuilder.cpp:273-282: centroid*2 along sort_dim, ties
 * by triangle index.  A total order inside one node, so any sort gives the
 * reference's bespoke quicksort result. */
static int cmp_ref(const void* pa, const void* pb, void* ctx)
{
    int dim = *(const int*)ctx;
    const ref_t* a = (const ref_t*)pa;
    const ref_t* b = (const ref_t*)pb;
    float ca = get3(a->b.lo, dim) + get3(a->b.hi, dim);
    float cb = get3(b->b.lo, dim) + get3(b->b.hi, dim);
    if (ca < cb) return -1;
    if (cb < ca) return 1;
    if (a->tri < b->tri) return -1;
    if (a->tri > b->tri) return 1;
    return 0;
}

static void sort_range(builder_t* B, size_t start, size_t end, int dim)
{
    if (end - start < 2) return;
    qsort_r(B->st + start, end - start, sizeof(ref_t), cmp_ref, &dim);
}

typedef struct { float sah; int dim; uint32_t num_left; aabb lb, rb; } obj_split;
typedef struct { float sah; int dim; float pos; } sp_split;

/* find_object_split, splitbvhbuilder.cpp:369-408 */
static obj_split find_object_split(builder_t* B, uint32_t num, float nodeSAH)
{
    obj_split s; s.sah = FLT_MAX; s.dim = 0; s.num_left = 0;
    s.lb = aabb_empty(); s.rb = aabb_empty();
    size_t base = B->n - num;
    float bestTie = FLT_MAX;
    for (int dim = 0; dim < 3; dim++) {
        B->sort_dim = dim;
        sort_range(B, base, B->n, dim);
        const ref_t* r = B->st + base;
        aabb rb = aabb_empty();
        for (uint32_t i = num - 1; i > 0; i--) {
            aabb_grow(&rb, &r[i].b);
            B->right_bounds[i - 1] = rb;
        }
        aabb lb = aabb_empty();
        for (uint32_t i = 1; i < num; i++) {
            aabb_grow(&lb, &r[i - 1].b);
            float sah = nodeSAH + aabb_area(&lb) * tri_cost(B, i) + aabb_area(&B->right_bounds[i - 1]) * tri_cost(B, num - i);
            float fi = (float)i, fr = (float)(num - i);
            float tie = fi * fi + fr * fr;
            if (sah < s.sah || (sah == s.sah && tie < bestTie)) {
                s.sah = sah; s.dim = dim; s.num_left = i;
                s.lb = lb; s.rb = B->right_bounds[i - 1];
                bestTie = tie;
            }
        }
    }
    return s;
}

/* split_reference, splitbvhbuilder.cpp:586-627 */
static void split_reference(const builder_t* B, ref_t* L, ref_t* R, const ref_t* ref, int dim, float pos)
{
    L->tri = R->tri = ref->tri;
    L->b = aabb_empty(); R->b = aabb_empty();
    const uint32_t* ind = B->idx + 3 * (size_t)ref->tri;
    f3 v1 = vtx(B, ind[2]);
    for (int i = 0; i < 3; i++) {
        f3 v0 = v1;
        v1 = vtx(B, ind[i]);
        float v0p = get3(v0, dim), v1p = get3(v1, dim);
        if (v0p <= pos) aabb_insert(&L->b, v0);
        if (v0p >= pos) aabb_insert(&R->b, v0);
        if ((v0p < pos && v1p > pos) || (v0p > pos && v1p < pos)) {
            float t = (pos - v0p) / (v1p - v0p);
            t = fmaxf(0.0f, fminf(t, 1.0f));
            f3 p = add3(v0, mul3s(sub3(v1, v0), t));
            aabb_insert(&L->b, p);
            aabb_insert(&R->b, p);
        }
    }
    set3(&L->b.hi, dim, pos);
    set3(&R->b.lo, dim, pos);
    aabb_intersect(&L->b, &ref->b);
    aabb_intersect(&R->b, &ref->b);
}

/* find_spatial_split, splitbvhbuilder.cpp:420-497 */
static sp_split find_spatial_split(builder_t* B, const aabb* nb, uint32_t num, float nodeSAH)
{
    f3 origin = nb->lo;
    f3 binSize = mul3s(sub3(nb->hi, origin), 1.0f / (float)NUM_BINS);
    f3 inv = mk3(1.0f / binSize.x, 1.0f / binSize.y, 1.0f / binSize.z);
    for (int d = 0; d < 3; d++)
        for (int i = 0; i < NUM_BINS; i++) {
            B->bin_b[d][i] = aabb_empty(); B->bin_enter[d][i] = 0; B->bin_exit[d][i] = 0;
        }
    for (size_t ri = B->n - num; ri < B->n; ri++) {
        const ref_t* ref = &B->st[ri];
        f3 rmin = mk3((ref->b.lo.x - origin.x) * inv.x, (ref->b.lo.y - origin.y) * inv.y, (ref->b.lo.z - origin.z) * inv.z);
        f3 rmax = mk3((ref->b.hi.x - origin.x) * inv.x, (ref->b.hi.y - origin.y) * inv.y, (ref->b.hi.z - origin.z) * inv.z);
        int first[3], last[3];
        for (int d = 0; d < 3; d++) {
            int f = trunc_int(get3(rmin, d));
            f = f < 0 ? 0 : (f > NUM_BINS - 1 ? NUM_BINS - 1 : f);
            int l = trunc_int(get3(rmax, d));
            l = l > NUM_BINS - 1 ? NUM_BINS - 1 : l;
            l = l < f ? f : l;
            first[d] = f; last[d] = l;
        }
        for (int d = 0; d < 3; d++) {
            ref_t cur = *ref;
            for (int i = first[d]; i < last[d]; i++) {
                ref_t L, R;
                split_reference(B, &L, &R, &cur, d, get3(origin, d) + get3(binSize, d) * (float)(i + 1));
                aabb_grow(&B->bin_b[d][i], &L.b);
                cur = R;
            }
            aabb_grow(&B->bin_b[d][last[d]], &cur.b);
            B->bin_enter[d][first[d]]++;
            B->bin_exit[d][last[d]]++;
        }
    }
    sp_split s; s.sah = FLT_MAX; s.dim = 0; s.pos = 0.0f;
    for (int d = 0; d < 3; d++) {
        aabb rb = aabb_empty();
        for (int i = NUM_BINS - 1; i > 0; i--) {
            aabb_grow(&rb, &B->bin_b[d][i]);
            B->right_bounds[i - 1] = rb;
        }
        aabb lb = aabb_empty();
        uint32_t ln = 0, rn = num;
        for (int i = 1; i < NUM_BINS; i++) {
            aabb_grow(&lb, &B->bin_b[d][i - 1]);
            ln += B->bin_enter[d][i - 1];
            rn -= B->bin_exit[d][i - 1];
            float sah = nodeSAH + aabb_area(&lb) * tri_cost(B, ln) + aabb_area(&B->right_bounds[i - 1]) * tri_cost(B, rn);
            if (sah < s.sah) {
                s.sah = sah; s.dim = d;
                s.pos = get3(origin, d) + get3(binSize, d) * (float)i;
            }
        }
    }
    return s;
}

/* perform_spatial_split, splitbvhbuilder.cpp:499-584 */
static void perform_spatial_split(builder_t* B, uint32_t* lnum, aabb* lb, uint32_t* rnum, aabb* rb,
                                  uint32_t num, const sp_split* sp)
{
    size_t leftStart = B->n - num, leftEnd = leftStart, rightStart = B->n;
    *lb = aabb_empty(); *rb = aabb_empty();
    for (size_t i = leftEnd; i < rightStart; i++) {
        if (get3(B->st[i].b.hi, sp->dim) <= sp->pos) {
            aabb_grow(lb, &B->st[i].b);
            ref_t t = B->st[i]; B->st[i] = B->st[leftEnd]; B->st[leftEnd] = t; leftEnd++;
        } else if (get3(B->st[i].b.lo, sp->dim) >= sp->pos) {
            aabb_grow(rb, &B->st[i].b);
            rightStart--;
            ref_t t = B->st[i]; B->st[i] = B->st[rightStart]; B->st[rightStart] = t;
            i--;
        }
    }
    while (leftEnd < rightStart) {
        ref_t L, R;
        split_reference(B, &L, &R, &B->st[leftEnd], sp->dim, sp->pos);
        aabb lub = *lb, rub = *rb, ldb = *lb, rdb = *rb;
        aabb_grow(&lub, &B->st[leftEnd].b);
        aabb_grow(&rub, &B->st[leftEnd].b);
        aabb_grow(&ldb, &L.b);
        aabb_grow(&rdb, &R.b);
        float lac = tri_cost(B, (uint32_t)(leftEnd - leftStart));
        float rac = tri_cost(B, (uint32_t)(B->n - rightStart));
        float lbc = tri_cost(B, (uint32_t)(leftEnd - leftStart + 1));
        float rbc = tri_cost(B, (uint32_t)(B->n - rightStart + 1));
        float usl = aabb_area(&lub) * lbc + aabb_area(rb) * rac;
        float usr = aabb_area(lb) * lac + aabb_area(&rub) * rbc;
        float dup = aabb_area(&ldb) * lbc + aabb_area(&rdb) * rbc;
        float m = fminf(usl, fminf(usr, dup));
        if (m == usl) {
            *lb = lub; leftEnd++;
        } else if (m == usr) {
            *rb = rub;
            rightStart--;
            ref_t t = B->st[leftEnd]; B->st[leftEnd] = B->st[rightStart]; B->st[rightStart] = t;
        } else {
            *lb = ldb; *rb = rdb;
            B->st[leftEnd++] = L;
            if (!push_ref(B, R)) return;
        }
    }
    *lnum = (uint32_t)(leftEnd - leftStart);
    *rnum = (uint32_t)(B->n - rightStart);
}

/* create_leaf, splitbvhbuilder.cpp:359-367: refs are popped from the back. */
static int32_t create_leaf(builder_t* B, uint32_t num, const aabb* b)
{
    orc_bvh* T = B->bvh;
    int32_t id = new_node(B);
    if (id < 0) return -1;
    if (T->n_tris + num > T->cap_tris) {
        uint32_t nc = T->cap_tris ? T->cap_tris : 1024;
        while (nc < T->n_tris + num) nc *= 2;
        uint32_t* p = (uint32_t*)realloc(T->tris, (size_t)nc * 4);
        if (!p) { B->oom = 1; return -1; }
        T->tris = p; T->cap_tris = nc;
    }
    for (uint32_t i = 0; i < num; i++) T->tris[T->n_tris++] = B->st[--B->n].tri;
    node_t* nd = &T->nodes[id];
    nd->leaf = 1; nd->b = *b; nd->child[0] = nd->child[1] = -1;
    nd->hi = T->n_tris; nd->lo = T->n_tris - num;
    return id;
}

/* Decision log of the builder (checker support): when started, every
 * build_node call on this thread records its level, bounds, references
 * (after degenerate removal) and what it decided, so a test can recompute
 * every candidate independently and check that the decision is the
 * minimum the reference's rules select. */
static __thread int g_log_on;
static __thread orc_node_log* g_log_nodes;
static __thread size_t g_log_n, g_log_cap;
static __thread float* g_log_refs;   /* 7 floats per reference: tri id bits, lo.xyz, hi.xyz */
static __thread size_t g_log_nrefs, g_log_refcap;

void orc_build_log_start(void)
{
    g_log_on = 1;
    g_log_n = g_log_nrefs = 0;
}

void orc_build_log_stop(void)
{
    g_log_on = 0;
    free(g_log_nodes); free(g_log_refs);
    g_log_nodes = NULL; g_log_refs = NULL;
    g_log_n = g_log_cap = g_log_nrefs = g_log_refcap = 0;
}

size_t orc_build_log_nodes(orc_node_log* out)
{
    if (out && g_log_n) memcpy(out, g_log_nodes, g_log_n * sizeof(orc_node_log));
    return g_log_n;
}

size_t orc_build_log_refs(float* out7)
{
    if (out7 && g_log_nrefs) memcpy(out7, g_log_refs, g_log_nrefs * 7 * sizeof(float));
    return g_log_nrefs;
}

/* Appends a node entry with its references; returns its index or -1. */
static long log_node(const builder_t* B, uint32_t num, const aabb* bounds, int level)
{
    if (g_log_n == g_log_cap) {
        size_t nc = g_log_cap ? 2 * g_log_cap : 1024;
        orc_node_log* p = (orc_node_log*)realloc(g_log_nodes, nc * sizeof(orc_node_log));
        if (!p) return -1;
        g_log_nodes = p; g_log_cap = nc;
    }
    if (g_log_nrefs + num > g_log_refcap) {
        size_t nc = g_log_refcap ? 2 * g_log_refcap : 4096;
        while (nc < g_log_nrefs + num) nc *= 2;
        float* p = (float*)realloc(g_log_refs, nc * 7 * sizeof(float));
        if (!p) return -1;
        g_log_refs = p; g_log_refcap = nc;
    }
    orc_node_log* e = &g_log_nodes[g_log_n];
    memset(e, 0, sizeof(*e));
    e->level = level;
    e->num = num;
    e->bounds[0] = bounds->lo.x; e->bounds[1] = bounds->lo.y; e->bounds[2] = bounds->lo.z;
    e->bounds[3] = bounds->hi.x; e->bounds[4] = bounds->hi.y; e->bounds[5] = bounds->hi.z;
    e->ref_off = g_log_nrefs;
    for (size_t i = B->n - num; i < B->n; i++) {
        float* r = g_log_refs + 7 * g_log_nrefs++;
        memcpy(&r[0], &B->st[i].tri, 4);
        r[1] = B->st[i].b.lo.x; r[2] = B->st[i].b.lo.y; r[3] = B->st[i].b.lo.z;
        r[4] = B->st[i].b.hi.x; r[5] = B->st[i].b.hi.y; r[6] = B->st[i].b.hi.z;
    }
    return (long)g_log_n++;
}

/* build_node, splitbvhbuilder.cpp:290-357 */
static int32_t build_node(builder_t* B, uint32_t num, aabb bounds, int level)
{
    if (B->oom) return -1;
    if ((uint32_t)level > B->bvh->max_depth) B->bvh->max_depth = (uint32_t)level;
    /* Remove degenerates (:300-316). */
    {
        long firstRef = (long)B->n - (long)num;
        for (long i = (long)B->n - 1; i >= firstRef; i--) {
            f3 sz = sub3(B->st[i].b.hi, B->st[i].b.lo);
            float mn = sz.x; if (sz.y < mn) mn = sz.y; if (sz.z < mn) mn = sz.z;
            float mx = sz.x; if (sz.y > mx) mx = sz.y; if (sz.z > mx) mx = sz.z;
            float sum = sz.x + sz.y + sz.z;
            if (mn < 0.0f || sum == mx) {
                ref_t back = B->st[B->n - 1];
                B->n--;
                if ((size_t)i < B->n) B->st[i] = back;
            }
        }
        num = (uint32_t)((long)B->n - firstRef);
    }
    const long lg = g_log_on ? log_node(B, num, &bounds, level) : -1;
    if (lg >= 0) {
        g_log_nodes[lg].leaf_sah = g_log_nodes[lg].obj_sah = g_log_nodes[lg].sp_sah = FLT_MAX;
    }
    if (num <= B->min_leaf || level >= MAX_DEPTH) return create_leaf(B, num, &bounds);

    float area = aabb_area(&bounds);
    float leafSAH = area * tri_cost(B, num);
    float nodeSAH = area * node_cost(B, 2);
    obj_split obj = find_object_split(B, num, nodeSAH);

    sp_split sp; sp.sah = FLT_MAX; sp.dim = 0; sp.pos = 0.0f;
    int sp_tried = 0;
    if (level < MAX_SPATIAL_DEPTH) {
        aabb ov = obj.lb;
        aabb_intersect(&ov, &obj.rb);
        if (aabb_area(&ov) >= B->min_overlap) {
            sp = find_spatial_split(B, &bounds, num, nodeSAH);
            sp_tried = 1;
        }
    }
    if (lg >= 0) {
        orc_node_log* e = &g_log_nodes[lg];
        e->leaf_sah = leafSAH; e->obj_sah = obj.sah; e->obj_dim = obj.dim; e->obj_left = obj.num_left;
        e->sp_sah = sp.sah; e->sp_dim = sp.dim; e->sp_pos = sp.pos; e->sp_tried = sp_tried;
    }

    float minSAH = leafSAH;
    { float m2 = obj.sah < sp.sah ? obj.sah : sp.sah; if (m2 < minSAH) minSAH = m2; }
    if (minSAH == leafSAH && num <= B->max_leaf) return create_leaf(B, num, &bounds);

    uint32_t ln = 0, rn = 0;
    aabb lb = aabb_empty(), rb = aabb_empty();
    int kind = 1;
    if (minSAH == sp.sah) {
        perform_spatial_split(B, &ln, &lb, &rn, &rb, num, &sp);
        kind = 2;
    }
    if (B->oom) return -1;
    if (!ln || !rn) {
        /* perform_object_split, :409-418 */
        sort_range(B, B->n - num, B->n, obj.dim);
        ln = obj.num_left; lb = obj.lb;
        rn = num - obj.num_left; rb = obj.rb;
        kind = 1;
    }
    if (lg >= 0) {
        g_log_nodes[lg].kind = kind;
        g_log_nodes[lg].ln = ln;
        g_log_nodes[lg].rn = rn;
    }
    B->bvh->duplicates += (uint64_t)(ln + rn - num);
    int32_t right = build_node(B, rn, rb, level + 1);
    int32_t left = build_node(B, ln, lb, level + 1);
    if (right < 0 || left < 0) return -1;
    int32_t id = new_node(B);
    if (id < 0) return -1;
    node_t* nd = &B->bvh->nodes[id];
    nd->leaf = 0; nd->b = bounds; nd->child[0] = left; nd->child[1] = right; nd->lo = nd->hi = 0;
    return id;
}

orc_bvh* orc_bvh_build(const float* verts, uint32_t nv, const uint32_t* idx, uint32_t nt,
                       uint32_t min_leaf, uint32_t max_leaf, float split_alpha,
                       float ncost, float tcost)
{
    (void)nv;
    orc_bvh* T = (orc_bvh*)calloc(1, sizeof(orc_bvh));
    builder_t* B = (builder_t*)calloc(1, sizeof(builder_t));
    if (!T || !B) { free(T); free(B); return NULL; }
    T->node_cost = ncost; T->tri_cost = tcost;
    T->verts = verts; T->idx = idx;
    B->bvh = T; B->verts = verts; B->idx = idx;
    B->min_leaf = min_leaf; B->max_leaf = max_leaf;
    /* run(), :234-271 */
    aabb root = aabb_empty();
    for (uint32_t i = 0; i < nt; i++) {
        ref_t r; r.tri = i; r.b = aabb_empty();
        aabb_insert(&r.b, vtx(B, idx[3 * (size_t)i]));
        aabb_insert(&r.b, vtx(B, idx[3 * (size_t)i + 1]));
        aabb_insert(&r.b, vtx(B, idx[3 * (size_t)i + 2]));
        if (!push_ref(B, r)) break;
        aabb_grow(&root, &r.b);
    }
    B->min_overlap = aabb_area(&root) * split_alpha;
    size_t rbn = (nt > NUM_BINS ? nt : NUM_BINS) - 1;
    B->right_bounds = (aabb*)malloc(rbn * sizeof(aabb));
    if (!B->right_bounds) B->oom = 1;
    if (!B->oom) T->root = build_node(B, nt, root, 0);
    int oom = B->oom;
    free(B->st); free(B->right_bounds); free(B);
    if (oom || T->root < 0) { orc_bvh_free(T); return NULL; }
    return T;
}

void orc_bvh_free(orc_bvh* T)
{
    if (!T) return;
    free(T->nodes); free(T->tris); free(T);
}

/* get_subtree_size / compute_subtree_probabilities,
 * src/accelerators/sbvh/bvnode.cpp:64-100. */
static void stats_rec(const orc_bvh* T, int32_t id, float prob, float* sah, uint64_t st[6])
{
    const node_t* n = &T->nodes[id];
    uint32_t nchild = n->leaf ? 0 : 2;
    uint32_t ntri = n->leaf ? n->hi - n->lo : 0;
    *sah += prob * ((float)nchild * T->node_cost + (float)ntri * T->tri_cost);
    if (n->leaf) { st[1]++; st[2] += ntri; return; }
    st[0]++; st[3] += 2;
    for (int i = 0; i < 2; i++) {
        const node_t* c = &T->nodes[n->child[i]];
        float cp = 0.0f;
        if (prob > 0.0f) cp = prob * aabb_area(&c->b) / aabb_area(&n->b);
        stats_rec(T, n->child[i], cp, sah, st);
    }
}

void orc_bvh_stats(const orc_bvh* T, uint64_t st[6], float* sah)
{
    memset(st, 0, 6 * sizeof(uint64_t));
    float s = 0.0f;
    stats_rec(T, T->root, 1.0f, &s, st);
    st[4] = T->duplicates;
    st[5] = T->max_depth;
    if (sah) *sah = s;
}

static uint32_t dump_rec(const orc_bvh* T, int32_t id, int32_t* info, float* bnd, uint32_t* k)
{
    uint32_t me = (*k)++;
    const node_t* n = &T->nodes[id];
    if (info) {
        info[5 * me + 0] = n->leaf;
        info[5 * me + 1] = -1; info[5 * me + 2] = -1;
        info[5 * me + 3] = (int32_t)n->lo; info[5 * me + 4] = (int32_t)n->hi;
    }
    if (bnd) {
        bnd[6 * me + 0] = n->b.lo.x; bnd[6 * me + 1] = n->b.lo.y; bnd[6 * me + 2] = n->b.lo.z;
        bnd[6 * me + 3] = n->b.hi.x; bnd[6 * me + 4] = n->b.hi.y; bnd[6 * me + 5] = n->b.hi.z;
    }
    if (!n->leaf) {
        uint32_t c0 = dump_rec(T, n->child[0], info, bnd, k);
        uint32_t c1 = dump_rec(T, n->child[1], info, bnd, k);
        if (info) { info[5 * me + 1] = (int32_t)c0; info[5 * me + 2] = (int32_t)c1; }
    }
    return me;
}

uint32_t orc_bvh_nodes(const orc_bvh* T, int32_t* info5, float* bounds6)
{
    uint32_t k = 0;
    if (!info5 && !bounds6) return T->n_nodes;
    dump_rec(T, T->root, info5, bounds6, &k);
    return k;
}

uint32_t orc_bvh_tri_indices(const orc_bvh* T, uint32_t* out)
{
    if (out) memcpy(out, T->tris, (size_t)T->n_tris * 4);
    return T->n_tris;
}

/* ------------------------------------------------------------------------ */
/* Compact2 layout, CudaBvh::create_compact(bvh, 16),                        */
/* src/accelerators/sbvh/cudabvh.cpp:216-296                                 */
/* ------------------------------------------------------------------------ */

int orc_bvh_compact2(const orc_bvh* T, uint32_t* nodes, size_t* n_node_rows,
                     uint32_t* woop, size_t* n_woop_rows, uint32_t* tri_index, size_t* n_index)
{
    /* A root leaf has no Compact2 form (the reference asserts, :237); it is
     * wrapped in an inner node whose second child is an empty leaf with an
     * inverted (never entered) box. */
    node_t wrap, empty;
    const int root_leaf = T->nodes[T->root].leaf;
    if (root_leaf) {
        empty.leaf = 1; empty.b = aabb_empty(); empty.child[0] = empty.child[1] = -1; empty.lo = empty.hi = 0;
        wrap.leaf = 0; wrap.b = T->nodes[T->root].b; wrap.child[0] = T->root; wrap.child[1] = -2;
        wrap.lo = wrap.hi = 0;
    }
    size_t cap_stack = 256, sp = 0;
    struct ent { int32_t node; size_t row; } *stack = malloc(cap_stack * sizeof(*stack));
    if (!stack) return -2;
    size_t nrows = 4, wrows = 0, irows = 0;
    stack[sp].node = root_leaf ? -3 : T->root; stack[sp].row = 0; sp++;
    while (sp) {
        struct ent e = stack[--sp];
        const node_t* n = e.node == -3 ? &wrap : &T->nodes[e.node];
        const aabb* cb[2];
        int32_t cidx[2];
        for (int i = 0; i < 2; i++) {
            const node_t* c = n->child[i] == -2 ? &empty : &T->nodes[n->child[i]];
            cb[i] = &c->b;
            if (!c->leaf) {
                cidx[i] = (int32_t)(nrows * 16 / 16);
                if (sp == cap_stack) {
                    cap_stack *= 2;
                    void* p = realloc(stack, cap_stack * sizeof(*stack));
                    if (!p) { free(stack); return -2; }
                    stack = p;
                }
                stack[sp].node = n->child[i]; stack[sp].row = nrows; sp++;
                nrows += 4;
                continue;
            }
            cidx[i] = ~(int32_t)wrows;
            for (uint32_t j = c->lo; j < c->hi; j++) {
                uint32_t tri = T->tris[j];
                const uint32_t* ind = T->idx + 3 * (size_t)tri;
                float w[12];
                orc_woopify(T->verts + 3 * (size_t)ind[0], T->verts + 3 * (size_t)ind[1], T->verts + 3 * (size_t)ind[2], w);
                if (w[0] == 0.0f) w[0] = 0.0f;
                if (woop) for (int k = 0; k < 12; k++) woop[wrows * 4 + k] = f2u(w[k]);
                if (tri_index) { tri_index[irows] = tri; tri_index[irows + 1] = 0; tri_index[irows + 2] = 0; }
                wrows += 3; irows += 3;
            }
            if (woop) { woop[wrows * 4 + 0] = woop[wrows * 4 + 1] = woop[wrows * 4 + 2] = woop[wrows * 4 + 3] = 0x80000000u; }
            if (tri_index) tri_index[irows] = 0;
            wrows += 1; irows += 1;
        }
        if (nodes) {
            uint32_t* d = nodes + e.row * 4;
            d[0] = f2u(cb[0]->lo.x); d[1] = f2u(cb[0]->hi.x); d[2] = f2u(cb[0]->lo.y); d[3] = f2u(cb[0]->hi.y);
            d[4] = f2u(cb[1]->lo.x); d[5] = f2u(cb[1]->hi.x); d[6] = f2u(cb[1]->lo.y); d[7] = f2u(cb[1]->hi.y);
            d[8] = f2u(cb[0]->lo.z); d[9] = f2u(cb[0]->hi.z); d[10] = f2u(cb[1]->lo.z); d[11] = f2u(cb[1]->hi.z);
            d[12] = (uint32_t)cidx[0]; d[13] = (uint32_t)cidx[1]; d[14] = 0; d[15] = 0;
        }
    }
    free(stack);
    if (n_node_rows) *n_node_rows = nrows;
    if (n_woop_rows) *n_woop_rows = wrows;
    if (n_index) *n_index = irows;
    return 0;
}

/* ------------------------------------------------------------------------ */
/* Traversal: kernel_kepler_trace, kepler_dynamic_fetch.h:67-363             */
/* ------------------------------------------------------------------------ */

#define SENTINEL 0x76543210
#define ORC_STACK 96

typedef struct { float x, y, z, w; } f4;
static inline f4 row(const uint32_t* a, int32_t r)
{
    f4 v; const uint32_t* p = a + 4 * (size_t)r;
    v.x = u2f(p[0]); v.y = u2f(p[1]); v.z = u2f(p[2]); v.w = u2f(p[3]);
    return v;
}

static void trace_one(const uint32_t* nodes, const uint32_t* woop, const uint32_t* tidx,
                      const float* ray, int any_hit, int32_t* out_id, float* out_t, uint32_t* cnt)
{
    int32_t stack[ORC_STACK];
    int sp = 0;
    stack[0] = SENTINEL;
    const float ox = ray[0], oy = ray[1], oz = ray[2], tmin = ray[3];
    const float dx = ray[4], dy = ray[5], dz = ray[6];
    float hitT = ray[7];
    const float ooeps = 0x1p-80f;
    const float idx_ = 1.0f / (fabsf(dx) > ooeps ? dx : copysignf(ooeps, dx));
    const float idy_ = 1.0f / (fabsf(dy) > ooeps ? dy : copysignf(ooeps, dy));
    const float idz_ = 1.0f / (fabsf(dz) > ooeps ? dz : copysignf(ooeps, dz));
    const float oodx = ox * idx_, oody = oy * idy_, oodz = oz * idz_;
    int32_t leafAddr = 0, nodeAddr = 0, hitIndex = -1;
    uint32_t hitId = 0xFFFFFFFFu;
    uint32_t n_inner = 0, n_tri = 0, n_leaf = 0;

    while (nodeAddr != SENTINEL) {
        while ((uint32_t)nodeAddr < (uint32_t)SENTINEL) {
            n_inner++;
            const f4 n0 = row(nodes, nodeAddr + 0), n1 = row(nodes, nodeAddr + 1);
            const f4 nz = row(nodes, nodeAddr + 2);
            const uint32_t* c = nodes + 4 * (size_t)(nodeAddr + 3);
            int32_t cx = (int32_t)c[0], cy = (int32_t)c[1];
            /* slab tests, :173-188, with n*idir - ood written as one fma */
            const float c0lox = fmaf(n0.x, idx_, -oodx), c0hix = fmaf(n0.y, idx_, -oodx);
            const float c0loy = fmaf(n0.z, idy_, -oody), c0hiy = fmaf(n0.w, idy_, -oody);
            const float c0loz = fmaf(nz.x, idz_, -oodz), c0hiz = fmaf(nz.y, idz_, -oodz);
            const float c1loz = fmaf(nz.z, idz_, -oodz), c1hiz = fmaf(nz.w, idz_, -oodz);
            const float c1lox = fmaf(n1.x, idx_, -oodx), c1hix = fmaf(n1.y, idx_, -oodx);
            const float c1loy = fmaf(n1.z, idy_, -oody), c1hiy = fmaf(n1.w, idy_, -oody);
            /* spanBeginKepler / spanEndKepler, cudatracerkernels.h:255-256 */
            const float c0min = fmaxf(fmaxf(fminf(c0lox, c0hix), fminf(c0loy, c0hiy)), fmaxf(fminf(c0loz, c0hiz), tmin));
            const float c0max = fminf(fminf(fmaxf(c0lox, c0hix), fmaxf(c0loy, c0hiy)), fminf(fmaxf(c0loz, c0hiz), hitT));
            const float c1min = fmaxf(fmaxf(fminf(c1lox, c1hix), fminf(c1loy, c1hiy)), fmaxf(fminf(c1loz, c1hiz), tmin));
            const float c1max = fminf(fminf(fmaxf(c1lox, c1hix), fmaxf(c1loy, c1hiy)), fminf(fmaxf(c1loz, c1hiz), hitT));
            const int swp = c1min < c0min;
            const int t0 = c0max >= c0min, t1 = c1max >= c1min;
            if (!t0 && !t1) {
                nodeAddr = stack[sp--];
            } else {
                nodeAddr = t0 ? cx : cy;
                if (t0 && t1) {
                    if (swp) { int32_t tmp = nodeAddr; nodeAddr = cy; cy = tmp; }
                    stack[++sp] = cy;
                }
            }
            if (nodeAddr < 0 && leafAddr >= 0) {
                leafAddr = nodeAddr;
                nodeAddr = stack[sp--];
            }
            if (!(leafAddr >= 0)) break;
        }
        while (leafAddr < 0) {
            for (int32_t triAddr = ~leafAddr;; triAddr += 3) {
                const f4 v00 = row(woop, triAddr + 0);
                if (f2u(v00.x) == 0x80000000u) { n_leaf++; break; }
                const f4 v11 = row(woop, triAddr + 1), v22 = row(woop, triAddr + 2);
                n_tri++;
                /* Woop test, :260-282, dot products written as fma chains */
                const float Oz = fmaf(-oz, v00.z, fmaf(-oy, v00.y, fmaf(-ox, v00.x, v00.w)));
                const float invDz = 1.0f / fmaf(dz, v00.z, fmaf(dy, v00.y, dx * v00.x));
                const float t = Oz * invDz;
                if (t > tmin && (t < hitT || (hitIndex != -1 && t == hitT))) {
                    const float Ox = fmaf(oz, v11.z, fmaf(oy, v11.y, fmaf(ox, v11.x, v11.w)));
                    const float Dx = fmaf(dz, v11.z, fmaf(dy, v11.y, dx * v11.x));
                    const float u = fmaf(t, Dx, Ox);
                    if (u >= 0.0f) {
                        const float Oy = fmaf(oz, v22.z, fmaf(oy, v22.y, fmaf(ox, v22.x, v22.w)));
                        const float Dy = fmaf(dz, v22.z, fmaf(dy, v22.y, dx * v22.x));
                        const float v = fmaf(t, Dy, Oy);
                        if (v >= 0.0f && u + v <= 1.0f) {
                            const uint32_t id = tidx[triAddr];
                            if (t < hitT || id < hitId) {
                                hitT = t; hitIndex = triAddr; hitId = id;
                                if (any_hit) { nodeAddr = SENTINEL; break; }
                            }
                        }
                    }
                }
            }
            leafAddr = nodeAddr;
            if (nodeAddr < 0) nodeAddr = stack[sp--];
        }
    }
    *out_id = hitIndex == -1 ? -1 : (int32_t)hitId;
    *out_t = hitT;
    if (cnt) { cnt[0] = n_inner; cnt[1] = n_tri; cnt[2] = n_leaf; }
}

void orc_trace_compact2(const uint32_t* nodes, const uint32_t* woop, const uint32_t* tri_index,
                        const float* rays, uint32_t n, int any_hit,
                        int32_t* out_id, float* out_t, uint32_t* counts, int nthreads)
{
#ifdef _OPENMP
    int nt = nthreads > 0 ? nthreads : omp_get_max_threads();
#pragma omp parallel for schedule(dynamic, 256) num_threads(nt)
#else
    (void)nthreads;
#endif
    for (long i = 0; i < (long)n; i++)
        trace_one(nodes, woop, tri_index, rays + 8 * (size_t)i, any_hit, out_id + i, out_t + i,
                  counts ? counts + 3 * (size_t)i : NULL);
}

/* ------------------------------------------------------------------------ */
/* FP64 brute force                                                          */
/* ------------------------------------------------------------------------ */

void orc_bruteforce_f64(const float* verts, const uint32_t* idx, uint32_t nt,
                        const float* rays, uint32_t n, int32_t* out_id, double* out_t64,
                        double* out_margin, double* out_gap, int nthreads)
{
#ifdef _OPENMP
    int nth = nthreads > 0 ? nthreads : omp_get_max_threads();
#pragma omp parallel for schedule(dynamic, 16) num_threads(nth)
#else
    (void)nthreads;
#endif
    for (long r = 0; r < (long)n; r++) {
        const float* R = rays + 8 * (size_t)r;
        double o[3] = {R[0], R[1], R[2]}, d[3] = {R[4], R[5], R[6]};
        double tmin = R[3], tmax = R[7];
        double best = INFINITY, second = INFINITY, margin = 0.0;
        int32_t bid = -1;
        for (uint32_t k = 0; k < nt; k++) {
            const float* a = verts + 3 * (size_t)idx[3 * (size_t)k];
            const float* b = verts + 3 * (size_t)idx[3 * (size_t)k + 1];
            const float* c = verts + 3 * (size_t)idx[3 * (size_t)k + 2];
            double e1[3] = {(double)b[0] - a[0], (double)b[1] - a[1], (double)b[2] - a[2]};
            double e2[3] = {(double)c[0] - a[0], (double)c[1] - a[1], (double)c[2] - a[2]};
            double p[3] = {d[1] * e2[2] - d[2] * e2[1], d[2] * e2[0] - d[0] * e2[2], d[0] * e2[1] - d[1] * e2[0]};
            double det = e1[0] * p[0] + e1[1] * p[1] + e1[2] * p[2];
            if (det == 0.0) continue;
            double id = 1.0 / det;
            double tv[3] = {o[0] - a[0], o[1] - a[1], o[2] - a[2]};
            double u = (tv[0] * p[0] + tv[1] * p[1] + tv[2] * p[2]) * id;
            if (u < 0.0 || u > 1.0) continue;
            double q[3] = {tv[1] * e1[2] - tv[2] * e1[1], tv[2] * e1[0] - tv[0] * e1[2], tv[0] * e1[1] - tv[1] * e1[0]};
            double v = (d[0] * q[0] + d[1] * q[1] + d[2] * q[2]) * id;
            if (v < 0.0 || u + v > 1.0) continue;
            double t = (e2[0] * q[0] + e2[1] * q[1] + e2[2] * q[2]) * id;
            if (!(t > tmin && t < tmax)) continue;
            double m = u < v ? u : v;
            if (1.0 - u - v < m) m = 1.0 - u - v;
            if (t < best || (t == best && (int32_t)k < bid)) {
                if (t < best) second = best;
                best = t; bid = (int32_t)k; margin = m;
            } else if (t < second) {
                second = t;
            }
        }
        out_id[r] = bid;
        out_t64[r] = bid < 0 ? (double)tmax : best;
        if (out_margin) out_margin[r] = bid < 0 ? 1.0 : margin;
        if (out_gap) out_gap[r] = bid < 0 ? INFINITY : (second - best) / (best > 1e-30 ? best : 1e-30);
    }
}

/* ------------------------------------------------------------------------ */
/* Philox4x32-10 (Salmon, Moraes, Dror, Shaw, "Parallel random numbers: as   */
/* easy as 1, 2, 3", SC'11): multipliers 0xD2511F53 / 0xCD9E8D57, Weyl key   */
/* increments 0x9E3779B9 / 0xBB67AE85, ten rounds.                           */
/* ------------------------------------------------------------------------ */

void orc_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4])
{
    uint32_t c0 = ctr[0], c1 = ctr[1], c2 = ctr[2], c3 = ctr[3], k0 = key[0], k1 = key[1];
    for (int r = 0; r < 10; r++) {
        uint64_t p0 = (uint64_t)0xD2511F53u * c0;
        uint64_t p1 = (uint64_t)0xCD9E8D57u * c2;
        uint32_t n0 = (uint32_t)(p1 >> 32) ^ c1 ^ k0;
        uint32_t n1 = (uint32_t)p1;
        uint32_t n2 = (uint32_t)(p0 >> 32) ^ c3 ^ k1;
        uint32_t n3 = (uint32_t)p0;
        c0 = n0; c1 = n1; c2 = n2; c3 = n3;
        k0 += 0x9E3779B9u; k1 += 0xBB67AE85u;
    }
    out[0] = c0; out[1] = c1; out[2] = c2; out[3] = c3;
}

/* ------------------------------------------------------------------------ */
/* Sampling (shared conventions of the passes).  The reference draws from    */
/* Thrust default_random_engine / cuRAND (src/core/montecarlo.cu, integrators)*/
/* whose streams cannot be reproduced here; both sides use Philox4x32-10     */
/* keyed (seed, stream tag) with counters (index, sample, bounce, dim).      */
/* ------------------------------------------------------------------------ */

typedef struct { uint32_t k0, k1; } orc_rng;

static orc_rng rng_make(uint64_t seed, uint32_t tag)
{
    orc_rng r; r.k0 = (uint32_t)seed; r.k1 = (uint32_t)(seed >> 32) ^ (tag * 0x85EBCA6Bu);
    return r;
}

static void rng_draw(orc_rng r, uint32_t a, uint32_t b, uint32_t c, uint32_t d, uint32_t out[4])
{
    uint32_t ctr[4] = {a, b, c, d}, key[2] = {r.k0, r.k1};
    orc_philox4x32_10(ctr, key, out);
}

static inline float u01(uint32_t x) { return (float)(x >> 8) * 5.9604644775390625e-08f; }

/* sin/cos of 2*pi*u: quarter-turn reduction + degree-9/10 Taylor series. */
void orc_sincos_2pi(float u, float* s, float* c)
{
    const float a = u * 4.0f;
    const int k = (int)(a + 0.5f);
    const float r = a - (float)k;
    const float x = r * 1.57079632679489662f;
    const float x2 = x * x;
    float ps = fmaf(x2, 2.75573192e-06f, -1.98412698e-04f);
    ps = fmaf(x2, ps, 8.33333333e-03f);
    ps = fmaf(x2, ps, -1.66666667e-01f);
    const float sn = fmaf(x * x2, ps, x);
    float pc = fmaf(x2, -2.75573192e-07f, 2.48015873e-05f);
    pc = fmaf(x2, pc, -1.38888889e-03f);
    pc = fmaf(x2, pc, 4.16666667e-02f);
    pc = fmaf(x2, pc, -0.5f);
    const float cs = fmaf(x2, pc, 1.0f);
    switch (k & 3) {
    case 0: *s = sn; *c = cs; break;
    case 1: *s = cs; *c = -sn; break;
    case 2: *s = -sn; *c = -cs; break;
    default: *s = -cs; *c = sn; break;
    }
}

/* Duff et al. 2017 orthonormal basis + Malley cosine-weighted direction. */
static f3 cosine_dir(f3 n, float u1, float u2)
{
    float s, c;
    orc_sincos_2pi(u2, &s, &c);
    const float r = sqrtf(u1);
    const float x = r * c, y = r * s;
    const float zz = 1.0f - u1;
    const float z = sqrtf(zz > 0.0f ? zz : 0.0f);
    const float sign = copysignf(1.0f, n.z);
    const float a = -1.0f / (sign + n.z);
    const float b = n.x * n.y * a;
    f3 b1 = mk3(1.0f + sign * n.x * n.x * a, sign * b, -sign * n.x);
    f3 b2 = mk3(b, sign + n.y * n.y * a, -n.y);
    return add3(add3(mul3s(b1, x), mul3s(b2, y)), mul3s(n, z));
}

void orc_cosine_dir(const float* n, float u1, float u2, float* out)
{
    f3 d = cosine_dir(mk3(n[0], n[1], n[2]), u1, u2);
    out[0] = d.x; out[1] = d.y; out[2] = d.z;
}

/* ------------------------------------------------------------------------ */
/* Camera: PerspectiveCamera ctor + generate_ray, src/core/camera.cu:30-98,  */
/* transforms of src/core/transform.cu:28-126 and transform.h:168-197.       */
/* ------------------------------------------------------------------------ */

typedef struct { float m[16]; } m4;
typedef struct { m4 m, inv; } xf4;

static m4 m4_ident(void) { m4 r; memset(&r, 0, sizeof(r)); r.m[0] = r.m[5] = r.m[10] = r.m[15] = 1.0f; return r; }
static m4 m4_mul(const m4* a, const m4* b)
{
    m4 r;
    for (int i = 0; i < 4; i++)
        for (int j = 0; j < 4; j++)
            r.m[4 * i + j] = a->m[4 * i] * b->m[j] + a->m[4 * i + 1] * b->m[4 + j] + a->m[4 * i + 2] * b->m[8 + j] + a->m[4 * i + 3] * b->m[12 + j];
    return r;
}
static m4 m4_inv(const m4* a) { m4 r; orc_invert4x4(a->m, r.m); return r; }
static xf4 xf_of(m4 m) { xf4 t; t.m = m; t.inv = m4_inv(&m); return t; }
static xf4 xf_mul(const xf4* a, const xf4* b) { xf4 t; t.m = m4_mul(&a->m, &b->m); t.inv = m4_mul(&b->inv, &a->inv); return t; }
static xf4 xf_inverse(const xf4* a) { xf4 t; t.m = a->inv; t.inv = a->m; return t; }
static xf4 xf_scale(float x, float y, float z) { m4 r = m4_ident(); r.m[0] = x; r.m[5] = y; r.m[10] = z; r.m[15] = 1.0f; return xf_of(r); }
static xf4 xf_translate(float x, float y, float z) { m4 r = m4_ident(); r.m[3] = x; r.m[7] = y; r.m[11] = z; return xf_of(r); }

void orc_camera_matrices(const float* pos, const float* at, const float* up, float hfov,
                         uint32_t width, uint32_t height, float* r2c, float* c2w)
{
    f3 P = mk3(pos[0], pos[1], pos[2]), A = mk3(at[0], at[1], at[2]), U = mk3(up[0], up[1], up[2]);
    /* lookat, transform.cu:78-107 */
    f3 dir = normalize3(sub3(A, P));
    f3 left = normalize3(cross3(normalize3(U), dir));
    f3 nup = cross3(dir, left);
    m4 cw;
    float mm[16] = {left.x, nup.x, dir.x, P.x, left.y, nup.y, dir.y, P.y, left.z, nup.z, dir.z, P.z, 0.0f, 0.0f, 0.0f, 1.0f};
    memcpy(cw.m, mm, sizeof(mm));
    xf4 look; look.m = m4_inv(&cw); look.inv = cw;
    xf4 s = xf_scale(-1.0f, 1.0f, 1.0f);
    /* camera-space handedness flip (scale * lookat); the reference's
     * lookat * scale (camera.cu:39) mirrors the world about x = 0 instead */
    xf4 w2c = xf_mul(&s, &look);
    xf4 cam2world = xf_inverse(&w2c);
    /* perspective, transform.cu:114-124 */
    m4 pp; memset(&pp, 0, sizeof(pp));
    const float n = 1.0f, f = 1000.0f;
    pp.m[0] = 1.0f; pp.m[5] = 1.0f; pp.m[10] = f / (f - n); pp.m[11] = -f * n / (f - n); pp.m[14] = 1.0f;
    const float to_rad = 3.14159265358979323846f / 180.f;
    float inv_tan = 1.0f / tanf(to_rad * hfov / 2.0f);
    xf4 sc = xf_scale(inv_tan, inv_tan, 1.0f), pr = xf_of(pp);
    xf4 cam2screen = xf_mul(&sc, &pr);
    xf4 a = xf_scale((float)width, (float)height, 1.0f), b = xf_scale(0.5f, -0.5f, 1.0f), t = xf_translate(1.0f, -1.0f, 0.0f);
    xf4 ab = xf_mul(&a, &b);
    xf4 screen2raster = xf_mul(&ab, &t);
    xf4 raster2screen = xf_inverse(&screen2raster);
    xf4 screen2cam = xf_inverse(&cam2screen);
    xf4 raster2cam = xf_mul(&screen2cam, &raster2screen);
    memcpy(r2c, raster2cam.m.m, 64);
    memcpy(c2w, cam2world.m.m, 64);
}

static f3 xform_point(const float* m, float x, float y, float z)
{
    float r[4];
    for (int i = 0; i < 4; i++) r[i] = m[4 * i] * x + m[4 * i + 1] * y + m[4 * i + 2] * z + m[4 * i + 3] * 1.0f;
    return mk3(r[0] / r[3], r[1] / r[3], r[2] / r[3]);
}

static f3 xform_vector(const float* m, f3 v)
{
    float r[3];
    for (int i = 0; i < 3; i++) r[i] = m[4 * i] * v.x + m[4 * i + 1] * v.y + m[4 * i + 2] * v.z + m[4 * i + 3] * 0.0f;
    return mk3(r[0], r[1], r[2]);
}

/* kernel_generate_perspective_camera_samples, src/renderers/samplerrenderer.cu:48-96:
 * ray (pixel-major, sample-minor) = generate_ray(px + jx, py + jy). */
void orc_raygen_camera(const float* r2c, const float* c2w, uint32_t width, uint32_t x0, uint32_t y0,
                       uint32_t w, uint32_t h, uint32_t spp, uint64_t seed, int jitter, float* rays)
{
    orc_rng rng = rng_make(seed, 1u);
    for (uint64_t i = 0; i < (uint64_t)w * h * spp; i++) {
        uint32_t s = (uint32_t)(i % spp), p = (uint32_t)(i / spp);
        uint32_t px = x0 + p % w, py = y0 + p / w;
        float jx = 0.5f, jy = 0.5f;
        if (jitter) {
            uint32_t r[4];
            rng_draw(rng, py * width + px, s, 0, 0, r);
            jx = u01(r[0]); jy = u01(r[1]);
        }
        f3 pc = xform_point(r2c, (float)px + jx, (float)py + jy, 0.0f);
        f3 d = normalize3(pc);
        f3 o = xform_point(c2w, 0.0f, 0.0f, 0.0f);
        f3 dw = xform_vector(c2w, d);
        float* R = rays + 8 * i;
        R[0] = o.x; R[1] = o.y; R[2] = o.z; R[3] = 0.0f;
        R[4] = dw.x; R[5] = dw.y; R[6] = dw.z; R[7] = FLT_MAX;
    }
}

/* Incoherent diffuse rays from primary hits (the trace benchmark's second
 * set): origin = hit point of primary i (a Philox-redrawn primary when i
 * missed), direction cosine-distributed about the geometric normal turned
 * against the incoming ray, tmin = 1e-3 t (the epsilon convention of
 * kepler_dynamic_fetch.h:356). */
void orc_raygen_diffuse(const float* verts, const uint32_t* idx, const float* prim, const int32_t* hit_id,
                        const float* hit_t, uint32_t np, uint32_t nout, uint64_t seed, float* out)
{
    orc_rng rng = rng_make(seed, 3u), redraw = rng_make(seed, 4u);
    for (uint32_t i = 0; i < nout; i++) {
        uint32_t j = i % np;
        for (uint32_t att = 0; hit_id[j] < 0 && att < 64; att++) {
            uint32_t r[4];
            rng_draw(redraw, i, att, 0, 0, r);
            j = r[0] % np;
        }
        float* R = out + 8 * (size_t)i;
        if (hit_id[j] < 0) {
            R[0] = R[1] = R[2] = R[3] = 0.0f; R[4] = 0.0f; R[5] = 0.0f; R[6] = 1.0f; R[7] = 0.0f;
            continue;
        }
        const float* o = prim + 8 * (size_t)j;
        const float t = hit_t[j];
        f3 p = add3(mk3(o[0], o[1], o[2]), mul3s(mk3(o[4], o[5], o[6]), t));
        const uint32_t* ind = idx + 3 * (size_t)hit_id[j];
        f3 v0 = mk3(verts[3 * ind[0]], verts[3 * ind[0] + 1], verts[3 * ind[0] + 2]);
        f3 v1 = mk3(verts[3 * ind[1]], verts[3 * ind[1] + 1], verts[3 * ind[1] + 2]);
        f3 v2 = mk3(verts[3 * ind[2]], verts[3 * ind[2] + 1], verts[3 * ind[2] + 2]);
        f3 n = normalize3(cross3(sub3(v1, v0), sub3(v2, v0)));
        if (dot3(n, mk3(o[4], o[5], o[6])) > 0.0f) n = mk3(-n.x, -n.y, -n.z);
        uint32_t r[4];
        rng_draw(rng, i, 0, 0, 0, r);
        f3 d = cosine_dir(n, u01(r[0]), u01(r[1]));
        R[0] = p.x; R[1] = p.y; R[2] = p.z; R[3] = 1.0e-3f * t;
        R[4] = d.x; R[5] = d.y; R[6] = d.z; R[7] = FLT_MAX;
    }
}

/* The Woop ray/triangle test of kepler_dynamic_fetch.h:260-282 on one
 * triangle (Woop rows from orc_woopify), with the traversal's arithmetic.
 * Returns 1 on a hit with tmin < t < tmax, u, v >= 0, u + v <= 1. */
int orc_woop_hit(const float* v0, const float* v1, const float* v2, const float* ray, float* t_out)
{
    float w[12];
    orc_woopify(v0, v1, v2, w);
    if (w[0] == 0.0f) w[0] = 0.0f;
    const float ox = ray[0], oy = ray[1], oz = ray[2], tmin = ray[3], dx = ray[4], dy = ray[5], dz = ray[6];
    const float Oz = fmaf(-oz, w[2], fmaf(-oy, w[1], fmaf(-ox, w[0], w[3])));
    const float invDz = 1.0f / fmaf(dz, w[2], fmaf(dy, w[1], dx * w[0]));
    const float t = Oz * invDz;
    *t_out = t;
    if (!(t > tmin && t < ray[7])) return 0;
    const float u = fmaf(t, fmaf(dz, w[6], fmaf(dy, w[5], dx * w[4])), fmaf(oz, w[6], fmaf(oy, w[5], fmaf(ox, w[4], w[7]))));
    if (!(u >= 0.0f)) return 0;
    const float v = fmaf(t, fmaf(dz, w[10], fmaf(dy, w[9], dx * w[8])), fmaf(oz, w[10], fmaf(oy, w[9], fmaf(ox, w[8], w[11]))));
    return v >= 0.0f && u + v <= 1.0f;
}

void orc_woop_hits(const float* verts, const uint32_t* idx, const int32_t* ids, const float* rays, uint32_t n,
                   float* out_t, uint8_t* out_hit)
{
    for (uint32_t i = 0; i < n; i++) {
        if (ids[i] < 0) { out_hit[i] = 0; out_t[i] = 0.0f; continue; }
        const uint32_t* ind = idx + 3 * (size_t)ids[i];
        out_hit[i] = (uint8_t)orc_woop_hit(verts + 3 * (size_t)ind[0], verts + 3 * (size_t)ind[1],
                                           verts + 3 * (size_t)ind[2], rays + 8 * (size_t)i, out_t + i);
    }
}

/* ------------------------------------------------------------------------ */
/* Render passes: AO (src/integrators/integrator_ao.cu:42-85), path tracer   */
/* (src/integrators/integrator_path.cu:43-261), film average                 */
/* (src/renderers/samplerrenderer.cu:103-169 with the box filter).           */
/* ------------------------------------------------------------------------ */

typedef struct { uint32_t kind; float color[3]; float param[8]; } orc_material;

static orc_material* mats_from12(const float* m12, uint32_t nmats)
{
    orc_material* mats = (orc_material*)calloc(nmats ? nmats : 1, sizeof(orc_material));
    for (uint32_t i = 0; i < nmats; i++) {
        uint32_t k; memcpy(&k, &m12[12 * i], 4);
        mats[i].kind = k;
        for (int c = 0; c < 3; c++) mats[i].color[c] = m12[12 * i + 1 + c];
        for (int c = 0; c < 8; c++) mats[i].param[c] = m12[12 * i + 4 + c];
    }
    return mats;
}


typedef struct {
    const uint32_t *nodes, *woop, *tidx;      /* Compact2 */
    const float* verts; const uint32_t* idx;  /* scene */
    const uint32_t* tri_mat; const orc_material* mats;
    const uint32_t* lights; uint32_t nl;
    /* per-triangle shading normals (n0, n1, n2: 9 floats) and which
     * triangles have them; NULL: every triangle shades with its geometric
     * normal */
    const float* tri_n; const uint8_t* has_n;
} orc_scene;

static void tri_verts(const orc_scene* S, uint32_t t, f3* a, f3* b, f3* c)
{
    const uint32_t* ind = S->idx + 3 * (size_t)t;
    *a = mk3(S->verts[3 * (size_t)ind[0]], S->verts[3 * (size_t)ind[0] + 1], S->verts[3 * (size_t)ind[0] + 2]);
    *b = mk3(S->verts[3 * (size_t)ind[1]], S->verts[3 * (size_t)ind[1] + 1], S->verts[3 * (size_t)ind[1] + 2]);
    *c = mk3(S->verts[3 * (size_t)ind[2]], S->verts[3 * (size_t)ind[2] + 1], S->verts[3 * (size_t)ind[2] + 2]);
}

static f3 tri_norm(f3 a, f3 b, f3 c) { return normalize3(cross3(sub3(b, a), sub3(c, a))); }

/* Barycentric weights u (of a) and v (of b; c takes 1 - u - v) of a point p
 * of the triangle's plane, from the normal equations of
 * p - c = u (a - c) + v (b - c): the u, v convention of the Woop transform
 * (M = [v0 - v2 | v1 - v2 | n | v2], cudabvh.cpp:303-322). */
static void bary_uv(f3 p, f3 a, f3 b, f3 c, float* u, float* v)
{
    f3 e0 = sub3(a, c), e1 = sub3(b, c), q = sub3(p, c);
    float d00 = dot3(e0, e0), d01 = dot3(e0, e1), d11 = dot3(e1, e1);
    float d20 = dot3(q, e0), d21 = dot3(q, e1);
    float den = d00 * d11 - d01 * d01;
    *u = den != 0.0f ? (d11 * d20 - d01 * d21) / den : 0.0f;
    *v = den != 0.0f ? (d00 * d21 - d01 * d20) / den : 0.0f;
}

/* Shading normal of triangle t at p (the hit's IntersectionBuffer
 * m_shading_normal, kepler_dynamic_fetch.h:327-334): the vertex normals
 * weighted by p's barycentrics, u n0 + v n1 + (1 - u - v) n2, normalised
 * (the reference weights n0, n0, n1 and does not normalise -- a slip this
 * restatement does not keep).  *ns = ng and 0 returned for a triangle
 * without normals or where the weighted sum vanishes; 1 when *ns is the
 * interpolated normal.  Not oriented: it points where the mesh's normals do. */
static int interp_normal(const orc_scene* S, uint32_t t, f3 p, f3 a, f3 b, f3 c, f3 ng, f3* ns)
{
    *ns = ng;
    if (!S->tri_n || !S->has_n[t]) return 0;
    const float* n = S->tri_n + 9 * (size_t)t;
    float u, v;
    bary_uv(p, a, b, c, &u, &v);
    float w = 1.0f - u - v;
    f3 s = mk3(n[0] * u + n[3] * v + n[6] * w, n[1] * u + n[4] * v + n[7] * w, n[2] * u + n[5] * v + n[8] * w);
    float l2 = dot3(s, s);
    if (!(l2 > 0.0f)) return 0;
    *ns = mul3s(s, 1.0f / sqrtf(l2));
    return 1;
}

/* The shading normal the passes shade with: interp_normal turned to the
 * geometric normal's side (so "facing the ray" means the same for both). */
static int shade_normal(const orc_scene* S, uint32_t t, f3 p, f3 a, f3 b, f3 c, f3 ng, f3* ns)
{
    int sn = interp_normal(S, t, p, a, b, c, ng, ns);
    if (sn && dot3(*ns, ng) < 0.0f) *ns = mk3(-ns->x, -ns->y, -ns->z);
    return sn;
}

/* Importance transport's shading-normal factor for a light sub-path vertex
 * (Veach 1997, sec. 5.3): |wo.ns| |wi.ng| / (|wo.ng| |wi.ns|), wo towards the
 * light, wi the continuation; 0 where the denominator is. */
static float adjoint_factor(f3 wo, f3 wi, f3 ng, f3 ns)
{
    float num = fabsf(dot3(wo, ns)) * fabsf(dot3(wi, ng));
    float den = fabsf(dot3(wo, ng)) * fabsf(dot3(wi, ns));
    return den != 0.0f ? num / den : 0.0f;
}

static void trace_ray8(const orc_scene* S, const float* r, int any, int32_t* id, float* t)
{
    trace_one(S->nodes, S->woop, S->tidx, r, any, id, t, NULL);
}

static uint32_t sample_gid(uint32_t width, uint32_t x0, uint32_t y0, uint32_t w, uint32_t spp, uint32_t i)
{
    uint32_t p = i / spp, s = i % spp;
    return ((y0 + p / w) * width + (x0 + p % w)) * spp + s;
}

/* ------------------------------------------------------------------------ */
/* Film: Film::Film's filter table (src/core/film.cpp:45-58, x offsets from  */
/* the x index) and kernel_update_film's splat (src/renderers/               */
/* samplerrenderer.cu:103-169) restated per output pixel: pixel X gets the   */
/* sample at raster position d + 0.5 when ceil(d - w) <= X <= floor(d + w)  */
/* on both axes, with weight table[|X - d| / w * 16]; output                */
/* max(0, sum(wL) * (1 / sum(w))) (film.cpp:68-86).  Filters of              */
/* src/core/filter.h:69-132.  The box of half-width 0.5 is the mean of the  */
/* pixel's own samples.  Sums run over neighbour rows, columns, samples.     */
/* ------------------------------------------------------------------------ */

static float filter_eval(uint32_t kind, float xw, float yw, float alpha, float x, float y)
{
    if (kind == 1u) return fmaxf(0.0f, xw - fabsf(x)) * fmaxf(0.0f, yw - fabsf(y));
    if (kind == 2u) {
        float ex = expf(-alpha * xw * xw), ey = expf(-alpha * yw * yw);
        return fmaxf(0.0f, expf(-alpha * x * x) - ex) * fmaxf(0.0f, expf(-alpha * y * y) - ey);
    }
    return 1.0f;
}

void orc_film(const float* Ls, uint32_t width, uint32_t x0, uint32_t y0, uint32_t w, uint32_t h, uint32_t spp,
              uint64_t seed, uint32_t kind, float xw, float yw, float alpha, float* out_rgb)
{
    if (kind == 0u && xw == 0.5f && yw == 0.5f) {
        for (uint32_t px = 0; px < w * h; px++) {
            float r = 0.0f, g = 0.0f, b = 0.0f;
            for (uint32_t s = 0; s < spp; s++) {
                const float* L = Ls + 3 * ((size_t)px * spp + s);
                r += L[0]; g += L[1]; b += L[2];
            }
            float inv = (float)spp;
            out_rgb[3 * px] = r / inv; out_rgb[3 * px + 1] = g / inv; out_rgb[3 * px + 2] = b / inv;
        }
        return;
    }
    float table[256];
    for (int y = 0; y < 16; y++)
        for (int x = 0; x < 16; x++)
            table[y * 16 + x] = filter_eval(kind, xw, yw, alpha, ((float)x + 0.5f) * xw / 16.0f,
                                            ((float)y + 0.5f) * yw / 16.0f);
    const float ixw = 1.0f / xw, iyw = 1.0f / yw;
    const int rx = (int)ceilf(xw) + 1, ry = (int)ceilf(yw) + 1;
    orc_rng rng = rng_make(seed, 1u);
#pragma omp parallel for schedule(dynamic, 16)
    for (long px = 0; px < (long)(w * h); px++) {
        int lx = (int)(px % w), ly = (int)(px / w);
        float X = (float)(x0 + (uint32_t)lx), Y = (float)(y0 + (uint32_t)ly);
        float r = 0.0f, g = 0.0f, b = 0.0f, ws = 0.0f;
        for (int ny = ly - ry; ny <= ly + ry; ny++) {
            if (ny < 0 || ny >= (int)h) continue;
            uint32_t gy = y0 + (uint32_t)ny;
            for (int nx = lx - rx; nx <= lx + rx; nx++) {
                if (nx < 0 || nx >= (int)w) continue;
                uint32_t gx = x0 + (uint32_t)nx;
                for (uint32_t s = 0; s < spp; s++) {
                    uint32_t q[4];
                    rng_draw(rng, gy * width + gx, s, 0, 0, q);
                    float dx = ((float)gx + u01(q[0])) - 0.5f, dy = ((float)gy + u01(q[1])) - 0.5f;
                    if (ceilf(dx - xw) > X || X > floorf(dx + xw)) continue;
                    if (ceilf(dy - yw) > Y || Y > floorf(dy + yw)) continue;
                    int ix = (int)floorf(fabsf((X - dx) * ixw * 16.0f));
                    int iy = (int)floorf(fabsf((Y - dy) * iyw * 16.0f));
                    if (ix > 15) ix = 15;
                    if (iy > 15) iy = 15;
                    float wt = table[iy * 16 + ix];
                    const float* L = Ls + 3 * (((size_t)ny * w + (size_t)nx) * spp + s);
                    r += wt * L[0]; g += wt * L[1]; b += wt * L[2];
                    ws += wt;
                }
            }
        }
        float o[3] = {0.0f, 0.0f, 0.0f};
        if (ws != 0.0f) {
            float inv = 1.0f / ws;
            o[0] = fmaxf(0.0f, r * inv); o[1] = fmaxf(0.0f, g * inv); o[2] = fmaxf(0.0f, b * inv);
        }
        out_rgb[3 * px] = o[0]; out_rgb[3 * px + 1] = o[1]; out_rgb[3 * px + 2] = o[2];
    }
}

/* AO: per sample, nao cosine rays from the hit (normal turned towards the
 * viewer), tmin = 1e-3 t, tmax = radius; value = unoccluded / nao. */
void orc_render_ao(const uint32_t* nodes, const uint32_t* woop, const uint32_t* tidx, const float* verts,
                   const uint32_t* idx, const float* r2c, const float* c2w, uint32_t width, uint32_t height,
                   uint32_t x0, uint32_t y0, uint32_t w, uint32_t h, uint32_t spp, uint64_t seed, uint32_t nao,
                   float radius, uint32_t filter, float fxw, float fyw, float falpha, float* out_rgb,
                   const float* tri_n, const uint8_t* has_n)
{
    (void)height;
    orc_scene S = {nodes, woop, tidx, verts, idx, NULL, NULL, NULL, 0, tri_n, has_n};
    uint32_t n = w * h * spp;
    float* rays = (float*)malloc((size_t)n * 32);
    float* Ls = (float*)malloc((size_t)n * 12);
    orc_raygen_camera(r2c, c2w, width, x0, y0, w, h, spp, seed, 1, rays);
    orc_rng rng = rng_make(seed, 2u);
#pragma omp parallel for schedule(dynamic, 16)
    for (long px = 0; px < (long)(w * h); px++) {
        for (uint32_t s = 0; s < spp; s++) {
            uint32_t i = (uint32_t)px * spp + s;
            const float* R = rays + 8 * (size_t)i;
            int32_t id; float t;
            trace_ray8(&S, R, 0, &id, &t);
            float c = 0.0f;
            if (id >= 0) {
                f3 d = mk3(R[4], R[5], R[6]);
                f3 p = add3(mk3(R[0], R[1], R[2]), mul3s(d, t));
                f3 a, b, cc;
                tri_verts(&S, (uint32_t)id, &a, &b, &cc);
                /* kernel_ambient_occlusion samples about the shading normal
                 * (integrator_ao.cu:55), here turned to face the ray */
                f3 ng = tri_norm(a, b, cc), nn;
                shade_normal(&S, (uint32_t)id, p, a, b, cc, ng, &nn);
                if (dot3(ng, d) > 0.0f) nn = mk3(-nn.x, -nn.y, -nn.z);
                uint32_t open = 0;
                uint32_t gid = sample_gid(width, x0, y0, w, spp, i);
                for (uint32_t k = 0; k < nao; k++) {
                    uint32_t r[4];
                    rng_draw(rng, gid, k, 0, 0, r);
                    f3 dir = cosine_dir(nn, u01(r[0]), u01(r[1]));
                    float ar[8] = {p.x, p.y, p.z, 1.0e-3f * t, dir.x, dir.y, dir.z, radius};
                    int32_t aid; float at;
                    trace_ray8(&S, ar, 1, &aid, &at);
                    open += aid < 0 ? 1u : 0u;
                }
                c = (float)open / (float)nao;
            }
            Ls[3 * (size_t)i] = c; Ls[3 * (size_t)i + 1] = c; Ls[3 * (size_t)i + 2] = c;
        }
    }
    orc_film(Ls, width, x0, y0, w, h, spp, seed, filter, fxw, fyw, falpha, out_rgb);
    free(rays);
    free(Ls);
}

static float light_pdf_mt(f3 p, f3 d, f3 v0, f3 v1, f3 v2, f3 ln, float area)
{
    f3 e1 = sub3(v1, v0), e2 = sub3(v2, v0);
    f3 pv = cross3(d, e2);
    float det = dot3(e1, pv);
    if (det < 1.0e-6f && det > -1.0e-6f) return 0.0f;
    float inv = 1.0f / det;
    f3 tv = sub3(p, v0);
    float u = dot3(tv, pv) * inv;
    if (u < 0.0f || u > 1.0f) return 0.0f;
    f3 qv = cross3(tv, e1);
    float v = dot3(d, qv) * inv;
    if (v < 0.0f || u + v > 1.0f) return 0.0f;
    float t = dot3(e2, qv) * inv;
    if (!(t > 0.0f)) return 0.0f;
    f3 hp = add3(p, mul3s(d, t));
    f3 dd = sub3(hp, p);
    float r2 = dot3(dd, dd);
    float cth = fabsf(dot3(ln, mk3(-d.x, -d.y, -d.z)));
    return r2 / (area * cth);
}

static int black3(f3 c) { return c.x == 0.0f && c.y == 0.0f && c.z == 0.0f; }

/* ------------------------------------------------------------------------ */
/* pow for the BSDFs without a math library: 2^(y log2 x), log2 from        */
/* 2 atanh((m-1)/(m+1)) on m in [sqrt(1/2), sqrt(2)), 2^f from its Taylor   */
/* series on [-1/2, 1/2]; the renderer's kernels use the same operations.   */
/* ------------------------------------------------------------------------ */

static float log2_pos(float x)
{
    uint32_t u = f2u(x);
    int e = 0;
    if (u < 0x00800000u) { x = x * 16777216.0f; u = f2u(x); e = -24; }
    e += (int)(u >> 23) - 127;
    float m = u2f((u & 0x007FFFFFu) | 0x3F800000u);
    if (m > 1.41421356f) { m = m * 0.5f; e += 1; }
    const float t = (m - 1.0f) / (m + 1.0f);
    const float t2 = t * t;
    float q = fmaf(t2, 0.222222222f, 0.285714286f);
    q = fmaf(t2, q, 0.4f);
    q = fmaf(t2, q, 0.666666667f);
    q = fmaf(t2, q, 2.0f);
    return fmaf(t * q, 1.44269504f, (float)e);
}

static float exp2_series(float y)
{
    if (y != y) return y;
    if (y >= 128.0f) return u2f(0x7F800000u);
    if (y < -150.0f) return 0.0f;
    const float k = floorf(y + 0.5f);
    const float z = (y - k) * 0.693147181f;
    float q = fmaf(z, 1.98412698e-4f, 1.38888889e-3f);
    q = fmaf(z, q, 8.33333333e-3f);
    q = fmaf(z, q, 4.16666667e-2f);
    q = fmaf(z, q, 0.166666667f);
    q = fmaf(z, q, 0.5f);
    q = fmaf(z, q, 1.0f);
    q = fmaf(z, q, 1.0f);
    int ki = (int)k;
    if (ki > 127) { q = q * 2.0f; ki = 127; }
    if (ki < -126) { q = q * 5.96046448e-8f; ki += 24; }
    return q * u2f((uint32_t)(ki + 127) << 23);
}

float orc_powf(float x, float y)
{
    if (x > 0.0f) return exp2_series(y * log2_pos(x));
    if (x == 0.0f) return y > 0.0f ? 0.0f : (y == 0.0f ? 1.0f : u2f(0x7F800000u));
    return u2f(0x7FC00000u);
}

/* ------------------------------------------------------------------------ */
/* BSDFs (src/core/reflection.cu:42-300, one layer per material).  Non-delta */
/* ones reflect only, in the Duff et al. frame of the normal facing wo.      */
/* ------------------------------------------------------------------------ */

static int delta_mat(const orc_material* m) { return m->kind == 1u || m->kind == 3u; }

typedef struct { f3 b1, b2, n; } frame3;
static frame3 frame_of(f3 n)
{
    frame3 f;
    const float sign = copysignf(1.0f, n.z);
    const float a = -1.0f / (sign + n.z);
    const float b = n.x * n.y * a;
    f.b1 = mk3(1.0f + sign * n.x * n.x * a, sign * b, -sign * n.x);
    f.b2 = mk3(b, sign + n.y * n.y * a, -n.y);
    f.n = n;
    return f;
}
static f3 to_local(const frame3* f, f3 v) { return mk3(dot3(v, f->b1), dot3(v, f->b2), dot3(v, f->n)); }
static f3 to_world(const frame3* f, f3 l) { return add3(add3(mul3s(f->b1, l.x), mul3s(f->b2, l.y)), mul3s(f->n, l.z)); }

static float sin_theta_l(f3 w) { float s2 = 1.0f - w.z * w.z; return sqrtf(s2 > 0.0f ? s2 : 0.0f); }
static float cos_phi_l(f3 w)
{
    float st = sin_theta_l(w);
    if (st == 0.0f) return 1.0f;
    float c = w.x / st;
    return c < -1.0f ? -1.0f : (c > 1.0f ? 1.0f : c);
}
static float sin_phi_l(f3 w)
{
    float st = sin_theta_l(w);
    if (st == 0.0f) return 0.0f;
    float c = w.y / st;
    return c < -1.0f ? -1.0f : (c > 1.0f ? 1.0f : c);
}

/* evaluate_orennayar, reflection.cu:144-178 (A = param[1], B = param[2]) */
static f3 orennayar_local(f3 wi, f3 wo, const orc_material* m)
{
    float sti = sin_theta_l(wi), sto = sin_theta_l(wo);
    float maxcos = 0.0f;
    if (sti > 1.0e-4f && sto > 1.0e-4f) {
        float dcos = cos_phi_l(wi) * cos_phi_l(wo) + sin_phi_l(wi) * sin_phi_l(wo);
        maxcos = dcos > 0.0f ? dcos : 0.0f;
    }
    float sinalpha, tanbeta;
    if (fabsf(wi.z) > fabsf(wo.z)) { sinalpha = sto; tanbeta = sti / fabsf(wi.z); }
    else { sinalpha = sti; tanbeta = sto / fabsf(wo.z); }
    float k = m->param[1] + m->param[2] * maxcos * sinalpha * tanbeta;
    float ip = 0.318309886183790672f;
    return mk3(m->color[0] * ip * k, m->color[1] * ip * k, m->color[2] * ip * k);
}

/* Fresnel::evaluate, src/core/reflection.h:147-217 */
static f3 fresnel_layer(const orc_material* m, float cosi)
{
    if (m->param[1] == 1.0f) {
        float rc[3];
        for (int c = 0; c < 3; c++) {
            float eta = m->param[2 + c], k = m->param[5 + c];
            float temp = (eta * eta + k * k) * cosi * cosi;
            float rparl2 = (temp - 2.0f * eta * cosi + 1.0f) / (temp + 2.0f * eta * cosi + 1.0f);
            float tf = eta * eta + k * k;
            float rperp2 = (tf - 2.0f * eta * cosi + cosi * cosi) / (tf + 2.0f * eta * cosi + cosi * cosi);
            rc[c] = (rparl2 + rperp2) / 2.0f;
        }
        return mk3(rc[0], rc[1], rc[2]);
    }
    if (m->param[1] == 2.0f) {
        float ei = m->param[2], et = m->param[3];
        if (!(cosi > 0.0f)) { float t = ei; ei = et; et = t; }
        float c2 = 1.0f - cosi * cosi;
        float sint = ei / et * sqrtf(c2 > 0.0f ? c2 : 0.0f);
        if (sint >= 1.0f) return mk3(1.0f, 1.0f, 1.0f);
        float t2 = 1.0f - sint * sint;
        float cost = sqrtf(t2 > 0.0f ? t2 : 0.0f);
        float ci = fabsf(cosi);
        float rparl = ((et * ci) - (ei * cost)) / ((et * ci) + (ei * cost));
        float rperp = ((ei * ci) - (et * cost)) / ((ei * ci) + (et * cost));
        float f = (rparl * rparl + rperp * rperp) / 2.0f;
        return mk3(f, f, f);
    }
    return mk3(1.0f, 1.0f, 1.0f);
}

/* Blinn D, Torrance-Sparrow G with both cosines (reflection.cu:198-251) */
static float blinn_d(f3 wh, float e) { return (e + 2.0f) * 0.159154943091895336f * orc_powf(fabsf(wh.z), e); }
static float micro_g(f3 wi, f3 wo, f3 wh)
{
    float nh = fabsf(wh.z), no = fabsf(wo.z), ni = fabsf(wi.z), oh = fabsf(dot3(wo, wh));
    float g1 = 2.0f * nh * no / oh, g2 = 2.0f * nh * ni / oh;
    float g = g1 < g2 ? g1 : g2;
    return g < 1.0f ? g : 1.0f;
}
static f3 microfacet_local(f3 wi, f3 wo, const orc_material* m)
{
    float ci = fabsf(wi.z), co = fabsf(wo.z);
    if (ci == 0.0f || co == 0.0f) return mk3(0.0f, 0.0f, 0.0f);
    f3 wh = add3(wi, wo);
    if (wh.x == 0.0f && wh.y == 0.0f && wh.z == 0.0f) return mk3(0.0f, 0.0f, 0.0f);
    wh = normalize3(wh);
    f3 F = fresnel_layer(m, dot3(wi, wh));
    float D = blinn_d(wh, m->param[0]), G = micro_g(wi, wo, wh);
    float den = 4.0f * ci * co;
    return mk3(m->color[0] * D * G * F.x / den, m->color[1] * D * G * F.y / den, m->color[2] * D * G * F.z / den);
}
static float microfacet_pdf_local(f3 wi, f3 wo, float e)
{
    if (!(wi.z * wo.z > 0.0f)) return 0.0f;
    f3 wh = normalize3(add3(wo, wi));
    float d = dot3(wo, wh);
    if (d <= 0.0f) return 0.0f;
    return ((e + 1.0f) * orc_powf(fabsf(wh.z), e)) / (6.28318531f * 4.0f * d);
}


static f3 lambert_f3(f3 wo, f3 wi, f3 ng, const orc_material* m)
{
    if (!(dot3(wo, ng) * dot3(wi, ng) > 0.0f)) return mk3(0.0f, 0.0f, 0.0f);
    const float k = 0.318309886183790672f;
    return mk3(m->color[0] * k, m->color[1] * k, m->color[2] * k);
}

static f3 bsdf_f3(const orc_material* m, f3 wo, f3 wi, f3 ng, f3 nf)
{
    if (m->kind == 0u) return lambert_f3(wo, wi, ng, m);
    if (delta_mat(m) || m->kind == 2u) return mk3(0.0f, 0.0f, 0.0f);
    if (!(dot3(wo, ng) * dot3(wi, ng) > 0.0f)) return mk3(0.0f, 0.0f, 0.0f);
    frame3 fr = frame_of(nf);
    f3 wil = to_local(&fr, wi), wol = to_local(&fr, wo);
    return m->kind == 4u ? orennayar_local(wil, wol, m) : microfacet_local(wil, wol, m);
}

static float bsdf_pdf3(const orc_material* m, f3 wo, f3 wi, f3 nf)
{
    if (m->kind == 5u) {
        frame3 fr = frame_of(nf);
        return microfacet_pdf_local(to_local(&fr, wi), to_local(&fr, wo), m->param[0]);
    }
    return fabsf(dot3(wi, nf)) * 0.318309886183790672f;
}

static void bsdf_sample3(const orc_material* m, f3 wo, f3 ng, f3 nf, float u1, float u2, f3* wi, f3* f, float* pdf)
{
    if (m->kind != 5u) {
        *wi = cosine_dir(nf, u1, u2);
        *f = bsdf_f3(m, wo, *wi, ng, nf);
        *pdf = fabsf(dot3(*wi, nf)) * 0.318309886183790672f;
        return;
    }
    /* sample_blinn_microfacet, reflection.cu:208-225 */
    frame3 fr = frame_of(nf);
    f3 wol = to_local(&fr, wo);
    float e = m->param[0];
    float ct = orc_powf(u1, 1.0f / (e + 1.0f));
    float s2 = 1.0f - ct * ct;
    float st = sqrtf(s2 > 0.0f ? s2 : 0.0f);
    float sp, cp;
    orc_sincos_2pi(u2, &sp, &cp);
    f3 wh = mk3(st * cp, st * sp, ct);
    if (!(wol.z * wh.z > 0.0f)) wh = mk3(-wh.x, -wh.y, -wh.z);
    float d = dot3(wol, wh);
    f3 wil = add3(mk3(-wol.x, -wol.y, -wol.z), mul3s(wh, 2.0f * d));
    *pdf = d > 0.0f ? ((e + 1.0f) * orc_powf(ct, e)) / (6.28318531f * 4.0f * d) : 0.0f;
    *wi = normalize3(to_world(&fr, wil));
    *f = (wil.z * wol.z > 0.0f) ? bsdf_f3(m, wo, *wi, ng, nf) : mk3(0.0f, 0.0f, 0.0f);
}

/* Glass: Fresnel-weighted reflection / refraction about the outward normal
 * ng (sample_glass, reflection.cu:109-140, plus the reflection branch);
 * returns f |cos| / pdf; radiance transport scales refraction by
 * (eta_wo / eta_wi)^2. */
static f3 glass_sample3(const orc_material* m, f3 wo, f3 ng, float u, int radiance, f3* wi)
{
    frame3 fr = frame_of(ng);
    f3 wol = to_local(&fr, wo);
    float cosi = wol.z;
    int entering = cosi > 0.0f;
    float ei = entering ? m->param[0] : m->param[1], et = entering ? m->param[1] : m->param[0];
    orc_material fm = *m;
    fm.param[1] = 2.0f;
    fm.param[2] = m->param[0];
    fm.param[3] = m->param[1];
    float F = fresnel_layer(&fm, cosi).x;
    float s2 = 1.0f - cosi * cosi;
    float eta = ei / et;
    float st2 = eta * eta * (s2 > 0.0f ? s2 : 0.0f);
    f3 wil, w;
    if (u < F || st2 >= 1.0f) {
        wil = mk3(-wol.x, -wol.y, wol.z);
        w = mk3(m->color[0], m->color[1], m->color[2]);
    } else {
        float c2 = 1.0f - st2;
        float cost = sqrtf(c2 > 0.0f ? c2 : 0.0f);
        if (entering) cost = -cost;
        wil = mk3(eta * -wol.x, eta * -wol.y, cost);
        float sc = radiance ? (ei * ei) / (et * et) : 1.0f;
        w = mk3(m->color[0] * sc, m->color[1] * sc, m->color[2] * sc);
    }
    *wi = normalize3(to_world(&fr, wil));
    return w;
}

/* One sampled direction of a single-layer material (material_sample_f's
 * lobes, material.cu:98-135): a non-delta BSDF gives f and pdf at wi; glass
 * gives f |cos| / pdf in *f with *weight set; a mirror gives color / |cos|
 * with pdf 1 (sample_mirror, reflection.cu:76-92). */
static void layer_sample3(const orc_material* m, f3 wo, f3 ng, f3 ns, f3 nf, float u1, float u2, int radiance, f3* wi,
                          f3* f, float* pdf, int* spec, int* weight)
{
    *weight = 0;
    if (!delta_mat(m)) {
        bsdf_sample3(m, wo, ng, nf, u1, u2, wi, f, pdf);
        *spec = 0;
        return;
    }
    *spec = 1;
    *pdf = 1.0f;
    if (m->kind == 3u) {
        *f = glass_sample3(m, wo, ns, u1, radiance, wi);
        *weight = 1;
        return;
    }
    float c = 2.0f * dot3(nf, wo);
    *wi = sub3(mul3s(nf, c), wo);
    float ac = fabsf(dot3(*wi, nf));
    *f = mk3(m->color[0] / ac, m->color[1] / ac, m->color[2] / ac);
}

/* Layered materials (kind 6: Material's num_valid_layers layers,
 * src/core/material.h:44-52; layer choice material.cu:80-95).  The material
 * is the equal mixture of its n single-layer materials, f = (1/n) sum f_i.
 * A sampled direction comes from one layer picked uniformly; a specular
 * layer's sample keeps its own throughput factor, a non-specular one has
 * the mixture's f and pdf (1/n) sum pdf_i over the non-specular layers.
 * Light samples and connections evaluate the non-specular layers; the light
 * sample's MIS weight uses the non-specular technique's pdf
 * (1/n_ns) sum pdf_i.  The reference's material_f sums the layers while its
 * sampled continuation divides by the pick (material.cu:137-170, 264-326),
 * so its two estimators disagree for n > 1; this is the mean, which its
 * continuation estimates. */
static uint32_t mat_layers3(const orc_material* m) { return m->kind == 6u ? (uint32_t)m->param[0] : 1u; }
static const orc_material* mat_layer3(const orc_scene* S, const orc_material* m, uint32_t k)
{
    return &S->mats[(uint32_t)m->param[1 + k]];
}
static int mat_delta3(const orc_scene* S, const orc_material* m)
{
    if (m->kind != 6u) return delta_mat(m);
    for (uint32_t k = 0; k < mat_layers3(m); k++)
        if (!delta_mat(mat_layer3(S, m, k))) return 0;
    return 1;
}
static f3 mat_f3(const orc_scene* S, const orc_material* m, f3 wo, f3 wi, f3 ng, f3 nf)
{
    if (m->kind != 6u) return bsdf_f3(m, wo, wi, ng, nf);
    uint32_t n = mat_layers3(m);
    f3 s = mk3(0.0f, 0.0f, 0.0f);
    for (uint32_t k = 0; k < n; k++) s = add3(s, bsdf_f3(mat_layer3(S, m, k), wo, wi, ng, nf));
    float fn = (float)n;
    return mk3(s.x / fn, s.y / fn, s.z / fn);
}
static float mat_pdf_sum3(const orc_scene* S, const orc_material* m, f3 wo, f3 wi, f3 nf, uint32_t* count)
{
    float s = 0.0f;
    uint32_t c = 0;
    for (uint32_t k = 0; k < mat_layers3(m); k++) {
        const orc_material* L = mat_layer3(S, m, k);
        if (delta_mat(L)) continue;
        s += bsdf_pdf3(L, wo, wi, nf);
        c++;
    }
    *count = c;
    return s;
}
static float mat_pdf_ns3(const orc_scene* S, const orc_material* m, f3 wo, f3 wi, f3 nf)
{
    if (m->kind != 6u) return bsdf_pdf3(m, wo, wi, nf);
    uint32_t c;
    float s = mat_pdf_sum3(S, m, wo, wi, nf, &c);
    return c ? s / (float)c : 0.0f;
}
/* a direction from the non-specular layers (one picked uniformly) */
static void mat_sample_ns3(const orc_scene* S, const orc_material* m, f3 wo, f3 ng, f3 nf, float upick, float u1,
                           float u2, f3* wi, f3* f, float* pdf)
{
    if (m->kind != 6u) {
        bsdf_sample3(m, wo, ng, nf, u1, u2, wi, f, pdf);
        return;
    }
    uint32_t n = mat_layers3(m), c = 0;
    for (uint32_t k = 0; k < n; k++) c += delta_mat(mat_layer3(S, m, k)) ? 0u : 1u;
    *f = mk3(0.0f, 0.0f, 0.0f);
    *pdf = 0.0f;
    if (!c) return;
    uint32_t j = (uint32_t)(upick * (float)c);
    if (j > c - 1) j = c - 1;
    const orc_material* L = m;
    for (uint32_t k = 0, i = 0; k < n; k++) {
        if (delta_mat(mat_layer3(S, m, k))) continue;
        if (i++ == j) { L = mat_layer3(S, m, k); break; }
    }
    f3 fl;
    float pl;
    bsdf_sample3(L, wo, ng, nf, u1, u2, wi, &fl, &pl);
    if (!(pl > 0.0f)) return;
    *f = mat_f3(S, m, wo, *wi, ng, nf);
    *pdf = mat_pdf_ns3(S, m, wo, *wi, nf);
}
/* the next direction from all layers (one picked uniformly) */
static void mat_sample3(const orc_scene* S, const orc_material* m, f3 wo, f3 ng, f3 ns, f3 nf, float upick, float u1,
                        float u2, int radiance, f3* wi, f3* f, float* pdf, int* spec, int* weight)
{
    if (m->kind != 6u) {
        layer_sample3(m, wo, ng, ns, nf, u1, u2, radiance, wi, f, pdf, spec, weight);
        return;
    }
    uint32_t n = mat_layers3(m);
    uint32_t k = (uint32_t)(upick * (float)n);
    if (k > n - 1) k = n - 1;
    layer_sample3(mat_layer3(S, m, k), wo, ng, ns, nf, u1, u2, radiance, wi, f, pdf, spec, weight);
    if (*spec || !(*pdf > 0.0f)) return;
    *f = mat_f3(S, m, wo, *wi, ng, nf);
    uint32_t c;
    *pdf = mat_pdf_sum3(S, m, wo, *wi, nf, &c) / (float)n;
}

/* BSDF value of a material for the tests: wo, wi, ng world; nf = ng facing wo. */
void orc_bsdf_eval(const float* m12, const float* wo, const float* wi, const float* ng, float* out)
{
    orc_material* m = mats_from12(m12, 1);
    f3 o = mk3(wo[0], wo[1], wo[2]), i = mk3(wi[0], wi[1], wi[2]), n = mk3(ng[0], ng[1], ng[2]);
    f3 nf = dot3(n, o) >= 0.0f ? n : mk3(-n.x, -n.y, -n.z);
    f3 f = bsdf_f3(m, o, i, n, nf);
    out[0] = f.x; out[1] = f.y; out[2] = f.z;
    out[3] = bsdf_pdf3(m, o, i, nf);
    free(m);
}

/* One BSDF sample: out = wi.xyz, f.rgb, pdf (delta materials: f = weight, pdf = 1). */
void orc_bsdf_sample(const float* m12, const float* wo, const float* ng, float u1, float u2, float* out)
{
    orc_material* m = mats_from12(m12, 1);
    f3 o = mk3(wo[0], wo[1], wo[2]), n = mk3(ng[0], ng[1], ng[2]);
    f3 nf = dot3(n, o) >= 0.0f ? n : mk3(-n.x, -n.y, -n.z);
    f3 wi, f;
    float pdf = 1.0f;
    if (m->kind == 3u) f = glass_sample3(m, o, n, u1, 1, &wi);
    else bsdf_sample3(m, o, n, nf, u1, u2, &wi, &f, &pdf);
    out[0] = wi.x; out[1] = wi.y; out[2] = wi.z; out[3] = f.x; out[4] = f.y; out[5] = f.z; out[6] = pdf;
    free(m);
}

/* ------------------------------------------------------------------------ */
/* BDPT with a light-vertex cache: IntegratorBdptLvc,                        */
/* src/integrators/integrator_bdpt_lvc_sk.cu:47-435, in the form the design  */
/* fixes: light vertices at fixed slots (path, depth), a connection picks a  */
/* path uniformly and one of its n_j vertices uniformly and is scaled by     */
/* n_j / connections, and every path is weighted by 1 / (number of          */
/* techniques able to build it at the given depth limits).                   */
/* ------------------------------------------------------------------------ */

typedef struct {
    uint32_t paths, depth, nconn;
    const uint32_t* count;                   /* vertices stored per light path */
    const float *pos, *ng, *wo, *alpha;      /* 3 floats per slot j * depth + d */
    const uint32_t *mid, *mask;              /* material, specular bits of y_1..y_(d+1) */
    orc_rng crng;
    const float* ns;                         /* shading normal per slot (NULL: the scene has none) */
    const uint8_t* snf;                      /* 1: ns is an interpolated normal */
} orc_lvc;

/* Techniques that build the path z_1..z_n (z_0 camera, z_n on the light,
 * mask bit i-1 = z_i specular) when camera vertices exist for bounces
 * 0..max_depth and the cache holds light vertices of depth 0..lvc_depth-1:
 * the direct-light technique (light sample + MIS BSDF sample at a diffuse
 * z_(n-1) on bounce n-2, or the emitter hit on bounce n-1 after a specular
 * z_(n-1) or straight from the camera) and each connection edge
 * (z_i, z_(i+1)) whose ends are both diffuse, camera end on bounce i-1,
 * light end stored at depth n-i-2. */
uint32_t orc_bdpt_techniques(uint32_t mask, uint32_t n, uint32_t max_depth, uint32_t lvc_depth)
{
    uint32_t c;
    if (n <= 1) c = 1;
    else if ((mask >> (n - 2)) & 1u) c = n - 1 <= max_depth ? 1u : 0u;
    else c = n - 2 <= max_depth ? 1u : 0u;
    for (uint32_t i = 1; i + 2 <= n; i++) {
        if ((mask >> (i - 1)) & 1u) continue;   /* camera end z_i specular */
        if ((mask >> i) & 1u) continue;         /* light end z_(i+1) specular */
        if (i - 1 > max_depth) continue;        /* camera sub-path too long */
        if (n - i - 2 >= lvc_depth) continue;   /* light sub-path not cached */
        c++;
    }
    return c;
}

static float bdpt_weight(uint32_t mask, uint32_t n, uint32_t max_depth, uint32_t lvc_depth)
{
    uint32_t c = orc_bdpt_techniques(mask, n, max_depth, lvc_depth);
    return c ? 1.0f / (float)c : 0.0f;
}

/* uniform_sample_hemisphere (src/core/montecarlo.cu:116-124) about unit n. */
static f3 uniform_dir(f3 n, float u1, float u2)
{
    float s, c;
    orc_sincos_2pi(u2, &s, &c);
    const float z = u1;
    const float rr = 1.0f - z * z;
    const float r = sqrtf(rr > 0.0f ? rr : 0.0f);
    const float sign = copysignf(1.0f, n.z);
    const float a = -1.0f / (sign + n.z);
    const float b = n.x * n.y * a;
    f3 b1 = mk3(1.0f + sign * n.x * n.x * a, sign * b, -sign * n.x);
    f3 b2 = mk3(b, sign + n.y * n.y * a, -n.y);
    return add3(add3(mul3s(b1, r * c), mul3s(b2, r * s)), mul3s(n, z));
}

/* Connections of the diffuse camera vertex p (bounce `bounce`, throughput
 * T) to the cache (integrator_bdpt_lvc_sk.cu:343-386); each unoccluded one
 * is added to L in connection order. */
static f3 lvc_connect(const orc_scene* S, const orc_lvc* B, uint32_t gid, uint32_t bounce, uint32_t max_depth, f3 p,
                      float eps, f3 wo, f3 ng, f3 ns, f3 nf, const orc_material* m, f3 T, uint32_t cm, f3 L)
{
    for (uint32_t i = 0; i < B->nconn; i++) {
        uint32_t rc[4];
        rng_draw(B->crng, gid, bounce, 2 + i / 2, 0, rc);
        uint32_t ua = (i & 1) ? rc[2] : rc[0], ub = (i & 1) ? rc[3] : rc[1];
        uint32_t j = (uint32_t)(u01(ua) * (float)B->paths);
        if (j > B->paths - 1) j = B->paths - 1;
        uint32_t nj = B->count[j];
        if (!nj) continue;
        uint32_t k = (uint32_t)(u01(ub) * (float)nj);
        if (k > nj - 1) k = nj - 1;
        size_t sl = (size_t)j * B->depth + k;
        f3 xp = mk3(B->pos[3 * sl], B->pos[3 * sl + 1], B->pos[3 * sl + 2]);
        f3 xn = mk3(B->ng[3 * sl], B->ng[3 * sl + 1], B->ng[3 * sl + 2]);
        f3 xw = mk3(B->wo[3 * sl], B->wo[3 * sl + 1], B->wo[3 * sl + 2]);
        f3 xa = mk3(B->alpha[3 * sl], B->alpha[3 * sl + 1], B->alpha[3 * sl + 2]);
        const orc_material* mx = &S->mats[B->mid[sl]];
        f3 seg = sub3(xp, p);
        float d2 = dot3(seg, seg);
        float dist = sqrtf(d2);
        f3 dir = normalize3(seg);
        f3 ndir = mk3(-dir.x, -dir.y, -dir.z);
        f3 fy = mat_f3(S, m, wo, dir, ng, nf);
        /* the cache vertex's shading normal; importance transport there
         * (light towards camera) takes the adjoint factor */
        int xsn = B->ns && B->snf[sl];
        f3 xs = xsn ? mk3(B->ns[3 * sl], B->ns[3 * sl + 1], B->ns[3 * sl + 2]) : xn;
        f3 nfx = dot3(xn, xw) >= 0.0f ? xs : mk3(-xs.x, -xs.y, -xs.z);
        f3 fx = mat_f3(S, mx, xw, ndir, xn, nfx);
        if (xsn) fx = mul3s(fx, adjoint_factor(xw, ndir, xn, xs));
        if (black3(fy) || black3(fx)) continue;
        float G = fabsf(dot3(ns, dir)) * fabsf(dot3(xs, ndir)) / d2;
        /* full path: camera vertices of bounces 0..bounce, then the light
         * vertices of depths k..0 */
        uint32_t M = cm;
        for (uint32_t e = 0; e <= k; e++)
            if ((B->mask[sl] >> (k - e)) & 1u) M |= 1u << (bounce + 1 + e);
        float wgt = bdpt_weight(M, bounce + k + 3, max_depth, B->nconn ? B->depth : 0u);
        float sc = (float)nj / (float)B->nconn;
        f3 c = mk3(T.x * xa.x * fy.x * G * fx.x * wgt * sc, T.y * xa.y * fy.y * G * fx.y * wgt * sc,
                   T.z * xa.z * fy.z * G * fx.z * wgt * sc);
        float sr[8] = {p.x, p.y, p.z, eps, dir.x, dir.y, dir.z, dist * (1.0f - 1.0e-3f)};
        int32_t sid; float st;
        trace_ray8(S, sr, 1, &sid, &st);
        if (sid < 0) L = add3(L, c);
    }
    return L;
}

/* One path of the path tracer, integrator_path.cu:52-254 with the fixes
 * listed in the design notes (roulette divides by the continue probability,
 * a fresh random number per bounce and dimension, true triangle areas). */
static f3 pt_path(const orc_scene* S, const float* R0, uint32_t gid, orc_rng rng, uint32_t max_depth,
                  uint32_t rr_start, float rr_prob, const orc_lvc* B)
{
    uint32_t cm = 0; /* BDPT: bit b = camera vertex of bounce b is specular */
    f3 L = mk3(0.0f, 0.0f, 0.0f), T = mk3(1.0f, 1.0f, 1.0f);
    float ray[8];
    memcpy(ray, R0, 32);
    int32_t id; float t;
    trace_ray8(S, ray, 0, &id, &t);
    if (id < 0) return L;
    int spec = 0;
    for (uint32_t bounce = 0;; ) {
        f3 d = mk3(ray[4], ray[5], ray[6]);
        f3 p = add3(mk3(ray[0], ray[1], ray[2]), mul3s(d, t));
        float eps = 1.0e-3f * t;
        f3 v0, v1, v2;
        tri_verts(S, (uint32_t)id, &v0, &v1, &v2);
        const orc_material* m = &S->mats[S->tri_mat[id]];
        f3 ng = tri_norm(v0, v1, v2), ns;
        shade_normal(S, (uint32_t)id, p, v0, v1, v2, ng, &ns);
        f3 wo = mk3(-d.x, -d.y, -d.z);
        f3 nf = dot3(ng, wo) >= 0.0f ? ns : mk3(-ns.x, -ns.y, -ns.z);
        if (m->kind == 2u) {
            if (bounce == 0 || spec) {
                f3 le = dot3(ng, wo) > 0.0f ? mk3(m->color[0], m->color[1], m->color[2]) : mk3(0.0f, 0.0f, 0.0f);
                if (B) le = mul3s(le, bdpt_weight(cm, bounce + 1, max_depth, B->nconn ? B->depth : 0u));
                L = mk3(L.x + T.x * le.x, L.y + T.y * le.y, L.z + T.z * le.z);
            }
            break;
        }
        uint32_t r0[4], r1[4], r2[4] = {0u, 0u, 0u, 0u};
        rng_draw(rng, gid, bounce, 0, 0, r0);
        rng_draw(rng, gid, bounce, 1, 0, r1);
        if (m->kind == 6u) rng_draw(rng, gid, bounce, 2, 0, r2); /* layer picks */
        int delta = mat_delta3(S, m);
        float wdir = B ? bdpt_weight(cm, bounce + 2, max_depth, B->nconn ? B->depth : 0u) : 1.0f;
        f3 Ld = mk3(0.0f, 0.0f, 0.0f);
        if (S->nl > 0) {
            uint32_t lid = (uint32_t)(u01(r0[0]) * (float)S->nl);
            if (lid > S->nl - 1) lid = S->nl - 1;
            uint32_t lt = S->lights[lid];
            f3 l0, l1, l2;
            tri_verts(S, lt, &l0, &l1, &l2);
            const orc_material* lm = &S->mats[S->tri_mat[lt]];
            f3 emit = mk3(lm->color[0], lm->color[1], lm->color[2]);
            f3 ln = tri_norm(l0, l1, l2);
            float area = 0.5f * sqrtf(dot3(cross3(sub3(l1, l0), sub3(l2, l0)), cross3(sub3(l1, l0), sub3(l2, l0))));
            /* light sample */
            /* uniform_sample_triangle, src/core/montecarlo.cu:157-164 */
            float su1 = sqrtf(u01(r0[1]));
            float uu = 1.0f - su1, vv = u01(r0[2]) * su1;
            f3 pl = add3(add3(mul3s(l0, uu), mul3s(l1, vv)), mul3s(l2, 1.0f - uu - vv));
            f3 wi = normalize3(sub3(pl, p));
            float lpdf = light_pdf_mt(p, wi, l0, l1, l2, ln, area);
            f3 Li = dot3(ln, mk3(-wi.x, -wi.y, -wi.z)) > 0.0f ? emit : mk3(0.0f, 0.0f, 0.0f);
            if (!delta && !black3(Li) && lpdf > 0.0f) {
                f3 f = mat_f3(S, m, wo, wi, ng, nf);
                if (!black3(f)) {
                    float bpdf = mat_pdf_ns3(S, m, wo, wi, nf);
                    float w = lpdf / (lpdf + bpdf);
                    float s = fabsf(dot3(wi, ns)) * w;
                    f3 ca = mk3(f.x * Li.x * s / lpdf, f.y * Li.y * s / lpdf, f.z * Li.z * s / lpdf);
                    f3 seg = sub3(pl, p);
                    float dist = sqrtf(dot3(seg, seg));
                    f3 sd = normalize3(seg);
                    float sr[8] = {p.x, p.y, p.z, eps, sd.x, sd.y, sd.z, dist * (1.0f - 1.0e-3f)};
                    int32_t sid; float st;
                    trace_ray8(S, sr, 1, &sid, &st);
                    if (sid < 0) Ld = add3(Ld, ca);
                }
            }
            if (!delta) {
                f3 wi2, f2;
                float pdf2;
                mat_sample_ns3(S, m, wo, ng, nf, u01(r2[1]), u01(r0[3]), u01(r1[0]), &wi2, &f2, &pdf2);
                if (!black3(f2) && pdf2 > 0.0f) {
                    float lpdf2 = light_pdf_mt(p, wi2, l0, l1, l2, ln, area);
                    if (lpdf2 > 0.0f) {
                        float w2 = pdf2 / (pdf2 + lpdf2);
                        f3 Li2 = dot3(ln, mk3(-wi2.x, -wi2.y, -wi2.z)) > 0.0f ? emit : mk3(0.0f, 0.0f, 0.0f);
                        float s2 = fabsf(dot3(wi2, ns)) * w2;
                        f3 cb = mk3(f2.x * Li2.x * s2 / pdf2, f2.y * Li2.y * s2 / pdf2, f2.z * Li2.z * s2 / pdf2);
                        float mr[8] = {p.x, p.y, p.z, eps, wi2.x, wi2.y, wi2.z, FLT_MAX};
                        int32_t mid2; float mt;
                        trace_ray8(S, mr, 0, &mid2, &mt);
                        if (mid2 == (int32_t)lt) Ld = add3(Ld, cb);
                    }
                }
            }
        }
        Ld = mul3s(Ld, (float)S->nl);
        if (B) Ld = mul3s(Ld, wdir);
        L = mk3(L.x + T.x * Ld.x, L.y + T.y * Ld.y, L.z + T.z * Ld.z);
        if (B && !delta && S->nl > 0) L = lvc_connect(S, B, gid, bounce, max_depth, p, eps, wo, ng, ns, nf, m, T, cm, L);
        /* next bounce */
        f3 wi3, f3v;
        float pdf3;
        int spec3, weight3; /* weight3: f3v already holds f |cos| / pdf */
        mat_sample3(S, m, wo, ng, ns, nf, u01(r2[0]), u01(r1[1]), u01(r1[2]), 1, &wi3, &f3v, &pdf3, &spec3, &weight3);
        /* the vertex is specular for the paths continuing through a
         * specular lobe */
        if (B && spec3) cm |= 1u << bounce;
        if (black3(f3v) || !(pdf3 > 0.0f)) break;
        float s3 = weight3 ? 1.0f : fabsf(dot3(wi3, ns));
        f3 Tn = mk3(T.x * (f3v.x * s3 / pdf3), T.y * (f3v.y * s3 / pdf3), T.z * (f3v.z * s3 / pdf3));
        int alive = 1;
        if (bounce > rr_start) {
            float lum = 0.212671f * Tn.x + 0.715160f * Tn.y + 0.072169f * Tn.z;
            float q = rr_prob < lum ? rr_prob : lum;
            if (u01(r1[3]) > q) alive = 0;
            else Tn = mk3(Tn.x / q, Tn.y / q, Tn.z / q);
        }
        if (!alive || bounce == max_depth) break;
        float nr[8] = {p.x, p.y, p.z, eps, wi3.x, wi3.y, wi3.z, FLT_MAX};
        trace_ray8(S, nr, 0, &id, &t);
        if (id < 0) break;
        memcpy(ray, nr, 32);
        T = Tn;
        spec = spec3;
        bounce++;
    }
    return L;
}


void orc_render_pt(const uint32_t* nodes, const uint32_t* woop, const uint32_t* tidx, const float* verts,
                   const uint32_t* idx, const uint32_t* tri_mat, const float* mats4, uint32_t nmats,
                   const uint32_t* lights, uint32_t nl, const float* r2c, const float* c2w, uint32_t width,
                   uint32_t x0, uint32_t y0, uint32_t w, uint32_t h, uint32_t spp, uint64_t seed,
                   uint32_t max_depth, uint32_t rr_start, float rr_prob, uint32_t filter, float fxw, float fyw,
                   float falpha, float* out_rgb, const float* tri_n, const uint8_t* has_n)
{
    orc_material* mats = mats_from12(mats4, nmats);
    orc_scene S = {nodes, woop, tidx, verts, idx, tri_mat, mats, lights, nl, tri_n, has_n};
    uint32_t n = w * h * spp;
    float* rays = (float*)malloc((size_t)n * 32);
    float* Ls = (float*)malloc((size_t)n * 12);
    orc_raygen_camera(r2c, c2w, width, x0, y0, w, h, spp, seed, 1, rays);
    orc_rng rng = rng_make(seed, 5u);
#pragma omp parallel for schedule(dynamic, 8)
    for (long px = 0; px < (long)(w * h); px++) {
        for (uint32_t s = 0; s < spp; s++) {
            uint32_t i = (uint32_t)px * spp + s;
            f3 L = pt_path(&S, rays + 8 * (size_t)i, sample_gid(width, x0, y0, w, spp, i), rng, max_depth, rr_start, rr_prob, NULL);
            Ls[3 * (size_t)i] = L.x; Ls[3 * (size_t)i + 1] = L.y; Ls[3 * (size_t)i + 2] = L.z;
        }
    }
    orc_film(Ls, width, x0, y0, w, h, spp, seed, filter, fxw, fyw, falpha, out_rgb);
    free(rays);
    free(Ls);
    free(mats);
}


/* One light sub-path (kernel_sk_light_trace, integrator_bdpt_lvc_sk.cu:47-168):
 * light chosen uniformly, point uniform on it, direction uniform over the
 * emitting hemisphere (DiffuseAreaLight, src/core/lights.cu:56-79); each
 * diffuse or mirror hit is stored at slot (j, depth) with the throughput
 * arriving there, then the path continues by a BSDF sample with roulette
 * (continue probability min(1, luminance)); emitters end it. */
static void light_path(const orc_scene* S, orc_rng rng, uint32_t j, uint32_t D, uint32_t* count, float* pos, float* ngo,
                       float* woo, float* alpha, uint32_t* mido, uint32_t* masko, float* nso, uint8_t* snfo)
{
    count[j] = 0;
    if (S->nl == 0) return;
    uint32_t r[4], r2[4];
    rng_draw(rng, j, 0, 0, 0, r);
    rng_draw(rng, j, 0, 1, 0, r2);
    uint32_t lid = (uint32_t)(u01(r[0]) * (float)S->nl);
    if (lid > S->nl - 1) lid = S->nl - 1;
    uint32_t lt = S->lights[lid];
    f3 l0, l1, l2;
    tri_verts(S, lt, &l0, &l1, &l2);
    const orc_material* lm = &S->mats[S->tri_mat[lt]];
    f3 ln = tri_norm(l0, l1, l2);
    f3 cr = cross3(sub3(l1, l0), sub3(l2, l0));
    float area = 0.5f * sqrtf(dot3(cr, cr));
    float su1 = sqrtf(u01(r[1]));
    float uu = 1.0f - su1, vv = u01(r[2]) * su1;
    f3 pl = add3(add3(mul3s(l0, uu), mul3s(l1, vv)), mul3s(l2, 1.0f - uu - vv));
    f3 dir = uniform_dir(ln, u01(r[3]), u01(r2[0]));
    float cth = fabsf(dot3(ln, dir));
    float pdf = (1.0f / area) * 0.159154943091895336f / cth;
    float kk = (float)S->nl / pdf;
    f3 a = mk3(lm->color[0] * kk, lm->color[1] * kk, lm->color[2] * kk);
    float ray[8] = {pl.x, pl.y, pl.z, 1.0e-4f, dir.x, dir.y, dir.z, FLT_MAX};
    uint32_t mask = 0;
    for (uint32_t d = 0; d < D; d++) {
        int32_t id; float t;
        trace_ray8(S, ray, 0, &id, &t);
        if (id < 0) return;
        f3 dv = mk3(ray[4], ray[5], ray[6]);
        f3 p = add3(mk3(ray[0], ray[1], ray[2]), mul3s(dv, t));
        f3 v0, v1, v2;
        tri_verts(S, (uint32_t)id, &v0, &v1, &v2);
        uint32_t mid = S->tri_mat[id];
        const orc_material* m = &S->mats[mid];
        if (m->kind == 2u) return;
        f3 ng = tri_norm(v0, v1, v2), ns;
        int sn = shade_normal(S, (uint32_t)id, p, v0, v1, v2, ng, &ns);
        f3 wo = mk3(-dv.x, -dv.y, -dv.z);
        size_t sl = (size_t)j * D + d;
        pos[3 * sl] = p.x; pos[3 * sl + 1] = p.y; pos[3 * sl + 2] = p.z;
        ngo[3 * sl] = ng.x; ngo[3 * sl + 1] = ng.y; ngo[3 * sl + 2] = ng.z;
        woo[3 * sl] = wo.x; woo[3 * sl + 1] = wo.y; woo[3 * sl + 2] = wo.z;
        alpha[3 * sl] = a.x; alpha[3 * sl + 1] = a.y; alpha[3 * sl + 2] = a.z;
        mido[sl] = mid;
        masko[sl] = mask | (mat_delta3(S, m) ? 1u << d : 0u);
        if (nso) {
            nso[3 * sl] = ns.x; nso[3 * sl + 1] = ns.y; nso[3 * sl + 2] = ns.z;
            snfo[sl] = (uint8_t)sn;
        }
        count[j] = d + 1;
        if (d + 1 >= D) return;
        uint32_t rr[4];
        rng_draw(rng, j, d + 1, 0, 0, rr);
        f3 nf = dot3(ng, wo) >= 0.0f ? ns : mk3(-ns.x, -ns.y, -ns.z);
        f3 wi, f;
        float bp;
        int spec, weight; /* weight: f already holds f |cos| / pdf */
        mat_sample3(S, m, wo, ng, ns, nf, u01(rr[3]), u01(rr[0]), u01(rr[1]), 0, &wi, &f, &bp, &spec, &weight);
        if (spec) mask |= 1u << d;
        if (black3(f) || !(bp > 0.0f)) return;
        float s = weight ? 1.0f : fabsf(dot3(ns, wi));
        f3 tmp = mk3(f.x * s / bp, f.y * s / bp, f.z * s / bp);
        if (sn) tmp = mul3s(tmp, adjoint_factor(wo, wi, ng, ns));
        float lum = 0.212671f * tmp.x + 0.715160f * tmp.y + 0.072169f * tmp.z;
        float cp = lum < 1.0f ? lum : 1.0f;
        if (u01(rr[2]) > cp) return;
        f3 na = mk3(a.x * (tmp.x / cp), a.y * (tmp.y / cp), a.z * (tmp.z / cp));
        if (black3(na)) return;
        a = na;
        float nr[8] = {p.x, p.y, p.z, 1.0e-3f * t, wi.x, wi.y, wi.z, FLT_MAX};
        memcpy(ray, nr, 32);
    }
}

void orc_render_bdpt(const uint32_t* nodes, const uint32_t* woop, const uint32_t* tidx, const float* verts,
                     const uint32_t* idx, const uint32_t* tri_mat, const float* mats4, uint32_t nmats,
                     const uint32_t* lights, uint32_t nl, const float* r2c, const float* c2w, uint32_t width,
                     uint32_t x0, uint32_t y0, uint32_t w, uint32_t h, uint32_t spp, uint64_t seed,
                     uint32_t max_depth, uint32_t rr_start, float rr_prob, uint32_t light_paths,
                     uint32_t light_max_depth, uint32_t connections, uint32_t filter, float fxw, float fyw,
                     float falpha, float* out_rgb, const float* tri_n, const uint8_t* has_n)
{
    orc_material* mats = mats_from12(mats4, nmats);
    orc_scene S = {nodes, woop, tidx, verts, idx, tri_mat, mats, lights, nl, tri_n, has_n};
    const uint32_t NL = light_paths, D = light_max_depth;
    size_t slots = (size_t)NL * D;
    uint32_t* count = (uint32_t*)calloc(NL ? NL : 1, 4);
    float* pos = (float*)calloc(slots * 3 + 1, 4);
    float* ngs = (float*)calloc(slots * 3 + 1, 4);
    float* wos = (float*)calloc(slots * 3 + 1, 4);
    float* alpha = (float*)calloc(slots * 3 + 1, 4);
    uint32_t* mid = (uint32_t*)calloc(slots + 1, 4);
    uint32_t* mask = (uint32_t*)calloc(slots + 1, 4);
    float* nss = tri_n ? (float*)calloc(slots * 3 + 1, 4) : NULL;
    uint8_t* snf = tri_n ? (uint8_t*)calloc(slots + 1, 1) : NULL;
    orc_rng lrng = rng_make(seed, 6u);
#pragma omp parallel for schedule(dynamic, 64)
    for (long j = 0; j < (long)NL; j++)
        light_path(&S, lrng, (uint32_t)j, D, count, pos, ngs, wos, alpha, mid, mask, nss, snf);
    orc_lvc B = {NL, D, nl ? connections : 0u, count, pos, ngs, wos, alpha, mid, mask, rng_make(seed, 8u), nss, snf};
    uint32_t n = w * h * spp;
    float* rays = (float*)malloc((size_t)n * 32);
    float* Ls = (float*)malloc((size_t)n * 12);
    orc_raygen_camera(r2c, c2w, width, x0, y0, w, h, spp, seed, 1, rays);
    orc_rng rng = rng_make(seed, 5u);
#pragma omp parallel for schedule(dynamic, 8)
    for (long px = 0; px < (long)(w * h); px++) {
        for (uint32_t s = 0; s < spp; s++) {
            uint32_t i = (uint32_t)px * spp + s;
            f3 L = pt_path(&S, rays + 8 * (size_t)i, sample_gid(width, x0, y0, w, spp, i), rng, max_depth, rr_start,
                           rr_prob, &B);
            Ls[3 * (size_t)i] = L.x; Ls[3 * (size_t)i + 1] = L.y; Ls[3 * (size_t)i + 2] = L.z;
        }
    }
    orc_film(Ls, width, x0, y0, w, h, spp, seed, filter, fxw, fyw, falpha, out_rgb);
    free(rays); free(Ls); free(count); free(pos); free(ngs); free(wos); free(alpha); free(mid); free(mask);
    free(nss); free(snf);
    free(mats);
}

/* ---- ray-cast debug integrator ------------------------------------------ */

/* kernel_ray_cast_update's colour table (src/integrators/integrator_raycast.cu:47-61) */
static const float kPalette[15][3] = {
    {1.0f, 0.0f, 0.0f}, {0.0f, 1.0f, 0.0f}, {0.0f, 0.0f, 1.0f}, {0.0f, 1.0f, 1.0f}, {1.0f, 0.0f, 1.0f},
    {1.0f, 1.0f, 0.0f}, {0.1f, 0.1f, 0.1f}, {0.5f, 0.5f, 0.5f}, {1.0f, 1.0f, 1.0f}, {0.0f, 0.25f, 0.5f},
    {1.0f, 0.5f, 0.0f}, {0.5f, 0.25f, 0.0f}, {0.2f, 0.9f, 1.0f}, {0.7f, 0.3f, 0.1f}, {0.34f, 0.89f, 0.45f}};

/* kernel_ray_cast_update (integrator_raycast.cu:37-128) per camera sample,
 * IntegratorRayCast::compute (:136-160) for the trace. */
void orc_render_raycast(const uint32_t* nodes, const uint32_t* woop, const uint32_t* tidx, const float* verts,
                        const uint32_t* idx, const uint32_t* tri_mat, const float* m12, uint32_t nmats,
                        const float* r2c, const float* c2w, uint32_t width, uint32_t x0, uint32_t y0, uint32_t w,
                        uint32_t h, uint32_t spp, uint64_t seed, uint32_t shade, uint32_t filter, float fxw,
                        float fyw, float falpha, float* out_rgb, const float* tri_n, const uint8_t* has_n)
{
    orc_material* mats = mats_from12(m12, nmats);
    orc_scene S = {nodes, woop, tidx, verts, idx, tri_mat, mats, NULL, 0, tri_n, has_n};
    uint32_t n = w * h * spp;
    float* rays = (float*)malloc((size_t)n * 32);
    float* Ls = (float*)malloc((size_t)n * 12);
    orc_raygen_camera(r2c, c2w, width, x0, y0, w, h, spp, seed, 1, rays);
#pragma omp parallel for schedule(dynamic, 64)
    for (long i = 0; i < (long)n; i++) {
        const float* R = rays + 8 * (size_t)i;
        int32_t id; float t;
        trace_ray8(&S, R, 0, &id, &t);
        f3 c = mk3(0.0f, 0.0f, 0.0f);
        if (id >= 0) {
            f3 a, b, cc;
            tri_verts(&S, (uint32_t)id, &a, &b, &cc);
            if (shade == 0) {
                const float* q = kPalette[(uint32_t)id % 15u];
                c = mk3(q[0], q[1], q[2]);
            } else if (shade == 1) {
                /* a layered material shows its first layer's colour
                 * (integrator_raycast.cu:71-76) */
                const orc_material* m = &mats[tri_mat[id]];
                if (m->kind == 6u) m = &mats[(uint32_t)m->param[1]];
                c = mk3(m->color[0], m->color[1], m->color[2]);
            } else if (shade == 2) {
                /* m_shading_normal (integrator_raycast.cu:109-115) */
                f3 p = add3(mk3(R[0], R[1], R[2]), mul3s(mk3(R[4], R[5], R[6]), t));
                interp_normal(&S, (uint32_t)id, p, a, b, cc, tri_norm(a, b, cc), &c);
            } else if (shade == 3) {
                /* barycentric weights of v0, v1 at the hit point */
                f3 p = add3(mk3(R[0], R[1], R[2]), mul3s(mk3(R[4], R[5], R[6]), t));
                f3 e0 = sub3(a, cc), e1 = sub3(b, cc), q = sub3(p, cc);
                float d00 = dot3(e0, e0), d01 = dot3(e0, e1), d11 = dot3(e1, e1);
                float d20 = dot3(q, e0), d21 = dot3(q, e1);
                float den = d00 * d11 - d01 * d01;
                c.x = den != 0.0f ? (d11 * d20 - d01 * d21) / den : 0.0f;
                c.y = den != 0.0f ? (d00 * d21 - d01 * d20) / den : 0.0f;
            } else {
                c = mk3(1.0f, 1.0f, 1.0f);
            }
        }
        Ls[3 * (size_t)i] = c.x; Ls[3 * (size_t)i + 1] = c.y; Ls[3 * (size_t)i + 2] = c.z;
    }
    orc_film(Ls, width, x0, y0, w, h, spp, seed, filter, fxw, fyw, falpha, out_rgb);
    free(rays);
    free(Ls);
    free(mats);
}

/* ---- intersection records ----------------------------------------------- */

/* The result-write block of kernel_kepler_trace (kepler_dynamic_fetch.h:312-360)
 * as its own pass over (id, t): each hit's position o + t d, Ng (unit
 * cross(v1 - v0, v2 - v0)), Ns (interp_normal), -d, barycentric (u, v) of
 * the position (the reference stores the u, v of the last triangle tested,
 * not of the hit -- a slip not kept), material id, front = Ng.d < 0,
 * epsilon = 1e-3 t; t itself. */
void orc_fill(const float* rays, const int32_t* ids, const float* ts, uint32_t n, const float* verts,
              const uint32_t* idx, const uint32_t* tri_mat, const float* tri_n, const uint8_t* has_n, float* out,
              uint32_t* mid_front)
{
    orc_scene S = {NULL, NULL, NULL, verts, idx, tri_mat, NULL, NULL, 0, tri_n, has_n};
#pragma omp parallel for schedule(static)
    for (long i = 0; i < (long)n; i++) {
        const int32_t id = ids[i];
        if (id < 0) continue;
        const float* R = rays + 8 * (size_t)i;
        const float t = ts[i];
        f3 d = mk3(R[4], R[5], R[6]);
        f3 p = add3(mk3(R[0], R[1], R[2]), mul3s(d, t));
        f3 a, b, c;
        tri_verts(&S, (uint32_t)id, &a, &b, &c);
        f3 ng = tri_norm(a, b, c), ns;
        interp_normal(&S, (uint32_t)id, p, a, b, c, ng, &ns);
        float u, v;
        bary_uv(p, a, b, c, &u, &v);
        float* o = out + 16 * (size_t)i;
        o[0] = p.x; o[1] = p.y; o[2] = p.z;
        o[3] = ng.x; o[4] = ng.y; o[5] = ng.z;
        o[6] = ns.x; o[7] = ns.y; o[8] = ns.z;
        o[9] = -d.x; o[10] = -d.y; o[11] = -d.z;
        o[12] = u; o[13] = v;
        o[14] = 1.0e-3f * t;
        o[15] = t;
        mid_front[2 * (size_t)i] = tri_mat[id];
        mid_front[2 * (size_t)i + 1] = dot3(ng, d) < 0.0f ? 1u : 0u;
    }
}

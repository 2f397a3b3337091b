// rb_layout.cpp -- GPU node layouts of the host SBVH.
//
//  * Compact2: byte-compatible with CudaBvh::create_compact(bvh, 16)
//    (renderbox2 src/accelerators/sbvh/cudabvh.cpp:216-296): kept so that
//    BVH files and byte streams can be compared with the reference format.
//  * MI355X: the traversal layout of this library (see export_mi355x).
#include <algorithm>
#include <cstdio>
#include <fstream>

#include "rb_internal.h"

namespace rb {

// Matrix4x4::invert (src/core/transform.h:205-333): cofactor expansion,
// evaluated term by term from left to right.
void invert4x4_ref(const float* m, float* out) {
    float a[16];
    a[0] = m[5] * m[10] * m[15] - m[5] * m[11] * m[14] - m[9] * m[6] * m[15] + m[9] * m[7] * m[14] + m[13] * m[6] * m[11] - m[13] * m[7] * m[10];
    a[4] = -m[4] * m[10] * m[15] + m[4] * m[11] * m[14] + m[8] * m[6] * m[15] - m[8] * m[7] * m[14] - m[12] * m[6] * m[11] + m[12] * m[7] * m[10];
    a[8] = m[4] * m[9] * m[15] - m[4] * m[11] * m[13] - m[8] * m[5] * m[15] + m[8] * m[7] * m[13] + m[12] * m[5] * m[11] - m[12] * m[7] * m[9];
    a[12] = -m[4] * m[9] * m[14] + m[4] * m[10] * m[13] + m[8] * m[5] * m[14] - m[8] * m[6] * m[13] - m[12] * m[5] * m[10] + m[12] * m[6] * m[9];
    a[1] = -m[1] * m[10] * m[15] + m[1] * m[11] * m[14] + m[9] * m[2] * m[15] - m[9] * m[3] * m[14] - m[13] * m[2] * m[11] + m[13] * m[3] * m[10];
    a[5] = m[0] * m[10] * m[15] - m[0] * m[11] * m[14] - m[8] * m[2] * m[15] + m[8] * m[3] * m[14] + m[12] * m[2] * m[11] - m[12] * m[3] * m[10];
    a[9] = -m[0] * m[9] * m[15] + m[0] * m[11] * m[13] + m[8] * m[1] * m[15] - m[8] * m[3] * m[13] - m[12] * m[1] * m[11] + m[12] * m[3] * m[9];
    a[13] = m[0] * m[9] * m[14] - m[0] * m[10] * m[13] - m[8] * m[1] * m[14] + m[8] * m[2] * m[13] + m[12] * m[1] * m[10] - m[12] * m[2] * m[9];
    a[2] = m[1] * m[6] * m[15] - m[1] * m[7] * m[14] - m[5] * m[2] * m[15] + m[5] * m[3] * m[14] + m[13] * m[2] * m[7] - m[13] * m[3] * m[6];
    a[6] = -m[0] * m[6] * m[15] + m[0] * m[7] * m[14] + m[4] * m[2] * m[15] - m[4] * m[3] * m[14] - m[12] * m[2] * m[7] + m[12] * m[3] * m[6];
    a[10] = m[0] * m[5] * m[15] - m[0] * m[7] * m[13] - m[4] * m[1] * m[15] + m[4] * m[3] * m[13] + m[12] * m[1] * m[7] - m[12] * m[3] * m[5];
    a[14] = -m[0] * m[5] * m[14] + m[0] * m[6] * m[13] + m[4] * m[1] * m[14] - m[4] * m[2] * m[13] - m[12] * m[1] * m[6] + m[12] * m[2] * m[5];
    a[3] = -m[1] * m[6] * m[11] + m[1] * m[7] * m[10] + m[5] * m[2] * m[11] - m[5] * m[3] * m[10] - m[9] * m[2] * m[7] + m[9] * m[3] * m[6];
    a[7] = m[0] * m[6] * m[11] - m[0] * m[7] * m[10] - m[4] * m[2] * m[11] + m[4] * m[3] * m[10] + m[8] * m[2] * m[7] - m[8] * m[3] * m[6];
    a[11] = -m[0] * m[5] * m[11] + m[0] * m[7] * m[9] + m[4] * m[1] * m[11] - m[4] * m[3] * m[9] - m[8] * m[1] * m[7] + m[8] * m[3] * m[5];
    a[15] = m[0] * m[5] * m[10] - m[0] * m[6] * m[9] - m[4] * m[1] * m[10] + m[4] * m[2] * m[9] + m[8] * m[1] * m[6] - m[8] * m[2] * m[5];
    float det = m[0] * a[0] + m[1] * a[4] + m[2] * a[8] + m[3] * a[12];
    det = 1.0f / det;
    for (int i = 0; i < 16; i++) out[i] = a[i] * det;
}

// CudaBvh::woopify_tri (src/accelerators/sbvh/cudabvh.cpp:303-322): the
// unit-triangle transform, rows (M^-1 row 2 with w negated, row 0, row 1).
// A leading -0 is stored as +0 so it cannot alias the Compact2 terminator.
void woopify(const float* a, const float* b, const float* c, float out[12]) {
    F3 v0 = f3(a[0], a[1], a[2]), v1 = f3(b[0], b[1], b[2]), v2 = f3(c[0], c[1], c[2]);
    F3 e0 = v0 - v2, e1 = v1 - v2;
    F3 n = normalize(cross(v0 - v2, v1 - v2));
    const float m[16] = {e0.x, e1.x, n.x, v2.x, e0.y, e1.y, n.y, v2.y, e0.z, e1.z, n.z, v2.z, 0.0f, 0.0f, 0.0f, 1.0f};
    float inv[16];
    invert4x4_ref(m, inv);
    out[0] = inv[8];
    out[1] = inv[9];
    out[2] = inv[10];
    out[3] = -inv[11];
    for (int i = 0; i < 4; i++) out[4 + i] = inv[i];
    for (int i = 0; i < 4; i++) out[8 + i] = inv[4 + i];
    if (out[0] == 0.0f) out[0] = 0.0f;
}

// A root leaf is wrapped in an inner node whose second child is an empty
// leaf with an inverted box (never entered); the reference asserts instead
// (cudabvh.cpp:237).
struct RootView {
    BNode wrap;
    BNode empty;
    const BNode* root;
    explicit RootView(const BNode* r) {
        root = r;
        if (r->leaf) {
            empty.leaf = true;
            wrap.leaf = false;
            wrap.b = r->b;
            wrap.c[0] = const_cast<BNode*>(r);
            wrap.c[1] = &empty;
            root = &wrap;
        }
    }
};

rb_status export_compact2(const HostBvh& H, std::vector<uint32_t>& nodes, std::vector<uint32_t>& woop,
                          std::vector<uint32_t>& index) {
    RootView rv(H.root);
    nodes.assign(16, 0);
    woop.clear();
    index.clear();
    struct Ent {
        const BNode* n;
        size_t row;
    };
    std::vector<Ent> stack;
    stack.push_back({rv.root, 0});
    while (!stack.empty()) {
        Ent e = stack.back();
        stack.pop_back();
        const Box* cb[2];
        int32_t cidx[2];
        for (int i = 0; i < 2; i++) {
            const BNode* ch = e.n->c[i];
            cb[i] = &ch->b;
            if (!ch->leaf) {
                size_t row = nodes.size() / 4;
                cidx[i] = (int32_t)row;  // byte offset / 16
                stack.push_back({ch, row});
                nodes.resize(nodes.size() + 16, 0);
                continue;
            }
            cidx[i] = ~(int32_t)(woop.size() / 4);
            for (uint32_t j = 0; j < ch->ntris; j++) {
                const uint32_t t = ch->tris[j];
                const uint32_t* ind = &H.idx[3 * (size_t)t];
                float w[12];
                woopify(&H.verts[3 * (size_t)ind[0]], &H.verts[3 * (size_t)ind[1]], &H.verts[3 * (size_t)ind[2]], w);
                for (int k = 0; k < 12; k++) woop.push_back(fbits(w[k]));
                index.push_back(t);
                index.push_back(0);
                index.push_back(0);
            }
            for (int k = 0; k < 4; k++) woop.push_back(0x80000000u);
            index.push_back(0);
        }
        uint32_t* d = &nodes[e.row * 4];
        d[0] = fbits(cb[0]->lo.x); d[1] = fbits(cb[0]->hi.x); d[2] = fbits(cb[0]->lo.y); d[3] = fbits(cb[0]->hi.y);
        d[4] = fbits(cb[1]->lo.x); d[5] = fbits(cb[1]->hi.x); d[6] = fbits(cb[1]->lo.y); d[7] = fbits(cb[1]->hi.y);
        d[8] = fbits(cb[0]->lo.z); d[9] = fbits(cb[0]->hi.z); d[10] = fbits(cb[1]->lo.z); d[11] = fbits(cb[1]->hi.z);
        d[12] = (uint32_t)cidx[0];
        d[13] = (uint32_t)cidx[1];
        d[14] = 0;
        d[15] = 0;
    }
    return RB_OK;
}

// MI355X layout.  One 64-byte record per inner node (64-byte aligned, so a
// node is four 16-byte loads of one half of a 128-byte cache line):
//   word 0-3   c0.lo.x c0.hi.x c0.lo.y c0.hi.y
//   word 4-7   c1.lo.x c1.hi.x c1.lo.y c1.hi.y
//   word 8-11  c0.lo.z c0.hi.z c1.lo.z c1.hi.z
//   word 12-15 link0 link1 count0 count1
// link >= 0: inner child record index.  link < 0: leaf, link =
// ~(first_reference << 5 | count) with count <= 31, so a leaf is fully
// described by its 32-bit link (no terminator rows, nothing to fetch before
// the triangles); words 14/15 repeat the counts for readers.  Leaves of more
// than 31 references are split into a balanced subtree of inner records
// that repeat the leaf's box (results are unchanged: every reference is
// still tested).  Triangles: 3 float4 Woop rows (48 bytes) per reference,
// leaf-contiguous; index: original triangle id per reference.  Records are
// allocated in the Compact2 order (both inner children of a node side by side).
static const BNode* split_big_leaf(const BNode* leaf, std::vector<std::unique_ptr<BNode>>& pool) {
    if (!leaf->leaf || leaf->ntris <= 31) return leaf;
    uint32_t h = leaf->ntris / 2;
    pool.emplace_back(new BNode());
    BNode* a = pool.back().get();
    a->leaf = true; a->b = leaf->b; a->tris = leaf->tris; a->ntris = h;
    pool.emplace_back(new BNode());
    BNode* b = pool.back().get();
    b->leaf = true; b->b = leaf->b; b->tris = leaf->tris + h; b->ntris = leaf->ntris - h;
    pool.emplace_back(new BNode());
    BNode* in = pool.back().get();
    in->leaf = false; in->b = leaf->b;
    in->c[0] = const_cast<BNode*>(split_big_leaf(a, pool));
    in->c[1] = const_cast<BNode*>(split_big_leaf(b, pool));
    return in;
}

// Moves the first `k` inner records in breadth-first order (root first) to
// the front of the node array and renumbers every inner link; the other
// records keep their relative order, so subtrees below the front stay
// contiguous.  The traversal reads the front from LDS (lds_nodes).
static void breadth_first_front(std::vector<uint32_t>& nodes, uint32_t k) {
    const uint32_t n = (uint32_t)(nodes.size() / 16);
    if (k > n) k = n;
    std::vector<uint32_t> order;  // old record ids, breadth-first
    order.reserve(k);
    std::vector<uint8_t> front(n, 0);
    order.push_back(0);
    front[0] = 1;
    for (size_t h = 0; h < order.size() && order.size() < k; h++) {
        const uint32_t* d = &nodes[(size_t)order[h] * 16];
        for (int c = 0; c < 2 && order.size() < k; c++) {
            const int32_t link = (int32_t)d[12 + c];
            if (link >= 0 && !front[link]) {
                front[link] = 1;
                order.push_back((uint32_t)link);
            }
        }
    }
    std::vector<uint32_t> newid(n);
    uint32_t next = 0;
    for (uint32_t r : order) newid[r] = next++;
    for (uint32_t r = 0; r < n; r++)
        if (!front[r]) newid[r] = next++;
    std::vector<uint32_t> out(nodes.size());
    for (uint32_t r = 0; r < n; r++) {
        uint32_t* d = &out[(size_t)newid[r] * 16];
        std::memcpy(d, &nodes[(size_t)r * 16], 64);
        for (int c = 0; c < 2; c++)
            if ((int32_t)d[12 + c] >= 0) d[12 + c] = newid[d[12 + c]];
    }
    nodes.swap(out);
}

rb_status export_mi355x(const HostBvh& H, Mi355Layout& L) {
    RootView rv(H.root);
    std::vector<std::unique_ptr<BNode>> pool;
    L.nodes.assign(16, 0);
    L.tris.clear();
    L.index.clear();
    struct Ent {
        const BNode* n;
        size_t rec;
    };
    std::vector<Ent> stack;
    stack.push_back({rv.root, 0});
    while (!stack.empty()) {
        Ent e = stack.back();
        stack.pop_back();
        const Box* cb[2];
        int32_t link[2];
        uint32_t cnt[2];
        for (int i = 0; i < 2; i++) {
            const BNode* ch = split_big_leaf(e.n->c[i], pool);
            cb[i] = &ch->b;
            if (!ch->leaf) {
                size_t rec = L.nodes.size() / 16;
                if (rec >= 0x7FFFFFFFu) return fail(RB_ERR_UNSUPPORTED, "too many BVH nodes");
                link[i] = (int32_t)rec;
                cnt[i] = 0;
                stack.push_back({ch, rec});
                L.nodes.resize(L.nodes.size() + 16, 0);
                continue;
            }
            if (L.index.size() >= (1u << 26))
                return fail(RB_ERR_UNSUPPORTED, "more than 2^26 triangle references");
            link[i] = ~(int32_t)(((uint32_t)L.index.size() << 5) | ch->ntris);
            cnt[i] = ch->ntris;
            for (uint32_t j = 0; j < ch->ntris; j++) {
                const uint32_t t = ch->tris[j];
                const uint32_t* ind = &H.idx[3 * (size_t)t];
                float w[12];
                woopify(&H.verts[3 * (size_t)ind[0]], &H.verts[3 * (size_t)ind[1]], &H.verts[3 * (size_t)ind[2]], w);
                L.tris.insert(L.tris.end(), w, w + 12);
                L.index.push_back(t);
            }
        }
        uint32_t* d = &L.nodes[e.rec * 16];
        d[0] = fbits(cb[0]->lo.x); d[1] = fbits(cb[0]->hi.x); d[2] = fbits(cb[0]->lo.y); d[3] = fbits(cb[0]->hi.y);
        d[4] = fbits(cb[1]->lo.x); d[5] = fbits(cb[1]->hi.x); d[6] = fbits(cb[1]->lo.y); d[7] = fbits(cb[1]->hi.y);
        d[8] = fbits(cb[0]->lo.z); d[9] = fbits(cb[0]->hi.z); d[10] = fbits(cb[1]->lo.z); d[11] = fbits(cb[1]->hi.z);
        d[12] = (uint32_t)link[0];
        d[13] = (uint32_t)link[1];
        d[14] = cnt[0];
        d[15] = cnt[1];
    }
    breadth_first_front(L.nodes, kLdsNodesMax);
    return RB_OK;
}

// Wide records from the binary ones: the children of a binary record, every
// inner child replaced by its own two children, so one fetch holds up to
// four boxes and a ray reads about half as many records.  Binary records
// that became wide records are numbered breadth-first from the root, so the
// records the traversal keeps in LDS are the top of the tree.
//
// Child boxes are stored as 8-bit steps from the record's origin p (the
// smallest corner of its children) in units of s = 2^e per axis; the
// traversal decodes a plane as fmaf(q, s, p), rounded once.  Each lo step is
// the floor of (lo - p) / s, lowered while its decoded plane lies above lo,
// and each hi step the ceiling of (hi - p) / s, raised while its plane lies
// below hi, both checked with that same fmaf, so the decoded box holds the
// exact one.  The slab distances fmaf(plane, 1/d, -o/d)
// are monotone in the plane, so a ray passes a decoded box whenever it
// passes the exact one: a ray tests every leaf the binary walk would at the
// same hitT, and the closest hit (ties to the lowest triangle id) is the
// same.  Empty leaves (no triangles) are left out; an inverted pair of
// planes is held as the box between them, which is what the slab test of
// the binary walk sees.
namespace {

float decode_plane(uint32_t q, float s, float p) { return std::fma((float)q, s, p); }

// Steps of one axis for `k` children (lo[i] <= hi[i]) from origin p: the
// smallest exponent (biased, 1..254) whose steps all fit in 8 bits.
bool quantize_axis(const float* lo, const float* hi, uint32_t k, float p, uint32_t* qlo, uint32_t* qhi,
                   uint32_t* ebias) {
    double ext = 0.0;
    for (uint32_t i = 0; i < k; i++) ext = std::max(ext, (double)hi[i] - (double)p);
    int e = -126;
    if (ext > 0.0) {
        int ex;
        std::frexp(ext / 250.0, &ex);  // 2^(ex-1) <= ext/250 < 2^ex
        e = std::max(-126, ex);
    }
    for (; e <= 127; e++) {
        const float s = std::ldexp(1.0f, e);
        bool fits = true;
        for (uint32_t i = 0; i < k && fits; i++) {
            double ql = std::floor(((double)lo[i] - (double)p) / (double)s);
            uint32_t a = ql <= 0.0 ? 0u : (ql >= 255.0 ? 255u : (uint32_t)ql);
            while (a > 0 && decode_plane(a, s, p) > lo[i]) a--;
            double qh = std::ceil(((double)hi[i] - (double)p) / (double)s);
            uint32_t b = qh <= (double)a ? a : (qh >= 256.0 ? 256u : (uint32_t)qh);
            while (b <= 255 && decode_plane(b, s, p) < hi[i]) b++;
            if (b > 255 || decode_plane(a, s, p) > lo[i]) {
                fits = false;
                break;
            }
            qlo[i] = a;
            qhi[i] = b;
        }
        if (fits) {
            *ebias = (uint32_t)(e + 127);
            return true;
        }
    }
    return false;
}

}  // namespace

rb_status export_wide(const std::vector<uint32_t>& bin, std::vector<uint32_t>& wide) {
    const uint32_t n = (uint32_t)(bin.size() / 16);
    wide.clear();
    if (n == 0) return RB_OK;
    std::vector<uint32_t> order;  // binary record of each wide record
    std::vector<int32_t> wid(n, -1);
    order.push_back(0);
    wid[0] = 0;
    for (size_t h = 0; h < order.size(); h++) {
        const uint32_t* d = &bin[(size_t)order[h] * 16];
        float box[4][6];  // lo.x, hi.x, lo.y, hi.y, lo.z, hi.z
        int32_t link[4];
        uint32_t k = 0;
        bool bad = false;
        auto add = [&](const uint32_t* rec, int c) {
            const int32_t l = (int32_t)rec[12 + c];
            if (l < 0 && (~(uint32_t)l & 31u) == 0) return;  // empty leaf
            const uint32_t* b = rec + 4 * c;
            const uint32_t w[6] = {b[0], b[1], b[2], b[3], rec[8 + 2 * c], rec[9 + 2 * c]};
            for (int r = 0; r < 6; r++) std::memcpy(&box[k][r], &w[r], 4);
            for (int a = 0; a < 3; a++) {
                float& lo = box[k][2 * a];
                float& hi = box[k][2 * a + 1];
                if (lo != lo || hi != hi) bad = true;  // NaN
                // the slab test takes the min and max of a pair of planes, so
                // an inverted pair (lo > hi) is the box between them: hold that
                if (lo > hi) std::swap(lo, hi);
            }
            if (l >= 0) {
                if ((uint32_t)l >= n) {
                    bad = true;
                    return;
                }
                if (wid[l] < 0) {
                    wid[l] = (int32_t)order.size();
                    order.push_back((uint32_t)l);
                }
                link[k] = wid[l];
            } else {
                link[k] = l;
            }
            k++;
        };
        for (int c = 0; c < 2; c++) {
            const int32_t l = (int32_t)d[12 + c];
            if (l >= 0) {  // an inner child: its two children instead
                if ((uint32_t)l >= n) return fail(RB_ERR_INVALID, "bad link in the node array");
                const uint32_t* g = &bin[(size_t)l * 16];
                add(g, 0);
                add(g, 1);
            } else {
                add(d, c);
            }
        }
        if (bad) return fail(RB_ERR_UNSUPPORTED, "a BVH box is NaN or a link leaves the array");
        const size_t base = wide.size();
        wide.resize(base + kWideWords, 0);
        uint32_t* w = &wide[base];
        uint32_t ebias[3] = {127, 127, 127};
        for (int a = 0; a < 3 && k; a++) {
            float lo[4], hi[4];
            uint32_t qlo[4], qhi[4];
            float p = box[0][2 * a];
            for (uint32_t i = 0; i < k; i++) {
                lo[i] = box[i][2 * a];
                hi[i] = box[i][2 * a + 1];
                p = std::min(p, lo[i]);
            }
            if (!quantize_axis(lo, hi, k, p, qlo, qhi, &ebias[a]))
                return fail(RB_ERR_UNSUPPORTED, "a BVH box cannot be stored in 8-bit steps");
            std::memcpy(&w[a], &p, 4);
            for (uint32_t i = 0; i < k; i++) {
                w[4 + 2 * a] |= qlo[i] << (8 * i);
                w[5 + 2 * a] |= qhi[i] << (8 * i);
            }
        }
        w[3] = ebias[0] | ebias[1] << 8 | ebias[2] << 16;
        for (uint32_t i = 0; i < 4; i++) w[10 + i] = i < k ? (uint32_t)link[i] : (uint32_t)kWideEmpty;
        if (order.size() > 0x7FFFFFF0u) return fail(RB_ERR_UNSUPPORTED, "too many BVH nodes");
    }
    return RB_OK;
}

uint32_t layout_depth(const std::vector<uint32_t>& nodes) {
    const uint32_t n = (uint32_t)(nodes.size() / 16);
    if (n == 0) return 0;
    uint32_t deepest = 0;
    std::vector<std::pair<uint32_t, uint32_t>> stack;  // (record, depth)
    stack.push_back({0u, 0u});
    while (!stack.empty()) {
        const auto e = stack.back();
        stack.pop_back();
        if (e.second > deepest) deepest = e.second;
        if (e.second > n) return UINT32_MAX;  // a cycle: no tree is deeper than its record count
        for (int c = 0; c < 2; c++) {
            const int32_t link = (int32_t)nodes[(size_t)e.first * 16 + 12 + c];
            if (link >= 0) {
                if ((uint32_t)link >= n) return UINT32_MAX;
                stack.push_back({(uint32_t)link, e.second + 1});
            }
        }
    }
    return deepest;
}

}  // namespace rb

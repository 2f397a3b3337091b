// rb_sbvh.cpp -- spatial-split BVH builder (Stich et al. SBVH) on the host.
//
// Same decisions as SplitBVHBuilder (renderbox2 src/accelerators/sbvh/
// splitbvhbuilder.cpp:234-627), reorganised for a many-core host:
//  * every node owns its references in a vector (the reference shares one
//    stack); the per-node operations (degenerate removal by swap-with-last,
//    the z-sorted order the spatial partition walks, the swap-based
//    partition, the reversed leaf order) are applied to that vector in the
//    same sequence, so the tree is identical;
//  * the three centroid sorts per node are key sorts (centroid bits, triangle
//    id), a total order inside a node, so their result equals the
//    reference's quicksort; big nodes sort and bin in parallel;
//  * subtrees above a size threshold are built on their own threads.
// The result does not depend on the thread count.
#include <algorithm>
#include <chrono>
#include <climits>
#include <mutex>
#include <thread>

#include "rb_internal.h"

namespace rb {
namespace {

enum { kMaxDepth = kMaxTreeDepth, kMaxSpatialDepth = 48, kBins = 128 };

struct Ref {
    uint32_t tri;
    Box b;
};

// (int)f as x86 cvttss2si computes it: INT_MIN for NaN or out of range.
inline int trunc_int(float f) {
    if (!(f > -2147483648.0f && f < 2147483648.0f)) return INT_MIN;
    return (int)f;
}

// Orderable key of a centroid: -0 folded onto +0 so that equal floats give
// equal keys, then the usual sign-magnitude to unsigned mapping.
inline uint64_t sort_key(const Ref& r, int dim) {
    float c = get(r.b.lo, dim) + get(r.b.hi, dim);
    c = c + 0.0f;
    uint32_t u = fbits(c);
    u = (u & 0x80000000u) ? ~u : (u | 0x80000000u);
    return ((uint64_t)u << 32) | r.tri;
}

struct Arena {
    std::vector<std::unique_ptr<std::vector<BNode>>> nblocks;
    std::vector<std::unique_ptr<std::vector<uint32_t>>> tblocks;
    size_t nused = 0, tused = 0;
    uint32_t* cur = nullptr;
    static constexpr size_t kNodeBlock = 8192, kTriBlock = 65536;
    BNode* node() {
        if (nblocks.empty() || nused == kNodeBlock) {
            nblocks.emplace_back(new std::vector<BNode>(kNodeBlock));
            nused = 0;
        }
        return &(*nblocks.back())[nused++];
    }
    uint32_t* tris(size_t n) {
        if (n == 0) return nullptr;
        if (n > kTriBlock / 4) {  // large leaves get a block of their own
            tblocks.emplace_back(new std::vector<uint32_t>(n));
            return tblocks.back()->data();
        }
        if (!cur || tused + n > kTriBlock) {
            tblocks.emplace_back(new std::vector<uint32_t>(kTriBlock));
            cur = tblocks.back()->data();
            tused = 0;
        }
        uint32_t* p = cur + tused;
        tused += n;
        return p;
    }
};

struct Ctx {
    const float* verts;
    const uint32_t* idx;
    float min_overlap;
    uint32_t min_leaf, max_leaf;
    float ncost, tcost;
    std::atomic<int> spare_threads{0};
    std::atomic<uint64_t> duplicates{0};
    std::atomic<uint32_t> max_depth{0};
    std::atomic<bool> oom{false};
    std::mutex mtx;
    HostBvh* out;
    int total_threads = 1;

    F3 v(uint32_t i) const { return f3(verts[3 * (size_t)i], verts[3 * (size_t)i + 1], verts[3 * (size_t)i + 2]); }
    float tri_cost(uint32_t n) const { return (float)n * tcost; }
    float node_cost(uint32_t n) const { return (float)n * ncost; }

    void adopt(Arena& a) {
        std::lock_guard<std::mutex> g(mtx);
        for (auto& b : a.nblocks) out->node_blocks.push_back(std::move(b));
        for (auto& b : a.tblocks) out->tri_blocks.push_back(std::move(b));
        a.nblocks.clear();
        a.tblocks.clear();
    }
};

// split_reference, splitbvhbuilder.cpp:586-627.
inline void split_tri(const F3 v[3], const Box& within, int dim, float pos, Box& L, Box& R) {
    L = Box();
    R = Box();
    F3 v1 = v[2];
    for (int i = 0; i < 3; i++) {
        F3 v0 = v1;
        v1 = v[i];
        float v0p = get(v0, dim), v1p = get(v1, dim);
        if (v0p <= pos) L.insert(v0);
        if (v0p >= pos) R.insert(v0);
        if ((v0p < pos && v1p > pos) || (v0p > pos && v1p < pos)) {
            float t = (pos - v0p) / (v1p - v0p);
            t = fmax_(0.0f, fmin_(t, 1.0f));
            F3 p = v0 + (v1 - v0) * t;
            L.insert(p);
            R.insert(p);
        }
    }
    getr(L.hi, dim) = pos;
    getr(R.lo, dim) = pos;
    L.intersect(within);
    R.intersect(within);
}

inline void tri_verts(const Ctx& C, uint32_t tri, F3 v[3]) {
    const uint32_t* ind = C.idx + 3 * (size_t)tri;
    for (int i = 0; i < 3; i++) v[i] = C.v(ind[i]);
}

void split_reference(const Ctx& C, Ref& L, Ref& R, const Ref& ref, int dim, float pos) {
    L.tri = R.tri = ref.tri;
    F3 v[3];
    tri_verts(C, ref.tri, v);
    split_tri(v, ref.b, dim, pos, L.b, R.b);
}

// Sort (key, position) pairs; parallel merge sort above a size threshold.
void sort_keys(std::vector<std::pair<uint64_t, uint32_t>>& k, int threads) {
    size_t n = k.size();
    if (threads <= 1 || n < (1u << 16)) {
        std::sort(k.begin(), k.end());
        return;
    }
    int parts = 1;
    while (parts * 2 <= threads && n / (parts * 2) >= (1u << 14)) parts *= 2;
    std::vector<size_t> cut(parts + 1);
    for (int i = 0; i <= parts; i++) cut[i] = n * i / parts;
    {
        std::vector<std::thread> th;
        for (int i = 0; i < parts; i++)
            th.emplace_back([&, i] { std::sort(k.begin() + cut[i], k.begin() + cut[i + 1]); });
        for (auto& t : th) t.join();
    }
    std::vector<std::pair<uint64_t, uint32_t>> tmp(n);
    auto* src = &k;
    auto* dst = &tmp;
    for (int w = 1; w < parts; w *= 2) {
        std::vector<std::thread> th;
        for (int i = 0; i < parts; i += 2 * w) {
            size_t a = cut[i], m = cut[std::min(i + w, parts)], e = cut[std::min(i + 2 * w, parts)];
            th.emplace_back([=] {
                std::merge(src->begin() + a, src->begin() + m, src->begin() + m, src->begin() + e,
                           dst->begin() + a);
            });
        }
        for (auto& t : th) t.join();
        std::swap(src, dst);
    }
    if (src != &k) k.swap(tmp);
}

struct ObjSplit {
    float sah = FLT_MAX;
    int dim = 0;
    uint32_t num_left = 0;
    Box lb, rb;
};
struct SpSplit {
    float sah = FLT_MAX;
    int dim = 0;
    float pos = 0.0f;
};

struct Builder {
    Ctx& C;
    Arena arena;
    explicit Builder(Ctx& c) : C(c) {}

    BNode* leaf(const std::vector<Ref>& refs, const Box& bounds) {
        BNode* n = arena.node();
        n->leaf = true;
        n->b = bounds;
        n->ntris = (uint32_t)refs.size();
        n->tris = arena.tris(refs.size());
        for (size_t i = 0; i < refs.size(); i++) n->tris[i] = refs[refs.size() - 1 - i].tri;
        return n;
    }

    // find_object_split, splitbvhbuilder.cpp:369-408.  order[d] receives the
    // positions of refs sorted along d.
    ObjSplit object_split(const std::vector<Ref>& refs, float nodeSAH, std::vector<uint32_t> order[3],
                          int threads) {
        ObjSplit s;
        const uint32_t n = (uint32_t)refs.size();
        float bestTie = FLT_MAX;
        std::vector<Box> right(n > 0 ? n - 1 : 0);
        std::vector<std::pair<uint64_t, uint32_t>> keys(n);
        for (int dim = 0; dim < 3; dim++) {
            for (uint32_t i = 0; i < n; i++) keys[i] = {sort_key(refs[i], dim), i};
            sort_keys(keys, threads);
            order[dim].resize(n);
            for (uint32_t i = 0; i < n; i++) order[dim][i] = keys[i].second;
            const uint32_t* o = order[dim].data();
            Box rb;
            for (uint32_t i = n - 1; i > 0; i--) {
                rb.grow(refs[o[i]].b);
                right[i - 1] = rb;
            }
            Box lb;
            for (uint32_t i = 1; i < n; i++) {
                lb.grow(refs[o[i - 1]].b);
                float sah = nodeSAH + lb.area() * C.tri_cost(i) + right[i - 1].area() * C.tri_cost(n - i);
                float fi = (float)i, fr = (float)(n - i);
                float tie = fi * fi + fr * fr;
                if (sah < s.sah || (sah == s.sah && tie < bestTie)) {
                    s.sah = sah;
                    s.dim = dim;
                    s.num_left = i;
                    s.lb = lb;
                    s.rb = right[i - 1];
                    bestTie = tie;
                }
            }
        }
        return s;
    }

    // used[d][i]: bin i of axis d was grown at least once.  Box::grow inserts
    // the other box's corners, so growing with an untouched (empty) bin would
    // make the result infinite: a parallel merge must skip those.
    struct Bins {
        Box b[3][kBins];
        uint32_t enter[3][kBins] = {}, exit[3][kBins] = {};
        bool used[3][kBins] = {};
    };

    void chop(const std::vector<Ref>& refs, size_t lo, size_t hi, F3 origin, F3 binSize, F3 inv, Bins& B) {
        for (size_t ri = lo; ri < hi; ri++) {
            const Ref& ref = refs[ri];
            F3 rmin = f3((ref.b.lo.x - origin.x) * inv.x, (ref.b.lo.y - origin.y) * inv.y, (ref.b.lo.z - origin.z) * inv.z);
            F3 rmax = f3((ref.b.hi.x - origin.x) * inv.x, (ref.b.hi.y - origin.y) * inv.y, (ref.b.hi.z - origin.z) * inv.z);
            F3 tv[3];
            bool have_tv = false;
            for (int d = 0; d < 3; d++) {
                int f = trunc_int(get(rmin, d));
                f = f < 0 ? 0 : (f > kBins - 1 ? kBins - 1 : f);
                int l = trunc_int(get(rmax, d));
                l = l > kBins - 1 ? kBins - 1 : l;
                l = l < f ? f : l;
                Box cur = ref.b;
                if (f < l && !have_tv) {  // the triangle's vertices, loaded once per reference
                    tri_verts(C, ref.tri, tv);
                    have_tv = true;
                }
                for (int i = f; i < l; i++) {
                    Box L, R;
                    split_tri(tv, cur, d, get(origin, d) + get(binSize, d) * (float)(i + 1), L, R);
                    B.b[d][i].grow(L);
                    B.used[d][i] = true;
                    cur = R;
                }
                B.b[d][l].grow(cur);
                B.used[d][l] = true;
                B.enter[d][f]++;
                B.exit[d][l]++;
            }
        }
    }

    // find_spatial_split, splitbvhbuilder.cpp:420-497.  Bin bounds merge by
    // min/max over the bins each range touched and counts by sums, so chopping
    // in parallel gives the serial bins exactly.
    SpSplit spatial_split(const std::vector<Ref>& refs, const Box& nb, float nodeSAH, int threads) {
        F3 origin = nb.lo;
        F3 binSize = (nb.hi - origin) * (1.0f / (float)kBins);
        F3 inv = f3(1.0f / binSize.x, 1.0f / binSize.y, 1.0f / binSize.z);
        const uint32_t n = (uint32_t)refs.size();
        std::unique_ptr<Bins> B(new Bins());
        if (threads > 1 && n >= (1u << 14)) {
            std::vector<std::unique_ptr<Bins>> part(threads);
            std::vector<std::thread> th;
            for (int t = 0; t < threads; t++) {
                part[t].reset(new Bins());
                size_t lo = (size_t)n * t / threads, hi = (size_t)n * (t + 1) / threads;
                th.emplace_back([&, t, lo, hi] { chop(refs, lo, hi, origin, binSize, inv, *part[t]); });
            }
            for (auto& x : th) x.join();
            for (int t = 0; t < threads; t++)
                for (int d = 0; d < 3; d++)
                    for (int i = 0; i < kBins; i++) {
                        if (part[t]->used[d][i]) B->b[d][i].grow(part[t]->b[d][i]);
                        B->enter[d][i] += part[t]->enter[d][i];
                        B->exit[d][i] += part[t]->exit[d][i];
                    }
        } else {
            chop(refs, 0, n, origin, binSize, inv, *B);
        }
        SpSplit s;
        Box right[kBins];
        for (int d = 0; d < 3; d++) {
            Box rb;
            for (int i = kBins - 1; i > 0; i--) {
                rb.grow(B->b[d][i]);
                right[i - 1] = rb;
            }
            Box lb;
            uint32_t ln = 0, rn = n;
            for (int i = 1; i < kBins; i++) {
                lb.grow(B->b[d][i - 1]);
                ln += B->enter[d][i - 1];
                rn -= B->exit[d][i - 1];
                float sah = nodeSAH + lb.area() * C.tri_cost(ln) + right[i - 1].area() * C.tri_cost(rn);
                if (sah < s.sah) {
                    s.sah = sah;
                    s.dim = d;
                    s.pos = get(origin, d) + get(binSize, d) * (float)i;
                }
            }
        }
        return s;
    }

    // perform_spatial_split, splitbvhbuilder.cpp:499-584, on the node's refs
    // (already in z order).  Duplicates are appended at the end as the
    // reference appends them to its stack.
    void spatial_partition(std::vector<Ref>& refs, const SpSplit& sp, std::vector<Ref>& left, Box& lb,
                           std::vector<Ref>& right, Box& rb) {
        size_t leftEnd = 0, rightStart = refs.size();
        lb = Box();
        rb = Box();
        for (size_t i = leftEnd; i < rightStart; i++) {
            if (get(refs[i].b.hi, sp.dim) <= sp.pos) {
                lb.grow(refs[i].b);
                std::swap(refs[i], refs[leftEnd++]);
            } else if (get(refs[i].b.lo, sp.dim) >= sp.pos) {
                rb.grow(refs[i].b);
                std::swap(refs[i], refs[--rightStart]);
                i--;
            }
        }
        while (leftEnd < rightStart) {
            Ref L, R;
            split_reference(C, L, R, refs[leftEnd], sp.dim, sp.pos);
            Box lub = lb, rub = rb, ldb = lb, rdb = rb;
            lub.grow(refs[leftEnd].b);
            rub.grow(refs[leftEnd].b);
            ldb.grow(L.b);
            rdb.grow(R.b);
            float lac = C.tri_cost((uint32_t)leftEnd);
            float rac = C.tri_cost((uint32_t)(refs.size() - rightStart));
            float lbc = C.tri_cost((uint32_t)(leftEnd + 1));
            float rbc = C.tri_cost((uint32_t)(refs.size() - rightStart + 1));
            float usl = lub.area() * lbc + rb.area() * rac;
            float usr = lb.area() * lac + rub.area() * rbc;
            float dup = ldb.area() * lbc + rdb.area() * rbc;
            float m = std::min(usl, std::min(usr, dup));
            if (m == usl) {
                lb = lub;
                leftEnd++;
            } else if (m == usr) {
                rb = rub;
                std::swap(refs[leftEnd], refs[--rightStart]);
            } else {
                lb = ldb;
                rb = rdb;
                refs[leftEnd++] = L;
                refs.push_back(R);
            }
        }
        left.assign(refs.begin(), refs.begin() + leftEnd);
        right.assign(refs.begin() + rightStart, refs.end());
    }

    // build_node, splitbvhbuilder.cpp:290-357.
    BNode* build(std::vector<Ref>& refs, Box bounds, int level) {
        if (C.oom.load(std::memory_order_relaxed)) return nullptr;
        uint32_t md = C.max_depth.load();
        while ((uint32_t)level > md && !C.max_depth.compare_exchange_weak(md, (uint32_t)level)) {}
        // Remove degenerates (:300-316).
        for (long i = (long)refs.size() - 1; i >= 0; i--) {
            F3 sz = refs[i].b.hi - refs[i].b.lo;
            float mn = sz.x;
            if (sz.y < mn) mn = sz.y;
            if (sz.z < mn) mn = sz.z;
            float mx = sz.x;
            if (sz.y > mx) mx = sz.y;
            if (sz.z > mx) mx = sz.z;
            if (mn < 0.0f || sz.x + sz.y + sz.z == mx) {
                Ref back = refs.back();
                refs.pop_back();
                if ((size_t)i < refs.size()) refs[i] = back;
            }
        }
        const uint32_t num = (uint32_t)refs.size();
        if (num <= C.min_leaf || level >= kMaxDepth) return leaf(refs, bounds);

        const int threads = std::max(1, num >= (1u << 16) ? C.total_threads : 1);
        float area = bounds.area();
        float leafSAH = area * C.tri_cost(num);
        float nodeSAH = area * C.node_cost(2);
        std::vector<uint32_t> order[3];
        ObjSplit obj = object_split(refs, nodeSAH, order, threads);
        {  // leave refs in the last sort's (z) order, as the reference does
            std::vector<Ref> z(num);
            for (uint32_t i = 0; i < num; i++) z[i] = refs[order[2][i]];
            refs.swap(z);
        }
        // order[obj.dim] now indexes the pre-sort array; rebase it on z order.
        if (obj.dim != 2) {
            // positions in z-order of each original position
            std::vector<uint32_t> zpos(num);
            for (uint32_t i = 0; i < num; i++) zpos[order[2][i]] = i;
            for (uint32_t i = 0; i < num; i++) order[obj.dim][i] = zpos[order[obj.dim][i]];
        }

        SpSplit sp;
        if (level < kMaxSpatialDepth) {
            Box ov = obj.lb;
            ov.intersect(obj.rb);
            if (ov.area() >= C.min_overlap) sp = spatial_split(refs, bounds, nodeSAH, threads);
        }
        float minSAH = std::min(leafSAH, std::min(obj.sah, sp.sah));
        if (minSAH == leafSAH && num <= C.max_leaf) return leaf(refs, bounds);

        std::vector<Ref> left, right;
        Box lb, rb;
        bool done = false;
        if (minSAH == sp.sah) {
            std::vector<Ref> work(refs);
            spatial_partition(work, sp, left, lb, right, rb);
            done = !left.empty() && !right.empty();
        }
        if (!done) {
            // perform_object_split (:409-418): sorted along obj.dim.
            left.resize(obj.num_left);
            right.resize(num - obj.num_left);
            if (obj.dim == 2) {
                std::copy(refs.begin(), refs.begin() + obj.num_left, left.begin());
                std::copy(refs.begin() + obj.num_left, refs.end(), right.begin());
            } else {
                const uint32_t* o = order[obj.dim].data();
                for (uint32_t i = 0; i < obj.num_left; i++) left[i] = refs[o[i]];
                for (uint32_t i = obj.num_left; i < num; i++) right[i - obj.num_left] = refs[o[i]];
            }
            lb = obj.lb;
            rb = obj.rb;
        }
        C.duplicates += (uint64_t)left.size() + right.size() - num;
        for (auto& o : order) std::vector<uint32_t>().swap(o);
        std::vector<Ref>().swap(refs);

        BNode* node = arena.node();
        node->leaf = false;
        node->b = bounds;
        BNode* rnode = nullptr;
        BNode* lnode = nullptr;
        bool spawned = false;
        std::thread th;
        if (right.size() >= 4096 && left.size() >= 4096) {
            int spare = C.spare_threads.load();
            while (spare > 0 && !C.spare_threads.compare_exchange_weak(spare, spare - 1)) {}
            if (spare > 0) {
                spawned = true;
                th = std::thread([&] {
                    Builder sub(C);
                    try {
                        rnode = sub.build(right, rb, level + 1);
                    } catch (const std::bad_alloc&) {
                        C.oom = true;
                    }
                    C.adopt(sub.arena);
                    C.spare_threads++;
                });
            }
        }
        if (!spawned) rnode = build(right, rb, level + 1);
        lnode = build(left, lb, level + 1);
        if (spawned) th.join();
        if (!rnode || !lnode) {
            C.oom = true;
            return nullptr;
        }
        node->c[0] = lnode;
        node->c[1] = rnode;
        return node;
    }
};

// compute_subtree_probabilities, bvnode.cpp:81-100 (pre-order accumulation).
void sah_rec(const BNode* n, float prob, float nc, float tc, float& sah, rb_bvh_stats& st) {
    uint32_t nchild = n->leaf ? 0 : 2;
    uint32_t ntri = n->leaf ? n->ntris : 0;
    sah += prob * ((float)nchild * nc + (float)ntri * tc);
    if (n->leaf) {
        st.n_leaf++;
        st.n_tri_refs += ntri;
        return;
    }
    st.n_inner++;
    st.n_child_nodes += 2;
    for (int i = 0; i < 2; i++) {
        float cp = 0.0f;
        if (prob > 0.0f) cp = prob * n->c[i]->b.area() / n->b.area();
        sah_rec(n->c[i], cp, nc, tc, sah, st);
    }
}

}  // namespace

HostBvh* build_sbvh(const float* verts, uint32_t nv, const uint32_t* idx, uint32_t nt,
                    const rb_build_params& p, rb_status* st) {
    auto t0 = std::chrono::steady_clock::now();
    for (uint64_t i = 0; i < 3ull * nt; i++)
        if (idx[i] >= nv) {
            *st = fail(RB_ERR_INVALID, "triangle %llu indexes vertex %u of %u", (unsigned long long)(i / 3),
                       idx[i], nv);
            return nullptr;
        }
    std::unique_ptr<HostBvh> H(new HostBvh());
    H->verts.assign(verts, verts + 3 * (size_t)nv);
    H->idx.assign(idx, idx + 3 * (size_t)nt);
    H->params = p;
    Ctx C;
    C.verts = H->verts.data();
    C.idx = H->idx.data();
    C.min_leaf = p.min_leaf;
    C.max_leaf = p.max_leaf;
    C.ncost = p.node_cost;
    C.tcost = p.tri_cost;
    C.out = H.get();
    int hw = (int)std::thread::hardware_concurrency();
    C.total_threads = p.threads ? (int)p.threads : std::max(1, std::min(hw, 16));
    C.spare_threads = C.total_threads - 1;

    std::vector<Ref> refs(nt);
    Box root;
    for (uint32_t i = 0; i < nt; i++) {
        refs[i].tri = i;
        refs[i].b = Box();
        refs[i].b.insert(C.v(idx[3 * (size_t)i]));
        refs[i].b.insert(C.v(idx[3 * (size_t)i + 1]));
        refs[i].b.insert(C.v(idx[3 * (size_t)i + 2]));
        root.grow(refs[i].b);
    }
    C.min_overlap = root.area() * p.split_alpha;
    Builder B(C);
    try {
        H->root = B.build(refs, root, 0);
    } catch (const std::bad_alloc&) {
        C.oom = true;
    }
    C.adopt(B.arena);
    if (C.oom || !H->root) {
        *st = fail(RB_ERR_OOM, "SBVH build ran out of host memory");
        return nullptr;
    }
    rb_bvh_stats& S = H->stats;
    std::memset(&S, 0, sizeof(S));
    float sah = 0.0f;
    sah_rec(H->root, 1.0f, p.node_cost, p.tri_cost, sah, S);
    S.sah = sah;
    S.duplicates = C.duplicates.load();
    S.max_depth = C.max_depth.load();
    S.build_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    *st = RB_OK;
    return H.release();
}

}  // namespace rb

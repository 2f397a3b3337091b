// rb_bvh_cache.cpp -- binary cache of built SBVHs.
//
// The reference caches its device layout as decimal text
// (CudaBvh::serialize / CudaBvh(filename), src/accelerators/sbvh/cudabvh.cpp:108-214),
// chosen by a yes/no flag in the config (src/core/application.cpp:131-161) with
// nothing tying the file to the scene it was built from.  This cache stores
// the host tree itself, so every layout (Compact2, MI355X) can be exported
// from it, and keys it by a hash of the geometry and of the build
// parameters that shape the tree: a file built from other triangles or
// other costs is refused, not used.
//
// File: header (magic "RBVHBIN1", version, key, sizes, parameters, stats),
// nodes in preorder (box, then either leaf count + offset into the
// reference list, or the index of the second child -- the first child is
// the next record), the reference list, and a checksum of everything
// before it.  Little-endian, as the host.
#include <cstdio>
#include <fstream>
#include <string>

#include "rb_internal.h"

namespace rb {
namespace {

constexpr char kMagic[8] = {'R', 'B', 'V', 'H', 'B', 'I', 'N', '1'};
constexpr uint32_t kVersion = 1;

// 64-bit hash over bytes: 8-byte words mixed with multiply / rotate rounds
// (the tail zero-padded), length folded in at the end.
struct Hash64 {
    uint64_t h = 0x9E3779B97F4A7C15ull;
    uint64_t len = 0;
    static uint64_t rotl(uint64_t x, int r) { return (x << r) | (x >> (64 - r)); }
    void word(uint64_t w) {
        h ^= rotl(w * 0xC2B2AE3D27D4EB4Full, 31) * 0x9E3779B185EBCA87ull;
        h = rotl(h, 27) * 0x9E3779B185EBCA87ull + 0x85EBCA77C2B2AE63ull;
    }
    void bytes(const void* p, size_t n) {
        const unsigned char* c = static_cast<const unsigned char*>(p);
        size_t i = 0;
        for (; i + 8 <= n; i += 8) {
            uint64_t w;
            std::memcpy(&w, c + i, 8);
            word(w);
        }
        if (i < n) {
            uint64_t w = 0;
            std::memcpy(&w, c + i, n - i);
            word(w);
        }
        len += n;
    }
    uint64_t done() const {
        uint64_t x = h ^ len;
        x ^= x >> 33;
        x *= 0xFF51AFD7ED558CCDull;
        x ^= x >> 33;
        x *= 0xC4CEB9FE1A85EC53ull;
        return x ^ (x >> 33);
    }
};

#pragma pack(push, 1)
struct FileHeader {
    char magic[8];
    uint32_t version;
    uint32_t reserved;
    uint64_t key;
    uint32_t nv, nt, n_nodes, n_refs;
    uint32_t min_leaf, max_leaf;
    float split_alpha, node_cost, tri_cost;
    float sah;
    uint32_t n_inner, n_leaf, n_tri_refs, n_child_nodes, max_depth;
    uint64_t duplicates;
    double build_ms;
};
struct FileNode {
    float lo[3], hi[3];
    uint32_t leaf_count;  // bit 31: leaf; low bits: references in the leaf
    uint32_t link;        // leaf: first reference; inner: index of the second child
};
#pragma pack(pop)

void count(const BNode* n, uint32_t& nodes, uint32_t& refs) {
    nodes++;
    if (n->leaf) {
        refs += n->ntris;
        return;
    }
    count(n->c[0], nodes, refs);
    count(n->c[1], nodes, refs);
}

void flatten(const BNode* n, std::vector<FileNode>& out, std::vector<uint32_t>& refs) {
    const size_t me = out.size();
    out.emplace_back();
    FileNode& f = out[me];
    std::memcpy(f.lo, &n->b.lo, 12);
    std::memcpy(f.hi, &n->b.hi, 12);
    if (n->leaf) {
        f.leaf_count = 0x80000000u | n->ntris;
        f.link = (uint32_t)refs.size();
        refs.insert(refs.end(), n->tris, n->tris + n->ntris);
        return;
    }
    f.leaf_count = 0;
    flatten(n->c[0], out, refs);
    out[me].link = (uint32_t)out.size();
    flatten(n->c[1], out, refs);
}

}  // namespace

uint64_t bvh_cache_key(const float* verts, uint32_t nv, const uint32_t* idx, uint32_t nt, const rb_build_params& p) {
    Hash64 h;
    h.bytes(&nv, 4);
    h.bytes(&nt, 4);
    h.bytes(verts, (size_t)nv * 12);
    h.bytes(idx, (size_t)nt * 12);
    // thread count does not shape the tree
    h.bytes(&p.min_leaf, 4);
    h.bytes(&p.max_leaf, 4);
    h.bytes(&p.split_alpha, 4);
    h.bytes(&p.node_cost, 4);
    h.bytes(&p.tri_cost, 4);
    return h.done();
}

rb_status save_bvh_binary(const HostBvh& b, const char* path) {
    const uint32_t nv = (uint32_t)(b.verts.size() / 3), nt = (uint32_t)(b.idx.size() / 3);
    uint32_t n_nodes = 0, n_refs = 0;
    count(b.root, n_nodes, n_refs);
    std::vector<FileNode> nodes;
    std::vector<uint32_t> refs;
    nodes.reserve(n_nodes);
    refs.reserve(n_refs);
    flatten(b.root, nodes, refs);
    FileHeader hd{};
    std::memcpy(hd.magic, kMagic, 8);
    hd.version = kVersion;
    hd.key = bvh_cache_key(b.verts.data(), nv, b.idx.data(), nt, b.params);
    hd.nv = nv;
    hd.nt = nt;
    hd.n_nodes = n_nodes;
    hd.n_refs = n_refs;
    hd.min_leaf = b.params.min_leaf;
    hd.max_leaf = b.params.max_leaf;
    hd.split_alpha = b.params.split_alpha;
    hd.node_cost = b.params.node_cost;
    hd.tri_cost = b.params.tri_cost;
    hd.sah = b.stats.sah;
    hd.n_inner = b.stats.n_inner;
    hd.n_leaf = b.stats.n_leaf;
    hd.n_tri_refs = b.stats.n_tri_refs;
    hd.n_child_nodes = b.stats.n_child_nodes;
    hd.max_depth = b.stats.max_depth;
    hd.duplicates = b.stats.duplicates;
    hd.build_ms = b.stats.build_ms;
    Hash64 sum;
    sum.bytes(&hd, sizeof(hd));
    sum.bytes(nodes.data(), nodes.size() * sizeof(FileNode));
    sum.bytes(refs.data(), refs.size() * 4);
    const uint64_t check = sum.done();
    // write to a temporary name, then rename: readers never see a partial file
    const std::string tmp = std::string(path) + ".tmp";
    {
        std::ofstream out(tmp, std::ios::binary | std::ios::trunc);
        if (!out) return fail(RB_ERR_INVALID, "cannot open %s", tmp.c_str());
        out.write(reinterpret_cast<const char*>(&hd), sizeof(hd));
        out.write(reinterpret_cast<const char*>(nodes.data()), (std::streamsize)(nodes.size() * sizeof(FileNode)));
        out.write(reinterpret_cast<const char*>(refs.data()), (std::streamsize)(refs.size() * 4));
        out.write(reinterpret_cast<const char*>(&check), 8);
        if (!out) {
            out.close();
            std::remove(tmp.c_str());
            return fail(RB_ERR_INVALID, "write to %s failed", tmp.c_str());
        }
    }
    if (std::rename(tmp.c_str(), path) != 0) {
        std::remove(tmp.c_str());
        return fail(RB_ERR_INVALID, "cannot move %s to %s", tmp.c_str(), path);
    }
    return RB_OK;
}

rb_status load_bvh_binary(const char* path, const float* verts, uint32_t nv, const uint32_t* idx, uint32_t nt,
                          const rb_build_params& p, HostBvh** out) {
    *out = nullptr;
    std::ifstream in(path, std::ios::binary);
    if (!in) return fail(RB_ERR_INVALID, "cannot open %s", path);
    FileHeader hd{};
    in.read(reinterpret_cast<char*>(&hd), sizeof(hd));
    if (!in || std::memcmp(hd.magic, kMagic, 8) != 0) return fail(RB_ERR_INVALID, "%s: not a binary BVH file", path);
    if (hd.version != kVersion) return fail(RB_ERR_INVALID, "%s: version %u, expected %u", path, hd.version, kVersion);
    if (hd.nv != nv || hd.nt != nt || hd.key != bvh_cache_key(verts, nv, idx, nt, p))
        return fail(RB_ERR_INVALID, "%s was built from other geometry or build parameters", path);
    if (hd.n_nodes == 0 || hd.n_nodes > (1u << 31) || hd.n_refs > (1u << 31))
        return fail(RB_ERR_INVALID, "%s: bad sizes", path);
    // the header's sizes must account for the file exactly before anything
    // is allocated from them
    const uint64_t want = sizeof(FileHeader) + (uint64_t)hd.n_nodes * sizeof(FileNode) + (uint64_t)hd.n_refs * 4 + 8;
    in.seekg(0, std::ios::end);
    const std::streamoff have = in.tellg();
    if (have < 0 || (uint64_t)have != want)
        return fail(RB_ERR_INVALID, "%s: %lld bytes, the header describes %llu", path, (long long)have,
                    (unsigned long long)want);
    in.seekg((std::streamoff)sizeof(FileHeader), std::ios::beg);
    std::vector<FileNode> nodes(hd.n_nodes);
    std::vector<uint32_t> refs(hd.n_refs);
    uint64_t check = 0;
    in.read(reinterpret_cast<char*>(nodes.data()), (std::streamsize)(nodes.size() * sizeof(FileNode)));
    in.read(reinterpret_cast<char*>(refs.data()), (std::streamsize)(refs.size() * 4));
    in.read(reinterpret_cast<char*>(&check), 8);
    if (!in) return fail(RB_ERR_INVALID, "%s: truncated", path);
    Hash64 sum;
    sum.bytes(&hd, sizeof(hd));
    sum.bytes(nodes.data(), nodes.size() * sizeof(FileNode));
    sum.bytes(refs.data(), refs.size() * 4);
    if (sum.done() != check) return fail(RB_ERR_INVALID, "%s: checksum mismatch (corrupt file)", path);

    std::unique_ptr<HostBvh> H(new HostBvh());
    H->verts.assign(verts, verts + 3 * (size_t)nv);
    H->idx.assign(idx, idx + 3 * (size_t)nt);
    H->params = p;
    H->stats.sah = hd.sah;
    H->stats.n_inner = hd.n_inner;
    H->stats.n_leaf = hd.n_leaf;
    H->stats.n_tri_refs = hd.n_tri_refs;
    H->stats.n_child_nodes = hd.n_child_nodes;
    H->stats.max_depth = hd.max_depth;
    H->stats.duplicates = hd.duplicates;
    H->stats.build_ms = hd.build_ms;
    H->node_blocks.emplace_back(new std::vector<BNode>(hd.n_nodes));
    H->tri_blocks.emplace_back(new std::vector<uint32_t>(std::move(refs)));
    std::vector<BNode>& bn = *H->node_blocks.back();
    uint32_t* tri = H->tri_blocks.back()->data();
    // Children always follow their parent, so one forward pass gives every
    // node's depth (set by the parent before the node is reached) and how
    // often it is linked: a tree links every node but the root exactly once,
    // and the builder's trees are at most kMaxTreeDepth deep (the traversal
    // stack is sized from that bound).  The checksum detects corruption; these
    // checks refuse a well-formed file that no build wrote.
    std::vector<uint8_t> depth(hd.n_nodes, 0), links(hd.n_nodes, 0);
    for (uint32_t i = 0; i < hd.n_nodes; i++) {
        const FileNode& f = nodes[i];
        BNode& n = bn[i];
        if (i > 0 && links[i] != 1)
            return fail(RB_ERR_INVALID, "%s: node %u is linked %u times (not a tree)", path, i, links[i]);
        if (depth[i] > kMaxTreeDepth)
            return fail(RB_ERR_INVALID, "%s: node %u at depth %u, deeper than a build makes (%u)", path, i, depth[i],
                        kMaxTreeDepth);
        std::memcpy(&n.b.lo, f.lo, 12);
        std::memcpy(&n.b.hi, f.hi, 12);
        if (f.leaf_count & 0x80000000u) {
            const uint32_t cnt = f.leaf_count & 0x7FFFFFFFu;
            if ((uint64_t)f.link + cnt > hd.n_refs) return fail(RB_ERR_INVALID, "%s: leaf %u out of range", path, i);
            for (uint32_t k = 0; k < cnt; k++)
                if (tri[f.link + k] >= nt) return fail(RB_ERR_INVALID, "%s: triangle id out of range", path);
            n.leaf = true;
            n.ntris = cnt;
            n.tris = cnt ? tri + f.link : nullptr;
        } else {
            if (i + 1 >= hd.n_nodes || f.link <= i + 1 || f.link >= hd.n_nodes)
                return fail(RB_ERR_INVALID, "%s: node %u links out of range", path, i);
            n.leaf = false;
            n.c[0] = &bn[i + 1];
            n.c[1] = &bn[f.link];
            for (uint32_t c : {i + 1, f.link}) {
                depth[c] = (uint8_t)(depth[i] + 1);
                if (links[c] < 255) links[c]++;
            }
        }
    }
    H->root = &bn[0];
    *out = H.release();
    return RB_OK;
}

}  // namespace rb

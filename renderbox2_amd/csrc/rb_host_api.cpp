// rb_host_api.cpp -- host half of the C-ABI: errors, scenes, seeded scene
// generators, SBVH build/export and the reference's text BVH format.
#include <cstdarg>
#include <cstdio>
#include <fstream>
#include <map>
#include <mutex>
#include <sstream>

#include "rb_internal.h"
#include "rb_rng.h"

namespace rb {

static thread_local std::string g_err;

void set_error(const char* fmt, ...) {
    char buf[1024];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    g_err = buf;
}

rb_status fail(rb_status s, const char* fmt, ...) {
    char buf[1024];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    g_err = buf;
    return s;
}

// One index array of a mesh's corners: every index of a used triple below
// `count`, or the whole triple kNoIndex (a triangle without that data).
static rb_status check_corners(const uint32_t* c, uint32_t nt, uint32_t count, const char* what) {
    if (!c) return RB_OK;
    for (uint64_t t = 0; t < nt; t++) {
        const uint32_t a = c[3 * t], b = c[3 * t + 1], d = c[3 * t + 2];
        if (a == kNoIndex && b == kNoIndex && d == kNoIndex) continue;
        if (a >= count || b >= count || d >= count)
            return fail(RB_ERR_INVALID, "mesh triangle %llu: %s index out of range (%u %u %u of %u; a triangle "
                        "without %ss takes %u %u %u)", (unsigned long long)t, what, a, b, d, count, what, kNoIndex,
                        kNoIndex, kNoIndex);
    }
    return RB_OK;
}

// Appends one per-corner attribute (normals: 3 floats, texcoords: 2) with
// its index triples, offset into the scene's arrays; keeps `sidx` empty
// while no triangle of the scene has the attribute.
static void append_corners(std::vector<float>& sdata, std::vector<uint32_t>& sidx, uint32_t tris_before,
                           const float* data, uint32_t count, uint32_t width, const uint32_t* c, uint32_t nt) {
    bool any = false;
    for (uint64_t i = 0; c && i < 3ull * nt && !any; i++) any = c[i] != kNoIndex;
    if (!any) {
        if (!sidx.empty()) sidx.insert(sidx.end(), 3 * (size_t)nt, kNoIndex);
        return;
    }
    if (sidx.empty()) sidx.assign(3 * (size_t)tris_before, kNoIndex);
    const uint32_t base = (uint32_t)(sdata.size() / width);
    sdata.insert(sdata.end(), data, data + (size_t)width * count);
    for (uint64_t i = 0; i < 3ull * nt; i++) sidx.push_back(c[i] == kNoIndex ? kNoIndex : c[i] + base);
}

rb_status append_mesh(Scene& S, const float* verts, uint32_t nv, const uint32_t* idx, uint32_t nt,
                      const float* normals, uint32_t nn, const uint32_t* nidx, const float* texcoords, uint32_t ntc,
                      const uint32_t* tidx, const uint32_t* tri_mat, uint32_t material_id) {
    rb_status st = scene_room(S, nv, nt);  // before any array is read
    if (st != RB_OK) return st;
    if ((uint64_t)S.normals.size() / 3 + nn > kMaxSceneVerts || (uint64_t)S.texcoords.size() / 2 + ntc > kMaxSceneVerts)
        return fail(RB_ERR_INVALID, "scene would exceed 2^32 - 1 normals or texture coordinates");
    if (nt && (!verts || !idx)) return fail(RB_ERR_INVALID, "NULL mesh arrays");
    if ((nn && !normals) || (ntc && !texcoords)) return fail(RB_ERR_INVALID, "NULL normal or texcoord array");
    for (uint64_t i = 0; i < 3ull * nt; i++)
        if (idx[i] >= nv) return fail(RB_ERR_INVALID, "mesh triangle %llu indexes vertex %u of %u",
                                      (unsigned long long)(i / 3), idx[i], nv);
    // normals without their own indices: one per vertex, indexed like the vertices
    if (normals && !nidx) {
        if (nn != nv) return fail(RB_ERR_INVALID, "%u normals without normal indices need one per vertex (%u)", nn, nv);
        nidx = idx;
    }
    if (texcoords && !tidx) {
        if (ntc != nv) return fail(RB_ERR_INVALID, "%u texcoords without indices need one per vertex (%u)", ntc, nv);
        tidx = idx;
    }
    if (!normals) nidx = nullptr;
    if (!texcoords) tidx = nullptr;
    if ((st = check_corners(nidx, nt, nn, "normal")) != RB_OK) return st;
    if ((st = check_corners(tidx, nt, ntc, "texcoord")) != RB_OK) return st;
    for (uint32_t t = 0; tri_mat && t < nt; t++)
        if (tri_mat[t] >= S.mats.size()) return fail(RB_ERR_INVALID, "material id %u out of range", tri_mat[t]);
    const uint32_t base = S.num_verts();
    const uint32_t tbase = S.num_tris();
    try {
        S.verts.insert(S.verts.end(), verts, verts + 3 * (size_t)nv);
        for (uint64_t i = 0; i < 3ull * nt; i++) S.idx.push_back(idx[i] + base);
        append_corners(S.normals, S.nidx, tbase, normals, nn, 3, nidx, nt);
        append_corners(S.texcoords, S.tidx, tbase, texcoords, ntc, 2, tidx, nt);
        for (uint32_t t = 0; t < nt; t++) {
            const uint32_t m = tri_mat ? tri_mat[t] : material_id;
            S.tri_mat.push_back(m);
            if (S.mats[m].kind == RB_MAT_EMITTER) S.light_tris.push_back(tbase + t);
        }
    } catch (...) {
        return fail(RB_ERR_OOM, "scene growth failed");
    }
    return RB_OK;
}

}  // namespace rb

using namespace rb;

#define RB_CHECK_ARG(c, msg)                            \
    do {                                                \
        if (!(c)) return fail(RB_ERR_INVALID, "%s", msg); \
    } while (0)

namespace rb {
// Checks a material and fills its derived parameters; the layers of a
// layered material are looked up in `mats`.
rb_status prepare_material(const std::vector<rb_material>& mats, rb_material* m) {
    RB_CHECK_ARG(m->kind <= RB_MAT_LAYERED, "unknown material kind");
    rb_material& c = *m;
    if (c.kind == RB_MAT_LAYERED) {
        const float nf = c.param[0];
        RB_CHECK_ARG(nf >= 1.0f && nf <= 4.0f && nf == (float)(uint32_t)nf, "a layered material has 1..4 layers");
        const uint32_t n = (uint32_t)nf;
        for (uint32_t k = 0; k < n; k++) {
            const float f = c.param[1 + k];
            if (!(f >= 0.0f && f == (float)(uint32_t)f && (uint32_t)f < mats.size()))
                return fail(RB_ERR_INVALID, "layer %u: material id %g is not a material of the scene", k, f);
            const uint32_t kind = mats[(uint32_t)f].kind;
            if (kind == RB_MAT_EMITTER || kind == RB_MAT_LAYERED)
                return fail(RB_ERR_INVALID, "layer %u: material %u is an emitter or layered (layers are single BSDFs)",
                            k, (uint32_t)f);
        }
        for (uint32_t k = n + 1; k < 8; k++) c.param[k] = 0.0f;
        c.color[0] = c.color[1] = c.color[2] = 0.0f;
    }
    if (c.kind == RB_MAT_GLASS) RB_CHECK_ARG(c.param[0] > 0.0f && c.param[1] > 0.0f, "glass needs etai, etat > 0");
    if (c.kind == RB_MAT_ORENNAYAR) {
        // A and B from sigma (CustomFileReader::process, src/io/customfilereader.cpp:137-146)
        float sigma = c.param[0] < 0.0f ? 0.0f : (c.param[0] > 90.0f ? 90.0f : c.param[0]);
        sigma = (3.14159265358979f / 180.0f) * sigma;
        const float s2 = sigma * sigma;
        c.param[1] = 1.0f - (s2 / (2.0f * (s2 + 0.33f)));
        c.param[2] = 0.45f * s2 / (s2 + 0.09f);
    }
    if (c.kind == RB_MAT_MICROFACET) {
        RB_CHECK_ARG(c.param[0] > 0.0f, "microfacet exponent must be > 0");
        const float ft = c.param[1];
        RB_CHECK_ARG(ft == RB_FRESNEL_NONE || ft == RB_FRESNEL_CONDUCTOR || ft == RB_FRESNEL_DIELECTRIC,
                     "microfacet Fresnel type must be 0, 1 or 2");
        if (ft == RB_FRESNEL_DIELECTRIC)
            RB_CHECK_ARG(c.param[2] > 0.0f && c.param[3] > 0.0f, "dielectric etas must be > 0");
    }
    return RB_OK;
}

}  // namespace rb

extern "C" {

size_t rb_last_error(char* buf, size_t n) {
    if (buf && n) {
        size_t k = g_err.size() < n - 1 ? g_err.size() : n - 1;
        std::memcpy(buf, g_err.data(), k);
        buf[k] = 0;
    }
    return g_err.size();
}

const char* rb_status_string(rb_status s) {
    switch (s) {
        case RB_OK: return "RB_OK";
        case RB_ERR_INVALID: return "RB_ERR_INVALID";
        case RB_ERR_OOM: return "RB_ERR_OOM";
        case RB_ERR_DEVICE: return "RB_ERR_DEVICE";
        case RB_ERR_BUSY: return "RB_ERR_BUSY";
        case RB_ERR_UNSUPPORTED: return "RB_ERR_UNSUPPORTED";
        case RB_ERR_NOT_READY: return "RB_ERR_NOT_READY";
    }
    return "RB_ERR_?";
}

const char* rb_version(void) { return "renderbox2_amd 0.1 (gfx950)"; }

// ------------------------------------------------------------------ scene --

rb_status rb_scene_create(rb_scene_t* out) {
    RB_CHECK_ARG(out, "out is NULL");
    try {
        *out = new rb_scene_s();
    } catch (...) {
        return fail(RB_ERR_OOM, "scene allocation failed");
    }
    return RB_OK;
}

rb_status rb_scene_destroy(rb_scene_t s) {
    delete s;
    return RB_OK;
}

rb_status rb_scene_add_material(rb_scene_t s, const rb_material* m, uint32_t* id) {
    RB_CHECK_ARG(s && m, "NULL scene or material");
    rb_material c = *m;
    const rb_status st = prepare_material(s->s.mats, &c);
    if (st != RB_OK) return st;
    s->s.mats.push_back(c);
    if (id) *id = (uint32_t)(s->s.mats.size() - 1);
    return RB_OK;
}

rb_status rb_scene_add_layered_material(rb_scene_t s, const uint32_t* layer_ids, uint32_t n, uint32_t* id) {
    RB_CHECK_ARG(s && layer_ids, "NULL argument");
    RB_CHECK_ARG(n >= 1 && n <= 4, "a layered material has 1..4 layers");
    rb_material m{};
    m.kind = RB_MAT_LAYERED;
    m.param[0] = (float)n;
    for (uint32_t k = 0; k < n; k++) {
        RB_CHECK_ARG(layer_ids[k] < (1u << 24), "layer material id out of range");
        m.param[1 + k] = (float)layer_ids[k];
    }
    return rb_scene_add_material(s, &m, id);
}

rb_status rb_scene_add_mesh(rb_scene_t s, const float* verts, uint32_t nv, const uint32_t* idx, uint32_t nt,
                            uint32_t material_id) {
    RB_CHECK_ARG(s, "NULL scene");
    RB_CHECK_ARG(material_id < s->s.mats.size(), "material id out of range (add the material first)");
    return append_mesh(s->s, verts, nv, idx, nt, nullptr, 0, nullptr, nullptr, 0, nullptr, nullptr, material_id);
}

rb_status rb_scene_add_mesh_ex(rb_scene_t s, const float* verts, uint32_t nv, const uint32_t* idx, uint32_t nt,
                               const float* normals, uint32_t nn, const uint32_t* nidx, const float* texcoords,
                               uint32_t ntc, const uint32_t* tidx, uint32_t material_id) {
    RB_CHECK_ARG(s, "NULL scene");
    RB_CHECK_ARG(material_id < s->s.mats.size(), "material id out of range (add the material first)");
    return append_mesh(s->s, verts, nv, idx, nt, normals, nn, nidx, texcoords, ntc, tidx, nullptr, material_id);
}

rb_status rb_scene_counts(rb_scene_t s, uint32_t* nv, uint32_t* nt, uint32_t* nl, uint32_t* nm) {
    RB_CHECK_ARG(s, "NULL scene");
    if (nv) *nv = s->s.num_verts();
    if (nt) *nt = s->s.num_tris();
    if (nl) *nl = (uint32_t)s->s.light_tris.size();
    if (nm) *nm = (uint32_t)s->s.mats.size();
    return RB_OK;
}

rb_status rb_scene_geometry(rb_scene_t s, float* verts, uint32_t* idx, uint32_t* tri_mat, uint32_t* lights) {
    RB_CHECK_ARG(s, "NULL scene");
    const Scene& S = s->s;
    copy_vec(verts, S.verts);
    copy_vec(idx, S.idx);
    copy_vec(tri_mat, S.tri_mat);
    copy_vec(lights, S.light_tris);
    return RB_OK;
}

rb_status rb_scene_shading_counts(rb_scene_t s, uint32_t* nn, uint32_t* ntc, int* has_nidx, int* has_tidx) {
    RB_CHECK_ARG(s, "NULL scene");
    const Scene& S = s->s;
    if (nn) *nn = (uint32_t)(S.normals.size() / 3);
    if (ntc) *ntc = (uint32_t)(S.texcoords.size() / 2);
    if (has_nidx) *has_nidx = S.nidx.empty() ? 0 : 1;
    if (has_tidx) *has_tidx = S.tidx.empty() ? 0 : 1;
    return RB_OK;
}

rb_status rb_scene_shading(rb_scene_t s, float* normals, uint32_t* nidx, float* texcoords, uint32_t* tidx) {
    RB_CHECK_ARG(s, "NULL scene");
    const Scene& S = s->s;
    copy_vec(normals, S.normals);
    copy_vec(nidx, S.nidx);
    copy_vec(texcoords, S.texcoords);
    copy_vec(tidx, S.tidx);
    return RB_OK;
}

rb_status rb_scene_materials(rb_scene_t s, rb_material* out) {
    RB_CHECK_ARG(s && out, "NULL argument");
    copy_vec(out, s->s.mats);
    return RB_OK;
}

rb_status rb_scene_bounds(rb_scene_t s, float lo[3], float hi[3]) {
    RB_CHECK_ARG(s && lo && hi, "NULL argument");
    Box b;
    const Scene& S = s->s;
    for (uint32_t i = 0; i < S.num_verts(); i++) b.insert(f3(S.verts[3 * i], S.verts[3 * i + 1], S.verts[3 * i + 2]));
    lo[0] = b.lo.x; lo[1] = b.lo.y; lo[2] = b.lo.z;
    hi[0] = b.hi.x; hi[1] = b.hi.y; hi[2] = b.hi.z;
    return RB_OK;
}

}  // extern "C"

// --------------------------------------------------------------- generators --

namespace {

struct Mesh {
    std::vector<double> v;  // xyz
    std::vector<uint32_t> t;
};

// Icosahedron subdivided `s` times, vertices on the unit sphere; 20*4^s
// triangles wound counter-clockwise seen from outside.
const Mesh& icosphere(uint32_t s) {
    static std::mutex mu;
    static std::map<uint32_t, Mesh> cache;
    std::lock_guard<std::mutex> g(mu);
    auto it = cache.find(s);
    if (it != cache.end()) return it->second;
    Mesh m;
    const double p = (1.0 + std::sqrt(5.0)) / 2.0;
    const double base[12][3] = {{-1, p, 0}, {1, p, 0}, {-1, -p, 0}, {1, -p, 0}, {0, -1, p}, {0, 1, p},
                                {0, -1, -p}, {0, 1, -p}, {p, 0, -1}, {p, 0, 1}, {-p, 0, -1}, {-p, 0, 1}};
    const uint32_t faces[20][3] = {{0, 11, 5}, {0, 5, 1}, {0, 1, 7}, {0, 7, 10}, {0, 10, 11}, {1, 5, 9}, {5, 11, 4},
                                   {11, 10, 2}, {10, 7, 6}, {7, 1, 8}, {3, 9, 4}, {3, 4, 2}, {3, 2, 6}, {3, 6, 8},
                                   {3, 8, 9}, {4, 9, 5}, {2, 4, 11}, {6, 2, 10}, {8, 6, 7}, {9, 8, 1}};
    auto push = [&](double x, double y, double z) {
        double l = std::sqrt(x * x + y * y + z * z);
        m.v.push_back(x / l);
        m.v.push_back(y / l);
        m.v.push_back(z / l);
        return (uint32_t)(m.v.size() / 3 - 1);
    };
    for (auto& b : base) push(b[0], b[1], b[2]);
    for (auto& f : faces) m.t.insert(m.t.end(), f, f + 3);
    for (uint32_t l = 0; l < s; l++) {
        std::map<uint64_t, uint32_t> mid;
        auto midpoint = [&](uint32_t a, uint32_t b) {
            uint64_t key = a < b ? ((uint64_t)a << 32 | b) : ((uint64_t)b << 32 | a);
            auto f = mid.find(key);
            if (f != mid.end()) return f->second;
            uint32_t r = push((m.v[3 * a] + m.v[3 * b]) * 0.5, (m.v[3 * a + 1] + m.v[3 * b + 1]) * 0.5,
                              (m.v[3 * a + 2] + m.v[3 * b + 2]) * 0.5);
            mid[key] = r;
            return r;
        };
        std::vector<uint32_t> nt;
        for (size_t i = 0; i < m.t.size(); i += 3) {
            uint32_t a = m.t[i], b = m.t[i + 1], c = m.t[i + 2];
            uint32_t ab = midpoint(a, b), bc = midpoint(b, c), ca = midpoint(c, a);
            uint32_t add[12] = {a, ab, ca, b, bc, ab, c, ca, bc, ab, bc, ca};
            nt.insert(nt.end(), add, add + 12);
        }
        m.t.swap(nt);
    }
    return cache.emplace(s, std::move(m)).first->second;
}

struct Gen {
    rb::Rng rng;
    uint32_t ctr = 0;
    rb::U4 buf{};
    int left = 0;
    explicit Gen(uint64_t seed, uint32_t kind) { rng = rb::Rng::make(seed, rb::RNG_SCENE + 16 * kind); }
    double uni() {  // [0,1) with 24 bits
        if (left == 0) {
            buf = rng.draw(ctr++, 0, 0, 0);
            left = 4;
        }
        uint32_t v = left == 4 ? buf.x : left == 3 ? buf.y : left == 2 ? buf.z : buf.w;
        left--;
        return (double)(v >> 8) * (1.0 / 16777216.0);
    }
    double range(double a, double b) { return a + (b - a) * uni(); }
};

rb_status add_material_mix(rb_scene_t s, Gen& g, const rb_gen_params* p, uint32_t* mat) {
    if (p->material_id >= 0) {
        *mat = (uint32_t)p->material_id;
        return RB_OK;
    }
    rb_material m{};
    if (p->mix_materials && g.uni() >= 0.7) {
        m.kind = RB_MAT_MIRROR;
        m.color[0] = m.color[1] = m.color[2] = 0.9f;
    } else {
        m.kind = RB_MAT_LAMBERT;
        for (int i = 0; i < 3; i++) m.color[i] = (float)g.range(0.2, 0.8);
    }
    return rb_scene_add_material(s, &m, mat);
}

}  // namespace

extern "C" rb_status rb_scene_generate(rb_scene_t s, const rb_gen_params* p) {
    RB_CHECK_ARG(s && p, "NULL argument");
    Gen g(p->seed, p->kind);
    if (p->material_id >= (int32_t)s->s.mats.size()) return fail(RB_ERR_INVALID, "material id out of range");
    rb_status st;
    std::vector<float> v;
    std::vector<uint32_t> t;
    uint32_t mat = 0;
    switch (p->kind) {
        case RB_GEN_ICOSPHERES: {
            RB_CHECK_ARG(p->subdiv <= 7 && p->subdiv_hi <= 7, "subdivision too deep");
            // which spheres use the finer subdivision: a seeded choice of hi_count
            std::vector<uint8_t> hi(p->count, 0);
            if (p->subdiv_hi > p->subdiv && p->hi_count) {
                uint32_t k = 0;
                while (k < p->hi_count && k < p->count) {
                    uint32_t j = (uint32_t)(g.uni() * p->count);
                    if (j < p->count && !hi[j]) { hi[j] = 1; k++; }
                }
            }
            for (uint32_t i = 0; i < p->count; i++) {
                const Mesh& m = icosphere(hi[i] ? p->subdiv_hi : p->subdiv);
                double r = g.range(p->r0, p->r1);
                double c[3] = {g.range(p->c0, p->c1), g.range(p->c0, p->c1), g.range(p->c0, p->c1)};
                if ((st = add_material_mix(s, g, p, &mat)) != RB_OK) return st;
                v.resize(m.v.size());
                for (size_t k = 0; k < m.v.size(); k++) v[k] = (float)(c[k % 3] + r * m.v[k]);
                if (p->smooth) {
                    // the unit-sphere positions are the sphere's outward normals
                    std::vector<float> nrm(m.v.size());
                    for (size_t k = 0; k < m.v.size(); k++) nrm[k] = (float)m.v[k];
                    st = append_mesh(s->s, v.data(), (uint32_t)(v.size() / 3), m.t.data(), (uint32_t)(m.t.size() / 3),
                                     nrm.data(), (uint32_t)(nrm.size() / 3), nullptr, nullptr, 0, nullptr, nullptr, mat);
                } else {
                    st = rb_scene_add_mesh(s, v.data(), (uint32_t)(v.size() / 3), m.t.data(), (uint32_t)(m.t.size() / 3),
                                           mat);
                }
                if (st != RB_OK) return st;
            }
            return RB_OK;
        }
        case RB_GEN_THIN_QUADS: {
            for (uint32_t i = 0; i < p->count; i++) {
                double L = g.range(p->r0, p->r1), W = L / 50.0;
                double c[3] = {g.range(p->c0, p->c1), g.range(p->c0, p->c1), g.range(p->c0, p->c1)};
                // uniform rotation (Shoemake's unit quaternion)
                double u1 = g.uni(), u2 = g.uni(), u3 = g.uni();
                double a = std::sqrt(1 - u1), b = std::sqrt(u1);
                double qx = a * std::sin(2 * M_PI * u2), qy = a * std::cos(2 * M_PI * u2);
                double qz = b * std::sin(2 * M_PI * u3), qw = b * std::cos(2 * M_PI * u3);
                double R[3][3] = {{1 - 2 * (qy * qy + qz * qz), 2 * (qx * qy - qz * qw), 2 * (qx * qz + qy * qw)},
                                  {2 * (qx * qy + qz * qw), 1 - 2 * (qx * qx + qz * qz), 2 * (qy * qz - qx * qw)},
                                  {2 * (qx * qz - qy * qw), 2 * (qy * qz + qx * qw), 1 - 2 * (qx * qx + qy * qy)}};
                const double corner[4][2] = {{-0.5, -0.5}, {0.5, -0.5}, {0.5, 0.5}, {-0.5, 0.5}};
                v.resize(12);
                for (int k = 0; k < 4; k++) {
                    double lx = corner[k][0] * L, ly = corner[k][1] * W;
                    for (int d = 0; d < 3; d++) v[3 * k + d] = (float)(c[d] + R[d][0] * lx + R[d][1] * ly);
                }
                const uint32_t qt[6] = {0, 1, 2, 0, 2, 3};
                if ((st = add_material_mix(s, g, p, &mat)) != RB_OK) return st;
                if ((st = rb_scene_add_mesh(s, v.data(), 4, qt, 2, mat)) != RB_OK) return st;
            }
            return RB_OK;
        }
        case RB_GEN_ROOM: {
            // closed box [c0,c1]^3, normals pointing inward; floor and
            // ceiling 16x16 quads, walls 16x8 quads: 2048 triangles.
            const double lo = p->c0, hi = p->c1;
            struct Face { int axis; double at; int sgn; int nu, nv; };
            const Face faces[6] = {{1, lo, +1, 16, 16}, {1, hi, -1, 16, 16}, {0, lo, +1, 16, 8},
                                   {0, hi, -1, 16, 8},  {2, lo, +1, 16, 8},  {2, hi, -1, 16, 8}};
            if ((st = add_material_mix(s, g, p, &mat)) != RB_OK) return st;
            for (const Face& f : faces) {
                int ua = (f.axis + 1) % 3, va = (f.axis + 2) % 3;
                v.clear();
                t.clear();
                for (int j = 0; j <= f.nv; j++)
                    for (int i = 0; i <= f.nu; i++) {
                        double x[3];
                        x[f.axis] = f.at;
                        x[ua] = lo + (hi - lo) * i / f.nu;
                        x[va] = lo + (hi - lo) * j / f.nv;
                        for (int d = 0; d < 3; d++) v.push_back((float)x[d]);
                    }
                for (int j = 0; j < f.nv; j++)
                    for (int i = 0; i < f.nu; i++) {
                        uint32_t a = j * (f.nu + 1) + i, b = a + 1, c = a + f.nu + 1, d = c + 1;
                        // (ua x va) = +axis; flip winding for a negative inward normal
                        if (f.sgn > 0) {
                            uint32_t add[6] = {a, b, d, a, d, c};
                            t.insert(t.end(), add, add + 6);
                        } else {
                            uint32_t add[6] = {a, d, b, a, c, d};
                            t.insert(t.end(), add, add + 6);
                        }
                    }
                if ((st = rb_scene_add_mesh(s, v.data(), (uint32_t)(v.size() / 3), t.data(), (uint32_t)(t.size() / 3),
                                            mat)) != RB_OK)
                    return st;
            }
            return RB_OK;
        }
        case RB_GEN_EMITTER_QUADS: {
            rb_material m{};
            m.kind = RB_MAT_EMITTER;
            for (int i = 0; i < 3; i++) m.color[i] = p->emit[i];
            if ((st = rb_scene_add_material(s, &m, &mat)) != RB_OK) return st;
            const double side = std::sqrt((double)p->r0);
            for (uint32_t i = 0; i < p->count; i++) {
                double cx = g.range(p->c0 + side, p->c1 - side), cz = g.range(p->c0 + side, p->c1 - side);
                double y = p->c1 - 0.01;
                float q[12] = {(float)(cx - side / 2), (float)y, (float)(cz - side / 2),
                               (float)(cx + side / 2), (float)y, (float)(cz - side / 2),
                               (float)(cx + side / 2), (float)y, (float)(cz + side / 2),
                               (float)(cx - side / 2), (float)y, (float)(cz + side / 2)};
                // facing down (-y): (v1-v0) x (v2-v0) = (s,0,0) x (s,0,s) = -s^2 y
                const uint32_t qt[6] = {0, 1, 2, 0, 2, 3};
                if ((st = rb_scene_add_mesh(s, q, 4, qt, 2, mat)) != RB_OK) return st;
            }
            return RB_OK;
        }
    }
    return fail(RB_ERR_INVALID, "unknown generator kind %u", p->kind);
}

// -------------------------------------------------------------------- BVH --

extern "C" {

void rb_build_params_default(rb_build_params* p) {
    if (!p) return;
    p->min_leaf = 1;
    p->max_leaf = 0x7FFFFFF;
    p->split_alpha = 1.0e-5f;
    p->node_cost = 1.0f;
    p->tri_cost = 1.0f;
    p->threads = 0;
}

rb_status rb_bvh_build_arrays(const float* verts, uint32_t nv, const uint32_t* idx, uint32_t nt,
                              const rb_build_params* p, rb_host_bvh_t* out) {
    RB_CHECK_ARG(out, "out is NULL");
    RB_CHECK_ARG(nt > 0 && verts && idx, "empty geometry");
    rb_build_params d;
    rb_build_params_default(&d);
    if (!p) p = &d;
    RB_CHECK_ARG(p->max_leaf >= 1 && p->min_leaf <= p->max_leaf, "bad leaf preferences");
    rb_status st = RB_OK;
    HostBvh* h = nullptr;
    try {
        h = build_sbvh(verts, nv, idx, nt, *p, &st);
    } catch (...) {
        return fail(RB_ERR_OOM, "SBVH build failed (host allocation)");
    }
    if (!h) return st;
    *out = new rb_host_bvh_s{h};
    return RB_OK;
}

rb_status rb_bvh_build_host(rb_scene_t s, const rb_build_params* p, rb_host_bvh_t* out) {
    RB_CHECK_ARG(s, "NULL scene");
    return rb_bvh_build_arrays(s->s.verts.data(), s->s.num_verts(), s->s.idx.data(), s->s.num_tris(), p, out);
}

rb_status rb_bvh_host_stats(rb_host_bvh_t b, rb_bvh_stats* out) {
    RB_CHECK_ARG(b && out, "NULL argument");
    *out = b->b->stats;
    return RB_OK;
}

rb_status rb_bvh_host_release(rb_host_bvh_t b) {
    if (b) {
        delete b->b;
        delete b;
    }
    return RB_OK;
}

static rb_status copy_out(const void* src, size_t bytes, void* dst, size_t* cap) {
    if (!cap) return RB_OK;
    if (dst) {
        if (*cap < bytes) return fail(RB_ERR_INVALID, "output buffer too small (%zu < %zu bytes)", *cap, bytes);
        if (bytes) std::memcpy(dst, src, bytes);
    }
    *cap = bytes;
    return RB_OK;
}

rb_status rb_bvh_export(rb_host_bvh_t b, uint32_t layout, void* nodes, size_t* nb, void* tris, size_t* tb,
                        void* index, size_t* ib) {
    RB_CHECK_ARG(b, "NULL bvh");
    try {
        if (layout == RB_LAYOUT_REF_COMPACT2) {
            std::vector<uint32_t> n, w, x;
            export_compact2(*b->b, n, w, x);
            rb_status st;
            if ((st = copy_out(n.data(), n.size() * 4, nodes, nb)) != RB_OK) return st;
            if ((st = copy_out(w.data(), w.size() * 4, tris, tb)) != RB_OK) return st;
            return copy_out(x.data(), x.size() * 4, index, ib);
        }
        if (layout == RB_LAYOUT_MI355X || layout == RB_LAYOUT_MI355X_WIDE) {
            Mi355Layout L;
            rb_status st = export_mi355x(*b->b, L);
            if (st != RB_OK) return st;
            if (layout == RB_LAYOUT_MI355X_WIDE) {
                if (!wide_fits_stack(layout_depth(L.nodes)))
                    return fail(RB_ERR_UNSUPPORTED, "the tree is too deep for the wide walk's stack");
                std::vector<uint32_t> wide;
                if ((st = export_wide(L.nodes, wide)) != RB_OK) return st;
                L.nodes.swap(wide);
            }
            if ((st = copy_out(L.nodes.data(), L.nodes.size() * 4, nodes, nb)) != RB_OK) return st;
            if ((st = copy_out(L.tris.data(), L.tris.size() * 4, tris, tb)) != RB_OK) return st;
            return copy_out(L.index.data(), L.index.size() * 4, index, ib);
        }
    } catch (...) {
        return fail(RB_ERR_OOM, "layout export failed (host allocation)");
    }
    return fail(RB_ERR_INVALID, "unknown layout %u", layout);
}

rb_status rb_bvh_save_binary(rb_host_bvh_t b, const char* path) {
    RB_CHECK_ARG(b && path, "NULL argument");
    try {
        return save_bvh_binary(*b->b, path);
    } catch (...) {
        return fail(RB_ERR_OOM, "BVH save failed (host allocation)");
    }
}

rb_status rb_bvh_load_binary(const char* path, rb_scene_t s, const rb_build_params* p, rb_host_bvh_t* out) {
    RB_CHECK_ARG(path && s && out, "NULL argument");
    rb_build_params d;
    rb_build_params_default(&d);
    HostBvh* h = nullptr;
    rb_status st;
    try {
        st = load_bvh_binary(path, s->s.verts.data(), s->s.num_verts(), s->s.idx.data(), s->s.num_tris(), p ? *p : d, &h);
    } catch (...) {
        return fail(RB_ERR_OOM, "BVH load failed (host allocation)");
    }
    if (st != RB_OK) return st;
    *out = new rb_host_bvh_s{h};
    return RB_OK;
}

rb_status rb_bvh_build_cached(rb_scene_t s, const rb_build_params* p, const char* cache_dir, rb_host_bvh_t* out,
                              int* from_cache) {
    RB_CHECK_ARG(s && cache_dir && out, "NULL argument");
    RB_CHECK_ARG(s->s.num_tris() > 0, "empty geometry");
    rb_build_params d;
    rb_build_params_default(&d);
    if (!p) p = &d;
    char name[64];
    std::snprintf(name, sizeof name, "/rbvh_%016llx.bin",
                  (unsigned long long)bvh_cache_key(s->s.verts.data(), s->s.num_verts(), s->s.idx.data(),
                                                     s->s.num_tris(), *p));
    const std::string path = std::string(cache_dir) + name;
    if (from_cache) *from_cache = 0;
    {
        std::ifstream probe(path, std::ios::binary);
        if (probe && rb_bvh_load_binary(path.c_str(), s, p, out) == RB_OK) {
            if (from_cache) *from_cache = 1;
            return RB_OK;
        }
    }
    rb_status st = rb_bvh_build_host(s, p, out);
    if (st != RB_OK) return st;
    if ((st = rb_bvh_save_binary(*out, path.c_str())) != RB_OK) {
        rb_bvh_host_release(*out);
        *out = nullptr;
        return st;
    }
    return RB_OK;
}

// CudaBvh::serialize (cudabvh.cpp:176-214): magic, three sizes, layout
// (1 = BVHLayout_Compact2), decimal uint4 rows, decimal indices, EOF marker.
rb_status rb_bvh_save_text(rb_host_bvh_t b, const char* path) {
    RB_CHECK_ARG(b && path, "NULL argument");
    std::vector<uint32_t> n, w, x;
    export_compact2(*b->b, n, w, x);
    std::ofstream out(path, std::ios::binary);
    if (!out) return fail(RB_ERR_INVALID, "cannot open %s", path);
    out << 0x76543210u << "\n" << n.size() / 4 << "\n" << w.size() / 4 << "\n" << x.size() << "\n" << 1u << "\n";
    for (size_t i = 0; i < n.size(); i += 4) out << n[i] << " " << n[i + 1] << " " << n[i + 2] << " " << n[i + 3] << "\n";
    for (size_t i = 0; i < w.size(); i += 4) out << w[i] << " " << w[i + 1] << " " << w[i + 2] << " " << w[i + 3] << "\n";
    for (size_t i = 0; i < x.size(); i++) out << x[i] << "\n";
    out << 0x1234567u << "\n";
    return out ? RB_OK : fail(RB_ERR_INVALID, "write to %s failed", path);
}

// CudaBvh(filename, ok) (cudabvh.cpp:108-165), with the layout check of
// Application::prepare (src/core/application.cpp:154-161).
rb_status rb_bvh_load_text_compact2(const char* path, void* nodes, size_t* nb, void* tris, size_t* tb, void* index,
                                    size_t* ib) {
    RB_CHECK_ARG(path, "NULL path");
    std::ifstream in(path);
    if (!in) return fail(RB_ERR_INVALID, "cannot open %s", path);
    uint32_t magic = 0, layout = 0, eof = 0;
    size_t nn = 0, nw = 0, ni = 0;
    in >> magic;
    if (magic != 0x76543210u) return fail(RB_ERR_INVALID, "%s: not a BVH file (magic %u)", path, magic);
    in >> nn >> nw >> ni >> layout;
    if (!in || layout != 1) return fail(RB_ERR_INVALID, "%s: layout %u is not Compact2", path, layout);
    if (nn > (1ull << 32) || nw > (1ull << 32) || ni > (1ull << 32)) return fail(RB_ERR_INVALID, "%s: sizes", path);
    std::vector<uint32_t> n(nn * 4), w(nw * 4), x(ni);
    for (auto& e : n) in >> e;
    for (auto& e : w) in >> e;
    for (auto& e : x) in >> e;
    in >> eof;
    if (!in || eof != 0x1234567u) return fail(RB_ERR_INVALID, "%s: truncated or missing EOF marker", path);
    rb_status st;
    if ((st = copy_out(n.data(), n.size() * 4, nodes, nb)) != RB_OK) return st;
    if ((st = copy_out(w.data(), w.size() * 4, tris, tb)) != RB_OK) return st;
    return copy_out(x.data(), x.size() * 4, index, ib);
}

}  // extern "C"

// rb_render.hip -- AO and path-tracing passes around the persistent tracer.
//
// Wavefront forms of the reference megakernels:
//  * AO: kernel_ambient_occlusion (src/integrators/integrator_ao.cu:42-85)
//    becomes spawn (n rays per primary hit) -> one any-hit trace launch ->
//    resolve;
//  * PT: kernel_pathtracer_single_monolithic (src/integrators/integrator_path.cu:43-261)
//    becomes, per bounce: shade (emission, light sample, MIS BSDF sample,
//    next-bounce sample, roulette) -> any-hit trace of the shadow rays and one
//    closest-hit trace of the MIS + bounce rays -> connect (direct light,
//    next hit) -> compaction of the live paths into the next queue.  Queue
//    lengths live in device memory, so the bounce loop needs no host
//    synchronisation (the reference counts live paths with thrust on the
//    host every bounce, integrator_bdpt_lvc_mk.cu:529-544).  Queue appends
//    are reserved per workgroup (block_reserve): one global atomic per 512
//    threads;
//  * BDPT: the light pass and the camera pass with cache connections
//    (rb_bdpt.h);
//  * ray-cast debug shading: camera rays -> closest hit -> colour;
//  * film: the reference's filtered splat restated as a gather (box,
//    triangle or Gaussian; kernel_update_film,
//    src/renderers/samplerrenderer.cu:103-169), in a fixed order and without
//    float atomics, so images are bit-reproducible.
#include "rb_shade.h"

namespace rb {

struct TileArgs {
    uint32_t width, x0, y0, w, spp;
};

// Global sample id of primary ray i of a tile: (image pixel) * spp + sample.
__device__ __forceinline__ uint32_t sample_gid(const TileArgs& T, uint32_t i) {
    const uint32_t p = i / T.spp, s = i % T.spp;
    return ((T.y0 + p / T.w) * T.width + (T.x0 + p % T.w)) * T.spp + s;
}

// ---------------------------------------------------------------- AO ----

__global__ void k_ao_spawn(const float4* __restrict__ prim, const int2* __restrict__ hits, uint32_t n, TileArgs T,
                           SceneDev S, uint32_t nao, float radius, RngArg rng_a, float4* __restrict__ rays,
                           uint8_t* __restrict__ mask, unsigned long long* ray_counter) {
    const Rng rng = rng_a.get();
    const uint64_t total = (uint64_t)n * nao;
    for (uint64_t j = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; j < total; j += (uint64_t)gridDim.x * blockDim.x) {
        const uint32_t i = (uint32_t)(j / nao), k = (uint32_t)(j % nao);
        const int2 h = hits[i];
        if (ray_counter) {  // AO rays traced: nao per primary hit, one atomic per wave
            const uint64_t m = __ballot(k == 0 && h.x >= 0);
            if (m && __lane_id() == (uint32_t)__builtin_ctzll(m))
                atomicAdd(ray_counter, (unsigned long long)__popcll(m) * nao);
        }
        if (h.x < 0) {
            mask[j] = 0;
            rays[2 * j] = make_float4(0.f, 0.f, 0.f, 0.f);
            rays[2 * j + 1] = make_float4(0.f, 0.f, 1.f, 0.f);
            continue;
        }
        const float4 o = prim[2 * (size_t)i], d = prim[2 * (size_t)i + 1];
        const float t = __int_as_float(h.y);
        const V3 dv = v3(d.x, d.y, d.z);
        const V3 p = vadd(v3(o.x, o.y, o.z), vmul(dv, t));
        V3 v0, v1, v2;
        uint32_t tf;
        tri_fetch(S, (uint32_t)h.x, &v0, &v1, &v2, nullptr, &tf);
        // about the shading normal (integrator_ao.cu:55), turned to face the ray
        const V3 ng = tri_normal(v0, v1, v2);
        V3 n;
        shade_normal(S, (uint32_t)h.x, tf, p, v0, v1, v2, ng, &n);
        if (vdot(ng, dv) > 0.0f) n = vneg(n);
        const U4 r = rng.draw(sample_gid(T, i), k, 0, 0);
        const V3 dir = cosine_hemisphere(n, u01(r.x), u01(r.y));
        rays[2 * j] = make_float4(p.x, p.y, p.z, 1.0e-3f * t);
        rays[2 * j + 1] = make_float4(dir.x, dir.y, dir.z, radius);
        mask[j] = 1;
    }
}

// Per sample: unoccluded AO rays / nao (misses give 0), for the film.
__global__ void k_ao_resolve(const int2* __restrict__ prim_hits, const int2* __restrict__ ao_hits, uint32_t n,
                             uint32_t nao, float4* __restrict__ L) {
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        float c = 0.0f;
        if (prim_hits[i].x >= 0) {
            uint32_t open = 0;
            for (uint32_t k = 0; k < nao; k++) open += ao_hits[(size_t)i * nao + k].x < 0 ? 1u : 0u;
            c = (float)open / (float)nao;
        }
        L[i] = make_float4(c, c, c, 0.f);
    }
}

// ---------------------------------------------------------------- PT ----

struct PathBufs {
    float4* T;       // throughput at this bounce (w unused; bounce 0's is 1 and not stored)
    float4* Tn;      // throughput after this bounce's sample: the next bounce's T (the two swap per bounce)
    const float4* Tprev;  // the previous bounce's T, for its contributions (the buffer Tn reuses)
    float4* L;       // radiance
    // The current vertex of a path: ray 2 * cur[pid] of pray and hit
    // cur[pid] of phit -- the camera rays at bounce 0, then the previous
    // bounce's closest-hit queue, read in place (the MIS + bounce queues
    // alternate between two buffers, so this bounce's rays do not overwrite
    // them)
    const float4* pray;
    const int2* phit;
    uint32_t* cur;
    float4* ca;      // light-sample contribution if unoccluded; w = direct-light technique weight
    float4* cb;      // MIS contribution if the MIS ray hits light `lprim`
    int4* slot;      // x: shadow slot, y: MIS slot, z: bounce slot, w: light triangle
    uint32_t* info;  // bounce | specular << 8 | next specular << 9 | emitter-done << 10
    uint32_t* cmask; // BDPT: specular camera vertices (bit b = vertex of bounce b)
    float4* cc;      // BDPT: connection contributions, kMaxConn per path
    int* cslot;      // BDPT: connection shadow-ray slots, kMaxConn per path
    uint32_t* q[2];  // path queues
    uint32_t* cnt;   // [0],[1] queue lengths, [2] shadow rays, [3] closest rays
    float4* ra;      // shadow rays
    int2* ha;
    float4* rbc;     // MIS + bounce rays
    int2* hbc;
};

constexpr uint32_t kMaxConn = 8;

}  // namespace rb

#include "rb_bdpt.h"

namespace rb {

struct BdArgs {
    LvcBufs C;
    uint32_t nconn;
    uint32_t tdepth;  // cache depth the technique count may use: 0 when no connections are made
    RngArg crng;
};

__global__ void k_reset(uint32_t* c, uint32_t a, uint32_t b, uint32_t d) {
    if (threadIdx.x == 0 && blockIdx.x == 0) {
        c[a] = 0;
        c[b] = 0;
        c[d] = 0;
    }
}

__global__ void k_set_u32(uint32_t* p, uint32_t v) {
    if (threadIdx.x == 0 && blockIdx.x == 0) *p = v;
}

__global__ void k_add64(unsigned long long* total, unsigned long long v) {
    if (threadIdx.x == 0 && blockIdx.x == 0) atomicAdd(total, v);
}

__global__ void k_add_count(const uint32_t* c, unsigned long long* total) {
    if (threadIdx.x == 0 && blockIdx.x == 0) atomicAdd(total, (unsigned long long)*c);
}

// Adds the rays of this bounce (shadow + closest queues) to the caller's counter.
__global__ void k_count_rays(const uint32_t* c, unsigned long long* total, unsigned long long extra) {
    if (threadIdx.x == 0 && blockIdx.x == 0 && total)
        atomicAdd(total, (unsigned long long)c[2] + (unsigned long long)c[3] + extra);
}

// Rays asked for by one shaded camera vertex: light sample (A, shadow
// queue), MIS BSDF sample (B) and next bounce (C, closest queue).
struct ShadeRays {
    bool wantA, wantB, wantC;
    float4 ra0, ra1, rb1, rc1;
    V3 p;
    float eps;
    int lt;  // light triangle of the samples
};

// Shading of camera path pid at its vertex (ray 2 * c of pray, hit c of
// phit; info: bounce | specular << 8; throughput Tp; radiance L so far).
// An emitter adds its emission when seen directly or through specular
// bounces (integrator_path.cu:94-109) and ends the path: writes L, returns
// false.  Otherwise light sample + MIS BSDF sample + next-bounce sample with
// roulette (integrator_path.cu:111-249); writes L, Tn, ca, cb, info (and
// cmask with BD) and fills R; returns true.  BD: with the technique weights;
// the light-vertex-cache connections follow in k_lvc_connect.
template <bool BD>
__device__ __forceinline__ bool shade_vertex(const PathBufs& P, uint32_t pid, uint32_t c, uint32_t info, float4 Tp,
                                             float4 L, const TileArgs& T, const SceneDev& S, const Rng& rng,
                                             uint32_t max_depth, uint32_t rr_start, float rr_prob, const BdArgs& B,
                                             ShadeRays& R) {
    const float4 o = P.pray[2 * (size_t)c], d = P.pray[2 * (size_t)c + 1];
    const int2 h = P.phit[c];
    const uint32_t bounce = info & 255u;
    const bool spec = (info >> 8) & 1u;
    const float t = __int_as_float(h.y);
    const V3 dv = v3(d.x, d.y, d.z);
    const V3 p = vadd(v3(o.x, o.y, o.z), vmul(dv, t));
    const float eps = 1.0e-3f * t;
    V3 v0, v1, v2;
    uint32_t mid, tf;
    tri_fetch(S, (uint32_t)h.x, &v0, &v1, &v2, &mid, &tf);
    const rb_material m = S.mats[mid];
    const V3 ng = tri_normal(v0, v1, v2);
    V3 ns;  // shading normal on ng's side (ng itself for a flat triangle)
    shade_normal(S, (uint32_t)h.x, tf, p, v0, v1, v2, ng, &ns);
    const V3 wo = vneg(dv);
    const V3 nf = vdot(ng, wo) >= 0.0f ? ns : vneg(ns);
    const V3 Tv = v3(Tp.x, Tp.y, Tp.z);
    uint32_t cm = BD ? (bounce ? P.cmask[pid] : 0u) : 0u;
    if (m.kind == RB_MAT_EMITTER) {
        if (bounce == 0 || spec) {
            V3 le = vdot(ng, wo) > 0.0f ? v3(m.color[0], m.color[1], m.color[2]) : v3(0.f, 0.f, 0.f);
            if (BD) le = vmul(le, technique_weight(cm, bounce + 1, max_depth, B.tdepth));
            L = make_float4(L.x + Tv.x * le.x, L.y + Tv.y * le.y, L.z + Tv.z * le.z, 0.f);
        }
        P.L[pid] = L;
        return false;
    }
    P.L[pid] = L;
    R.p = p;
    R.eps = eps;
    const uint32_t gid = sample_gid(T, pid);
    const U4 r0 = rng.draw(gid, bounce, 0, 0);
    const U4 r1 = rng.draw(gid, bounce, 1, 0);
    // layer picks of a layered material: next bounce, MIS sample
    const bool layered = m.kind == RB_MAT_LAYERED;
    const U4 r2 = layered ? rng.draw(gid, bounce, 2, 0) : U4{0u, 0u, 0u, 0u};
    V3 ca = v3(0.f, 0.f, 0.f), cb = v3(0.f, 0.f, 0.f);
    const bool delta = mat_delta(S, m);
    const float wdir = BD ? technique_weight(cm, bounce + 2, max_depth, B.tdepth) : 1.0f;
    if (S.n_lights > 0) {
        uint32_t lid = (uint32_t)(u01(r0.x) * (float)S.n_lights);
        if (lid > S.n_lights - 1) lid = S.n_lights - 1;
        const uint32_t lt = S.light_tris[lid];
        V3 l0, l1, l2;
        uint32_t lmid;
        tri_fetch(S, lt, &l0, &l1, &l2, &lmid);
        const rb_material lm = S.mats[lmid];
        const V3 emit = v3(lm.color[0], lm.color[1], lm.color[2]);
        const V3 ln = tri_normal(l0, l1, l2);
        const float area = tri_area(l0, l1, l2);
        R.lt = (int)lt;
        // light sample (DiffuseAreaLight::sample_L, src/core/lights.cu:56-65)
        const V3 pl = sample_triangle(l0, l1, l2, u01(r0.y), u01(r0.z));
        const V3 wi = vnorm(vsub(pl, p));
        const float lpdf = light_pdf(p, wi, l0, l1, l2, ln, area);
        const V3 Li = vdot(ln, vneg(wi)) > 0.0f ? emit : v3(0.f, 0.f, 0.f);
        if (!delta && !is_black(Li) && lpdf > 0.0f) {
            const V3 f = mat_f(S, m, wo, wi, ng, nf);
            if (!is_black(f)) {
                const float bpdf = mat_pdf_ns(S, m, wo, wi, nf);
                const float w = lpdf / (lpdf + bpdf);
                const float s = fabsf(vdot(wi, ns)) * w;
                ca = v3(f.x * Li.x * s / lpdf, f.y * Li.y * s / lpdf, f.z * Li.z * s / lpdf);
                const V3 seg = vsub(pl, p);
                const float dist = sqrtf(vdot(seg, seg));
                const V3 sd = vnorm(seg);
                R.ra0 = make_float4(p.x, p.y, p.z, eps);
                R.ra1 = make_float4(sd.x, sd.y, sd.z, dist * (1.0f - 1.0e-3f));
                R.wantA = true;
            }
        }
        // BSDF sample with MIS (non-specular components only)
        if (!delta) {
            V3 wi2, f2;
            float pdf2;
            mat_sample_ns(S, m, wo, ng, nf, u01(r2.y), u01(r0.w), u01(r1.x), &wi2, &f2, &pdf2);
            if (!is_black(f2) && pdf2 > 0.0f) {
                const float lpdf2 = light_pdf(p, wi2, l0, l1, l2, ln, area);
                if (lpdf2 > 0.0f) {
                    const float w2 = pdf2 / (pdf2 + lpdf2);
                    const V3 Li2 = vdot(ln, vneg(wi2)) > 0.0f ? emit : v3(0.f, 0.f, 0.f);
                    const float s2 = fabsf(vdot(wi2, ns)) * w2;
                    cb = v3(f2.x * Li2.x * s2 / pdf2, f2.y * Li2.y * s2 / pdf2, f2.z * Li2.z * s2 / pdf2);
                    R.rb1 = make_float4(wi2.x, wi2.y, wi2.z, 3.402823466e+38f);
                    R.wantB = true;
                }
            }
        }
    }
    // next bounce (material_sample_f with every component)
    V3 wi3, f3;
    float pdf3;
    bool spec3;
    bool weight3;  // f3 already holds f |cos| / pdf
    mat_sample(S, m, wo, ng, ns, nf, u01(r2.x), u01(r1.y), u01(r1.z), true, &wi3, &f3, &pdf3, &spec3, &weight3);
    // the vertex counts as specular for the technique weights of the paths
    // that continue through a specular lobe
    if (BD && spec3) cm |= 1u << bounce;
    uint32_t ninfo = info & 0x1FFu;
    if (!is_black(f3) && pdf3 > 0.0f) {
        const float s3 = weight3 ? 1.0f : fabsf(vdot(wi3, ns));
        V3 Tn = v3(Tv.x * (f3.x * s3 / pdf3), Tv.y * (f3.y * s3 / pdf3), Tv.z * (f3.z * s3 / pdf3));
        bool alive = true;
        if (bounce > rr_start) {
            const float lum = luminance(Tn);
            const float qq = rr_prob < lum ? rr_prob : lum;
            if (u01(r1.w) > qq) alive = false;
            else Tn = v3(Tn.x / qq, Tn.y / qq, Tn.z / qq);
        }
        if (alive && bounce != max_depth) {
            P.Tn[pid] = make_float4(Tn.x, Tn.y, Tn.z, 0.f);
            R.rc1 = make_float4(wi3.x, wi3.y, wi3.z, 3.402823466e+38f);
            R.wantC = true;
            if (spec3) ninfo |= 1u << 9;
        }
    }
    if (BD) P.cmask[pid] = cm;
    P.ca[pid] = make_float4(ca.x, ca.y, ca.z, wdir);
    P.cb[pid] = make_float4(cb.x, cb.y, cb.z, 0.f);
    P.info[pid] = ninfo;
    return true;
}

// The contributions of path pid's last shaded vertex once its rays are
// traced: light sample if unoccluded, MIS sample if it hit the sampled
// light (times the light count and, with BD, the direct-light technique
// weight), then BD's cache connections whose shadow rays are unoccluded
// (integrator_path.cu:136-200).  Returns the new radiance in *L; true, with
// the next vertex (bounce ray cur, info), when the bounce ray hit.
template <bool BD>
__device__ __forceinline__ bool connect_vertex(const PathBufs& P, uint32_t pid, float nlights, uint32_t nconn,
                                               float4* L, uint32_t* cur, uint32_t* ninfo) {
    const uint32_t info = P.info[pid];
    const int4 s = P.slot[pid];
    V3 Ld = v3(0.f, 0.f, 0.f);
    const float4 ca = P.ca[pid];
    if (s.x >= 0 && P.ha[s.x].x < 0) Ld = vadd(Ld, v3(ca.x, ca.y, ca.z));
    if (s.y >= 0 && P.phit[s.y].x == s.w) {
        const float4 c = P.cb[pid];
        Ld = vadd(Ld, v3(c.x, c.y, c.z));
    }
    Ld = vmul(Ld, nlights);
    if (BD) Ld = vmul(Ld, ca.w);
    // the vertex's throughput: 1 at bounce 0 (not stored)
    const float4 Tp = (info & 255u) ? P.Tprev[pid] : make_float4(1.f, 1.f, 1.f, 0.f);
    float4 l = P.L[pid];
    l = make_float4(l.x + Tp.x * Ld.x, l.y + Tp.y * Ld.y, l.z + Tp.z * Ld.z, 0.f);
    if (BD) {
        for (uint32_t i = 0; i < nconn; i++) {
            const int cs = P.cslot[(size_t)pid * kMaxConn + i];
            if (cs >= 0 && P.ha[cs].x < 0) {
                const float4 c = P.cc[(size_t)pid * kMaxConn + i];
                l = make_float4(l.x + c.x, l.y + c.y, l.z + c.z, 0.f);
            }
        }
    }
    *L = l;
    if (s.z >= 0) {
        const int2 h = P.phit[s.z];
        if (h.x >= 0) {
            *cur = (uint32_t)s.z;  // the next vertex: bounce ray s.z, read in place
            *ninfo = ((info & 255u) + 1u) | (((info >> 9) & 1u) << 8);
            return true;
        }
    }
    return false;
}

// One step of the camera pass over a queue of paths, with no host
// synchronisation between steps.
//  MODE 0: the n0 camera rays' hits start the paths (L = 0, T = 1) and
//          bounce 0 is shaded;
//  MODE 1: each path of queue qi (shaded at the previous bounce, its rays
//          traced) takes that bounce's contributions (connect_vertex) and,
//          when its bounce ray hit, is shaded at this bounce by the same
//          thread (shade_vertex), so its state is read once per bounce.  (A
//          per-workgroup compaction of the continuing paths before shading
//          measured slower: its barrier stalls the workgroup on the slowest
//          connect; the paths that end leave idle lanes only at the bounces
//          roulette thins, where the queue is short);
//  MODE 2: contributions only (after the last bounce).
// Shaded non-emitter vertices are appended to queue qi ^ 1 and their rays to
// the shadow (cnt[2]) and closest (cnt[3]) queues.
template <bool BD, int MODE>
__global__ __launch_bounds__(kQueueBlock) __attribute__((amdgpu_waves_per_eu(4))) void k_pt_step(
    PathBufs P, uint32_t qi, uint32_t n0, TileArgs T, SceneDev S, RngArg rng_a, uint32_t max_depth,
    uint32_t rr_start, float rr_prob, float nlights, BdArgs B) {
    const Rng rng = rng_a.get();
    const uint32_t count = MODE == 0 ? n0 : P.cnt[qi];
    const uint32_t stride = gridDim.x * blockDim.x;
    const uint32_t end = uniform_end(count);  // whole workgroups iterate together
    for (uint32_t q = blockIdx.x * blockDim.x + threadIdx.x; q < end; q += stride) {
        const bool valid = q < count;
        bool go = false;
        uint32_t pid = 0, cur = 0, info = 0;
        float4 L = make_float4(0.f, 0.f, 0.f, 0.f);
        if (valid) {
            if (MODE == 0) {
                pid = cur = q;
                go = P.phit[q].x >= 0;
                if (!go) P.L[q] = L;
            } else {
                pid = P.q[qi][q];
                go = connect_vertex<BD>(P, pid, nlights, B.nconn, &L, &cur, &info);
                if (MODE == 2 || !go) P.L[pid] = L;
            }
        }
        if (MODE == 2) continue;
        ShadeRays R;
        R.wantA = R.wantB = R.wantC = false;
        R.lt = -1;
        bool shaded = false;
        if (go) {
            const float4 Tp = (info & 255u) ? P.T[pid] : make_float4(1.f, 1.f, 1.f, 0.f);
            shaded = shade_vertex<BD>(P, pid, cur, info, Tp, L, T, S, rng, max_depth, rr_start, rr_prob, B, R);
            if (BD && shaded) P.cur[pid] = cur;  // k_lvc_connect reads the vertex
        }
        // workgroup-aggregated appends (every thread of the workgroup reaches
        // these): shadow queue = light sample; closest queue = MIS ray, then
        // the bounce ray; the shaded vertices to the next path queue.  BDPT
        // connection rays are appended by k_lvc_connect.
        const uint32_t sa = block_reserve(&P.cnt[2], R.wantA ? 1u : 0u);
        const uint32_t sb = block_reserve(&P.cnt[3], (R.wantB ? 1u : 0u) + (R.wantC ? 1u : 0u));
        const uint32_t sq = block_reserve(&P.cnt[qi ^ 1], shaded ? 1u : 0u);
        const uint32_t sc = sb + (R.wantB ? 1u : 0u);
        if (R.wantA) {
            P.ra[2 * (size_t)sa] = R.ra0;
            P.ra[2 * (size_t)sa + 1] = R.ra1;
        }
        if (R.wantB) {
            P.rbc[2 * (size_t)sb] = make_float4(R.p.x, R.p.y, R.p.z, R.eps);
            P.rbc[2 * (size_t)sb + 1] = R.rb1;
        }
        if (R.wantC) {
            P.rbc[2 * (size_t)sc] = make_float4(R.p.x, R.p.y, R.p.z, R.eps);
            P.rbc[2 * (size_t)sc + 1] = R.rc1;
        }
        if (shaded) {
            P.slot[pid] = make_int4(R.wantA ? (int)sa : -1, R.wantB ? (int)sb : -1, R.wantC ? (int)sc : -1, R.lt);
            P.q[qi ^ 1][sq] = pid;
        }
    }
}

// BDPT connections of the queued camera vertices to the light-vertex cache
// (integrator_bdpt_lvc_sk.cu:343-386), one thread per (path, connection):
// thread t serves queue entry t / nconn, connection t % nconn.  Runs after
// the k_pt_step that shaded the vertices (queue qi), which leaves the camera
// vertex (cur, hit, throughput, earlier specular mask, bounce) in place.
// Writes the contribution and the shadow-ray slot of every connection (-1:
// no ray) and appends the shadow rays to the shadow queue.  The nconn
// threads of a vertex read it from the same cache lines; making the
// connections in the shading thread instead (its state already in
// registers) measured slower, since one thread then waits on nconn cache
// records in turn and the kernel needs more registers than four waves per
// SIMD allow.
__global__ __launch_bounds__(kQueueBlock) void k_lvc_connect(PathBufs P, uint32_t qi, TileArgs T, SceneDev S,
                                                             uint32_t max_depth, BdArgs B) {
    const Rng crng = B.crng.get();
    const uint32_t nconn = B.nconn;
    const uint32_t total = P.cnt[qi] * nconn;
    const uint32_t stride = gridDim.x * blockDim.x;
    const uint32_t end = uniform_end(total);  // whole workgroups iterate together (block_reserve)
    for (uint32_t t = blockIdx.x * blockDim.x + threadIdx.x; t < end; t += stride) {
        const bool valid = t < total;
        const uint32_t q = valid ? t / nconn : 0u;
        const uint32_t i = valid ? t - q * nconn : 0u;
        const uint32_t pid = valid ? P.q[qi][q] : 0u;
        bool want = false;
        float4 r0, r1;
        if (valid) {
            const uint32_t c = P.cur[pid];
            const float4 o = P.pray[2 * (size_t)c], d = P.pray[2 * (size_t)c + 1];
            const int2 h = P.phit[c];
            const uint32_t bounce = P.info[pid] & 255u;
            const float th = __int_as_float(h.y);
            const V3 dv = v3(d.x, d.y, d.z);
            const V3 p = vadd(v3(o.x, o.y, o.z), vmul(dv, th));
            V3 v0, v1, v2;
            uint32_t mid, tf;
            tri_fetch(S, (uint32_t)h.x, &v0, &v1, &v2, &mid, &tf);
            const rb_material m = S.mats[mid];
            if (m.kind != RB_MAT_EMITTER && !mat_delta(S, m)) {
                const V3 ng = tri_normal(v0, v1, v2);
                V3 ns;
                shade_normal(S, (uint32_t)h.x, tf, p, v0, v1, v2, ng, &ns);
                const V3 wo = vneg(dv);
                const V3 nf = vdot(ng, wo) >= 0.0f ? ns : vneg(ns);
                const float4 Tp = bounce ? P.T[pid] : make_float4(1.f, 1.f, 1.f, 0.f);  // bounce 0's is not stored
                const V3 Tv = v3(Tp.x, Tp.y, Tp.z);
                // earlier vertices only: k_pt_shade has already marked this
                // vertex when its next direction took a specular lobe
                const uint32_t cm = P.cmask[pid] & ((1u << bounce) - 1u);
                const uint32_t gid = sample_gid(T, pid);
                const U4 rc = crng.draw(gid, bounce, 2 + i / 2, 0);
                const uint32_t ua = (i & 1) ? rc.z : rc.x, ub = (i & 1) ? rc.w : rc.y;
                float4 contrib = make_float4(0.f, 0.f, 0.f, 0.f);
                uint32_t j = (uint32_t)(u01(ua) * (float)B.C.paths);
                if (j > B.C.paths - 1) j = B.C.paths - 1;
                const uint32_t nj = B.C.count[j];
                if (nj) {
                    uint32_t k = (uint32_t)(u01(ub) * (float)nj);
                    if (k > nj - 1) k = nj - 1;
                    const size_t sl = (size_t)j * B.C.depth + k;
                    const float4* rec = B.C.v + B.C.rows * sl;
                    const float4 xp = rec[0], xn = rec[1], xw = rec[2], xa = rec[3];
                    const rb_material mx = S.mats[__float_as_uint(xn.w)];
                    const V3 seg = vsub(v3(xp.x, xp.y, xp.z), p);
                    const float d2 = vdot(seg, seg);
                    const float dist = sqrtf(d2);
                    const V3 dir = vnorm(seg);
                    const V3 ngx = v3(xn.x, xn.y, xn.z);
                    const V3 fy = mat_f(S, m, wo, dir, ng, nf);
                    const V3 xwv = v3(xw.x, xw.y, xw.z);
                    // the cache vertex's shading normal (row 4, when the scene
                    // has normals); importance transport there takes the
                    // adjoint factor
                    const bool xsn = B.C.rows > 4 && __float_as_uint(xp.w) != 0u;
                    V3 xs = ngx;
                    if (xsn) {
                        const float4 r4 = rec[4];
                        xs = v3(r4.x, r4.y, r4.z);
                    }
                    const V3 nfx = vdot(ngx, xwv) >= 0.0f ? xs : vneg(xs);
                    V3 fx = mat_f(S, mx, xwv, vneg(dir), ngx, nfx);
                    if (xsn) fx = vmul(fx, adjoint_factor(xwv, vneg(dir), ngx, xs));
                    if (!is_black(fy) && !is_black(fx)) {
                        const float G = fabsf(vdot(ns, dir)) * fabsf(vdot(xs, vneg(dir))) / d2;
                        // full path: camera vertices 0..bounce, then light vertices k+1..1
                        const uint32_t lmask = __float_as_uint(xw.w);
                        uint32_t M = cm;
                        for (uint32_t e = 0; e <= k; e++)
                            if ((lmask >> (k - e)) & 1u) M |= 1u << (bounce + 1 + e);
                        const float w = technique_weight(M, bounce + k + 3, max_depth, B.tdepth);
                        const float sc = (float)nj / (float)B.nconn;
                        contrib = make_float4(Tv.x * xa.x * fy.x * G * fx.x * w * sc,
                                              Tv.y * xa.y * fy.y * G * fx.y * w * sc,
                                              Tv.z * xa.z * fy.z * G * fx.z * w * sc, 0.f);
                        r0 = make_float4(p.x, p.y, p.z, 1.0e-3f * th);
                        r1 = make_float4(dir.x, dir.y, dir.z, dist * (1.0f - 1.0e-3f));
                        want = true;
                    }
                }
                P.cc[(size_t)pid * kMaxConn + i] = contrib;
            }
        }
        const uint32_t si = block_reserve(&P.cnt[2], want ? 1u : 0u);
        if (valid) P.cslot[(size_t)pid * kMaxConn + i] = want ? (int)si : -1;
        if (want) {
            P.ra[2 * (size_t)si] = r0;
            P.ra[2 * (size_t)si + 1] = r1;
        }
    }
}

// ------------------------------------------------------------ ray cast --

// Debug colours of kernel_ray_cast_update (src/integrators/integrator_raycast.cu:47-61).
__constant__ float kRaycastPalette[15][3] = {
    {1.0f, 0.0f, 0.0f}, {0.0f, 1.0f, 0.0f}, {0.0f, 0.0f, 1.0f},   {0.0f, 1.0f, 1.0f},  {1.0f, 0.0f, 1.0f},
    {1.0f, 1.0f, 0.0f}, {0.1f, 0.1f, 0.1f}, {0.5f, 0.5f, 0.5f},   {1.0f, 1.0f, 1.0f},  {0.0f, 0.25f, 0.5f},
    {1.0f, 0.5f, 0.0f}, {0.5f, 0.25f, 0.0f}, {0.2f, 0.9f, 1.0f}, {0.7f, 0.3f, 0.1f}, {0.34f, 0.89f, 0.45f}};

// Per sample debug colour of the closest hit (kernel_ray_cast_update,
// integrator_raycast.cu:37-128); misses are black.  PRIMITIVE_ID: palette
// entry of the triangle id mod 15; MATERIAL_ID: the material's colour;
// NORMALS: the shading normal (interpolated vertex normals; for a triangle
// without them the geometric normal cross(v1 - v0, v2 - v0), normalised);
// UVS: the hit's barycentric weights of v0 and v1; UNIFORM: white.
__global__ void k_raycast_shade(const float4* __restrict__ rays, const int2* __restrict__ hits, uint32_t n,
                                SceneDev S, uint32_t mode, float4* __restrict__ L) {
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        const int2 h = hits[i];
        V3 c = v3(0.0f, 0.0f, 0.0f);
        if (h.x >= 0) {
            const uint32_t t = (uint32_t)h.x;
            V3 v0, v1, v2;
            uint32_t mat, tf;
            tri_fetch(S, t, &v0, &v1, &v2, &mat, &tf);
            if (mode == RB_SHADE_PRIMITIVE_ID) {
                const uint32_t k = t % 15u;
                c = v3(kRaycastPalette[k][0], kRaycastPalette[k][1], kRaycastPalette[k][2]);
            } else if (mode == RB_SHADE_MATERIAL_ID) {
                // a layered material shows its first layer's colour
                // (integrator_raycast.cu:71-76)
                const rb_material& ml = S.mats[mat];
                const rb_material& m = ml.kind == RB_MAT_LAYERED ? S.mats[(uint32_t)ml.param[1]] : ml;
                c = v3(m.color[0], m.color[1], m.color[2]);
            } else if (mode == RB_SHADE_NORMALS) {
                // m_shading_normal (integrator_raycast.cu:109-115)
                const float4 o = rays[2 * (size_t)i], d = rays[2 * (size_t)i + 1];
                const float tt = __int_as_float(h.y);
                const V3 p = vadd(v3(o.x, o.y, o.z), vmul(v3(d.x, d.y, d.z), tt));
                interp_normal(S, t, tf, p, v0, v1, v2, tri_normal(v0, v1, v2), &c);
            } else if (mode == RB_SHADE_UVS) {
                const float4 o = rays[2 * (size_t)i], d = rays[2 * (size_t)i + 1];
                const float tt = __int_as_float(h.y);
                const V3 p = vadd(v3(o.x, o.y, o.z), vmul(v3(d.x, d.y, d.z), tt));
                const V3 e0 = vsub(v0, v2), e1 = vsub(v1, v2), q = vsub(p, v2);
                const float d00 = vdot(e0, e0), d01 = vdot(e0, e1), d11 = vdot(e1, e1);
                const float d20 = vdot(q, e0), d21 = vdot(q, e1);
                const float den = d00 * d11 - d01 * d01;
                c.x = den != 0.0f ? (d11 * d20 - d01 * d21) / den : 0.0f;
                c.y = den != 0.0f ? (d00 * d21 - d01 * d20) / den : 0.0f;
            } else {
                c = v3(1.0f, 1.0f, 1.0f);
            }
        }
        L[i] = make_float4(c.x, c.y, c.z, 0.0f);
    }
}

__global__ void k_film(const float4* __restrict__ L, uint32_t npix, uint32_t spp, float* __restrict__ out) {
    for (uint32_t px = blockIdx.x * blockDim.x + threadIdx.x; px < npix; px += gridDim.x * blockDim.x) {
        float r = 0.f, g = 0.f, b = 0.f;
        for (uint32_t s = 0; s < spp; s++) {
            const float4 v = L[(size_t)px * spp + s];
            r += v.x;
            g += v.y;
            b += v.z;
        }
        const float inv = (float)spp;
        out[3 * (size_t)px] = r / inv;
        out[3 * (size_t)px + 1] = g / inv;
        out[3 * (size_t)px + 2] = b / inv;
    }
}

// Filtered film (kernel_update_film, src/renderers/samplerrenderer.cu:103-169,
// as a gather): pixel X takes sample s of tile pixel g when the sample's splat
// window [ceil(d - w), floor(d + w)] (d = sample position - 0.5) contains X;
// the weight is the filter table entry of (|X - d| / w * 16) per axis.
// Sample positions are those of the camera rays (same Philox draw), so
// nothing is stored; sums run in a fixed order (neighbour row, column,
// sample), so images are reproducible.
struct FilmFilter {
    float xw, yw, inv_xw, inv_yw;
    float table[256];  // [y * 16 + x]
};

__global__ void k_film_filter(const float4* __restrict__ L, TileArgs T, uint32_t h, FilmFilter F, RngArg rng_a,
                              float* __restrict__ out) {
    const Rng rng = rng_a.get();
    const uint32_t npix = T.w * h;
    const int rx = (int)ceilf(F.xw) + 1, ry = (int)ceilf(F.yw) + 1;
    for (uint32_t px = blockIdx.x * blockDim.x + threadIdx.x; px < npix; px += gridDim.x * blockDim.x) {
        const int lx = (int)(px % T.w), ly = (int)(px / T.w);
        const float X = (float)(T.x0 + (uint32_t)lx), Y = (float)(T.y0 + (uint32_t)ly);
        float r = 0.f, g = 0.f, b = 0.f, ws = 0.f;
        const int ny0 = ly - ry < 0 ? 0 : ly - ry, ny1 = ly + ry > (int)h - 1 ? (int)h - 1 : ly + ry;
        const int nx0 = lx - rx < 0 ? 0 : lx - rx, nx1 = lx + rx > (int)T.w - 1 ? (int)T.w - 1 : lx + rx;
        for (int ny = ny0; ny <= ny1; ny++) {
            const uint32_t gy = T.y0 + (uint32_t)ny;
            for (int nx = nx0; nx <= nx1; nx++) {
                const uint32_t gx = T.x0 + (uint32_t)nx;
                for (uint32_t s = 0; s < T.spp; s++) {
                    const U4 q = rng.draw(gy * T.width + gx, s, 0, 0);
                    const float dx = ((float)gx + u01(q.x)) - 0.5f;
                    const float dy = ((float)gy + u01(q.y)) - 0.5f;
                    if (!(ceilf(dx - F.xw) <= X && X <= floorf(dx + F.xw))) continue;
                    if (!(ceilf(dy - F.yw) <= Y && Y <= floorf(dy + F.yw))) continue;
                    int ix = (int)floorf(fabsf((X - dx) * F.inv_xw * 16.0f));
                    int iy = (int)floorf(fabsf((Y - dy) * F.inv_yw * 16.0f));
                    ix = ix < 15 ? ix : 15;
                    iy = iy < 15 ? iy : 15;
                    const float wgt = F.table[iy * 16 + ix];
                    const float4 v = L[((size_t)ny * T.w + (size_t)nx) * T.spp + s];
                    r += wgt * v.x;
                    g += wgt * v.y;
                    b += wgt * v.z;
                    ws += wgt;
                }
            }
        }
        float o0 = 0.f, o1 = 0.f, o2 = 0.f;
        if (ws != 0.0f) {
            const float inv = 1.0f / ws;
            o0 = fmaxf(0.0f, r * inv);
            o1 = fmaxf(0.0f, g * inv);
            o2 = fmaxf(0.0f, b * inv);
        }
        out[3 * (size_t)px] = o0;
        out[3 * (size_t)px + 1] = o1;
        out[3 * (size_t)px + 2] = o2;
    }
}

// Filter table of Film::Film (src/core/film.cpp:45-58; x offsets from the x
// index, where the reference reuses y): entry (x, y) = filter at
// ((x + 0.5) w_x / 16, (y + 0.5) w_y / 16).  Filters of src/core/filter.h:69-132.
static rb_status make_filter(const rb_render_params* p, FilmFilter* f, bool* plain) {
    const float xw = p->filter_width[0], yw = p->filter_width[1];
    if (p->filter > RB_FILTER_GAUSSIAN) return fail(RB_ERR_INVALID, "unknown filter %u", p->filter);
    if (!(xw > 0.0f && xw <= 8.0f && yw > 0.0f && yw <= 8.0f))
        return fail(RB_ERR_INVALID, "filter half-widths must be in (0, 8]");
    *plain = p->filter == RB_FILTER_BOX && xw == 0.5f && yw == 0.5f;
    f->xw = xw;
    f->yw = yw;
    f->inv_xw = 1.0f / xw;
    f->inv_yw = 1.0f / yw;
    const float a = p->filter_alpha;
    const float ex = expf(-a * xw * xw), ey = expf(-a * yw * yw);
    for (int y = 0; y < 16; y++) {
        const float fy = ((float)y + 0.5f) * yw / 16.0f;
        for (int x = 0; x < 16; x++) {
            const float fx = ((float)x + 0.5f) * xw / 16.0f;
            float v = 1.0f;
            if (p->filter == RB_FILTER_TRIANGLE) v = fmaxf(0.0f, xw - fabsf(fx)) * fmaxf(0.0f, yw - fabsf(fy));
            else if (p->filter == RB_FILTER_GAUSSIAN)
                v = fmaxf(0.0f, expf(-a * fx * fx) - ex) * fmaxf(0.0f, expf(-a * fy * fy) - ey);
            f->table[y * 16 + x] = v;
        }
    }
    return RB_OK;
}

// Film of a tile from per-sample radiance: the pixel mean for the default
// box of half-width 0.5, the filtered gather otherwise.
static rb_status film(const float4* L, const rb_camera* cam, uint32_t x0, uint32_t y0, uint32_t w, uint32_t h,
                      const rb_render_params* p, float* out, hipStream_t s);

static uint32_t grid_of(uint64_t n) {
    uint64_t g = (n + 255) / 256;
    const uint64_t cap = (uint64_t)dev().num_cus * 16;
    if (g > cap) g = cap;
    return (uint32_t)(g ? g : 1);
}

// Grid of the queue kernels (kQueueBlock-thread workgroups, same thread cap).
static uint32_t grid_q(uint64_t n) {
    uint64_t g = (n + kQueueBlock - 1) / kQueueBlock;
    const uint64_t cap = (uint64_t)dev().num_cus * 16 * 256 / kQueueBlock;
    if (g > cap) g = cap;
    return (uint32_t)(g ? g : 1);
}

// Stream-ordered scratch allocations released at the end of a render call;
// while a render graph is captured, persistent allocations the graph owns;
// inside the server, carved from the launching slot's arena.
struct Scratch {
    hipStream_t s;
    std::vector<void*> ptrs;
    rb_status st = RB_OK;
    explicit Scratch(hipStream_t st_) : s(st_) {}
    template <class T>
    T* get(size_t count) {
        void* p = nullptr;
        if (st != RB_OK) return nullptr;
        if (t_capture) {
            hipError_t e = hipMalloc(&p, count * sizeof(T) + 16);
            if (e != hipSuccess) {
                (void)hipGetLastError();
                st = fail(RB_ERR_OOM, "render graph scratch (%zu bytes): %s", count * sizeof(T), hipGetErrorString(e));
                return nullptr;
            }
            t_capture->allocs.push_back(p);
            return (T*)p;
        }
        if (t_arena && (p = t_arena->take(count * sizeof(T) + 16)) != nullptr) return (T*)p;
        hipError_t e = hipMallocAsync(&p, count * sizeof(T) + 16, s);
        if (e != hipSuccess) {
            (void)hipGetLastError();
            st = fail(e == hipErrorOutOfMemory ? RB_ERR_OOM : RB_ERR_DEVICE, "render scratch (%zu bytes): %s",
                      count * sizeof(T), hipGetErrorString(e));
            return nullptr;
        }
        ptrs.push_back(p);
        return (T*)p;
    }
    ~Scratch() {
        for (void* p : ptrs) (void)hipFreeAsync(p, s);
    }
};

static rb_status check_render(rb_bvh_t bvh, const rb_camera* cam, uint32_t x0, uint32_t y0, uint32_t w, uint32_t h,
                              const rb_render_params* p, float* out) {
    if (!bvh || !cam || !p || !out) return fail(RB_ERR_INVALID, "NULL argument");
    if (!bvh->tri_v) return fail(RB_ERR_INVALID, "bvh was uploaded without its scene");
    if (w == 0 || h == 0 || (uint64_t)x0 + w > cam->width || (uint64_t)y0 + h > cam->height)
        return fail(RB_ERR_INVALID, "bad tile");
    if (p->spp == 0) return fail(RB_ERR_INVALID, "spp must be >= 1");
    if ((uint64_t)w * h * p->spp > (1ull << 30)) return fail(RB_ERR_INVALID, "tile x spp too large for one call");
    FilmFilter f;
    bool plain;
    return make_filter(p, &f, &plain);
}

static rb_status film(const float4* L, const rb_camera* cam, uint32_t x0, uint32_t y0, uint32_t w, uint32_t h,
                      const rb_render_params* p, float* out, hipStream_t s) {
    FilmFilter f;
    bool plain;
    rb_status st = make_filter(p, &f, &plain);
    if (st != RB_OK) return st;
    if (plain) {
        hipLaunchKernelGGL(k_film, dim3(grid_of(w * h)), dim3(256), 0, s, L, w * h, p->spp, out);
    } else {
        TileArgs T{cam->width, x0, y0, w, p->spp};
        hipLaunchKernelGGL(k_film_filter, dim3(grid_of(w * h)), dim3(256), 0, s, L, T, h, f,
                           rng_arg(p->seed, RNG_CAMERA), out);
    }
    hipError_t e = hipGetLastError();
    return e == hipSuccess ? RB_OK : hip_fail(e, "film launch");
}

static SceneDev scene_of(rb_bvh_t b) {
    SceneDev S;
    S.tri_v = b->tri_v;
    S.mats = b->mats;
    S.light_tris = b->light_tris;
    S.n_lights = b->n_lights;
    S.tri_n = b->tri_n;
    return S;
}

uint64_t bdpt_scratch_bytes(uint32_t w, uint32_t h, const rb_render_params* p) {
    const uint64_t n = (uint64_t)w * h * (p->spp ? p->spp : 1);
    const uint64_t D = p->light_max_depth ? p->light_max_depth : 1;
    const uint64_t nl = p->light_paths ? p->light_paths : p->lvc_capacity / D;
    const uint64_t nc = p->connections;
    // cache records are 64 bytes, 80 in scenes with shading normals: the
    // estimate (made without the scene) takes the larger
    return n * 352 + n * (4 + kMaxConn * 20) + n * nc * 40 + nl * D * 80 + nl * (32 + 16 + 4 + 8 + 4 + 8 + 32);
}

// ------------------------------------------------------------- jobs ----
// A render call as a sequence of steps.  Each step enqueues the integrator
// kernels that lead up to a trace and names the ray sets that trace takes
// (one or two); the last step writes the film.  render_* run a job's steps
// with one launch per step (a pair launch for two sets); the server's
// lockstep mode traces the sets of every in-flight job in one batched launch
// per step.  Every buffer is allocated when the job is made.

struct AoJob : RenderJob {
    rb_camera cam;
    rb_render_params p;
    uint32_t x0, y0, w, h, n;
    uint64_t na;
    float* out;
    Scratch M;
    float4 *rays, *arays, *Ls;
    int2 *hits, *ahits;
    uint8_t* amask;
    int state = 0;
    explicit AoJob(hipStream_t s) : M(s) {}
    rb_status step(hipStream_t s, TraceSet* sets, uint32_t* nsets, bool* done) override {
        rb_status st;
        *nsets = 0;
        *done = false;
        if (state == 0) {
            if ((st = rb_raygen_camera(&cam, x0, y0, w, h, p.spp, p.seed, 1, (rb_ray*)rays, s)) != RB_OK) return st;
            sets[0] = TraceSet{rays, hits, n, nullptr, RB_CLOSEST, nullptr, dev().cfg};
            *nsets = 1;
        } else if (state == 1) {
            TileArgs T{cam.width, x0, y0, w, p.spp};
            hipLaunchKernelGGL(k_ao_spawn, dim3(grid_of(na)), dim3(256), 0, s, rays, hits, n, T, scene_of(bvh),
                               p.ao_samples, p.ao_radius, rng_arg(p.seed, RNG_AO), arays, amask, p.ray_counter);
            if (p.ray_counter) hipLaunchKernelGGL(k_add64, dim3(1), dim3(64), 0, s, p.ray_counter, (unsigned long long)n);
            // AO rays are short (tmax = the AO radius): a wave finishes a
            // claim of rays quickly and the launch's work-counter atomics set
            // the pace; claims of 128 rays halve them (AO frame -19% at 2 Mi
            // rays) once the launch is large enough for every resident wave
            // to get several claims (small tiles keep the finer claims).
            rb_trace_config ao_cfg = dev().cfg;
            const uint64_t resident = (uint64_t)dev().num_cus * ao_cfg.waves_per_cu;
            if (ao_cfg.chunk < 128 && na >= resident * 128 * 2) ao_cfg.chunk = 128;
            sets[0] = TraceSet{arays, ahits, (uint32_t)na, nullptr, RB_ANY | kOcclusion | RB_MASKED, amask, ao_cfg};
            *nsets = 1;
        } else {
            hipLaunchKernelGGL(k_ao_resolve, dim3(grid_of(n)), dim3(256), 0, s, hits, ahits, n, p.ao_samples, Ls);
            hipError_t e = hipGetLastError();
            if (e != hipSuccess) return hip_fail(e, "AO launch");
            if ((st = film(Ls, &cam, x0, y0, w, h, &p, out, s)) != RB_OK) return st;
            *done = true;
        }
        state++;
        return RB_OK;
    }
};

struct RaycastJob : RenderJob {
    rb_camera cam;
    rb_render_params p;
    uint32_t x0, y0, w, h, n;
    float* out;
    Scratch M;
    float4 *rays, *Ls;
    int2* hits;
    int state = 0;
    explicit RaycastJob(hipStream_t s) : M(s) {}
    rb_status step(hipStream_t s, TraceSet* sets, uint32_t* nsets, bool* done) override {
        rb_status st;
        *nsets = 0;
        *done = false;
        if (state == 0) {
            if ((st = rb_raygen_camera(&cam, x0, y0, w, h, p.spp, p.seed, 1, (rb_ray*)rays, s)) != RB_OK) return st;
            sets[0] = TraceSet{rays, hits, n, nullptr, RB_CLOSEST, nullptr, dev().cfg};
            *nsets = 1;
        } else {
            if (p.ray_counter) hipLaunchKernelGGL(k_add64, dim3(1), dim3(64), 0, s, p.ray_counter, (unsigned long long)n);
            hipLaunchKernelGGL(k_raycast_shade, dim3(grid_of(n)), dim3(256), 0, s, rays, hits, n, scene_of(bvh), p.shade,
                               Ls);
            hipError_t e = hipGetLastError();
            if (e != hipSuccess) return hip_fail(e, "ray-cast launch");
            if ((st = film(Ls, &cam, x0, y0, w, h, &p, out, s)) != RB_OK) return st;
            *done = true;
        }
        state++;
        return RB_OK;
    }
};

// The path tracer (bd = false) and BDPT (bd = true): BDPT's light pass
// (kernel_sk_light_trace, integrator_bdpt_lvc_sk.cu:47-168), its first
// trace paired with the camera rays (independent work, one launch tail
// instead of two), then the camera pass (integrator_path.cu:43-261;
// kernel_sk_camera_trace, integrator_bdpt_lvc_sk.cu:170-435): step b takes
// bounce b - 1's contributions and shades bounce b -- its input queue (the
// vertices shaded at b - 1) is (b + 1) & 1, its output queue b & 1 -- and
// each bounce's shadow rays (any hit) and MIS + bounce rays (closest hit)
// are traced between the steps.
struct CameraJob : RenderJob {
    rb_camera cam;
    rb_render_params p;
    uint32_t x0, y0, w, h, n, na;
    float* out;
    bool bd;
    Scratch M;
    PathBufs P;
    BdArgs B;
    float4* prim;
    int2* phits;
    float4* Tq[2];
    float4* rbcq[2];
    int2* hbcq[2];
    // BDPT light pass
    LightPath LP;
    float4* lrays = nullptr;
    uint32_t* ncam = nullptr;
    uint32_t NL = 0, D = 0;
    // progress: 0 start, 1 light pass, 2 camera bounces, 3 last contributions
    int phase = 0;
    uint32_t b = 0;
    explicit CameraJob(hipStream_t s) : M(s) {}

    TileArgs tile() const { return TileArgs{cam.width, x0, y0, w, p.spp}; }

    void launch_step(hipStream_t s, int mode, uint32_t qin) {
        const dim3 g(grid_q(n)), blk(kQueueBlock);
        const TileArgs T = tile();
        const SceneDev S = scene_of(bvh);
        const RngArg rng = rng_arg(p.seed, RNG_PATH);
        const uint32_t md = p.max_depth, rs = p.rr_start;
        const float rp = p.rr_prob, nl = (float)bvh->n_lights;
        if (bd) {
            if (mode == 0) hipLaunchKernelGGL((k_pt_step<true, 0>), g, blk, 0, s, P, qin, n, T, S, rng, md, rs, rp, nl, B);
            else if (mode == 1) hipLaunchKernelGGL((k_pt_step<true, 1>), g, blk, 0, s, P, qin, n, T, S, rng, md, rs, rp, nl, B);
            else hipLaunchKernelGGL((k_pt_step<true, 2>), g, blk, 0, s, P, qin, n, T, S, rng, md, rs, rp, nl, B);
        } else {
            if (mode == 0) hipLaunchKernelGGL((k_pt_step<false, 0>), g, blk, 0, s, P, qin, n, T, S, rng, md, rs, rp, nl, B);
            else if (mode == 1) hipLaunchKernelGGL((k_pt_step<false, 1>), g, blk, 0, s, P, qin, n, T, S, rng, md, rs, rp, nl, B);
            else hipLaunchKernelGGL((k_pt_step<false, 2>), g, blk, 0, s, P, qin, n, T, S, rng, md, rs, rp, nl, B);
        }
    }

    rb_status step(hipStream_t s, TraceSet* sets, uint32_t* nsets, bool* done) override {
        rb_status st;
        *nsets = 0;
        *done = false;
        const TileArgs T = tile();
        const SceneDev S = scene_of(bvh);
        if (phase == 0) {
            if ((st = rb_raygen_camera(&cam, x0, y0, w, h, p.spp, p.seed, 1, (rb_ray*)prim, s)) != RB_OK) return st;
            if (!bd) {
                sets[0] = TraceSet{prim, phits, n, nullptr, RB_CLOSEST, nullptr, dev().cfg};
                *nsets = 1;
                phase = 2;
                return RB_OK;
            }
            hipLaunchKernelGGL(k_set_u32, dim3(1), dim3(64), 0, s, ncam, n);
            hipLaunchKernelGGL(k_reset, dim3(1), dim3(64), 0, s, LP.cnt, 0u, 1u, 2u);
            hipLaunchKernelGGL(k_light_emit, dim3(grid_q(NL)), dim3(kQueueBlock), 0, s, NL, S,
                               rng_arg(p.seed, RNG_LIGHT), LP, B.C);
            hipLaunchKernelGGL(k_light_gather, dim3(grid_of(NL)), dim3(256), 0, s, LP, lrays);
            sets[0] = TraceSet{lrays, LP.hit, NL, LP.cnt + 0, RB_CLOSEST, nullptr, dev().cfg};
            sets[1] = TraceSet{prim, phits, n, ncam, RB_CLOSEST, nullptr, dev().cfg};
            *nsets = 2;
            phase = 1;
            return RB_OK;
        }
        if (phase == 1) {  // light depth 0 has been traced: the rest of the light pass, then the camera pass
            if (p.ray_counter) hipLaunchKernelGGL(k_add_count, dim3(1), dim3(64), 0, s, LP.cnt + 0, p.ray_counter);
            if (p.light_path_counter)
                hipLaunchKernelGGL(k_add_count, dim3(1), dim3(64), 0, s, LP.cnt + 0, p.light_path_counter);
            const rb_bvh_dev bv = device_view(bvh);
            hipLaunchKernelGGL(k_light_walk, dim3(grid_of(NL)), dim3(256), 0, s, S, rng_arg(p.seed, RNG_LIGHT), LP, B.C,
                               bv, p.ray_counter, p.light_path_counter);
            phase = 2;
        }
        if (phase == 2) {
            if (b > 0 && p.ray_counter)  // the previous bounce's rays
                hipLaunchKernelGGL(k_count_rays, dim3(1), dim3(64), 0, s, (const uint32_t*)P.cnt, p.ray_counter,
                                   b == 1 ? (unsigned long long)n : 0ull);
            if (b > p.max_depth) {
                phase = 3;
            } else {
                const uint32_t qin = (b + 1) & 1, qout = b & 1;
                P.pray = b == 0 ? prim : rbcq[(b - 1) & 1];
                P.phit = b == 0 ? phits : hbcq[(b - 1) & 1];
                P.rbc = rbcq[b & 1];
                P.hbc = hbcq[b & 1];
                P.T = Tq[b & 1];
                P.Tn = Tq[(b + 1) & 1];
                P.Tprev = Tq[(b + 1) & 1];  // bounce b - 1's T; this bounce's Tn overwrites it path by path
                hipLaunchKernelGGL(k_reset, dim3(1), dim3(64), 0, s, P.cnt, qout, 2u, 3u);
                launch_step(s, b == 0 ? 0 : 1, qin);
                if (bd && B.nconn)
                    hipLaunchKernelGGL(k_lvc_connect, dim3(grid_q((uint64_t)n * B.nconn)), dim3(kQueueBlock), 0, s, P,
                                       qout, T, S, p.max_depth, B);
                if (p.path_counter)
                    hipLaunchKernelGGL(k_add_count, dim3(1), dim3(64), 0, s, (const uint32_t*)P.cnt + qout,
                                       p.path_counter + b);
                sets[0] = TraceSet{P.ra, P.ha, na, P.cnt + 2, RB_ANY | kOcclusion, nullptr, dev().cfg};
                sets[1] = TraceSet{P.rbc, P.hbc, 2 * n, P.cnt + 3, RB_CLOSEST, nullptr, dev().cfg};
                *nsets = 2;
                b++;
                return RB_OK;
            }
        }
        // the last bounce's contributions, then the film
        P.phit = hbcq[p.max_depth & 1];
        P.Tprev = Tq[p.max_depth & 1];
        launch_step(s, 2, p.max_depth & 1);
        hipError_t e = hipGetLastError();
        if (e != hipSuccess) return hip_fail(e, "camera pass launch");
        if ((st = film(P.L, &cam, x0, y0, w, h, &p, out, s)) != RB_OK) return st;
        *done = true;
        return RB_OK;
    }
};

rb_status make_render_job(uint32_t kind, rb_bvh_t bvh, const rb_camera* cam, uint32_t x0, uint32_t y0, uint32_t w,
                          uint32_t h, const rb_render_params* p, float* out, hipStream_t s,
                          std::unique_ptr<RenderJob>* job) {
    rb_status st = check_render(bvh, cam, x0, y0, w, h, p, out);
    if (st != RB_OK) return st;
    const uint32_t n = w * h * p->spp;
    try {
        if (kind == RB_REQ_AO) {
            if (p->ao_samples == 0) return fail(RB_ERR_INVALID, "ao_samples must be >= 1");
            const uint64_t na = (uint64_t)n * p->ao_samples;
            if (na > 0x7FFFFFFFull) return fail(RB_ERR_INVALID, "too many AO rays for one call");
            std::unique_ptr<AoJob> J(new AoJob(s));
            J->bvh = bvh;
            J->cam = *cam;
            J->p = *p;
            J->x0 = x0, J->y0 = y0, J->w = w, J->h = h, J->n = n, J->na = na, J->out = out;
            Scratch& M = J->M;
            J->rays = M.get<float4>(2 * (size_t)n);
            J->hits = M.get<int2>(n);
            J->arays = M.get<float4>(2 * na);
            J->amask = M.get<uint8_t>(na);
            J->ahits = M.get<int2>(na);
            J->Ls = M.get<float4>(n);
            if (M.st != RB_OK) return M.st;
            *job = std::move(J);
            return RB_OK;
        }
        if (kind == RB_REQ_RAYCAST) {
            if (p->shade > RB_SHADE_UNIFORM) return fail(RB_ERR_INVALID, "unknown ray-cast shade mode %u", p->shade);
            std::unique_ptr<RaycastJob> J(new RaycastJob(s));
            J->bvh = bvh;
            J->cam = *cam;
            J->p = *p;
            J->x0 = x0, J->y0 = y0, J->w = w, J->h = h, J->n = n, J->out = out;
            Scratch& M = J->M;
            J->rays = M.get<float4>(2 * (size_t)n);
            J->hits = M.get<int2>(n);
            J->Ls = M.get<float4>(n);
            if (M.st != RB_OK) return M.st;
            *job = std::move(J);
            return RB_OK;
        }
        if (kind != RB_REQ_PT && kind != RB_REQ_BDPT) return fail(RB_ERR_INVALID, "unknown render kind %u", kind);
        const bool bd = kind == RB_REQ_BDPT;
        if (!bd && p->max_depth > 250) return fail(RB_ERR_INVALID, "max_depth must be <= 250");
        if (bd) {
            if (p->max_depth > 24) return fail(RB_ERR_INVALID, "BDPT max_depth must be <= 24");
            if (p->light_max_depth == 0 || p->light_max_depth > 8)
                return fail(RB_ERR_INVALID, "light_max_depth must be 1..8");
            // full-path specular masks are 32-bit: path length max_depth + light_max_depth + 2 <= 32
            if (p->max_depth + p->light_max_depth > 30)
                return fail(RB_ERR_INVALID, "max_depth + light_max_depth must be <= 30");
            if (p->connections > kMaxConn) return fail(RB_ERR_INVALID, "connections must be <= %u", kMaxConn);
            if (p->light_paths > (1u << 24) || (p->light_paths == 0 && p->lvc_capacity == 0))
                return fail(RB_ERR_INVALID, "light_paths must be 1..2^24 (or 0 with an lvc_capacity)");
        }
        // one launch traces at most 2^31 - 1 rays: the shadow queue (light
        // samples + connections) and the MIS + bounce queue (2 per path)
        const uint32_t nconn = bd && bvh->n_lights ? p->connections : 0;
        const uint64_t na64 = (uint64_t)n * (1 + (bd ? p->connections : 0));
        if (na64 > 0x7FFFFFFFull || 2ull * n > 0x7FFFFFFFull)
            return fail(RB_ERR_INVALID, "tile x spp too large: %u paths need %llu shadow rays and %llu bounce rays per "
                        "launch (at most 2^31 - 1)", n, (unsigned long long)na64, 2ull * n);
        std::unique_ptr<CameraJob> J(new CameraJob(s));
        J->bvh = bvh;
        J->cam = *cam;
        J->p = *p;
        J->x0 = x0, J->y0 = y0, J->w = w, J->h = h, J->n = n, J->out = out, J->bd = bd;
        J->na = (uint32_t)(bd ? (uint64_t)n * (1 + nconn) : n);
        Scratch& M = J->M;
        PathBufs& P = J->P;
        BdArgs& B = J->B;
        std::memset(&P, 0, sizeof(P));
        std::memset(&B, 0, sizeof(B));
        if (bd) {
            const uint32_t D = p->light_max_depth;
            // light sub-paths: given, or as many as fill a cache of lvc_capacity vertices
            const uint32_t NL = p->light_paths ? p->light_paths : (p->lvc_capacity / D ? p->lvc_capacity / D : 1);
            J->D = D;
            J->NL = NL;
            B.nconn = nconn;
            B.tdepth = B.nconn ? D : 0;
            B.crng = rng_arg(p->seed, RNG_CONNECT);
            B.C.depth = D;
            B.C.paths = NL;
            B.C.rows = bvh->tri_n ? 5u : 4u;
            B.C.v = M.get<float4>((size_t)B.C.rows * NL * D);
            B.C.count = M.get<uint32_t>(NL);
            LightPath& LP = J->LP;
            LP.ray = M.get<float4>(2 * (size_t)NL);
            LP.alpha = M.get<float4>(NL);
            LP.mask = M.get<uint32_t>(NL);
            LP.q = M.get<uint32_t>(NL);
            LP.cnt = M.get<uint32_t>(4);
            LP.hit = M.get<int2>(NL);
            J->lrays = M.get<float4>(2 * (size_t)NL);
            J->ncam = M.get<uint32_t>(1);  // camera ray count for the paired launch
        }
        J->prim = M.get<float4>(2 * (size_t)n);
        J->phits = M.get<int2>(n);
        J->Tq[0] = M.get<float4>(n);
        J->Tq[1] = M.get<float4>(n);
        P.L = M.get<float4>(n);
        P.cur = M.get<uint32_t>(n);
        P.ca = M.get<float4>(n);
        P.cb = M.get<float4>(n);
        P.slot = M.get<int4>(n);
        P.info = M.get<uint32_t>(n);
        P.q[0] = M.get<uint32_t>(n);
        P.q[1] = M.get<uint32_t>(n);
        P.cnt = M.get<uint32_t>(4);
        P.ra = M.get<float4>(2 * (size_t)J->na);
        P.ha = M.get<int2>(J->na);
        for (int k = 0; k < 2; k++) {
            J->rbcq[k] = M.get<float4>(4 * (size_t)n);
            J->hbcq[k] = M.get<int2>(2 * (size_t)n);
        }
        if (bd) {
            P.cmask = M.get<uint32_t>(n);
            P.cc = M.get<float4>((size_t)n * kMaxConn);
            P.cslot = M.get<int>((size_t)n * kMaxConn);
        }
        if (M.st != RB_OK) return M.st;
        *job = std::move(J);
        return RB_OK;
    } catch (...) {
        return fail(RB_ERR_OOM, "render job (host allocation)");
    }
}

// Runs a job on one stream: each step's ray sets in one launch (a pair
// launch for two sets).
rb_status run_job(RenderJob& J, hipStream_t s) {
    TraceSet sets[2];
    for (;;) {
        uint32_t ns = 0;
        bool done = false;
        rb_status st = J.step(s, sets, &ns, &done);
        if (st != RB_OK || done) return st;
        if (ns == 1)
            st = launch_trace(J.bvh, (const rb_ray*)sets[0].rays, sets[0].n, sets[0].flags, sets[0].mask,
                              (rb_hit*)sets[0].hits, nullptr, s, sets[0].cfg, sets[0].n_dev);
        else
            st = launch_trace_pair(J.bvh, (const rb_ray*)sets[0].rays, sets[0].n, sets[0].n_dev, (rb_hit*)sets[0].hits,
                                   sets[0].flags, (const rb_ray*)sets[1].rays, sets[1].n, sets[1].n_dev,
                                   (rb_hit*)sets[1].hits, s, sets[0].cfg);
        if (st != RB_OK) return st;
    }
}

static rb_status render_kind(uint32_t kind, rb_bvh_t bvh, const rb_camera* cam, uint32_t x0, uint32_t y0, uint32_t w,
                             uint32_t h, const rb_render_params* p, float* out, hipStream_t s) {
    std::unique_ptr<RenderJob> J;
    rb_status st = make_render_job(kind, bvh, cam, x0, y0, w, h, p, out, s, &J);
    return st == RB_OK ? run_job(*J, s) : st;
}

rb_status render_ao(rb_bvh_t bvh, const rb_camera* cam, uint32_t x0, uint32_t y0, uint32_t w, uint32_t h,
                    const rb_render_params* p, float* out, hipStream_t s) {
    return render_kind(RB_REQ_AO, bvh, cam, x0, y0, w, h, p, out, s);
}
rb_status render_raycast(rb_bvh_t bvh, const rb_camera* cam, uint32_t x0, uint32_t y0, uint32_t w, uint32_t h,
                         const rb_render_params* p, float* out, hipStream_t s) {
    return render_kind(RB_REQ_RAYCAST, bvh, cam, x0, y0, w, h, p, out, s);
}
rb_status render_pt(rb_bvh_t bvh, const rb_camera* cam, uint32_t x0, uint32_t y0, uint32_t w, uint32_t h,
                    const rb_render_params* p, float* out, hipStream_t s) {
    return render_kind(RB_REQ_PT, bvh, cam, x0, y0, w, h, p, out, s);
}
rb_status render_bdpt(rb_bvh_t bvh, const rb_camera* cam, uint32_t x0, uint32_t y0, uint32_t w, uint32_t h,
                      const rb_render_params* p, float* out, hipStream_t s) {
    return render_kind(RB_REQ_BDPT, bvh, cam, x0, y0, w, h, p, out, s);
}

uint64_t render_scratch_bytes(uint32_t kind, uint32_t w, uint32_t h, const rb_render_params* p) {
    const uint64_t n = (uint64_t)w * h * (p->spp ? p->spp : 1);
    const uint64_t slack = 64 * 40;  // per-allocation padding and alignment
    switch (kind) {
        case RB_REQ_AO: return n * 56 + n * (p->ao_samples ? p->ao_samples : 1) * 41 + slack;
        case RB_REQ_PT: return n * 352 + slack;
        case RB_REQ_BDPT: return bdpt_scratch_bytes(w, h, p) + slack;
        case RB_REQ_RAYCAST: return n * 56 + slack;
        default: return slack;
    }
}

}  // namespace rb

using namespace rb;

extern "C" {

void rb_render_params_default(rb_render_params* p) {
    if (!p) return;
    p->spp = 1;
    p->seed = 0;
    p->ao_samples = 8;
    p->ao_radius = 1.0f;
    p->max_depth = 8;
    p->rr_start = 3;
    p->rr_prob = 0.5f;
    p->light_paths = 1u << 18;
    p->lvc_capacity = 1u << 20;
    p->connections = 5;
    p->light_max_depth = 6;
    p->ray_counter = nullptr;
    p->path_counter = nullptr;
    p->light_path_counter = nullptr;
    p->filter = RB_FILTER_BOX;
    p->filter_width[0] = 0.5f;
    p->filter_width[1] = 0.5f;
    p->filter_alpha = 2.0f;
    p->shade = RB_SHADE_PRIMITIVE_ID;
}

rb_status rb_render_ao(rb_bvh_t bvh, const rb_camera* cam, uint32_t x0, uint32_t y0, uint32_t w, uint32_t h,
                       const rb_render_params* p, float* out, void* stream) {
    rb_status st = ensure_device();
    if (st != RB_OK) return st;
    return render_ao(bvh, cam, x0, y0, w, h, p, out, stream ? (hipStream_t)stream : dev().stream);
}

rb_status rb_render_pt(rb_bvh_t bvh, const rb_camera* cam, uint32_t x0, uint32_t y0, uint32_t w, uint32_t h,
                       const rb_render_params* p, float* out, void* stream) {
    rb_status st = ensure_device();
    if (st != RB_OK) return st;
    return render_pt(bvh, cam, x0, y0, w, h, p, out, stream ? (hipStream_t)stream : dev().stream);
}

rb_status rb_render_bdpt(rb_bvh_t bvh, const rb_camera* cam, uint32_t x0, uint32_t y0, uint32_t w, uint32_t h,
                         const rb_render_params* p, float* out, void* stream) {
    rb_status st = ensure_device();
    if (st != RB_OK) return st;
    return render_bdpt(bvh, cam, x0, y0, w, h, p, out, stream ? (hipStream_t)stream : dev().stream);
}

rb_status rb_render_raycast(rb_bvh_t bvh, const rb_camera* cam, uint32_t x0, uint32_t y0, uint32_t w, uint32_t h,
                            const rb_render_params* p, float* out, void* stream) {
    rb_status st = ensure_device();
    if (st != RB_OK) return st;
    return render_raycast(bvh, cam, x0, y0, w, h, p, out, stream ? (hipStream_t)stream : dev().stream);
}

}  // extern "C"

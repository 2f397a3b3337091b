// rb_trace.hip -- persistent while-while BVH traversal for gfx950 (wave64).
//
// The Aila-Laine persistent "while-while" kernel of renderbox2
// (src/accelerators/sbvh/kepler_dynamic_fetch.h:67-363) redesigned for CDNA4:
//  * work fetch: one ballot per wave, one atomic per `chunk` rays from the
//    wave's leader lane, rays handed to terminated lanes by ballot rank
//    (the reference does one atomic per warp refill, :99-153);
//  * traversal stack: entry 0..S-1 of each lane in LDS (lane-interleaved,
//    conflict-free), deeper entries in per-lane private memory (the
//    reference keeps all 64 entries in local memory, :62-71);
//  * nodes: the MI355X layout (4 x 16-byte loads per node, leaf size in the
//    32-bit link, no terminator rows);
//  * ties at equal t go to the lowest original triangle id; the reference
//    keeps the first one found (:264);
//  * the work counter lives in a per-launch slot that the last block resets
//    (the reference copies a zero into a global symbol before each launch,
//    cudatracer.cu:164-165), so concurrent launches and graph replays work;
//  * leaf order: by default a lane tests a leaf as soon as it reaches one,
//    so every ray visits nodes and leaves in the same order whatever the
//    other rays of its wave do -- the order of the single-ray restatement --
//    and results are the same run to run.  SPEC = the reference's
//    speculative form (a lane holding a leaf keeps walking inner nodes until
//    the whole wave holds leaves, kepler_dynamic_fetch.h:222-243): fewer
//    idle lanes, but the leaf order then depends on the wave, and where a
//    slab test rounds a box entry past a hit inside it, which hit survives
//    culling can change with it (rare; ids and t may then differ);
//  * batched mode: one launch serves up to kMaxBatch ray batches against
//    different resident BVHs; every lane carries the scene pointers and the
//    any/closest flag of the batch its current ray belongs to, so a wave can
//    mix rays of several scenes while it refills;
//  * paired mode: two queued ray sets of one BVH (an any-hit or closest-hit
//    set, then a closest-hit set) in one launch, so a render bounce pays
//    one launch's ramp-up and tail instead of two;
//  * top of the tree in LDS: each block copies the first lds_nodes node
//    records (the layout stores them breadth-first) into LDS at start and
//    reads them from there; every ray passes through them, often several
//    times (the reference reads every node through the texture path).
// Arithmetic (fp32, explicit fmaf, reciprocals equal to the IEEE division)
// matches the oracle's restatement bit for bit.
#include "rb_device.h"

namespace rb {

struct TraceArgs {
    const float4* __restrict__ nodes;
    const float4* __restrict__ tris;
    const uint32_t* __restrict__ index;
    const float4* __restrict__ rays;
    const uint8_t* __restrict__ mask;
    int2* __restrict__ hits;
    uint32_t* __restrict__ counts;
    uint32_t* counter;
    uint32_t* errors;  // RB_TRACE_ERR_* flags (device word), set on a traversal-stack overflow
    uint32_t n;
    const uint32_t* n_dev;  // when set, the ray count is read from device memory
    // paired launch: a second ray set (closest hit) after the first (any hit
    // when any0 is set, else closest hit)
    const float4* __restrict__ rays2;
    int2* __restrict__ hits2;
    const uint32_t* n_dev2;
    uint32_t any0;
    uint32_t chunk;
    uint32_t threshold;
    uint32_t leaf_gate;  // non-speculative: leave the inner-node phase once <= this many lanes still search
    uint32_t top_nodes;  // node records 0..top_nodes-1 are read from a copy in LDS
};

constexpr uint32_t kMaxBatch = 64;

struct BatchDesc {
    const float4* nodes;
    const float4* tris;
    const uint32_t* index;
    const float4* rays;
    int2* hits;
    const uint8_t* mask;    // rays with mask 0 report a miss untraced (NULL: none)
    const uint32_t* n_dev;  // ray count in device memory (NULL: n)
    uint32_t n;             // ray count, or the bound of *n_dev
    uint32_t start;         // first global ray index of the batch (set by the kernel)
    uint32_t any;           // 1: terminate on the first hit
};

struct BatchTable {
    uint32_t nb;
    BatchDesc d[kMaxBatch];
};

// Global-memory pointer types.  A pointer read from memory (the LDS batch
// table) is generic to the compiler and loads through it become flat loads,
// which wait on both the vector-memory and the LDS counters; the scene and
// ray pointers of a lane are held in these types so every access is a
// global one.
// Plain clang vectors (no constructors) so the same loads compile in the
// host pass of the source.
typedef float V4f __attribute__((ext_vector_type(4)));
typedef int V4i __attribute__((ext_vector_type(4)));
typedef int V2i __attribute__((ext_vector_type(2)));
typedef __attribute__((address_space(1))) const V4f* GF4;
typedef __attribute__((address_space(1))) const V4i* GI4;
typedef __attribute__((address_space(1))) const uint32_t* GU32;
typedef __attribute__((address_space(1))) V2i* GI2;

// 1/x rounded to nearest even, equal to the IEEE division 1.0f / x: the
// hardware reciprocal refined by one fused Newton step, r + r * (1 - x * r).
// Wherever the refined value is a normal number it is the correctly rounded
// quotient (checked against the division for every one of the 2^32
// single-precision inputs on gfx950: 4,227,858,434 take this path, none
// differs).  Zero, subnormal, infinite and NaN inputs and |x| > 2^126 give a
// non-normal refined value and take the division.  5 instructions instead of
// the division's 11 on the common path.
__device__ __forceinline__ float rcp_rn(float x) {
    const float r = __builtin_amdgcn_rcpf(x);
    const float y = fmaf(fmaf(-x, r, 1.0f), r, r);
    if (__builtin_expect(__builtin_isnormal(y), 1)) return y;
    return 1.0f / x;
}

// Slab-span min / max as single instructions.  fminf / fmaxf on values that
// come from another basic block (the ray's tmin and the current hit
// distance) are preceded by a quieting copy of those operands on every node
// visit (IEEE mode: minNum must not see a signalling NaN); both operands
// here are quiet by construction (fused products, and tmin / hit distance
// quieted once per ray), so the instructions give what fminf / fmaxf give.
__device__ __forceinline__ float vmin_(float a, float b) {
    float r;
    asm("v_min_f32 %0, %1, %2" : "=v"(r) : "v"(a), "v"(b));
    return r;
}
__device__ __forceinline__ float vmax_(float a, float b) {
    float r;
    asm("v_max_f32 %0, %1, %2" : "=v"(r) : "v"(a), "v"(b));
    return r;
}
__device__ __forceinline__ float vmin3_(float a, float b, float c) {
    float r;
    asm("v_min3_f32 %0, %1, %2, %3" : "=v"(r) : "v"(a), "v"(b), "v"(c));
    return r;
}
__device__ __forceinline__ float vmax3_(float a, float b, float c) {
    float r;
    asm("v_max3_f32 %0, %1, %2, %3" : "=v"(r) : "v"(a), "v"(b), "v"(c));
    return r;
}

// Every single-precision bit pattern through rcp_rn's fast path and through
// rcp_rn itself, against the division: out[0] += inputs whose refined
// reciprocal is taken, out[1] += inputs where either differs from 1.0f / x
// (NaN results count as equal whatever their payload).
__global__ __launch_bounds__(256) void k_rcp_check(unsigned long long* out) {
    __shared__ unsigned long long s_fast, s_bad;
    if (threadIdx.x == 0) {
        s_fast = 0;
        s_bad = 0;
    }
    __syncthreads();
    unsigned long long fast = 0, bad = 0;
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t x = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; x < (1ull << 32); x += stride) {
        const float f = __uint_as_float((uint32_t)x);
        const float q = 1.0f / f;
        const float r = __builtin_amdgcn_rcpf(f);
        const float y = fmaf(fmaf(-f, r, 1.0f), r, r);
        if (__builtin_isnormal(y)) {
            fast++;
            if (__float_as_uint(y) != __float_as_uint(q)) bad++;
        }
        const float z = rcp_rn(f);
        if (!(__float_as_uint(z) == __float_as_uint(q) || (z != z && q != q))) bad++;
    }
    atomicAdd(&s_fast, fast);
    atomicAdd(&s_bad, bad);
    __syncthreads();
    if (threadIdx.x == 0) {
        atomicAdd(&out[0], s_fast);
        atomicAdd(&out[1], s_bad);
    }
}

// PAIR: rays [0, n0) are the first set (a.rays / a.hits; any hit when
// a.any0), rays [n0, nrays) the closest-hit set (a.rays2 / a.hits2).
// WIDE: a.nodes holds wide (4-child) records (export_wide); for closest
// hit, whose (id, t) does not depend on the visit order, and for occlusion
// sets (any hit read only as hit / miss).
template <bool ANY_T, bool COUNT, int S, int BLOCK, bool MULTI, bool SPEC, bool PAIR = false, bool WIDE = false>
__device__ __forceinline__ void trace_body(const TraceArgs& a, const BatchDesc* tab, uint32_t nb, uint32_t nrays,
                                           uint32_t n0 = 0) {
    __shared__ int lstack[S * BLOCK];
    int spill[kMaxStack - S];
    const int tid = threadIdx.x;
    const uint32_t lane = __lane_id();
    const uint64_t lt_mask = (lane == 0) ? 0ull : (~0ull >> (64 - lane));

    // The upload refuses trees deep enough to fill the stack, so the else
    // branch is unreachable for a library BVH; it flags the launch instead of
    // writing past the stack.
#define RB_PUSH(v)                                            \
    do {                                                      \
        if (sp + 1 < kMaxStack) {                             \
            ++sp;                                             \
            if (sp < S)                                       \
                lstack[sp * BLOCK + tid] = (v);               \
            else                                              \
                spill[sp - S] = (v);                          \
        } else {                                              \
            atomicOr(a.errors, (uint32_t)RB_TRACE_ERR_STACK); \
        }                                                     \
    } while (0)
    // Pops happen only at sp >= 0 (the sentinel at entry 0 ends the ray).  The
    // LDS read is unconditional (the empty asm takes its value, so it cannot
    // be folded away) and the spill read a separate branch: a select between
    // the two arrays would become one generic (flat) load, which waits on
    // both the vector-memory and the LDS counters.
#define RB_POP(dst)                                                        \
    do {                                                                   \
        int v_ = lstack[(sp < S ? sp : S - 1) * BLOCK + tid];              \
        asm volatile("" : "+v"(v_));                                       \
        if (sp >= S) v_ = spill[sp - S];                                   \
        (dst) = v_;                                                        \
        --sp;                                                              \
    } while (0)

    lstack[tid] = kSentinel;
    // top of the tree (the layout's first records, breadth-first) in LDS
    extern __shared__ V4f lnodes[];
    const uint32_t top = MULTI ? 0u : a.top_nodes;
    if (top) {  // 4 float4 rows per record, binary or wide
        const GF4 src = (GF4)a.nodes;
        for (uint32_t i = tid; i < 4 * top; i += BLOCK) lnodes[i] = src[i];
        __syncthreads();
    }
    uint32_t pool_next = 0, pool_end = 0;
    // ray bands: kRayBands contiguous ranges of chunk-multiple length; a
    // wave claims from the band of its XCD until that runs out, then from
    // the next band in turn (so every ray is claimed exactly once).  A
    // batched launch claims in order from counter word 0: its batches come
    // from several scenes, and with bands the chip works on up to eight of
    // them at once, more than the Infinity Cache holds (config-4 batches
    // 2.97-3.01 against 3.48-3.49 Grays/s)
    const uint32_t band_len = MULTI ? 0u : ((nrays + kRayBands - 1) / kRayBands + a.chunk - 1) / a.chunk * a.chunk;
    uint32_t band = 0;
    if (!MULTI) {
        asm volatile("s_getreg_b32 %0, hwreg(HW_REG_XCC_ID)" : "=s"(band));
        band &= kRayBands - 1;
    }
    int nodeAddr = kSentinel, leafAddr = 0, sp = 0, hitIndex = -1;
    uint32_t rayidx = 0;
    float ox = 0.f, oy = 0.f, oz = 0.f, tmin = 0.f, dx = 0.f, dy = 0.f, dz = 0.f, hitT = 0.f, tmax_in = 0.f;
    float idx_ = 0.f, idy_ = 0.f, idz_ = 0.f, oodx = 0.f, oody = 0.f, oodz = 0.f;
    uint32_t c_inner = 0, c_tri = 0, c_leaf = 0;
    // scene of the lane's current ray (uniform in single-batch mode)
    GF4 nodes = (GF4)a.nodes;
    GF4 tris = (GF4)a.tris;
    GU32 index = (GU32)a.index;
    GI2 hits = (GI2)a.hits;
    bool any = ANY_T;

    while (true) {
        // ---- fetch rays for terminated lanes (wave-uniform control) ----
        const bool term = nodeAddr == kSentinel;
        const uint64_t m = __ballot(term);
        if (m) {
            const uint32_t nterm = __popcll(m);
            const uint32_t rank = __popcll(m & lt_mask);
            const uint32_t avail = pool_end - pool_next;
            uint32_t mine;
            if (avail >= nterm) {
                mine = pool_next + rank;
                pool_next += nterm;
            } else {
                const int leader = __builtin_ctzll(m);
                uint32_t base = 0;
                if (MULTI) {
                    if ((int)lane == leader) base = atomicAdd(a.counter, a.chunk);
                } else {
                    if ((int)lane == leader) {
                        base = nrays;  // every band claimed: the lanes finish
                        for (uint32_t t = 0; t < kRayBands; t++) {
                            const uint32_t o = atomicAdd(a.counter + kBandStride * (band + 1), a.chunk);
                            if (o < band_len && band * band_len + o < nrays) {  // the last band ends at nrays
                                base = band * band_len + o;
                                break;
                            }
                            band = (band + 1) & (kRayBands - 1);
                        }
                    }
                    band = __shfl(band, leader);
                }
                base = __shfl(base, leader);
                mine = rank < avail ? pool_next + rank : base + (rank - avail);
                pool_next = base + (nterm - avail);
                pool_end = base + a.chunk;
            }
            if (term) {
                if (mine >= nrays) break;
                GF4 rays = (GF4)a.rays;
                rayidx = mine;
                if (PAIR) {
                    const bool second = mine >= n0;
                    rays = second ? (GF4)a.rays2 : (GF4)a.rays;
                    hits = second ? (GI2)a.hits2 : (GI2)a.hits;
                    any = !second && a.any0 != 0;
                    rayidx = second ? mine - n0 : mine;
                }
                const uint8_t* mk = a.mask;
                if (MULTI) {
                    // the last batch starting at or before `mine` (empty
                    // batches share their successor's start)
                    uint32_t lo = 0, hi = nb;
                    while (hi - lo > 1) {
                        const uint32_t mid = (lo + hi) >> 1;
                        if (tab[mid].start <= mine) lo = mid;
                        else hi = mid;
                    }
                    const BatchDesc& B = tab[lo];
                    nodes = (GF4)B.nodes;
                    tris = (GF4)B.tris;
                    index = (GU32)B.index;
                    hits = (GI2)B.hits;
                    rays = (GF4)B.rays;
                    any = B.any != 0;
                    rayidx = mine - B.start;
                    mk = B.mask;
                }
                const V4f o = rays[2 * (size_t)rayidx];
                const V4f d = rays[2 * (size_t)rayidx + 1];
                ox = o.x; oy = o.y; oz = o.z;
                dx = d.x; dy = d.y; dz = d.z;
                // quiet copies for the slab min / max (comparisons treat every
                // NaN alike); a miss reports the caller's tmax bits unchanged
                tmin = __builtin_canonicalizef(o.w);
                hitT = __builtin_canonicalizef(d.w);
                tmax_in = d.w;
                const float ooeps = 0x1p-80f;
                idx_ = rcp_rn(fabsf(dx) > ooeps ? dx : copysignf(ooeps, dx));
                idy_ = rcp_rn(fabsf(dy) > ooeps ? dy : copysignf(ooeps, dy));
                idz_ = rcp_rn(fabsf(dz) > ooeps ? dz : copysignf(ooeps, dz));
                oodx = ox * idx_;
                oody = oy * idy_;
                oodz = oz * idz_;
                sp = 0;
                leafAddr = 0;
                hitIndex = -1;
                nodeAddr = (mk && mk[rayidx] == 0) ? kSentinel : 0;
                if (COUNT) { c_inner = 0; c_tri = 0; c_leaf = 0; }
            }
        }

        // ---- traversal ----
        while (nodeAddr != kSentinel) {
            while (WIDE && (uint32_t)nodeAddr < (uint32_t)kSentinel && leafAddr >= 0) {
                // a wide record: the four children's slab tests on their
                // decoded boxes, then the hit children nearest first (the
                // rest pushed, farthest deepest)
                V4f r0;
                V4i r1, r2, r3;
                if ((uint32_t)nodeAddr < top) {
                    const V4f* ln = lnodes + 4 * nodeAddr;
                    r0 = ln[0];
                    r1 = __builtin_bit_cast(V4i, ln[1]);
                    r2 = __builtin_bit_cast(V4i, ln[2]);
                    r3 = __builtin_bit_cast(V4i, ln[3]);
                } else {
                    const GF4 np = nodes + 4 * (size_t)nodeAddr;
                    r0 = np[0];
                    r1 = *(GI4)(np + 1);
                    r2 = *(GI4)(np + 2);
                    r3 = *(GI4)(np + 3);
                    asm volatile("" ::"v"(r0.x), "v"(r1.x), "v"(r2.x), "v"(r3.x));
                }
                if (COUNT) c_inner++;
                // step sizes 2^e from the biased exponents (bytes of word 3)
                const uint32_t eb = __float_as_uint(r0.w);
                const float sx = __uint_as_float((eb & 0xFFu) << 23);
                const float sy = __uint_as_float(((eb >> 8) & 0xFFu) << 23);
                const float sz = __uint_as_float(((eb >> 16) & 0xFFu) << 23);
                const int lk[4] = {r2.z, r2.w, r3.x, r3.y};
                float key[4];
                int lnk[4];
#pragma unroll
                for (int c = 0; c < 4; c++) {
#define RB_Q(w, s, p) fmaf((float)(((uint32_t)(w) >> (8 * c)) & 0xFFu), s, p)
                    const float lx = RB_Q(r1.x, sx, r0.x), hx = RB_Q(r1.y, sx, r0.x);
                    const float ly = RB_Q(r1.z, sy, r0.y), hy = RB_Q(r1.w, sy, r0.y);
                    const float lz = RB_Q(r2.x, sz, r0.z), hz = RB_Q(r2.y, sz, r0.z);
#undef RB_Q
                    const float clox = fmaf(lx, idx_, -oodx), chix = fmaf(hx, idx_, -oodx);
                    const float cloy = fmaf(ly, idy_, -oody), chiy = fmaf(hy, idy_, -oody);
                    const float cloz = fmaf(lz, idz_, -oodz), chiz = fmaf(hz, idz_, -oodz);
                    const float cmin = vmax3_(vmin_(clox, chix), vmin_(cloy, chiy), vmax_(vmin_(cloz, chiz), tmin));
                    const float cmax = vmin3_(vmax_(clox, chix), vmax_(cloy, chiy), vmin_(vmax_(cloz, chiz), hitT));
                    const bool h = cmax >= cmin && lk[c] != kWideEmpty;
                    // a hit box's key stays finite (misses sort last as inf)
                    key[c] = h ? fminf(cmin, 0x1.fffffep127f) : __builtin_inff();
                    lnk[c] = lk[c];
                }
                // sort the four (entry distance, link) pairs, misses last
#define RB_CS(i, j)                                   \
    if (key[j] < key[i]) {                            \
        const float tk = key[i];                      \
        key[i] = key[j];                              \
        key[j] = tk;                                  \
        const int tl = lnk[i];                        \
        lnk[i] = lnk[j];                              \
        lnk[j] = tl;                                  \
    }
                RB_CS(0, 1) RB_CS(2, 3) RB_CS(0, 2) RB_CS(1, 3) RB_CS(1, 2)
#undef RB_CS
                // the hit children are a prefix of the sorted four
                const int npush = (int)(key[1] != __builtin_inff()) + (int)(key[2] != __builtin_inff()) +
                                  (int)(key[3] != __builtin_inff());
                if (key[0] == __builtin_inff()) {
                    RB_POP(nodeAddr);
                } else {
                    nodeAddr = lnk[0];
                    if (sp + 3 < S) {
                        // entries sp+1..sp+npush take lnk[npush..1]
                        // (farthest deepest), written as three straight-line
                        // stores: those past sp+npush lie above the new top
                        const int e1 = npush == 3 ? lnk[3] : npush == 2 ? lnk[2] : lnk[1];
                        const int e2 = npush == 3 ? lnk[2] : lnk[1];
                        lstack[(sp + 1) * BLOCK + tid] = e1;
                        lstack[(sp + 2) * BLOCK + tid] = e2;
                        lstack[(sp + 3) * BLOCK + tid] = lnk[1];
                        sp += npush;
                    } else {
                        if (npush == 3) RB_PUSH(lnk[3]);
                        if (npush >= 2) RB_PUSH(lnk[2]);
                        if (npush >= 1) RB_PUSH(lnk[1]);
                    }
                }
                if (nodeAddr < 0 && leafAddr >= 0) {
                    leafAddr = nodeAddr;
                    RB_POP(nodeAddr);
                }
                if (__popcll(__ballot(1)) <= a.leaf_gate) break;
            }
            while (!WIDE && (uint32_t)nodeAddr < (uint32_t)kSentinel && (SPEC || leafAddr >= 0)) {
                V4f n0, n1, nz;
                V4i lk;
                if ((uint32_t)nodeAddr < top) {
                    const V4f* ln = lnodes + 4 * nodeAddr;
                    n0 = ln[0];
                    n1 = ln[1];
                    nz = ln[2];
                    lk = __builtin_bit_cast(V4i, ln[3]);
                } else {
                    const GF4 np = nodes + 4 * (size_t)nodeAddr;
                    n0 = np[0];
                    n1 = np[1];
                    nz = np[2];
                    lk = *(GI4)(np + 3);
                    // All four loads of the record in flight at once: without
                    // this the compiler sinks the link load below the slab tests
                    // (it is only read when a child is hit), a second memory
                    // round trip per visited node.  The empty asm consumes every
                    // loaded value.
                    asm volatile("" ::"v"(n0.x), "v"(n0.y), "v"(n0.z), "v"(n0.w), "v"(n1.x), "v"(n1.y), "v"(n1.z),
                                 "v"(n1.w), "v"(nz.x), "v"(nz.y), "v"(nz.z), "v"(nz.w), "v"(lk.x), "v"(lk.y));
                }
                if (COUNT) c_inner++;
                const float c0lox = fmaf(n0.x, idx_, -oodx), c0hix = fmaf(n0.y, idx_, -oodx);
                const float c0loy = fmaf(n0.z, idy_, -oody), c0hiy = fmaf(n0.w, idy_, -oody);
                const float c0loz = fmaf(nz.x, idz_, -oodz), c0hiz = fmaf(nz.y, idz_, -oodz);
                const float c1loz = fmaf(nz.z, idz_, -oodz), c1hiz = fmaf(nz.w, idz_, -oodz);
                const float c1lox = fmaf(n1.x, idx_, -oodx), c1hix = fmaf(n1.y, idx_, -oodx);
                const float c1loy = fmaf(n1.z, idy_, -oody), c1hiy = fmaf(n1.w, idy_, -oody);
                const float c0min = vmax3_(vmin_(c0lox, c0hix), vmin_(c0loy, c0hiy), vmax_(vmin_(c0loz, c0hiz), tmin));
                const float c0max = vmin3_(vmax_(c0lox, c0hix), vmax_(c0loy, c0hiy), vmin_(vmax_(c0loz, c0hiz), hitT));
                const float c1min = vmax3_(vmin_(c1lox, c1hix), vmin_(c1loy, c1hiy), vmax_(vmin_(c1loz, c1hiz), tmin));
                const float c1max = vmin3_(vmax_(c1lox, c1hix), vmax_(c1loy, c1hiy), vmin_(vmax_(c1loz, c1hiz), hitT));
                const bool swp = c1min < c0min;
                const bool t0 = c0max >= c0min;
                const bool t1 = c1max >= c1min;
                int far = lk.y;
                if (!t0 && !t1) {
                    RB_POP(nodeAddr);
                } else {
                    nodeAddr = t0 ? lk.x : lk.y;
                    if (t0 && t1) {
                        if (swp) {
                            const int tmp = nodeAddr;
                            nodeAddr = far;
                            far = tmp;
                        }
                        RB_PUSH(far);
                    }
                }
                if (nodeAddr < 0 && leafAddr >= 0) {
                    leafAddr = nodeAddr;
                    RB_POP(nodeAddr);
                }
                if (SPEC) {
                    if (!__any(leafAddr >= 0)) break;
                } else if (__popcll(__ballot(1)) <= a.leaf_gate) {
                    break;  // the rest keep their place and go on after the leaf phase
                }
            }

            while (leafAddr < 0) {
                const uint32_t l = ~(uint32_t)leafAddr;
                const uint32_t first = l >> 5;
                const uint32_t cnt = l & 31u;
                if (COUNT) c_leaf++;
                for (uint32_t k = 0; k < cnt; k++) {
                    const GF4 w = tris + 3 * (size_t)(first + k);
                    const V4f v0 = w[0];
                    const V4f v1 = w[1];
                    const V4f v2 = w[2];
                    if (COUNT) c_tri++;
                    const float Oz = fmaf(-oz, v0.z, fmaf(-oy, v0.y, fmaf(-ox, v0.x, v0.w)));
                    const float invDz = rcp_rn(fmaf(dz, v0.z, fmaf(dy, v0.y, dx * v0.x)));
                    const float t = Oz * invDz;
                    if (t > tmin && (t < hitT || (hitIndex != -1 && t == hitT))) {
                        const float Ox = fmaf(oz, v1.z, fmaf(oy, v1.y, fmaf(ox, v1.x, v1.w)));
                        const float Dx = fmaf(dz, v1.z, fmaf(dy, v1.y, dx * v1.x));
                        const float u = fmaf(t, Dx, Ox);
                        if (u >= 0.0f) {
                            const float Oy = fmaf(oz, v2.z, fmaf(oy, v2.y, fmaf(ox, v2.x, v2.w)));
                            const float Dy = fmaf(dz, v2.z, fmaf(dy, v2.y, dx * v2.x));
                            const float v = fmaf(t, Dy, Oy);
                            if (v >= 0.0f && u + v <= 1.0f) {
                                bool take = t < hitT;
                                if (!take) take = index[first + k] < index[hitIndex];
                                if (take) {
                                    hitT = t;
                                    hitIndex = (int)(first + k);
                                    if (any) {
                                        nodeAddr = kSentinel;
                                        break;
                                    }
                                }
                            }
                        }
                    }
                }
                leafAddr = nodeAddr;
                if (nodeAddr < 0) RB_POP(nodeAddr);
            }

            // Dynamic fetch: refill when too few lanes are still busy.
            if (__popcll(__ballot(1)) < a.threshold) break;
        }

        if (nodeAddr == kSentinel) {
            V2i r;
            r.x = hitIndex < 0 ? -1 : (int)index[hitIndex];
            r.y = __float_as_int(hitIndex < 0 ? tmax_in : hitT);
            hits[rayidx] = r;
            if (COUNT) {
                a.counts[3 * (size_t)rayidx + 0] = c_inner;
                a.counts[3 * (size_t)rayidx + 1] = c_tri;
                a.counts[3 * (size_t)rayidx + 2] = c_leaf;
            }
        }
    }
#undef RB_PUSH
#undef RB_POP

    __syncthreads();
    if (tid == 0) {
        __threadfence();
        if (atomicAdd(a.counter + 1, 1u) == gridDim.x - 1) {
            atomicExch(a.counter + 1, 0u);
            if (MULTI) {
                atomicExch(a.counter + 0, 0u);
            } else {
                // unrolled: as a loop it took the batched kernel from 71 to
                // 81 VGPRs (7 -> 5 waves per SIMD)
#pragma unroll
                for (uint32_t b = 0; b < kRayBands; b++) atomicExch(a.counter + kBandStride * (b + 1), 0u);
            }
        }
    }
}

template <bool ANY, bool COUNT, int S, int BLOCK, bool SPEC>
__global__ __launch_bounds__(BLOCK) void k_trace(TraceArgs a) {
    const uint32_t nrays = a.n_dev ? __builtin_nontemporal_load(a.n_dev) : a.n;
    trace_body<ANY, COUNT, S, BLOCK, false, SPEC>(a, nullptr, 0, nrays);
}

// Closest hit, or occlusion (ANY), over the wide records (COUNT: inner =
// wide records visited).
template <bool ANY, bool COUNT, int S, int BLOCK>
__global__ __launch_bounds__(BLOCK) void k_trace_wide(TraceArgs a) {
    const uint32_t nrays = a.n_dev ? __builtin_nontemporal_load(a.n_dev) : a.n;
    trace_body<ANY, COUNT, S, BLOCK, false, false, false, true>(a, nullptr, 0, nrays);
}

template <int S, int BLOCK, bool WIDE>
__global__ __launch_bounds__(BLOCK) void k_trace_pair(TraceArgs a) {
    const uint32_t n0 = __builtin_nontemporal_load(a.n_dev);
    const uint32_t n1 = __builtin_nontemporal_load(a.n_dev2);
    trace_body<true, false, S, BLOCK, false, false, true, WIDE>(a, nullptr, 0, n0 + n1, n0);
}

template <int S, int BLOCK, bool SPEC, bool WIDE>
__global__ __launch_bounds__(BLOCK) void k_trace_batch(TraceArgs a, BatchTable t) {
    __shared__ BatchDesc tab[kMaxBatch];
    __shared__ uint32_t s_total;
    for (uint32_t i = threadIdx.x; i < t.nb; i += BLOCK) tab[i] = t.d[i];
    __syncthreads();
    if (threadIdx.x == 0) {  // batch starts from the counts (device counts were final before the launch)
        uint32_t sum = 0;
        for (uint32_t i = 0; i < t.nb; i++) {
            uint32_t n = tab[i].n;
            if (tab[i].n_dev) {
                const uint32_t c = __builtin_nontemporal_load(tab[i].n_dev);
                n = c < n ? c : n;
            }
            tab[i].start = sum;
            sum += n;
        }
        s_total = sum;
    }
    __syncthreads();
    trace_body<false, false, S, BLOCK, true, SPEC, false, WIDE>(a, tab, t.nb, s_total);
}

// ---------------------------------------------------------------- launch --

typedef void (*TraceKernel)(TraceArgs);
typedef void (*BatchKernel)(TraceArgs, BatchTable);

#define RB_FOR_CONFIGS(X)                                                        \
    X(8, 64) X(8, 128) X(8, 256) X(8, 512) X(8, 1024) X(12, 64) X(12, 128)       \
    X(12, 256) X(12, 512) X(12, 1024) X(16, 64) X(16, 128) X(16, 256) X(16, 512) \
    X(20, 64) X(20, 128) X(20, 256) X(24, 64) X(24, 128) X(24, 256)

template <bool ANY, bool COUNT>
static TraceKernel pick(const rb_trace_config& c) {
#define RB_K(s, b)                             \
    if (c.lds_stack == s && c.block == b)      \
        return c.speculative ? k_trace<ANY, COUNT, s, b, true> : k_trace<ANY, COUNT, s, b, false>;
    RB_FOR_CONFIGS(RB_K)
#undef RB_K
    return nullptr;
}

template <bool ANY, bool COUNT>
static TraceKernel pick_wide(const rb_trace_config& c) {
#define RB_K(s, b) \
    if (c.lds_stack == s && c.block == b) return k_trace_wide<ANY, COUNT, s, b>;
    RB_FOR_CONFIGS(RB_K)
#undef RB_K
    return nullptr;
}

template <bool WIDE>
static TraceKernel pick_pair(const rb_trace_config& c) {
#define RB_K(s, b) \
    if (c.lds_stack == s && c.block == b) return k_trace_pair<s, b, WIDE>;
    RB_FOR_CONFIGS(RB_K)
#undef RB_K
    return nullptr;
}

static BatchKernel pick_batch(const rb_trace_config& c, bool wide) {
#define RB_K(s, b)                                                                    \
    if (c.lds_stack == s && c.block == b)                                             \
        return c.speculative ? k_trace_batch<s, b, true, false>                       \
                             : wide ? k_trace_batch<s, b, false, true> : k_trace_batch<s, b, false, false>;
    RB_FOR_CONFIGS(RB_K)
#undef RB_K
    return nullptr;
}

// Whether a set may walk the wide records: closest hit, or an any-hit set
// read only as hit / miss (kOcclusion), under a non-speculative config with
// the wide walk on, over a BVH that has wide records.
static bool wide_walk(const rb_trace_config& cfg, rb_bvh_t bvh, uint32_t flags) {
    return cfg.wide && !cfg.speculative && bvh->wnodes && (!(flags & RB_ANY) || (flags & kOcclusion));
}

static uint32_t persistent_blocks(const rb_trace_config& cfg, uint64_t n) {
    const uint32_t waves = dev().num_cus * cfg.waves_per_cu;
    uint32_t blocks = waves * 64 / cfg.block;
    const uint64_t need = (n + cfg.block - 1) / cfg.block;
    if (blocks > need) blocks = (uint32_t)need;
    return blocks ? blocks : 1;
}

rb_status launch_trace(rb_bvh_t bvh, const rb_ray* rays, uint32_t n, uint32_t flags, const uint8_t* mask,
                       rb_hit* hits, uint32_t* counts, hipStream_t stream, const rb_trace_config& cfg,
                       const uint32_t* n_dev, uint32_t* counter) {
    if (n == 0) return RB_OK;
    if (!bvh || !rays || !hits) return fail(RB_ERR_INVALID, "NULL bvh, rays or hits");
    if (n > 0x7FFFFFFFu) return fail(RB_ERR_INVALID, "too many rays in one call (%u)", n);
    if ((flags & RB_MASKED) && !mask) return fail(RB_ERR_INVALID, "RB_MASKED without a mask");
    const bool any = (flags & RB_ANY) != 0;
    // Closest hit (the same (id, t)) and occlusion walk the wide records;
    // visit counts follow the binary walk unless RB_WIDE_COUNTS asks for the
    // wide one.
    const bool wide_counts = (flags & RB_WIDE_COUNTS) != 0;
    if (wide_counts && (!counts || any || cfg.speculative))
        return fail(RB_ERR_INVALID, "RB_WIDE_COUNTS: visit counts of a closest-hit, non-speculative trace only");
    if (wide_counts && !bvh->wnodes)
        return fail(RB_ERR_UNSUPPORTED, "RB_WIDE_COUNTS: this BVH has no wide records (the upload kept the binary walk)");
    const bool wide = wide_counts || (!counts && wide_walk(cfg, bvh, flags));
    TraceKernel k = wide ? (counts ? pick_wide<false, true>(cfg) : any ? pick_wide<true, false>(cfg)
                                                                       : pick_wide<false, false>(cfg))
                  : counts ? (any ? pick<true, true>(cfg) : pick<false, true>(cfg))
                           : (any ? pick<true, false>(cfg) : pick<false, false>(cfg));
    if (!k) return fail(RB_ERR_INVALID, "unsupported trace config (lds_stack %u, block %u)", cfg.lds_stack, cfg.block);
    TraceArgs a;
    std::memset(&a, 0, sizeof(a));
    a.nodes = bvh->nodes;
    a.tris = bvh->tris;
    a.index = bvh->index;
    a.rays = reinterpret_cast<const float4*>(rays);
    a.mask = (flags & RB_MASKED) ? mask : nullptr;
    a.hits = reinterpret_cast<int2*>(hits);
    a.counts = counts;
    a.errors = dev().errors;
    a.n = n;
    a.n_dev = n_dev;
    a.chunk = cfg.chunk;
    a.threshold = cfg.fetch_threshold;
    a.leaf_gate = cfg.leaf_gate;
    a.top_nodes = cfg.lds_nodes < bvh->n_nodes ? cfg.lds_nodes : bvh->n_nodes;
    if (wide) {  // wide records are 64 bytes too
        a.nodes = bvh->wnodes;
        a.top_nodes = cfg.lds_nodes < bvh->n_wnodes ? cfg.lds_nodes : bvh->n_wnodes;
    }
    const size_t top_bytes = (size_t)a.top_nodes * 64;
    const size_t lds = (size_t)cfg.lds_stack * cfg.block * 4 + top_bytes;
    if (lds > dev().lds_per_block)
        return fail(RB_ERR_INVALID, "trace config needs %zu bytes of LDS per block (stack %u x %u lanes + %u node records); "
                    "the device allows %zu", lds, cfg.lds_stack, cfg.block, a.top_nodes, dev().lds_per_block);
    SlotLease L;
    if (counter) {
        L.counter = counter;
    } else {
        L = take_counter_slot(stream);
        if (L.st != RB_OK) return L.st;
        if (!L.counter) return fail(RB_ERR_INVALID, "render graph: out of work-counter slots");
    }
    a.counter = L.counter;
    hipLaunchKernelGGL(k, dim3(persistent_blocks(cfg, n)), dim3(cfg.block), top_bytes, stream, a);
    const hipError_t e = hipGetLastError();
    release_counter_slot(L, stream);
    if (e != hipSuccess) return hip_fail(e, "trace launch");
    return RB_OK;
}

rb_status launch_trace_pair(rb_bvh_t bvh, const rb_ray* rays0, uint32_t cap0, const uint32_t* n0_dev, rb_hit* hits0,
                            uint32_t flags0, const rb_ray* rays1, uint32_t cap1, const uint32_t* n1_dev, rb_hit* hits1,
                            hipStream_t stream, const rb_trace_config& cfg) {
    if (!bvh) return fail(RB_ERR_INVALID, "paired trace: NULL bvh");
    // both sets walk the wide records when set 0 may (set 1 is closest hit)
    const bool wide = wide_walk(cfg, bvh, flags0);
    TraceKernel k = cfg.speculative ? nullptr : wide ? pick_pair<true>(cfg) : pick_pair<false>(cfg);
    // the speculative order, or more rays than one counter range takes: two launches
    if (!k || (uint64_t)cap0 + cap1 > 0x7FFFFFFFull) {
        rb_status st = launch_trace(bvh, rays0, cap0, flags0 & (RB_ANY | kOcclusion), nullptr, hits0, nullptr, stream,
                                    cfg, n0_dev);
        if (st != RB_OK) return st;
        return launch_trace(bvh, rays1, cap1, RB_CLOSEST, nullptr, hits1, nullptr, stream, cfg, n1_dev);
    }
    if (cap0 + cap1 == 0) return RB_OK;
    if (!rays0 || !rays1 || !hits0 || !hits1 || !n0_dev || !n1_dev)
        return fail(RB_ERR_INVALID, "paired trace: NULL argument");
    TraceArgs a;
    std::memset(&a, 0, sizeof(a));
    a.nodes = wide ? bvh->wnodes : bvh->nodes;
    a.tris = bvh->tris;
    a.index = bvh->index;
    a.rays = reinterpret_cast<const float4*>(rays0);
    a.hits = reinterpret_cast<int2*>(hits0);
    a.n_dev = n0_dev;
    a.rays2 = reinterpret_cast<const float4*>(rays1);
    a.hits2 = reinterpret_cast<int2*>(hits1);
    a.n_dev2 = n1_dev;
    a.any0 = (flags0 & RB_ANY) ? 1u : 0u;
    a.errors = dev().errors;
    a.n = cap0 + cap1;
    a.chunk = cfg.chunk;
    a.threshold = cfg.fetch_threshold;
    a.leaf_gate = cfg.leaf_gate;
    const uint32_t nrec = wide ? bvh->n_wnodes : bvh->n_nodes;
    a.top_nodes = cfg.lds_nodes < nrec ? cfg.lds_nodes : nrec;
    const size_t lds = (size_t)cfg.lds_stack * cfg.block * 4 + (size_t)a.top_nodes * 64;
    if (lds > dev().lds_per_block)
        return fail(RB_ERR_INVALID, "trace config needs %zu bytes of LDS per block; the device allows %zu", lds,
                    dev().lds_per_block);
    const SlotLease L = take_counter_slot(stream);
    if (L.st != RB_OK) return L.st;
    if (!L.counter) return fail(RB_ERR_INVALID, "render graph: out of work-counter slots");
    a.counter = L.counter;
    hipLaunchKernelGGL(k, dim3(persistent_blocks(cfg, a.n)), dim3(cfg.block), (size_t)a.top_nodes * 64, stream, a);
    const hipError_t e = hipGetLastError();
    release_counter_slot(L, stream);
    if (e != hipSuccess) return hip_fail(e, "paired trace launch");
    return RB_OK;
}

static rb_status launch_trace_sets_pass(const BatchSet* b, uint32_t nb, hipStream_t stream, const rb_trace_config& cfg,
                                        bool wide);

// The sets that may walk the wide records go through launches of their own
// (pass 1), the rest through the binary kernel (pass 0).
rb_status launch_trace_sets(const BatchSet* b, uint32_t nb, hipStream_t stream, const rb_trace_config& cfg) {
    for (uint32_t i = 0; i < nb; i++) {
        if (b[i].n == 0) continue;
        if (!b[i].bvh || !b[i].rays || !b[i].hits) return fail(RB_ERR_INVALID, "batch %u: NULL bvh, rays or hits", i);
        if (b[i].n > 0x7FFFFFFFu) return fail(RB_ERR_INVALID, "batch %u: more than 2^31 - 1 rays", i);
    }
    for (int pass = 0; pass < 2; pass++) {
        const rb_status st = launch_trace_sets_pass(b, nb, stream, cfg, pass == 1);
        if (st != RB_OK) return st;
    }
    return RB_OK;
}

static rb_status launch_trace_sets_pass(const BatchSet* b, uint32_t nb, hipStream_t stream, const rb_trace_config& cfg,
                                        bool wide) {
    BatchKernel k = pick_batch(cfg, wide);
    if (!k) return fail(RB_ERR_INVALID, "unsupported trace config (lds_stack %u, block %u)", cfg.lds_stack, cfg.block);
    auto in_pass = [&](const BatchSet& x) {
        return wide_walk(cfg, x.bvh, (x.any ? RB_ANY : 0u) | (x.occlusion ? kOcclusion : 0u)) == wide;
    };
    for (uint32_t i0 = 0; i0 < nb;) {
        BatchTable t;
        std::memset(&t, 0, sizeof(t));
        uint64_t total = 0;  // bound of the launch's rays (sizes the grid)
        uint32_t i = i0;
        for (; i < nb && t.nb < kMaxBatch; i++) {
            if (b[i].n == 0 || !in_pass(b[i])) continue;
            if (total + b[i].n > 0x7FFFFFFFull) break;  // the next launch takes it
            BatchDesc& d = t.d[t.nb++];
            d.nodes = wide ? b[i].bvh->wnodes : b[i].bvh->nodes;
            d.tris = b[i].bvh->tris;
            d.index = b[i].bvh->index;
            d.rays = b[i].rays;
            d.hits = b[i].hits;
            d.mask = b[i].mask;
            d.n_dev = b[i].n_dev;
            d.n = b[i].n;
            d.any = b[i].any ? 1u : 0u;
            total += b[i].n;
        }
        i0 = i;
        if (t.nb == 0) continue;
        TraceArgs a;
        std::memset(&a, 0, sizeof(a));
        a.errors = dev().errors;
        a.n = (uint32_t)total;
        a.chunk = cfg.chunk;
        a.threshold = cfg.fetch_threshold;
        a.leaf_gate = cfg.leaf_gate;
        const SlotLease L = take_counter_slot(stream);
        if (L.st != RB_OK) return L.st;
        if (!L.counter) return fail(RB_ERR_INVALID, "render graph: out of work-counter slots");
        a.counter = L.counter;
        hipLaunchKernelGGL(k, dim3(persistent_blocks(cfg, total)), dim3(cfg.block), 0, stream, a, t);
        const hipError_t e = hipGetLastError();
        release_counter_slot(L, stream);
        if (e != hipSuccess) return hip_fail(e, "batched trace launch");
    }
    return RB_OK;
}

rb_status launch_trace_batch(const rb_batch_desc* b, uint32_t nb, hipStream_t stream, const rb_trace_config& cfg) {
    uint64_t total = 0;
    std::vector<BatchSet> sets;
    try {
        sets.reserve(nb);
    } catch (...) {
        return fail(RB_ERR_OOM, "batch table (host allocation)");
    }
    for (uint32_t i = 0; i < nb; i++) {
        if (b[i].n == 0) continue;
        if (!b[i].bvh || !b[i].rays || !b[i].hits) return fail(RB_ERR_INVALID, "batch %u: NULL bvh, rays or hits", i);
        if (b[i].flags & RB_MASKED) return fail(RB_ERR_INVALID, "batch %u: masks are not supported in batches", i);
        if (b[i].flags & ~(uint32_t)RB_ANY) return fail(RB_ERR_INVALID, "batch %u: unknown trace flags 0x%x", i, b[i].flags);
        // the rays of one table of kMaxBatch batches go through one launch
        total = (sets.size() % kMaxBatch == 0 ? 0 : total) + b[i].n;
        if (total > 0x7FFFFFFFull) return fail(RB_ERR_INVALID, "more than 2^31 rays in one batched call");
        sets.push_back(BatchSet{b[i].bvh, reinterpret_cast<const float4*>(b[i].rays), reinterpret_cast<int2*>(b[i].hits),
                                b[i].n, nullptr, (b[i].flags & RB_ANY) != 0, nullptr});
    }
    return launch_trace_sets(sets.data(), (uint32_t)sets.size(), stream, cfg);
}

static rb_status check_cfg(const rb_trace_config& c) {
    if (c.block != 64 && c.block != 128 && c.block != 256 && c.block != 512 && c.block != 1024)
        return fail(RB_ERR_INVALID, "block must be 64/128/256/512/1024");
    if (!pick<false, false>(c))
        return fail(RB_ERR_INVALID, "no kernel for lds_stack %u at block %u (512: stack 8-16; 1024: stack 8-12)",
                    c.lds_stack, c.block);
    if (c.lds_stack < 8 || c.lds_stack > 24 || c.lds_stack % 4) return fail(RB_ERR_INVALID, "lds_stack must be 8..24, step 4");
    if (c.chunk < 64 || c.chunk % 64 || c.chunk > 65536) return fail(RB_ERR_INVALID, "chunk must be a multiple of 64 up to 65536");
    if (c.fetch_threshold > 64) return fail(RB_ERR_INVALID, "fetch_threshold must be <= 64");
    if (c.waves_per_cu < 1 || c.waves_per_cu > 64) return fail(RB_ERR_INVALID, "waves_per_cu must be 1..64");
    if (c.speculative > 1) return fail(RB_ERR_INVALID, "speculative must be 0 or 1");
    if (c.leaf_gate > 63) return fail(RB_ERR_INVALID, "leaf_gate must be 0..63");
    if (c.lds_nodes > 512 || c.lds_nodes % 64) return fail(RB_ERR_INVALID, "lds_nodes must be 0..512, step 64");
    if (c.wide > 1) return fail(RB_ERR_INVALID, "wide must be 0 or 1");
    return RB_OK;
}

}  // namespace rb

using namespace rb;

extern "C" {

rb_status rb_trace_get_config(rb_trace_config* c) {
    if (!c) return fail(RB_ERR_INVALID, "NULL config");
    rb_status st = ensure_device();
    if (st != RB_OK) return st;
    *c = dev().cfg;
    return RB_OK;
}

rb_status rb_trace_set_config(const rb_trace_config* c) {
    if (!c) return fail(RB_ERR_INVALID, "NULL config");
    rb_status st = ensure_device();
    if (st != RB_OK) return st;
    if ((st = check_cfg(*c)) != RB_OK) return st;
    dev().cfg = *c;
    return RB_OK;
}

rb_status rb_trace(rb_bvh_t bvh, const rb_ray* rays, uint32_t n, uint32_t flags, const uint8_t* mask, rb_hit* hits,
                   void* stream) {
    rb_status st = ensure_device();
    if (st != RB_OK) return st;
    if (flags & ~(uint32_t)(RB_ANY | RB_MASKED)) return fail(RB_ERR_INVALID, "unknown trace flags 0x%x", flags);
    return launch_trace(bvh, rays, n, flags, mask, hits, nullptr, stream ? (hipStream_t)stream : dev().stream,
                        dev().cfg);
}

rb_status rb_trace_batch(const rb_batch_desc* batches, uint32_t nb, void* stream) {
    rb_status st = ensure_device();
    if (st != RB_OK) return st;
    if (nb && !batches) return fail(RB_ERR_INVALID, "NULL batch table");
    return launch_trace_batch(batches, nb, stream ? (hipStream_t)stream : dev().stream, dev().cfg);
}

rb_status rb_trace_autotune(rb_bvh_t bvh, const rb_ray* rays, uint32_t n, uint32_t flags, rb_hit* hits,
                            uint32_t reps, rb_trace_config* best, float* best_ms, void* stream) {
    rb_status st = ensure_device();
    if (st != RB_OK) return st;
    if (!bvh || !rays || !hits || n == 0) return fail(RB_ERR_INVALID, "autotune needs a bvh and a non-empty ray batch");
    if (flags & RB_MASKED) return fail(RB_ERR_INVALID, "autotune takes unmasked batches");
    if (flags & ~(uint32_t)RB_ANY) return fail(RB_ERR_INVALID, "unknown trace flags 0x%x", flags);
    if (reps == 0) reps = 3;
    hipStream_t s = stream ? (hipStream_t)stream : dev().stream;
    hipEvent_t e0, e1;
    if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess)
        return fail(RB_ERR_DEVICE, "hipEventCreate failed");
    auto measure = [&](const rb_trace_config& c, float* ms) -> rb_status {
        rb_status r = launch_trace(bvh, rays, n, flags, nullptr, hits, nullptr, s, c);  // warm-up
        if (r != RB_OK) return r;
        hipError_t e = hipEventRecord(e0, s);
        if (e != hipSuccess) return hip_fail(e, "autotune event");
        for (uint32_t i = 0; i < reps; i++)
            if ((r = launch_trace(bvh, rays, n, flags, nullptr, hits, nullptr, s, c)) != RB_OK) return r;
        if ((e = hipEventRecord(e1, s)) != hipSuccess) return hip_fail(e, "autotune event");
        e = hipEventSynchronize(e1);
        if (e != hipSuccess) return hip_fail(e, "autotune sync");
        if ((e = hipEventElapsedTime(ms, e0, e1)) != hipSuccess) return hip_fail(e, "autotune timing");
        *ms /= (float)reps;
        return RB_OK;
    };
    rb_trace_config cur = dev().cfg;
    float cur_ms = 0.f;
    if ((st = measure(cur, &cur_ms)) != RB_OK) goto done;
    {
        const uint32_t blocks[] = {64, 128, 256, 512, 1024};
        const uint32_t stacks[] = {8, 12, 16, 20, 24};
        const uint32_t chunks[] = {64, 128, 256, 1024};
        const uint32_t thresholds[] = {0, 16, 32, 40, 48, 56, 60};
        const uint32_t waves[] = {8, 12, 16, 24, 32};
        const uint32_t gates[] = {0, 4, 8, 16, 24, 32};
        const uint32_t tops[] = {0, 64, 128, 192, 256, 384};
        for (int pass = 0; pass < 2; pass++) {
            for (int dim = 0; dim < 7; dim++) {
                const uint32_t* vals;
                int nv;
                switch (dim) {
                    case 0: vals = blocks; nv = 5; break;
                    case 1: vals = stacks; nv = 5; break;
                    case 2: vals = chunks; nv = 4; break;
                    case 3: vals = thresholds; nv = 7; break;
                    case 4: vals = waves; nv = 5; break;
                    case 5: vals = gates; nv = 6; break;
                    default: vals = tops; nv = 6; break;
                }
                if (dim == 5 && cur.speculative) continue;  // the gate only acts on the non-speculative order
                for (int k = 0; k < nv; k++) {
                    rb_trace_config c = cur;
                    uint32_t* f = dim == 0 ? &c.block : dim == 1 ? &c.lds_stack : dim == 2 ? &c.chunk
                                : dim == 3 ? &c.fetch_threshold : dim == 4 ? &c.waves_per_cu
                                : dim == 5 ? &c.leaf_gate : &c.lds_nodes;
                    if (*f == vals[k]) continue;
                    *f = vals[k];
                    if (!pick<false, false>(c) ||
                        (size_t)c.lds_stack * c.block * 4 + (size_t)c.lds_nodes * 64 > dev().lds_per_block)
                        continue;  // no kernel for this shape, or more LDS than a block gets
                    float ms = 0.f;
                    if ((st = measure(c, &ms)) != RB_OK) goto done;
                    if (ms < cur_ms * 0.99f) {
                        cur = c;
                        cur_ms = ms;
                    }
                }
            }
        }
    }
    dev().cfg = cur;
    if (best) *best = cur;
    if (best_ms) *best_ms = cur_ms;
done:
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    return st;
}

rb_status rb_selftest_rcp(unsigned long long* fast_inputs, unsigned long long* mismatches) {
    rb_status st = ensure_device();
    if (st != RB_OK) return st;
    if (!fast_inputs || !mismatches) return fail(RB_ERR_INVALID, "NULL argument");
    unsigned long long* d = nullptr;
    if ((st = dalloc((void**)&d, 2 * sizeof(unsigned long long))) != RB_OK) return st;
    unsigned long long h[2] = {0, 0};
    const hipStream_t s = dev().stream;
    hipError_t e = hipMemsetAsync(d, 0, sizeof(h), s);
    if (e == hipSuccess) {
        hipLaunchKernelGGL(k_rcp_check, dim3(4096), dim3(256), 0, s, d);
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipMemcpyAsync(h, d, sizeof(h), hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    dfree(d, 2 * sizeof(unsigned long long));
    if (e != hipSuccess) return hip_fail(e, "reciprocal self-test");
    *fast_inputs = h[0];
    *mismatches = h[1];
    return RB_OK;
}

rb_status rb_trace_errors(uint32_t* flags) {
    rb_status st = ensure_device();
    if (st != RB_OK) return st;
    if (!flags) return fail(RB_ERR_INVALID, "NULL flags");
    uint32_t h = 0;
    hipError_t e = hipDeviceSynchronize();
    if (e == hipSuccess) e = hipMemcpy(&h, dev().errors, 4, hipMemcpyDeviceToHost);
    if (e == hipSuccess && h) e = hipMemset(dev().errors, 0, 4);
    if (e != hipSuccess) return hip_fail(e, "trace error flags");
    *flags = h;
    if (h & RB_TRACE_ERR_STACK)
        return fail(RB_ERR_INVALID, "a traversal stack overflowed: hits of the launches since the last check may be wrong");
    if (h) return fail(RB_ERR_INVALID, "trace error flags 0x%x", h);
    return RB_OK;
}

rb_status rb_trace_counts(rb_bvh_t bvh, const rb_ray* rays, uint32_t n, uint32_t flags, rb_hit* hits,
                          uint32_t* counts, void* stream) {
    rb_status st = ensure_device();
    if (st != RB_OK) return st;
    if (!counts) return fail(RB_ERR_INVALID, "NULL counts");
    if (flags & ~(uint32_t)(RB_ANY | RB_MASKED | RB_WIDE_COUNTS))
        return fail(RB_ERR_INVALID, "unknown trace flags 0x%x", flags);
    return launch_trace(bvh, rays, n, flags & ~RB_MASKED, nullptr, hits, counts,
                        stream ? (hipStream_t)stream : dev().stream, dev().cfg);
}

}  // extern "C"

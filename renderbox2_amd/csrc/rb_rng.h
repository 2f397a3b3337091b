// rb_rng.h -- counter-based random numbers and sampling shared by the host
// generators and the device kernels.
//
// Philox4x32-10 (Salmon, Moraes, Dror, Shaw, SC'11).  Every random number is
// a pure function of (key, counter), so a render is reproducible no matter
// which lane, launch or request mix computes it.  The reference keys
// Thrust's default_random_engine by simplehash(thread index)
// (src/renderers/samplerrenderer.cu:59-70) instead.
#pragma once
#include <stdint.h>

#if defined(__HIPCC__)
#define RB_HD __host__ __device__ __forceinline__
#else
#define RB_HD inline
#endif

namespace rb {

struct U4 {
    uint32_t x, y, z, w;
};

RB_HD uint32_t mulhi32(uint32_t a, uint32_t b) { return (uint32_t)(((uint64_t)a * b) >> 32); }

RB_HD U4 philox4x32_10(U4 c, uint32_t k0, uint32_t k1) {
#ifdef __HIPCC__
#pragma unroll
#endif
    for (int r = 0; r < 10; r++) {
        const uint32_t hi0 = mulhi32(0xD2511F53u, c.x), lo0 = 0xD2511F53u * c.x;
        const uint32_t hi1 = mulhi32(0xCD9E8D57u, c.z), lo1 = 0xCD9E8D57u * c.z;
        U4 n;
        n.x = hi1 ^ c.y ^ k0;
        n.y = lo1;
        n.z = hi0 ^ c.w ^ k1;
        n.w = lo0;
        c = n;
        k0 += 0x9E3779B9u;
        k1 += 0xBB67AE85u;
    }
    return c;
}

// 24 random bits -> float in [0, 1): exact (x >> 8) * 2^-24.
RB_HD float u01(uint32_t x) { return (float)(x >> 8) * 5.9604644775390625e-08f; }

// Streams: key = (seed lo, seed hi ^ stream tag); counter = (a, b, c, d).
struct Rng {
    uint32_t k0, k1;
    RB_HD static Rng make(uint64_t seed, uint32_t tag) {
        Rng r;
        r.k0 = (uint32_t)seed;
        r.k1 = (uint32_t)(seed >> 32) ^ (tag * 0x85EBCA6Bu);
        return r;
    }
    RB_HD U4 draw(uint32_t a, uint32_t b, uint32_t c, uint32_t d) const {
        U4 ctr;
        ctr.x = a;
        ctr.y = b;
        ctr.z = c;
        ctr.w = d;
        return philox4x32_10(ctr, k0, k1);
    }
};

#if defined(__HIPCC__)
// A kernel's stream key: known on the host, or -- inside a captured render
// graph, whose seed changes from launch to launch -- read from a device
// table of keys indexed by stream tag.
struct RngArg {
    Rng r;
    const Rng* dev;
    __device__ __forceinline__ Rng get() const { return dev ? *dev : r; }
};
#endif

// Stream tags.
enum : uint32_t {
    RNG_CAMERA = 1,
    RNG_AO = 2,
    RNG_DIFFUSE = 3,
    RNG_REDRAW = 4,
    RNG_PATH = 5,
    RNG_LIGHT = 6,
    RNG_SCENE = 7,
    RNG_CONNECT = 8
};

}  // namespace rb

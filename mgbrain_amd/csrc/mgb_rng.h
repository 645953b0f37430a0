// mgb_rng.h -- counter-based random streams shared by host and device code.
//
// Philox4x32-10 (Salmon, Moraes, Dror, Shaw, "Parallel random numbers: as easy
// as 1, 2, 3", SC'11).  Every draw is a pure function of (key, counter), so the
// GPU kernels, the host-side generator of the C++ facade and the device-side
// generator produce identical values without sharing state.  It replaces the
// reference's time-seeded cuRAND XORWOW streams (mgcore.cu:449-455, 665-669);
// no cuRAND algorithm is reproduced.
#pragma once
#include <stdint.h>

#if defined(__HIPCC__)
#define MGB_HD __host__ __device__ __forceinline__
#else
#define MGB_HD inline
#endif

namespace mgb {

// Stream tags (fourth counter word).
constexpr uint32_t kTagPoisson = 0x504F4953u;  // "POIS": Poisson input, per neuron and step
constexpr uint32_t kTagBern = 0x4245524Eu;     // "BERN": Bernoulli connector, per pair
constexpr uint32_t kTagBernW = 0x42455257u;
constexpr uint32_t kTagBernD = 0x42455244u;
constexpr uint32_t kTagFixK = 0x4649584Bu;     // "FIXK": fixed in-degree, per (target, k)
constexpr uint32_t kTagFixKW = 0x46495857u;
constexpr uint32_t kTagFixKD = 0x46495844u;
constexpr uint32_t kTagVInit = 0x56494E49u;    // "VINI": initial membrane potential

struct U32x4 {
    uint32_t v[4];
};

MGB_HD U32x4 philox4x32_10(uint32_t c0, uint32_t c1, uint32_t c2, uint32_t c3, uint32_t k0,
                           uint32_t k1) {
#pragma unroll
    for (int r = 0; r < 10; ++r) {
        if (r > 0) {
            k0 += 0x9E3779B9u;
            k1 += 0xBB67AE85u;
        }
        const uint64_t p0 = (uint64_t)0xD2511F53u * c0;
        const uint64_t p1 = (uint64_t)0xCD9E8D57u * c2;
        const uint32_t n0 = (uint32_t)(p1 >> 32) ^ c1 ^ k0;
        const uint32_t n2 = (uint32_t)(p0 >> 32) ^ c3 ^ k1;
        c1 = (uint32_t)p1;
        c3 = (uint32_t)p0;
        c0 = n0;
        c2 = n2;
    }
    U32x4 o;
    o.v[0] = c0;
    o.v[1] = c1;
    o.v[2] = c2;
    o.v[3] = c3;
    return o;
}

MGB_HD uint32_t pick4(const U32x4& o, uint32_t i) {
    // select without dynamic register indexing
    uint32_t a = (i & 1u) ? o.v[1] : o.v[0];
    uint32_t b = (i & 1u) ? o.v[3] : o.v[2];
    return (i & 2u) ? b : a;
}

// One 32-bit draw of a flat 64-bit index f inside stream (proj, tag).
MGB_HD uint32_t philox_pick(uint64_t seed, uint64_t f, uint32_t proj, uint32_t tag) {
    U32x4 o = philox4x32_10((uint32_t)(f >> 2), (uint32_t)(f >> 34), proj, tag, (uint32_t)seed,
                            (uint32_t)(seed >> 32));
    return pick4(o, (uint32_t)(f & 3u));
}

// u in (0, 1] with 24 random bits (curand_uniform's range, mgcore.cu:123).
MGB_HD float u01_open0(uint32_t x) { return (float)((x >> 8) + 1u) * 0x1p-24f; }
// u in [0, 1) with 24 random bits.
MGB_HD float u01_open1(uint32_t x) { return (float)(x >> 8) * 0x1p-24f; }

// Value grids of NetGen::rand_weight / rand_delay (model/netgen.cpp:19-38) with the
// grid index drawn from a Philox stream instead of glibc rand().
MGB_HD float grid_weight(uint32_t r, float lo, float hi, float dw) {
    if (hi == lo) return lo;
    const int range = (int)((hi - lo) / dw) + 1;
    const int k = (int)(r % (uint32_t)range);
    const float step = (float)k * dw;
    return lo < 0 ? lo + step : lo - step;
}

MGB_HD float grid_delay(uint32_t r, float lo, float hi, float dt) {
    if (hi == lo) return lo;
    const int range = (int)((hi - lo) / dt) + 1;
    const int k = (int)(r % (uint32_t)range);
    return lo + (float)k * dt;
}

// Initial membrane potential uniform on [lo, hi) for neuron `local_id` of a stream.
MGB_HD float vinit_draw(uint64_t seed, uint32_t local_id, float lo, float hi) {
    if (hi == lo) return lo;
    U32x4 o = philox4x32_10(local_id, 0u, 0u, kTagVInit, (uint32_t)seed, (uint32_t)(seed >> 32));
    const float span = hi - lo;
    const float t = u01_open1(o.v[0]) * span;
    return lo + t;
}

}  // namespace mgb

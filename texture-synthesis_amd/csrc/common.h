// Shared host/device definitions of the tsgpu patch-search engine.
//
// The PCG32 stream and the bounded-integer draws restate what
// texture-synthesis consumes through rand_pcg 0.3.1 (Lcg64Xsh32),
// rand_core 0.6.3 (seed_from_u64) and rand 0.8.5 (gen_range), at the call
// sites lib/src/ms.rs:386, 454-458, 549-564, 804.
#pragma once
#include <stdint.h>

#if defined(__HIPCC__)
#include <hip/hip_runtime.h>
#define TSG_HD __host__ __device__ inline
#else
#define TSG_HD inline
#endif

namespace tsg {

constexpr uint64_t kPcgMul = 6364136223846793005ULL;
constexpr uint64_t kSeedInc = 11634580027462260723ULL;

TSG_HD uint32_t rotr32(uint32_t x, uint32_t r) { return (x >> r) | (x << ((32u - r) & 31u)); }

// 64x64 -> high 64 bits from 32-bit partial products.
TSG_HD uint64_t mulhi64(uint64_t a, uint64_t b) {
    uint64_t a0 = (uint32_t)a, a1 = a >> 32, b0 = (uint32_t)b, b1 = b >> 32;
    uint64_t p00 = a0 * b0, p01 = a0 * b1, p10 = a1 * b0, p11 = a1 * b1;
    uint64_t mid = (p00 >> 32) + (uint32_t)p01 + (uint32_t)p10;
    return p11 + (p01 >> 32) + (p10 >> 32) + (mid >> 32);
}

// PCG-XSH-RR 64/32: output taken from the state before the advance.
struct Pcg32 {
    uint64_t state, inc;

    TSG_HD void step() { state = state * kPcgMul + inc; }
    TSG_HD uint32_t next_u32() {
        uint64_t s = state;
        step();
        return rotr32((uint32_t)(((s >> 18) ^ s) >> 27), (uint32_t)(s >> 59));
    }
    TSG_HD uint64_t next_u64() {  // low word drawn first
        uint64_t lo = next_u32();
        uint64_t hi = next_u32();
        return (hi << 32) | lo;
    }
    TSG_HD static Pcg32 from_state_incr(uint64_t st, uint64_t inc) {
        Pcg32 p{st + inc, inc};
        p.step();
        return p;
    }
    // SeedableRng::seed_from_u64: 16 seed bytes from four outputs of an
    // auxiliary PCG (advance first), little-endian, then from_seed with the
    // increment forced odd.
    TSG_HD static Pcg32 seed_from_u64(uint64_t st) {
        uint32_t w[4];
        for (int c = 0; c < 4; ++c) {
            st = st * kPcgMul + kSeedInc;
            w[c] = rotr32((uint32_t)(((st >> 18) ^ st) >> 27), (uint32_t)(st >> 59));
        }
        uint64_t s0 = (uint64_t)w[0] | ((uint64_t)w[1] << 32);
        uint64_t s1 = (uint64_t)w[2] | ((uint64_t)w[3] << 32);
        return from_state_incr(s0, s1 | 1u);
    }
};

// gen_range(lo..hi) for u32: widening multiply, rejection zone
// (range << leading_zeros(range)) - 1.
TSG_HD uint32_t range_u32(Pcg32& g, uint32_t lo, uint32_t hi) {
    uint32_t range = hi - lo;
    if (range == 0) return g.next_u32();
    uint32_t zone = (range << __builtin_clz(range)) - 1u;
    for (;;) {
        uint32_t v = g.next_u32();
        uint64_t m = (uint64_t)v * range;
        if ((uint32_t)m <= zone) return lo + (uint32_t)(m >> 32);
    }
}

// gen_range(lo..hi) for u8 and u16: 32-bit draws and a widening multiply,
// rejection zone u32::MAX - (2^32 - range) % range (types of at most 16 bits).
TSG_HD uint32_t range_small(Pcg32& g, uint32_t lo, uint32_t hi) {
    const uint32_t range = hi - lo;
    const uint32_t zone = 0xffffffffu - (uint32_t)((0x100000000ull - range) % range);
    for (;;) {
        const uint32_t v = g.next_u32();
        const uint64_t m = (uint64_t)v * range;
        if ((uint32_t)m <= zone) return lo + (uint32_t)(m >> 32);
    }
}

// gen_range(lo..hi) for u64 / usize: same method on next_u64 draws.
TSG_HD uint64_t range_u64(Pcg32& g, uint64_t lo, uint64_t hi) {
    uint64_t range = hi - lo;
    if (range == 0) return g.next_u64();
    uint64_t zone = (range << __builtin_clzll(range)) - 1u;
    for (;;) {
        uint64_t v = g.next_u64();
        if (v * range <= zone) return lo + mulhi64(v, range);
    }
}

// Per-map view at one pyramid level. Pixels are RGBA8 packed little-endian
// into u32 (R in the low byte), row-major.
struct MapDesc {
    const uint32_t* px;     // example level
    const uint32_t* gpx;    // example guide level (null when unguided)
    const uint8_t* smask;   // 1 = may be sampled; null = SampleMethod::All
    uint32_t w, h;          // example size (bounds of px and smask)
    uint32_t gw, gh;        // guide size (bounds of gpx)
};

// Packed per-output-pixel state: colour, source coordinate (x | y << 16),
// map ids (coord map id | id-map map id << 16), patch id.
// Mirrors color_map / coord_map / id_map of Generator (ms.rs:207-211).
constexpr uint32_t kOutsidePx = 0xFF000000u;  // [0, 0, 0, 255]
constexpr uint32_t kFallbackBit = 0x80000000u;

}  // namespace tsg

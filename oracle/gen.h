// TEST INFRASTRUCTURE (oracle side). Seeded synthetic graph generators used to
// check the device-side generator and to build CPU-side fixtures. Nothing in the
// product path includes this header.
//
// All randomness is a counter-based hash of (seed, item index), so the edge
// multiset does not depend on thread count or evaluation order and a device
// implementation of the same integer function yields the same edges.
#pragma once
#include <stdint.h>

// SplitMix64 finaliser (Steele, Lea, Flood 2014; "Fast splittable
// pseudorandom number generators"), used as a stateless 64-bit mixer.
static inline uint64_t orc_mix64(uint64_t z) {
    z += 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

// Stream base for a seed and a purpose tag (R-MAT, lattice, long-range).
static inline uint64_t orc_stream(uint64_t seed, uint64_t tag) {
    return orc_mix64(seed * 0x632BE59BD9B4E019ull + tag * 0xD1B54A32D192ED03ull + 1ull);
}

// R-MAT quadrant thresholds as 32-bit integers: draw r < t1 -> (0,0),
// r < t2 -> (0,1), r < t3 -> (1,0), else (1,1).
typedef struct { uint32_t t1, t2, t3; } orc_rmat_thr;

static inline orc_rmat_thr orc_rmat_thresholds(double a, double b, double c) {
    orc_rmat_thr t;
    const double s = 4294967296.0;
    t.t1 = (uint32_t)(a * s);
    t.t2 = (uint32_t)((a + b) * s);
    t.t3 = (uint32_t)((a + b + c) * s);
    return t;
}

// Endpoints of R-MAT draw e at `scale` bit levels (MSB first).
static inline void orc_rmat_edge(uint64_t base, uint64_t e, int scale, orc_rmat_thr t,
                                 uint32_t* src, uint32_t* dst) {
    uint32_t s = 0, d = 0;
    uint64_t h = 0;
    for (int k = 0; k < scale; ++k) {
        if ((k & 1) == 0) h = orc_mix64(base ^ ((e << 5) | (uint64_t)(k >> 1)));
        uint32_t r = (k & 1) ? (uint32_t)(h >> 32) : (uint32_t)h;
        uint32_t sb = r >= t.t2;                        // quadrants (1,0),(1,1)
        uint32_t db = (r >= t.t1 && r < t.t2) || r >= t.t3;  // quadrants (0,1),(1,1)
        s |= sb << (scale - 1 - k);
        d |= db << (scale - 1 - k);
    }
    *src = s;
    *dst = d;
}

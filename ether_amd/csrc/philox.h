// philox.h - Philox4x32-10 counter-based generator (Salmon, Moraes, Dror, Shaw,
// "Parallel random numbers: as easy as 1, 2, 3", SC'11, section 4.2), written
// from the paper.  Used on the device for every random draw of the sweep
// engine and on the host for set-up.
//
// Counter layout of the engine (the oracle follows the same contract):
//   key = {seed & 0xffffffff, seed >> 32}
//   ctr = {id, replica, step, (purpose << 16) | attempt}
// so a draw is a pure function of (seed, site slot, replica, sweep, purpose):
// no generator state lives in memory, and runs are bit-reproducible.
#pragma once
#include <stdint.h>

#if defined(__HIPCC__)
#define ETHER_HD __host__ __device__ __forceinline__
#else
#define ETHER_HD inline
#endif

namespace ether {

enum RngPurpose : uint32_t {
    RNG_METROPOLIS = 1,
    RNG_INIT = 2,
    RNG_OVRR = 3,
    RNG_PT = 4,
};

struct u32x4 {
    uint32_t x, y, z, w;
};

ETHER_HD void mulhilo32(uint32_t a, uint32_t b, uint32_t& hi, uint32_t& lo) {
    uint64_t p = (uint64_t)a * (uint64_t)b;
    hi = (uint32_t)(p >> 32);
    lo = (uint32_t)p;
}

ETHER_HD u32x4 philox4x32_10(u32x4 c, uint32_t k0, uint32_t k1) {
#if defined(__HIP_DEVICE_COMPILE__)
#pragma unroll
#endif
    for (int r = 0; r < 10; ++r) {
        if (r > 0) {
            k0 += 0x9E3779B9u;
            k1 += 0xBB67AE85u;
        }
        uint32_t hi0, lo0, hi1, lo1;
        mulhilo32(0xD2511F53u, c.x, hi0, lo0);
        mulhilo32(0xCD9E8D57u, c.z, hi1, lo1);
        u32x4 n;
        n.x = hi1 ^ c.y ^ k0;
        n.y = lo1;
        n.z = hi0 ^ c.w ^ k1;
        n.w = lo0;
        c = n;
    }
    return c;
}

ETHER_HD u32x4 rng_draw(uint64_t seed, uint32_t id, uint32_t replica, uint32_t step,
                        uint32_t purpose, uint32_t attempt) {
    u32x4 c = {id, replica, step, (purpose << 16) | (attempt & 0xffffu)};
    return philox4x32_10(c, (uint32_t)seed, (uint32_t)(seed >> 32));
}

}  // namespace ether

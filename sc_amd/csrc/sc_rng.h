// sc_rng.h -- counter-based random numbers for the device and host.
// Philox4x32-10 (Salmon, Moraes, Dror, Shaw, "Parallel random numbers: as
// easy as 1, 2, 3", SC'11): 10 rounds of two 32x32->64 multiplies with the
// published multipliers and Weyl key increments.  Every thread of a wave
// evaluates the same counter to obtain the same numbers without communication.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace scg {

struct U4 {
    uint32_t x, y, z, w;
};

__host__ __device__ __forceinline__ void mulhilo(uint32_t a, uint32_t b, uint32_t& hi, uint32_t& lo) {
    uint64_t p = (uint64_t)a * (uint64_t)b;
    hi = (uint32_t)(p >> 32);
    lo = (uint32_t)p;
}

__host__ __device__ __forceinline__ U4 philox(U4 c, uint32_t k0, uint32_t k1) {
    const uint32_t M0 = 0xD2511F53u, M1 = 0xCD9E8D57u, W0 = 0x9E3779B9u, W1 = 0xBB67AE85u;
#pragma unroll
    for (int r = 0; r < 10; r++) {
        uint32_t h0, l0, h1, l1;
        mulhilo(M0, c.x, h0, l0);
        mulhilo(M1, c.z, h1, l1);
        U4 n;
        n.x = h1 ^ c.y ^ k0;
        n.y = l1;
        n.z = h0 ^ c.w ^ k1;
        n.w = l0;
        c = n;
        k0 += W0;
        k1 += W1;
    }
    return c;
}

// uniform in [0,1) with 24 bits, as float; and 53 bits as double from two words
__host__ __device__ __forceinline__ float u01f(uint32_t x) { return (float)(x >> 8) * (1.0f / 16777216.0f); }
__host__ __device__ __forceinline__ double u01d(uint32_t a, uint32_t b) {
    uint64_t v = ((uint64_t)a << 21) ^ (uint64_t)(b >> 11);
    return (double)(v & ((1ull << 53) - 1)) * (1.0 / 9007199254740992.0);
}

}  // namespace scg

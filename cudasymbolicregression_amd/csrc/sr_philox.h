// Philox-4x32-10 counter-based generator (Salmon, Moraes, Dror, Shaw,
// "Parallel random numbers: as easy as 1, 2, 3", SC'11): a 128-bit counter
// and a 64-bit key go through ten rounds of two 32x32->64 multiplies, with
// the key bumped by the Weyl constants between rounds.  Shared by the host
// engine (libsr_host.so, g++) and the device kernels (libsr_hip.so, hipcc).
//
// Stream layout: the generator of row `row` under seed `seed` has key
// (seed_lo, seed_hi) and counter (n_lo, n_hi, row_lo, row_hi); block n
// gives four 32-bit words, consumed in order.  A word becomes a float in
// [0, 1) from its top 24 bits: (w >> 8) * 2^-24 (exact in FP32).
#ifndef SR_PHILOX_H
#define SR_PHILOX_H

#include <stdint.h>

#if defined(__HIPCC__)
#define SR_HD __host__ __device__
#else
#define SR_HD
#endif

namespace sr {

struct Philox4x32 {
    uint32_t c[4];
    uint32_t k[2];
};

SR_HD inline void philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]) {
    const uint32_t M0 = 0xD2511F53u, M1 = 0xCD9E8D57u;
    const uint32_t W0 = 0x9E3779B9u, W1 = 0xBB67AE85u;
    uint32_t x0 = ctr[0], x1 = ctr[1], x2 = ctr[2], x3 = ctr[3];
    uint32_t k0 = key[0], k1 = key[1];
    for (int r = 0; r < 10; ++r) {
        const uint64_t p0 = (uint64_t)M0 * x0;
        const uint64_t p1 = (uint64_t)M1 * x2;
        const uint32_t y0 = (uint32_t)(p1 >> 32) ^ x1 ^ k0;
        const uint32_t y1 = (uint32_t)p1;
        const uint32_t y2 = (uint32_t)(p0 >> 32) ^ x3 ^ k1;
        const uint32_t y3 = (uint32_t)p0;
        x0 = y0; x1 = y1; x2 = y2; x3 = y3;
        k0 += W0;
        k1 += W1;
    }
    out[0] = x0; out[1] = x1; out[2] = x2; out[3] = x3;
}

// Uniform floats of one row's stream, four per Philox block.
struct PhiloxUniform {
    uint32_t key[2];
    uint32_t ctr[4];
    uint32_t buf[4];
    int used;

    SR_HD PhiloxUniform(uint64_t seed, uint64_t row) : used(4) {
        key[0] = (uint32_t)seed;
        key[1] = (uint32_t)(seed >> 32);
        ctr[0] = 0u;
        ctr[1] = 0u;
        ctr[2] = (uint32_t)row;
        ctr[3] = (uint32_t)(row >> 32);
        buf[0] = buf[1] = buf[2] = buf[3] = 0u;
    }
    SR_HD uint32_t next_word() {
        if (used == 4) {
            philox4x32_10(ctr, key, buf);
            if (++ctr[0] == 0u) ++ctr[1];  // 64-bit block index; the row words stay fixed
            used = 0;
        }
        return buf[used++];
    }
    SR_HD float operator()() { return (float)(next_word() >> 8) * 5.9604644775390625e-8f; }
};

}  // namespace sr

#endif  // SR_PHILOX_H

/*
 * orc_rng.c -- counter-based replacement for the reference's libc rand() calls.
 *
 * TEST INFRASTRUCTURE (oracle). Independent implementation of Philox4x32-10
 * (Salmon, Moraes, Dror, Shaw, "Parallel random numbers: as easy as 1, 2, 3",
 * SC'11), written from the published round function. Pinned by the Random123
 * known-answer vectors in tests/test_oracle_primitives.py.
 *
 * Draw convention (shared by construction with the HIP library, checked on GPU):
 *   key = {seed_lo, seed_hi}
 *   ctr = {pixel, frame, (phase<<24) | ((event&0xFFFF)<<8) | group, stream}
 *   draw d uses group d/4 and output word d%4; value = word >> 1, i.e. a
 *   31-bit non-negative int like glibc rand() in [0, RAND_MAX=2^31-1], so every
 *   `rand()%n` of the reference keeps its meaning.
 * Call-site mapping (reference rand() sites -> draw index), SuBSENSE apply
 * (BackgroundSubtractorSuBSENSE.cpp):
 *   FG  :264 rand()%2 -> d0, :265 rand()%N -> d1
 *   BG  :278 rand()%L -> d0, :279 rand()%N -> d1, :286/288 neighbour -> d2,
 *       :289 n_rand -> d3, :297 rand()%N -> d4     (3ch: :406-407, :422-423, :432-435, :444)
 * LOBSTER apply (BackgroundSubtractorLOBSTER.cpp:499-508, :557-568):
 *   BG  rand()%L -> d0, rand()%N -> d1, rand()%L -> d2, neighbour -> d3, rand()%N -> d4
 * refreshModel (SuBSENSE :86,:93; LOBSTER :416,:422), phase REFRESH, event = refresh counter:
 *   start position rand()%N -> pixel 0xFFFFFFFF, d0; sample j of pixel p -> pixel p, draw j.
 */
#include "orc.h"

#define PHILOX_M0 0xD2511F53u
#define PHILOX_M1 0xCD9E8D57u
#define PHILOX_W0 0x9E3779B9u
#define PHILOX_W1 0xBB67AE85u

void orc_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]) {
    uint32_t c0 = ctr[0], c1 = ctr[1], c2 = ctr[2], c3 = ctr[3];
    uint32_t k0 = key[0], k1 = key[1];
    for (int r = 0; r < 10; ++r) {
        const uint64_t p0 = (uint64_t)PHILOX_M0 * c0;
        const uint64_t p1 = (uint64_t)PHILOX_M1 * c2;
        const uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
        const uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
        const uint32_t n0 = hi1 ^ c1 ^ k0;
        const uint32_t n2 = hi0 ^ c3 ^ k1;
        c0 = n0; c1 = lo1; c2 = n2; c3 = lo0;
        k0 += PHILOX_W0; k1 += PHILOX_W1;
    }
    out[0] = c0; out[1] = c1; out[2] = c2; out[3] = c3;
}

uint32_t orc_draw(uint64_t seed, uint32_t stream, uint32_t frame, uint32_t phase, uint32_t event,
                  uint32_t pixel, uint32_t draw) {
    const uint32_t key[2] = {(uint32_t)seed, (uint32_t)(seed >> 32)};
    const uint32_t ctr[4] = {pixel, frame, (phase << 24) | ((event & 0xFFFFu) << 8) | (draw >> 2), stream};
    uint32_t out[4];
    orc_philox4x32_10(ctr, key, out);
    return out[draw & 3u] >> 1;
}

/*
 * Oracle: Philox4x32-10 counter-based generator (J. Salmon, M. Moraes,
 * R. Dror, D. Shaw, "Parallel random numbers: as easy as 1, 2, 3", SC'11).
 * TEST INFRASTRUCTURE ONLY.  Pinned by the published known-answer vectors
 * in tests/test_oracle_rng.py.
 *
 * Chain draw layout (shared with the GPU chain kernel, which has its own
 * implementation): draw j of jump `step` of chain `roi` under `seed` is
 * word j%4 of philox(ctr = {step, j/4, hi32(seed), 0}, key = {roi, lo32(seed)}),
 * shifted right by one to a 31-bit integer r.  The uniform is then
 * (float)r / 2^31 exactly as urand() maps rand() (LibRJMCMCreturn.hpp:17).
 */
#include "oracle.h"

static void mulhilo(uint32_t a, uint32_t b, uint32_t* hi, uint32_t* lo) {
    uint64_t p = (uint64_t)a * (uint64_t)b;
    *hi = (uint32_t)(p >> 32);
    *lo = (uint32_t)p;
}

void orc_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]) {
    uint32_t c0 = ctr[0], c1 = ctr[1], c2 = ctr[2], c3 = ctr[3];
    uint32_t k0 = key[0], k1 = key[1];
    for (int r = 0; r < 10; r++) {
        uint32_t hi0, lo0, hi1, lo1;
        mulhilo(0xD2511F53u, c0, &hi0, &lo0);
        mulhilo(0xCD9E8D57u, c2, &hi1, &lo1);
        uint32_t n0 = hi1 ^ c1 ^ k0;
        uint32_t n1 = lo1;
        uint32_t n2 = hi0 ^ c3 ^ k1;
        uint32_t n3 = lo0;
        c0 = n0; c1 = n1; c2 = n2; c3 = n3;
        k0 += 0x9E3779B9u;
        k1 += 0xBB67AE85u;
    }
    out[0] = c0; out[1] = c1; out[2] = c2; out[3] = c3;
}

uint32_t orc_chain_raw(uint64_t seed, uint32_t roi, uint32_t step, uint32_t j) {
    /* one Philox block serves draw j of four consecutive steps */
    uint32_t ctr[4] = {step >> 2, j, (uint32_t)(seed >> 32), 0u};
    uint32_t key[2] = {roi, (uint32_t)seed};
    uint32_t out[4];
    orc_philox4x32_10(ctr, key, out);
    return out[step & 3u] >> 1;
}

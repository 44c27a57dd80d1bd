/*
 * srw_rng.h - counter-based normal draws for macro-electron sampling.
 *
 * Philox4x32-10 (Salmon et al., "Parallel random numbers: as easy as 1, 2,
 * 3", SC'11) keyed by the 64-bit seed, with counter (i, block, 0, 0) for
 * electron i; Box-Muller turns uniform pairs into normals.  Electron i uses
 * blocks 0 and 1 (8 uniforms) for its 6 normals, so a draw depends only on
 * (seed, i), not on batch sizes.  Plain C99, shared by the library and its
 * tests.
 */
#ifndef SRW_AMD_RNG_H
#define SRW_AMD_RNG_H
#include <math.h>
#include <stdint.h>

static inline void srw_philox4x32_10(const uint32_t ctr_in[4], uint32_t k0, uint32_t k1, uint32_t out[4])
{
    uint32_t c0 = ctr_in[0], c1 = ctr_in[1], c2 = ctr_in[2], c3 = ctr_in[3];
    for (int r = 0; r < 10; r++) {
        uint64_t p0 = (uint64_t)0xD2511F53u * c0, p1 = (uint64_t)0xCD9E8D57u * c2;
        uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
        uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
        uint32_t n0 = hi1 ^ c1 ^ k0, n1 = lo1, n2 = hi0 ^ c3 ^ k1, n3 = lo0;
        c0 = n0; c1 = n1; c2 = n2; c3 = n3;
        k0 += 0x9E3779B9u; k1 += 0xBB67AE85u;
    }
    out[0] = c0; out[1] = c1; out[2] = c2; out[3] = c3;
}

/* Uniform in (0, 1) from a 32-bit word. */
static inline double srw_u01(uint32_t x) { return ((double)x + 0.5) * 2.3283064365386963e-10; }

/* Six independent N(0,1) draws of electron i. */
static inline void srw_normals6(unsigned long long seed, unsigned long long i, double* g)
{
    uint32_t u[8];
    for (int b = 0; b < 2; b++) {
        uint32_t ctr[4] = {(uint32_t)i, (uint32_t)b, (uint32_t)(i >> 32), 0u};
        srw_philox4x32_10(ctr, (uint32_t)seed, (uint32_t)(seed >> 32), u + 4 * b);
    }
    for (int k = 0; k < 3; k++) {
        double r = sqrt(-2. * log(srw_u01(u[2 * k])));
        double a = 6.283185307179586 * srw_u01(u[2 * k + 1]);
        g[2 * k] = r * cos(a);
        g[2 * k + 1] = r * sin(a);
    }
}

#endif

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

/*
 * Randomized quasi-Monte Carlo points.  Sobol' sequence in 5 dimensions with
 * the direction numbers of S. Joe and F. Y. Kuo ("Constructing Sobol
 * sequences with better two-dimensional projections", SIAM J. Sci. Comput. 30
 * (2008) 2635-2654; dimensions 2-5: s = 1, 2, 3, 3; a = 0, 1, 1, 2; initial
 * m = {1}, {1, 3}, {1, 3, 1}, {1, 1, 1}), point k in natural order
 * (x_d(k) = XOR of the direction numbers of the set bits of k).  Replicate r
 * applies a random digital shift (XOR with a Philox word per dimension), which
 * keeps the net structure and makes every point uniform on the unit cube, so
 * each replicate's average is unbiased and the spread of the replicate
 * averages estimates the error.
 */
#define SRW_SOBOL_DIMS 5

static inline void srw_sobol_directions(uint32_t V[SRW_SOBOL_DIMS][32])
{
    static const int sdeg[SRW_SOBOL_DIMS] = {0, 1, 2, 3, 3};
    static const int apoly[SRW_SOBOL_DIMS] = {0, 0, 1, 1, 2};
    static const uint32_t minit[SRW_SOBOL_DIMS][3] = {{0, 0, 0}, {1, 0, 0}, {1, 3, 0}, {1, 3, 1}, {1, 1, 1}};
    for (int b = 0; b < 32; b++) V[0][b] = 1u << (31 - b);
    for (int d = 1; d < SRW_SOBOL_DIMS; d++) {
        int s = sdeg[d], a = apoly[d];
        for (int i = 0; i < s; i++) V[d][i] = minit[d][i] << (31 - i);
        for (int i = s; i < 32; i++) {
            uint32_t v = V[d][i - s] ^ (V[d][i - s] >> s);
            for (int k = 1; k < s; k++)
                if ((a >> (s - 1 - k)) & 1) v ^= V[d][i - k];
            V[d][i] = v;
        }
    }
}

static inline uint32_t srw_sobol_word(const uint32_t V[SRW_SOBOL_DIMS][32], int d, uint64_t k)
{
    uint32_t x = 0;
    for (int b = 0; k != 0 && b < 32; b++, k >>= 1)
        if (k & 1) x ^= V[d][b];
    return x;
}

/* Digital shift of replicate r, dimension d. */
static inline uint32_t srw_rqmc_shift(unsigned long long seed, unsigned r, int d)
{
    uint32_t ctr[4] = {r, (uint32_t)d, 0x51u, 0u}, u[4];
    srw_philox4x32_10(ctr, (uint32_t)seed, (uint32_t)(seed >> 32), u);
    return u[0];
}

/* Inverse of the standard normal distribution function: the rational start
 * of Abramowitz and Stegun 26.2.23, then Newton steps on
 * Phi(g) = erfc(-g / sqrt 2) / 2. */
static inline double srw_norm_inv(double u)
{
    double p = u < 0.5 ? u : 1. - u;
    double t = sqrt(-2. * log(p));
    double g = t - (2.515517 + t * (0.802853 + t * 0.010328)) / (1. + t * (1.432788 + t * (0.189269 + t * 0.001308)));
    if (u < 0.5) g = -g;
    for (int it = 0; it < 3; it++) {
        double cdf = 0.5 * erfc(-g * 0.7071067811865476);
        double pdf = 0.3989422804014327 * exp(-0.5 * g * g);
        g -= (cdf - u) / pdf;
    }
    return g;
}

/* Six normals of point k of replicate r (the sixth, unused by the beam
 * mapping, is 0). */
static inline void srw_rqmc_normals6(const uint32_t V[SRW_SOBOL_DIMS][32], unsigned long long seed, unsigned r,
                                     unsigned long long k, double* g)
{
    for (int d = 0; d < SRW_SOBOL_DIMS; d++) g[d] = srw_norm_inv(srw_u01(srw_sobol_word(V, d, k) ^ srw_rqmc_shift(seed, r, d)));
    g[5] = 0.;
}

#endif

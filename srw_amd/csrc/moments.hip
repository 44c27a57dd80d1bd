// moments.hip - statistical moments of wavefront slices (the quantities of
// SRW's ComputeRadMoments) for gfx950.
//
// Reference algorithms (paths relative to the reference source root):
//   moment sums                 cpp/src/core/sroptelm.cpp:1608-1941
//   90 % power rectangle        cpp/src/core/srmatsta.cpp:350-630
//   quadratic-phase treatment   cpp/src/core/sroptelm.cpp:4692-4990
//   float sin/cos of the optics cpp/src/core/sroptelm.h:105-115, 856-872
//
// Six launches over the whole slice set:
//   1. partial row sums of the intensity (one pass over the field)
//   2. row limits of the 90 % power rectangle (one wave per slice)
//   3. partial column sums inside the row limits (one pass)
//   4. column limits (one wave per slice)
//   5. moment sums per band of rows (one pass), 4 waves added in LDS
//   6. band partials added per slice
// Two mappings of lanes onto the data:
//   * lanes over x (LE = false): a wave walks rows of one slice; for slices
//     whose rows are contiguous in x (perX = 1);
//   * lanes over slices (LE = true): lane j of a wave owns slice 64 g + j and
//     walks a stretch of x serially; for SRW's photon-energy-fastest field
//     order, where the energies of one (x, z) point are adjacent.
// Every pass is split finely enough to give several thousand waves; all
// partial results are added in a fixed order, so results do not depend on
// scheduling.  The rectangle's running sums are found with a wave prefix scan
// in the reference's direction of accumulation.  The phase-treated field is
// formed on the fly; the stored field is not modified.
#include "srw_moments.h"
#include "srw_optics.h"

#include <math.h>

#include <algorithm>

namespace {

#define SRWM_T 256

__device__ __forceinline__ double intensity(const srwm_mslice& S, long long o)
{   // total polarisation, as the float extraction of the reference
    double v = 0.;
    if (S.ex) { float2 a = S.ex[o]; v += (double)a.x * a.x + (double)a.y * a.y; }
    if (S.ez) { float2 b = S.ez[o]; v += (double)b.x * b.x + (double)b.y * b.y; }
    return (double)(float)v;
}

__device__ __forceinline__ double wave_sum(double v)
{
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}

// First index i (in the order i = 0, 1, ... of the running sum) at which the
// running sum of val(i) exceeds stop; -1 when it never does.  Whole wave.
template <typename F>
__device__ int first_exceed(F val, int n, double stop)
{
    const int lane = threadIdx.x & 63;
    double carry = 0.;
    for (int base = 0; base < n; base += 64) {
        const int i = base + lane;
        double x = i < n ? val(i) : 0.;
        for (int off = 1; off < 64; off <<= 1) {
            double y = __shfl_up(x, off, 64);
            if (lane >= off) x += y;
        }
        x += carry;
        unsigned long long m = __ballot(i < n && x > stop);
        if (m) return base + __ffsll((long long)m) - 1;
        carry = __shfl(x, 63, 64);
    }
    return -1;
}

// Work split shared by the host launcher and the kernels.
struct mom_plan {
    int le, groups;
    int nSegR, segLenR;                    // pass 1: x stretches per row (lanes over slices)
    int nSegC, segLenC;                    // pass 3: z stretches per column
    int rpw, nSeg, segLen, gridY, nBands;  // pass 5
};

// ---- pass 1: partial row sums rowP[(q*n + s)*ldR + iz] (srmatsta.cpp:350-370) ----
template <bool LE>
__global__ void __launch_bounds__(SRWM_T) k_rows(const srwm_mslice* __restrict__ sl, int n, mom_plan P, double* rowP,
                                                 long long ldR)
{
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    if (LE) {
        const int s = blockIdx.x * 64 + lane;
        const int id = blockIdx.y * 4 + w, iz = id / P.nSegR, q = id % P.nSegR;
        if (s >= n) return;
        const srwm_mslice S = sl[s];
        if (iz >= S.nz) return;
        const int ix1 = min(S.nx, (q + 1) * P.segLenR);
        double r = 0.;
        for (int ix = q * P.segLenR; ix < ix1; ix++) r += intensity(S, iz * S.perZ + ix * S.perX);
        rowP[((long long)q * n + s) * ldR + iz] = r;
    } else {
        const int s = blockIdx.x, iz = blockIdx.y * 4 + w;
        const srwm_mslice S = sl[s];
        if (iz >= S.nz) return;
        double r = 0.;
        for (int ix = lane; ix < S.nx; ix += 64) r += intensity(S, iz * S.perZ + ix * S.perX);
        r = wave_sum(r);
        if (lane == 0) rowP[(long long)s * ldR + iz] = r;
    }
}

// ---- pass 2: row limits (srmatsta.cpp:512-546), one wave per slice ----
// rl[s] = {ok, iz0, iz1, 0}; totF[s] = float-rounded integral.
__global__ void __launch_bounds__(SRWM_T) k_row_limits(const srwm_mslice* __restrict__ sl, int n, mom_plan P,
                                                       const double* __restrict__ rowP, long long ldR, int4* rl,
                                                       double* totF)
{
    const int lane = threadIdx.x & 63, s = blockIdx.x * 4 + (threadIdx.x >> 6);
    if (s >= n) return;
    const srwm_mslice S = sl[s];
    const int nz = S.nz, nSeg = P.le ? P.nSegR : 1;
    auto row = [&](int i) {
        double r = 0.;
        for (int q = 0; q < nSeg; q++) r += rowP[((long long)q * n + s) * ldR + i];
        return r;
    };
    int4 o = make_int4(0, 0, nz - 1, 0);
    double tf = 0.;
    if (S.nx > 1) {
        double tot = 0.;
        for (int i = lane; i < nz; i += 64) tot += row(i);
        tot = wave_sum(tot);
        float totf = (float)(tot * S.xStep * S.zStep);
        double stop = (double)totf * (1. - 0.9) * 0.25 / S.zStep;
        int a = first_exceed([&](int i) { return S.xStep * row(i); }, nz, stop);
        int b = first_exceed([&](int i) { return S.xStep * row(nz - 1 - i); }, nz, stop);
        int i0 = a >= 0 ? a : 0, i1 = b >= 0 ? nz - 1 - b : nz - 1;
        if (i1 > i0) o = make_int4(1, i0, i1, 0);
        tf = (double)totf;
    }
    if (lane == 0) { rl[s] = o; totF[s] = tf; }
}

// ---- pass 3: partial column sums inside the row limits (srmatsta.cpp:547-560) ----
// colP[(q*n + s)*ldC + ix] over rows [q*segLenC, (q+1)*segLenC) of [iz0, iz1].
template <bool LE>
__global__ void __launch_bounds__(SRWM_T) k_cols(const srwm_mslice* __restrict__ sl, int n, mom_plan P,
                                                 const int4* __restrict__ rl, double* colP, long long ldC)
{
    int s, ix, q;
    if (LE) {
        const int id = blockIdx.y * 4 + (threadIdx.x >> 6);
        s = blockIdx.x * 64 + (threadIdx.x & 63);
        ix = id / P.nSegC; q = id % P.nSegC;
    } else {
        s = blockIdx.x;
        ix = blockIdx.y * SRWM_T + threadIdx.x;
        q = blockIdx.z;
    }
    if (s >= n) return;
    const int4 L = rl[s];
    if (!L.x) return;
    const srwm_mslice S = sl[s];
    if (ix >= S.nx) return;
    const int iz0 = max(L.y, q * P.segLenC), iz1 = min(L.z, (q + 1) * P.segLenC - 1);
    double c = 0.;
    for (int iz = iz0; iz <= iz1; iz++) c += intensity(S, iz * S.perZ + ix * S.perX);
    colP[((long long)q * n + s) * ldC + ix] = c;
}

// ---- pass 4: column limits and the rectangle's indices (srmatsta.cpp:561-630) ----
__global__ void __launch_bounds__(SRWM_T) k_col_limits(const srwm_mslice* __restrict__ sl, int n, mom_plan P,
                                                       const int4* __restrict__ rl, const double* __restrict__ totF,
                                                       const double* __restrict__ colP, long long ldC, int* lims)
{
    const int lane = threadIdx.x & 63, s = blockIdx.x * 4 + (threadIdx.x >> 6);
    if (s >= n) return;
    const srwm_mslice S = sl[s];
    const int nx = S.nx, nz = S.nz;
    const int4 R = rl[s];
    int L0 = 0, L1 = nx - 1, L2 = 0, L3 = nz - 1;
    if (R.x) {
        const int nSeg = P.nSegC;
        auto col = [&](int i) {
            double c = 0.;
            for (int q = 0; q < nSeg; q++) c += colP[((long long)q * n + s) * ldC + i];
            return S.zStep * c;
        };
        double stop = totF[s] * (1. - 0.9) * 0.25 / S.xStep;
        int a = first_exceed(col, nx, stop);
        int b = first_exceed([&](int i) { return col(nx - 1 - i); }, nx, stop);
        long long j0 = a >= 0 ? a : 0, j1 = b >= 0 ? nx - 1 - b : nx - 1;
        // positions pass through float as in srmatsta.cpp:616-625
        float fx0 = (float)(S.xStart + S.xStep * j0), fx1 = (float)(S.xStart + S.xStep * j1);
        float fz0 = (float)(S.zStart + S.zStep * R.y), fz1 = (float)(S.zStart + S.zStep * R.z);
        L0 = (int)((fx0 - S.xStart) * 1.0000001 / S.xStep); if (L0 < 0) L0 = 0;
        L1 = (int)((fx1 - S.xStart) * 1.0000001 / S.xStep); if (L1 >= nx) L1 = nx - 1;
        L2 = (int)((fz0 - S.zStart) * 1.0000001 / S.zStep); if (L2 < 0) L2 = 0;
        L3 = (int)((fz1 - S.zStart) * 1.0000001 / S.zStep); if (L3 >= nz) L3 = nz - 1;
    }
    if (lane == 0) {
        lims[s * 4 + 0] = L0; lims[s * 4 + 1] = L1; lims[s * 4 + 2] = L2; lims[s * 4 + 3] = L3;
    }
}

// ---- pass 5: moment sums (sroptelm.cpp:1688-1941) ----
struct mom_geo {
    double kx, kz, kx2, kz2;
    double cRxE, cRzE, x0, z0;   // quadratic phase removed before the sums
    bool fx, fz, any;
};

__device__ __forceinline__ mom_geo make_geo(const srwm_mslice& S)
{
    const double TwoPi = 3.141592653590 * 2., FourPi = TwoPi * 2.;
    const double lam4 = 1.239842E-06 / (FourPi * S.ePh), lamM = lam4 * FourPi;
    const double rx = S.Rx != 0. ? S.Rx : 100. * lamM, rz = S.Rz != 0. ? S.Rz : 100. * lamM;
    mom_geo g;
    g.kx = 0.5 * (S.xStep * ((1. / lam4) / rx));
    g.kz = 0.5 * (S.zStep * ((1. / lam4) / rz));
    g.kx2 = g.kx * g.kx; g.kz2 = g.kz * g.kz;
    g.cRxE = -(SRW_QUAD_CONST / S.Rx) * S.ePh; g.cRzE = -(SRW_QUAD_CONST / S.Rz) * S.ePh;
    g.x0 = S.xStart - S.xc; g.z0 = S.zStart - S.zc;
    g.fx = S.fx != 0; g.fz = S.fz != 0; g.any = g.fx || g.fz;
    return g;
}

// Phase-treated value of one component at (ix, iz).
__device__ __forceinline__ float2 treated(const float2* p, long long o, const mom_geo& g, const srwm_mslice& S,
                                          int ix, int iz)
{
    float2 a = p[o];
    if (!g.any) return a;
    double ph = 0.;
    if (g.fz) { double z = g.z0 + iz * S.zStep; ph = g.cRzE * z * z; }
    if (g.fx) { double x = g.x0 + ix * S.xStep; ph += g.cRxE * x * x; }
    float c, s;
    cos_sin_optics(ph, c, s);
    return rotate_f(a, c, s);
}

// Adds the 11 weighted terms of one component at one point.  l / u are the
// left (ix - 1) and upper (iz - 1) neighbours, used when hasL / hasU.
__device__ __forceinline__ void point_sums(double* acc, float2 e, float2 l, float2 u, bool hasL, bool hasU, double x,
                                           double z, bool in, double wAll, double wDer, double rAll, double rDer,
                                           const mom_geo& g)
{
    double f[11];
    double re = e.x, im = e.y;
    f[0] = re * re + im * im;
    f[1] = x * f[0];
    f[3] = z * f[0];
    f[5] = in ? x * f[1] : 0.;
    f[8] = in ? z * f[3] : 0.;
    if (hasL) {
        double dRe = re - (double)l.x, dIm = im - (double)l.y, cr = dIm * re - dRe * im;
        f[2] = cr + g.kx * x * f[0];
        f[6] = in ? x * f[2] : 0.;
        f[7] = in ? (dRe * dRe + dIm * dIm) + cr * g.kx * x + g.kx2 * x * x * f[0] : 0.;
    } else { f[2] = 0.; f[6] = 0.; f[7] = 0.; }
    if (hasU) {
        double dRe = re - (double)u.x, dIm = im - (double)u.y, cr = dIm * re - dRe * im;
        f[4] = cr + g.kz * z * f[0];
        f[9] = in ? z * f[4] : 0.;
        f[10] = in ? (dRe * dRe + dIm * dIm) + cr * g.kz * z + g.kz2 * z * z * f[0] : 0.;
    } else { f[4] = 0.; f[9] = 0.; f[10] = 0.; }
#pragma unroll
    for (int k = 0; k < 11; k++) {
        double wx = (k == 2 || k == 6 || k == 7) ? wDer : wAll;
        double wz = (k == 4 || k == 9 || k == 10) ? rDer : rAll;
        acc[k] += (wx * wz) * f[k];
    }
}

// Lanes over slices: wave item id = blockIdx.y*4 + w covers rows
// [r*rpw, r*rpw + rpw) and x in [q*segLen, q*segLen + segLen) with
// r = id / nSeg, q = id % nSeg, for the 64 slices of group blockIdx.x.  The 4
// waves' sums are added in order into part[(blockIdx.y*n + s)*24 + k].
__global__ void __launch_bounds__(SRWM_T) k_sums_le(const srwm_mslice* __restrict__ sl, int n, mom_plan P,
                                                    const int* __restrict__ lims, double* part)
{
    __shared__ double red[4 * SRWM_NSUM * 64];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int s = blockIdx.x * 64 + lane, id = blockIdx.y * 4 + w;
    double acc[SRWM_NSUM];
#pragma unroll
    for (int k = 0; k < SRWM_NSUM; k++) acc[k] = 0.;
    if (s < n) {
        const srwm_mslice S = sl[s];
        const mom_geo g = make_geo(S);
        const int nx = S.nx, nz = S.nz;
        const int L0 = lims[s * 4], L1 = lims[s * 4 + 1], L2 = lims[s * 4 + 2], L3 = lims[s * 4 + 3];
        const int r = id / P.nSeg, q = id % P.nSeg;
        const int izEnd = min(nz, (r + 1) * P.rpw);
        const int ixBeg = q * P.segLen, ixEnd = min(nx, ixBeg + P.segLen);
        for (int iz = r * P.rpw; iz < izEnd; iz++) {
            const bool inZ = iz >= L2 && iz <= L3;
            const double z = S.zStart + iz * S.zStep;
            const double rAll = (iz == 0 || iz == nz - 1) ? 0.5 : 1.;
            const double rDer = (iz == 1) ? 0.5 * rAll : rAll;
            float2 prev[2] = {make_float2(0.f, 0.f), make_float2(0.f, 0.f)};
            if (ixBeg > 0 && ixBeg < ixEnd) {
                const long long o = iz * S.perZ + (long long)(ixBeg - 1) * S.perX;
                if (S.ex) prev[0] = treated(S.ex, o, g, S, ixBeg - 1, iz);
                if (S.ez) prev[1] = treated(S.ez, o, g, S, ixBeg - 1, iz);
            }
            for (int ix = ixBeg; ix < ixEnd; ix++) {
                const long long o = iz * S.perZ + (long long)ix * S.perX;
                const bool in = inZ && ix >= L0 && ix <= L1;
                const double x = S.xStart + ix * S.xStep;
                const double wAll = (ix == 0 || ix == nx - 1) ? 0.5 : 1.;
                const double wDer = (ix == 1) ? 0.5 * wAll : wAll;
#pragma unroll
                for (int c = 0; c < 2; c++) {
                    const float2* p = c == 0 ? S.ex : S.ez;
                    if (!p) continue;
                    float2 e = treated(p, o, g, S, ix, iz);
                    float2 u = iz > 0 ? treated(p, o - S.perZ, g, S, ix, iz - 1) : make_float2(0.f, 0.f);
                    point_sums(acc + 11 * c, e, prev[c], u, ix > 0, iz > 0, x, z, in, wAll, wDer, rAll, rDer, g);
                    prev[c] = e;
                }
            }
        }
    }
#pragma unroll
    for (int k = 0; k < SRWM_NSUM; k++) red[(w * SRWM_NSUM + k) * 64 + lane] = acc[k];
    __syncthreads();
    // thread (w, lane) adds the 4 waves for sums k = w, w + 4, ... of slice blockIdx.x*64 + lane
    if (s < n) {
        for (int k = w; k < SRWM_NSUM; k += 4) {
            double v = 0.;
            for (int q = 0; q < 4; q++) v += red[(q * SRWM_NSUM + k) * 64 + lane];
            part[((long long)blockIdx.y * n + s) * 24 + k] = v;
        }
    }
}

// Lanes over x: block (s = blockIdx.x, band = blockIdx.y), wave w sums rows
// [(band*4 + w)*rpw, ... + rpw); the 4 waves are added in order into
// part[(band*n + s)*24 + k].
__global__ void __launch_bounds__(SRWM_T) k_sums_lx(const srwm_mslice* __restrict__ sl, int n, mom_plan P,
                                                    const int* __restrict__ lims, double* part)
{
    __shared__ double red[4 * SRWM_NSUM];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int s = blockIdx.x, band = blockIdx.y;
    const srwm_mslice S = sl[s];
    const mom_geo g = make_geo(S);
    const int nx = S.nx, nz = S.nz;
    const int L0 = lims[s * 4], L1 = lims[s * 4 + 1], L2 = lims[s * 4 + 2], L3 = lims[s * 4 + 3];
    double acc[SRWM_NSUM];
#pragma unroll
    for (int k = 0; k < SRWM_NSUM; k++) acc[k] = 0.;
    const int iz0 = (band * 4 + w) * P.rpw, izEnd = min(nz, iz0 + P.rpw);
    for (int iz = iz0; iz < izEnd; iz++) {
        const bool inZ = iz >= L2 && iz <= L3;
        const double z = S.zStart + iz * S.zStep;
        const double rAll = (iz == 0 || iz == nz - 1) ? 0.5 : 1.;
        const double rDer = (iz == 1) ? 0.5 * rAll : rAll;
        float2 carry[2] = {make_float2(0.f, 0.f), make_float2(0.f, 0.f)};
        for (int base = 0; base < nx; base += 64) {
            // wave-uniform trip count: every lane takes part in the shuffles
            const int ix = base + lane;
            const bool valid = ix < nx;
            const long long o = iz * S.perZ + (long long)ix * S.perX;
            const bool in = inZ && ix >= L0 && ix <= L1;
            const double x = S.xStart + ix * S.xStep;
            const double wAll = (ix == 0 || ix == nx - 1) ? 0.5 : 1.;
            const double wDer = (ix == 1) ? 0.5 * wAll : wAll;
#pragma unroll
            for (int c = 0; c < 2; c++) {
                const float2* p = c == 0 ? S.ex : S.ez;
                if (!p) continue;   // uniform across the wave
                float2 e = valid ? treated(p, o, g, S, ix, iz) : make_float2(0.f, 0.f);
                float lx = __shfl_up(e.x, 1, 64), ly = __shfl_up(e.y, 1, 64);
                float2 l = (lane == 0) ? carry[c] : make_float2(lx, ly);
                carry[c] = make_float2(__shfl(e.x, 63, 64), __shfl(e.y, 63, 64));
                if (!valid) continue;
                float2 u = iz > 0 ? treated(p, o - S.perZ, g, S, ix, iz - 1) : make_float2(0.f, 0.f);
                point_sums(acc + 11 * c, e, l, u, ix > 0, iz > 0, x, z, in, wAll, wDer, rAll, rDer, g);
            }
        }
    }
#pragma unroll
    for (int k = 0; k < SRWM_NSUM; k++) acc[k] = wave_sum(acc[k]);
    if (lane == 0)
        for (int k = 0; k < SRWM_NSUM; k++) red[w * SRWM_NSUM + k] = acc[k];
    __syncthreads();
    if (threadIdx.x < SRWM_NSUM) {
        double v = 0.;
        for (int q = 0; q < 4; q++) v += red[q * SRWM_NSUM + threadIdx.x];
        part[((long long)band * n + s) * 24 + threadIdx.x] = v;
    }
}

// ---- pass 6: sums[s*24 + k] = sum over bands; one block per slice ----
__global__ void __launch_bounds__(SRWM_T) k_sum_bands(const double* __restrict__ part, int n, int nBands,
                                                      double* sums)
{
    __shared__ double red[4 * SRWM_NSUM];
    const int s = blockIdx.x, lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    double acc[SRWM_NSUM];
#pragma unroll
    for (int k = 0; k < SRWM_NSUM; k++) acc[k] = 0.;
    for (int b = threadIdx.x; b < nBands; b += SRWM_T) {
        const double* p = part + ((long long)b * n + s) * 24;
#pragma unroll
        for (int k = 0; k < SRWM_NSUM; k++) acc[k] += p[k];
    }
#pragma unroll
    for (int k = 0; k < SRWM_NSUM; k++) acc[k] = wave_sum(acc[k]);
    if (lane == 0)
        for (int k = 0; k < SRWM_NSUM; k++) red[w * SRWM_NSUM + k] = acc[k];
    __syncthreads();
    if (threadIdx.x < 24) {
        double v = 0.;
        if (threadIdx.x < SRWM_NSUM)
            for (int q = 0; q < 4; q++) v += red[q * SRWM_NSUM + threadIdx.x];
        sums[s * 24 + threadIdx.x] = v;
    }
}

inline long long align_up(long long v, long long a) { return (v + a - 1) / a * a; }
inline long long cdiv(long long a, long long b) { return (a + b - 1) / b; }

mom_plan make_plan(long long n, int nxMax, int nzMax, int le)
{
    const long long TP = 8192, TS = 4096;   // target waves of the read passes / of the sum pass
    mom_plan p;
    p.le = le;
    p.groups = (int)cdiv(n, 64);
    if (le) {
        const long long g = p.groups;
        p.nSegR = (int)std::max<long long>(1, std::min(cdiv(TP, g * nzMax), cdiv(nxMax, 8)));
        p.segLenR = (int)cdiv(nxMax, p.nSegR);
        p.nSegC = (int)std::max<long long>(1, std::min(cdiv(TP, g * nxMax), cdiv(nzMax, 8)));
        if (g * nzMax >= TS) {
            p.rpw = (int)cdiv(g * nzMax, TS);
            p.nSeg = 1;
        } else {
            p.rpw = 1;
            p.nSeg = (int)std::max<long long>(1, std::min(cdiv(TS, g * nzMax), cdiv(nxMax, 16)));
        }
        p.segLen = (int)cdiv(nxMax, p.nSeg);
        p.gridY = (int)cdiv(cdiv(nzMax, p.rpw) * p.nSeg, 4);
    } else {
        p.nSegR = 1;
        p.segLenR = nxMax;
        p.nSegC = (int)std::max<long long>(1, std::min(cdiv(TP, n * cdiv(nxMax, 64)), cdiv(nzMax, 8)));
        p.rpw = (int)std::max<long long>(1, cdiv(n * nzMax, TS));
        p.nSeg = 1;
        p.segLen = nxMax;
        p.gridY = (int)cdiv(nzMax, 4LL * p.rpw);
    }
    p.segLenC = (int)cdiv(nzMax, p.nSegC);
    p.nBands = p.gridY;
    return p;
}

struct mom_layout {
    long long rowOff, colOff, rlOff, totOff, partOff, bytes;
};

mom_layout make_layout(long long n, int nxMax, int nzMax, const mom_plan& p)
{
    mom_layout l;
    long long b = 0;
    l.rowOff = b;  b += align_up(8 * (long long)p.nSegR * n * nzMax, 256);
    l.colOff = b;  b += align_up(8 * (long long)p.nSegC * n * nxMax, 256);
    l.rlOff = b;   b += align_up(16 * n, 256);
    l.totOff = b;  b += align_up(8 * n, 256);
    l.partOff = b; b += align_up(8 * 24 * n * (long long)p.nBands, 256);
    l.bytes = b;
    return l;
}

}  // namespace

extern "C" long long srwm_work_bytes(int nSlices, int nxMax, int nzMax, int lanesOverSlices)
{
    mom_plan p = make_plan(nSlices, nxMax, nzMax, lanesOverSlices ? 1 : 0);
    return make_layout(nSlices, nxMax, nzMax, p).bytes;
}

extern "C" int srwm_launch_moments(const srwm_mslice* dSlices, int nSlices, int nxMax, int nzMax, int lanesOverSlices,
                                   void* dWork, double* dSums, int* dLims, hipStream_t st)
{
    if (nSlices <= 0) return 0;
    const long long n = nSlices;
    const mom_plan P = make_plan(n, nxMax, nzMax, lanesOverSlices ? 1 : 0);
    const mom_layout L = make_layout(n, nxMax, nzMax, P);
    char* wp = (char*)dWork;
    double* rowP = (double*)(wp + L.rowOff);
    double* colP = (double*)(wp + L.colOff);
    int4* rl = (int4*)(wp + L.rlOff);
    double* totF = (double*)(wp + L.totOff);
    double* part = (double*)(wp + L.partOff);
    const unsigned perSlice = (unsigned)cdiv(n, 4);   // one wave per slice
    if (P.le) {
        hipLaunchKernelGGL(k_rows<true>, dim3(P.groups, (unsigned)cdiv((long long)nzMax * P.nSegR, 4)), dim3(SRWM_T),
                           0, st, dSlices, nSlices, P, rowP, (long long)nzMax);
    } else {
        hipLaunchKernelGGL(k_rows<false>, dim3(nSlices, (unsigned)cdiv(nzMax, 4)), dim3(SRWM_T), 0, st, dSlices,
                           nSlices, P, rowP, (long long)nzMax);
    }
    hipLaunchKernelGGL(k_row_limits, dim3(perSlice), dim3(SRWM_T), 0, st, dSlices, nSlices, P, rowP,
                       (long long)nzMax, rl, totF);
    if (P.le) {
        hipLaunchKernelGGL(k_cols<true>, dim3(P.groups, (unsigned)cdiv((long long)nxMax * P.nSegC, 4)), dim3(SRWM_T),
                           0, st, dSlices, nSlices, P, rl, colP, (long long)nxMax);
    } else {
        hipLaunchKernelGGL(k_cols<false>, dim3(nSlices, (unsigned)cdiv(nxMax, SRWM_T), P.nSegC), dim3(SRWM_T), 0, st,
                           dSlices, nSlices, P, rl, colP, (long long)nxMax);
    }
    hipLaunchKernelGGL(k_col_limits, dim3(perSlice), dim3(SRWM_T), 0, st, dSlices, nSlices, P, rl, totF, colP,
                       (long long)nxMax, dLims);
    if (P.le) {
        hipLaunchKernelGGL(k_sums_le, dim3(P.groups, P.gridY), dim3(SRWM_T), 0, st, dSlices, nSlices, P, dLims, part);
    } else {
        hipLaunchKernelGGL(k_sums_lx, dim3(nSlices, P.gridY), dim3(SRWM_T), 0, st, dSlices, nSlices, P, dLims, part);
    }
    hipLaunchKernelGGL(k_sum_bands, dim3(nSlices), dim3(SRWM_T), 0, st, part, nSlices, P.nBands, dSums);
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

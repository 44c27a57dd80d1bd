// moments.hip - statistical moments of wavefront slices (the quantities of
// SRW's ComputeRadMoments) for gfx950.
//
// Reference algorithms (paths relative to the reference source root):
//   moment sums                 cpp/src/core/sroptelm.cpp:1608-1941
//   90 % power rectangle        cpp/src/core/srmatsta.cpp:350-630
//   quadratic-phase treatment   cpp/src/core/sroptelm.cpp:4692-4990
//   float sin/cos of the optics cpp/src/core/sroptelm.h:105-115, 856-872
//
// One 256-thread workgroup per slice.  The slice is read three times: row
// sums of the intensity, column sums inside the row limits, and the moment
// sums; rows are contiguous in x (perX = 1) for coalesced reads.  The
// phase-treated field is formed on the fly; the stored field is not
// modified.
#include "srw_moments.h"
#include "srw_optics.h"

#include <math.h>

namespace {

#define SRWM_T 256

__device__ __forceinline__ double intensity(const srwm_mslice& S, long long o)
{   // total polarisation, as the float extraction of the reference
    double v = 0.;
    if (S.ex) { float2 a = S.ex[o]; v += (double)a.x * a.x + (double)a.y * a.y; }
    if (S.ez) { float2 b = S.ez[o]; v += (double)b.x * b.x + (double)b.y * b.y; }
    return (double)(float)v;
}

__device__ __forceinline__ double block_sum(double v, double* red)
{
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    __syncthreads();
    if (lane == 0) red[w] = v;
    __syncthreads();
    double t = 0.;
    for (int q = 0; q < SRWM_T / 64; q++) t += red[q];
    return t;
}

// Phase-treated value of one component at (ix, iz).
struct treat_ctx {
    double cRxE, cRzE, x0, z0, xStep, zStep;
    int fx, fz;
};

__device__ __forceinline__ float2 treated(const float2* p, long long o, const treat_ctx& t, int ix, int iz,
                                          bool any)
{
    float2 a = p[o];
    if (!any) return a;
    double ph = 0.;
    if (t.fz) { double z = t.z0 + iz * t.zStep; ph = t.cRzE * z * z; }
    if (t.fx) { double x = t.x0 + ix * t.xStep; ph += t.cRxE * x * x; }
    float c, s;
    cos_sin_optics(ph, c, s);
    return rotate_f(a, c, s);
}

__global__ void __launch_bounds__(SRWM_T) k_rad_moments(const srwm_mslice* __restrict__ slices, double* work,
                                                        long long workStride, double* sums, int* lims)
{
    __shared__ double red[SRWM_T / 64 * SRWM_NSUM];
    __shared__ int lim[4];
    const srwm_mslice S = slices[blockIdx.x];
    const int t = threadIdx.x, lane = t & 63, w = t >> 6;
    const int nx = S.nx, nz = S.nz;
    double* rowS = work + blockIdx.x * workStride;
    double* colS = rowS + nz;
    // ---- row sums and total of the intensity (srmatsta.cpp:350-370, 512-523)
    double tot = 0.;
    for (int iz = w; iz < nz; iz += SRWM_T / 64) {
        double r = 0.;
        for (int ix = lane; ix < nx; ix += 64) r += intensity(S, iz * S.perZ + ix * S.perX);
        for (int off = 32; off > 0; off >>= 1) r += __shfl_xor(r, off, 64);
        if (lane == 0) { rowS[iz] = (nx > 1) ? S.xStep * r : 0.; }
        tot += (lane == 0) ? r : 0.;
    }
    tot = block_sum(tot, red);
    __shared__ int shOk, shIz0, shIz1;
    __shared__ double shTotF;
    if (t == 0) {
        shOk = 0; shIz0 = 0; shIz1 = nz - 1; shTotF = 0.;
        if (nx > 1) {
            float totF = (float)(tot * S.xStep * S.zStep);
            double stop = (double)totF * (1. - 0.9) * 0.25 / S.zStep;
            int i0 = 0, i1 = nz - 1;
            double a = 0.;
            for (int i = 0; i < nz; i++) { a += rowS[i]; if (a > stop) { i0 = i; break; } }
            a = 0.;
            for (int i = 0; i < nz; i++) { a += rowS[nz - 1 - i]; if (a > stop) { i1 = nz - 1 - i; break; } }
            if (i1 > i0) { shOk = 1; shIz0 = i0; shIz1 = i1; }
            shTotF = (double)totF;
        }
    }
    __syncthreads();
    const bool rowsOk = shOk != 0;
    const int iz0 = shIz0, iz1 = shIz1;
    const double totF = shTotF;
    // ---- column sums inside the row limits (srmatsta.cpp:547-560)
    if (rowsOk) {
        for (int ix = t; ix < nx; ix += SRWM_T) {
            double c = 0.;
            for (int iz = iz0; iz <= iz1; iz++) c += intensity(S, iz * S.perZ + ix * S.perX);
            colS[ix] = S.zStep * c;
        }
    }
    __syncthreads();
    if (t == 0) {
        int L0 = 0, L1 = nx - 1, L2 = 0, L3 = nz - 1;
        if (rowsOk) {
            double stop = totF * (1. - 0.9) * 0.25 / S.xStep;
            long long j0 = 0, j1 = nx - 1;
            double a = 0.;
            for (int i = 0; i < nx; i++) { a += colS[i]; if (a > stop) { j0 = i; break; } }
            a = 0.;
            for (int i = 0; i < nx; i++) { a += colS[nx - 1 - i]; if (a > stop) { j1 = nx - 1 - i; break; } }
            // positions pass through float as in srmatsta.cpp:616-625
            float fx0 = (float)(S.xStart + S.xStep * j0), fx1 = (float)(S.xStart + S.xStep * j1);
            float fz0 = (float)(S.zStart + S.zStep * iz0), fz1 = (float)(S.zStart + S.zStep * iz1);
            L0 = (int)((fx0 - S.xStart) * 1.0000001 / S.xStep); if (L0 < 0) L0 = 0;
            L1 = (int)((fx1 - S.xStart) * 1.0000001 / S.xStep); if (L1 >= nx) L1 = nx - 1;
            L2 = (int)((fz0 - S.zStart) * 1.0000001 / S.zStep); if (L2 < 0) L2 = 0;
            L3 = (int)((fz1 - S.zStart) * 1.0000001 / S.zStep); if (L3 >= nz) L3 = nz - 1;
        }
        lim[0] = L0; lim[1] = L1; lim[2] = L2; lim[3] = L3;
        for (int k = 0; k < 4; k++) lims[blockIdx.x * 4 + k] = lim[k];
    }
    __syncthreads();
    const int L0 = lim[0], L1 = lim[1], L2 = lim[2], L3 = lim[3];
    // ---- moment sums (sroptelm.cpp:1688-1941)
    const double TwoPi = 3.141592653590 * 2., FourPi = TwoPi * 2.;
    const double lam4 = 1.239842E-06 / (FourPi * S.ePh), lamM = lam4 * FourPi;
    const double rx = S.Rx != 0. ? S.Rx : 100. * lamM, rz = S.Rz != 0. ? S.Rz : 100. * lamM;
    const double kx = 0.5 * (S.xStep * ((1. / lam4) / rx)), kz = 0.5 * (S.zStep * ((1. / lam4) / rz));
    const double kx2 = kx * kx, kz2 = kz * kz;
    treat_ctx tc;
    tc.cRxE = -(SRW_QUAD_CONST / S.Rx) * S.ePh; tc.cRzE = -(SRW_QUAD_CONST / S.Rz) * S.ePh;
    tc.x0 = S.xStart - S.xc; tc.z0 = S.zStart - S.zc; tc.xStep = S.xStep; tc.zStep = S.zStep;
    tc.fx = S.fx; tc.fz = S.fz;
    const bool any = S.fx || S.fz;
    double acc[SRWM_NSUM];
    for (int k = 0; k < SRWM_NSUM; k++) acc[k] = 0.;
    // one wave per row, lanes across x: the backward x difference comes from
    // the neighbouring lane (lane 0 from the previous 64-point chunk)
    for (int iz = w; iz < nz; iz += SRWM_T / 64) {
        const bool inZ = iz >= L2 && iz <= L3;
        const double z = S.zStart + iz * S.zStep;
        const double rAll = (iz == 0 || iz == nz - 1) ? 0.5 : 1.;
        const double rDer = (iz == 1) ? 0.5 * rAll : rAll;
        float2 carry[2] = {make_float2(0.f, 0.f), make_float2(0.f, 0.f)};
        for (int base = 0; base < nx; base += 64) {
            const int ix = base + lane;
            const bool valid = ix < nx;
            const long long o = iz * S.perZ + (long long)ix * S.perX;
            const bool in = inZ && ix >= L0 && ix <= L1;
            const double x = S.xStart + ix * S.xStep;
            float2 e[2], l[2];
            for (int c = 0; c < 2; c++) {
                const float2* p = c == 0 ? S.ex : S.ez;
                e[c] = (valid && p) ? treated(p, o, tc, ix, iz, any) : make_float2(0.f, 0.f);
                float lx = __shfl_up(e[c].x, 1, 64), ly = __shfl_up(e[c].y, 1, 64);
                l[c] = (lane == 0) ? carry[c] : make_float2(lx, ly);
                carry[c] = make_float2(__shfl(e[c].x, 63, 64), __shfl(e[c].y, 63, 64));
            }
            if (!valid) continue;
            const double wAll = (ix == 0 || ix == nx - 1) ? 0.5 : 1.;
            const double wDer = (ix == 1) ? 0.5 * wAll : wAll;
            for (int c = 0; c < 2; c++) {
                const float2* p = c == 0 ? S.ex : S.ez;
                double f[11];
                double re = e[c].x, im = e[c].y;
                f[0] = re * re + im * im;
                f[1] = x * f[0];
                f[3] = z * f[0];
                f[5] = in ? x * f[1] : 0.;
                f[8] = in ? z * f[3] : 0.;
                if (ix > 0) {
                    double dRe = re - (double)l[c].x, dIm = im - (double)l[c].y, cr = dIm * re - dRe * im;
                    f[2] = cr + kx * x * f[0];
                    f[6] = in ? x * f[2] : 0.;
                    f[7] = in ? (dRe * dRe + dIm * dIm) + cr * kx * x + kx2 * x * x * f[0] : 0.;
                } else { f[2] = 0.; f[6] = 0.; f[7] = 0.; }
                if (iz > 0) {
                    float2 u = p ? treated(p, o - S.perZ, tc, ix, iz - 1, any) : make_float2(0.f, 0.f);
                    double dRe = re - (double)u.x, dIm = im - (double)u.y, cr = dIm * re - dRe * im;
                    f[4] = cr + kz * z * f[0];
                    f[9] = in ? z * f[4] : 0.;
                    f[10] = in ? (dRe * dRe + dIm * dIm) + cr * kz * z + kz2 * z * z * f[0] : 0.;
                } else { f[4] = 0.; f[9] = 0.; f[10] = 0.; }
                for (int k = 0; k < 11; k++) {
                    double wx = (k == 2 || k == 6 || k == 7) ? wDer : wAll;
                    double wz = (k == 4 || k == 9 || k == 10) ? rDer : rAll;
                    acc[11 * c + k] += (wx * wz) * f[k];
                }
            }
        }
    }
    // block reduction of the 22 sums
    for (int k = 0; k < SRWM_NSUM; k++) {
        double v = acc[k];
        for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
        acc[k] = v;
    }
    __syncthreads();
    if (lane == 0) for (int k = 0; k < SRWM_NSUM; k++) red[w * SRWM_NSUM + k] = acc[k];
    __syncthreads();
    if (t < SRWM_NSUM) {
        double v = 0.;
        for (int q = 0; q < SRWM_T / 64; q++) v += red[q * SRWM_NSUM + t];
        sums[blockIdx.x * 24 + t] = v;
    }
}

}  // namespace

extern "C" int srwm_launch_moments(const srwm_mslice* dSlices, int nSlices, double* dWork, long long workStride,
                                   double* dSums, int* dLims, hipStream_t st)
{
    if (nSlices <= 0) return 0;
    hipLaunchKernelGGL(k_rad_moments, dim3(nSlices), dim3(SRWM_T), 0, st, dSlices, dWork, workStride, dSums, dLims);
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

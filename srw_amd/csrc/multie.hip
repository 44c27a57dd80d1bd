// multie.hip - intensity extraction and partially coherent accumulation.
//
// Reference algorithms (paths relative to the reference source root):
//   intensity            cpp/src/core/srradmnp.h:367-430, srradmnp.cpp:681
//   beam sampling        env/python/srwpy/srwlib.py:9000-9029, 9793-9803
//   averaging on a mesh  env/python/srwpy/srwlib.py:1083-1170 (avg_update_*)
#include "srw_prop.h"
#include "srw_multie.h"

#include <math.h>

namespace {

__device__ __forceinline__ float int_comp(float xr, float xi, float zr, float zi, int pol)
{   // srradmnp.h:367-430 (polarisation components and Stokes parameters)
    switch (pol) {
    case 0: return xr * xr + xi * xi;
    case 1: return zr * zr + zi * zi;
    case 2: { float a = xr + zr, b = xi + zi; return 0.5f * (a * a + b * b); }
    case 3: { float a = xr - zr, b = xi - zi; return 0.5f * (a * a + b * b); }
    case 5: { float a = xr - zi, b = xi + zr; return 0.5f * (a * a + b * b); }
    case 4: { float a = xr + zi, b = xi - zr; return 0.5f * (a * a + b * b); }
    case -1: return xr * xr + xi * xi + zr * zr + zi * zi;
    case -2: return xr * xr + xi * xi - (zr * zr + zi * zi);
    case -3: return 2.f * (xr * zr + xi * zi);
    case -4: return 2.f * (xr * zi - xi * zr);
    default: return xr * xr + xi * xi + zr * zr + zi * zi;
    }
}

__global__ void k_intensity(const float2* __restrict__ ex, const float2* __restrict__ ez, int ne, int ie0, int ie1,
                            float t, int pol, long long nxy, float* out)
{
    long long p = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= nxy) return;
    float2 a = ex[p * ne + ie0], b = ez[p * ne + ie0];
    float v = int_comp(a.x, a.y, b.x, b.y, pol);
    if (ne > 1) {
        float2 a1 = ex[p * ne + ie1], b1 = ez[p * ne + ie1];
        float v1 = int_comp(a1.x, a1.y, b1.x, b1.y, pol);
        v = v + (v1 - v) * t;
    }
    out[p] = v;
}

}  // namespace

extern "C" {
int srwp_launch_intensity(const float2* ex, const float2* ez, int ne, int ie0, int ie1, float t, int pol,
                          long long nxy, float* out, hipStream_t st)
{
    hipLaunchKernelGGL(k_intensity, dim3((unsigned)((nxy + 255) / 256)), dim3(256), 0, st, ex, ez, ne, ie0, ie1, t,
                       pol, nxy, out);
    return hipGetLastError() == hipSuccess ? 0 : -1;
}
}

namespace {

// Bilinear re-gridding of per-electron intensities onto the result mesh and
// accumulation of the weighted sums S1 = sum w*I and S2 = sum (w*I)^2
// (srwlib.py avg_update_interp, _ord = 1: points outside the electron's mesh
// contribute zero).  One thread per result point; loops over the batch.
__global__ void __launch_bounds__(256) k_accum_interp(const float* __restrict__ I, long long iStride,
                                                      const srwm_emesh* __restrict__ em, const float* __restrict__ w,
                                                      int nb, srwm_rmesh rm, double* S1, double* S2, double* norm2,
                                                      double* Srep, int perRep)
{
    long long r = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    long long nr = (long long)rm.nx * rm.nz;
    bool valid = r < nr;
    int ix = valid ? (int)(r % rm.nx) : 0, iz = valid ? (int)(r / rm.nx) : 0;
    double xr = rm.xStart + ix * rm.xStep, zr = rm.zStart + iz * rm.zStep;
    double s1 = 0., s2 = 0., sg = 0.;
    for (int j = 0; j < nb; j++) {
        const srwm_emesh E = em[j];
        float f = 0.f;
        int ix0 = (int)trunc((xr - E.xStart) / E.xStep + 1.e-09);
        int iz0 = (int)trunc((zr - E.zStart) / E.zStep + 1.e-09);
        if (valid && ix0 >= 0 && ix0 < E.nx - 1 && iz0 >= 0 && iz0 < E.nz - 1) {
            double tx = (xr - (E.xStart + E.xStep * ix0)) / E.xStep;
            double tz = (zr - (E.zStart + E.zStep * iz0)) / E.zStep;
            const float* P = I + j * iStride + (long long)iz0 * E.nx + ix0;
            float a00 = P[0], f10 = P[1], f01 = P[E.nx], f11 = P[E.nx + 1];
            float a10 = f10 - a00, a01 = f01 - a00, a11 = a00 - f01 - f10 + f11;
            f = (float)(a00 + tx * (a10 + tz * a11) + tz * a01);
        }
        double v = (double)w[j] * (double)f;
        s1 += v;
        s2 += v * v;
        if (Srep) {
            sg += v;
            if ((j + 1) % perRep == 0) {
                if (valid) Srep[(long long)(j / perRep) * nr + r] += sg;
                sg = 0.;
            }
        }
        if (norm2) {
            // per-electron squared L2 norm of the re-gridded intensity (training signal)
            double q = (double)f * (double)f;
            for (int off = 32; off > 0; off >>= 1) q += __shfl_down(q, off);
            if ((threadIdx.x & 63) == 0) atomicAdd(&norm2[j], q);
        }
    }
    if (!valid) return;
    S1[r] += s1;
    S2[r] += s2;
}

__global__ void k_stokes_out(const double* __restrict__ S1, double scale, long long n, float* out)
{
    long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[i] = (float)(S1[i] * scale);
}

}  // namespace

extern "C" {
int srwm_launch_accum(const float* I, long long iStride, const srwm_emesh* em, const float* w, int nb,
                      srwm_rmesh rm, double* S1, double* S2, double* norm2, double* Srep, int perRep,
                      hipStream_t st)
{
    long long nr = (long long)rm.nx * rm.nz;
    if (Srep && (perRep <= 0 || nb % perRep != 0)) return -1;
    hipLaunchKernelGGL(k_accum_interp, dim3((unsigned)((nr + 255) / 256)), dim3(256), 0, st, I, iStride, em, w, nb,
                       rm, S1, S2, norm2, Srep, perRep);
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

int srwm_launch_stokes_out(const double* S1, double scale, long long n, float* out, hipStream_t st)
{
    hipLaunchKernelGGL(k_stokes_out, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, S1, scale, n, out);
    return hipGetLastError() == hipSuccess ? 0 : -1;
}
}

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

// extract.hip - single-electron intensity / field-component extraction
// (srwlCalcIntFromElecField), for gfx950.
//
// Reference algorithms (paths relative to the reference source root):
//   component of E          cpp/src/core/srradmnp.h:367-430 (IntensityComponent)
//   off-mesh interpolation  cpp/src/core/srradmnp.h:338-363, srradmnp.cpp:35-79
//   loops per dependence    cpp/src/core/srradmnp.cpp:224-320 (vs e), 324-498 (vs x),
//                           500-679 (vs y), 681-955 (vs x, y), 957-1030 (vs e, x),
//                           1032-1105 (vs e, y), 1107-1180 (vs e, x, y)
//
// One thread per output value; the wavefront is [ny][nx][ne] complex float.
// The arithmetic repeats the reference's exactly (float component, float
// differences scaled by double interpolation weights, one rounding to float)
// with contraction off, so the result is bit-identical to a CPU evaluation.
#include "srw_prop.h"

namespace {

#pragma clang fp contract(off)

__device__ __forceinline__ float comp(const float2* ex, const float2* ez, long long o, int pol, int what)
{
    float xr = 0.f, xi = 0.f, zr = 0.f, zi = 0.f;   // a missing component reads as zero
    if (ex) { float2 v = ex[o]; xr = v.x; xi = v.y; }
    if (ez) { float2 v = ez[o]; zr = v.x; zi = v.y; }
    const bool inten = (what == 0);
    switch (pol) {
    case 0: return inten ? xr * xr + xi * xi : (what == 3 ? xr : xi);
    case 1: return inten ? zr * zr + zi * zi : (what == 3 ? zr : zi);
    case 2: { float a = xr + zr, b = xi + zi; return (float)(inten ? 0.5 * (double)(a * a + b * b) : 0.70710678 * (double)(what == 3 ? a : b)); }
    case 3: { float a = xr - zr, b = xi - zi; return (float)(inten ? 0.5 * (double)(a * a + b * b) : 0.70710678 * (double)(what == 3 ? a : b)); }
    case 5: { float a = xr - zi, b = xi + zr; return (float)(inten ? 0.5 * (double)(a * a + b * b) : 0.70710678 * (double)(what == 3 ? a : b)); }
    case 4: { float a = xr + zi, b = xi - zr; return (float)(inten ? 0.5 * (double)(a * a + b * b) : 0.70710678 * (double)(what == 3 ? a : b)); }
    case -1: return xr * xr + xi * xi + zr * zr + zi * zi;
    case -2: return xr * xr + xi * xi - (zr * zr + zi * zi);
    case -3: return (float)(2. * (double)(xr * zr + xi * zi));
    case -4: return (float)(2. * (double)(xr * zi - xi * zr));
    default: return inten ? xr * xr + xi * xi + zr * zr + zi * zi : (what == 3 ? xr : xi);
    }
}

__device__ __forceinline__ float interp1(float a, float b, double t) { return (float)((double)(b - a) * t + (double)a); }

__device__ __forceinline__ float interp2(float i00, float i10, float i01, float i11, double t1, double t2)
{
    const double t12 = t1 * t2;
    return (float)((double)(i00 - i01 - i10 + i11) * t12 + (double)(i10 - i00) * t1 + (double)(i01 - i00) * t2 +
                   (double)i00);
}

__global__ void __launch_bounds__(256) k_extract(const float2* __restrict__ ex, const float2* __restrict__ ez,
                                                 srwx_extract P, float* __restrict__ out)
{
    const long long k = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= P.nout) return;
    const long long ne = P.ne, nx = P.nx;
    auto at = [&](long long ie, long long ix, long long iz) { return (iz * nx + ix) * ne + ie; };
    const int npol = P.pol == -5 ? 4 : 1;
    for (int s = 0; s < npol; s++) {
        const int pol = P.pol == -5 ? -1 - s : P.pol;
        auto C = [&](long long ie, long long ix, long long iz) { return comp(ex, ez, at(ie, ix, iz), pol, P.what); };
        float v;
        switch (P.dep) {
        case 0:
            v = interp2(C(k, P.ix0, P.iz0), C(k, P.ix1, P.iz0), C(k, P.ix0, P.iz1), C(k, P.ix1, P.iz1), P.tx, P.tz);
            break;
        case 1:
            v = interp2(C(P.ie0, k, P.iz0), C(P.ie1, k, P.iz0), C(P.ie0, k, P.iz1), C(P.ie1, k, P.iz1), P.te, P.tz);
            break;
        case 2:
            v = interp2(C(P.ie0, P.ix0, k), C(P.ie1, P.ix0, k), C(P.ie0, P.ix1, k), C(P.ie1, P.ix1, k), P.te, P.tx);
            break;
        case 3: {
            const long long iz = k / nx, ix = k - iz * nx;
            v = interp1(C(P.ie0, ix, iz), C(P.ie1, ix, iz), P.te);
            break;
        }
        case 4: {
            const long long ix = k / ne, ie = k - ix * ne;
            v = interp1(C(ie, ix, P.iz0), C(ie, ix, P.iz1), P.tz);
            break;
        }
        case 5: {
            const long long iz = k / ne, ie = k - iz * ne;
            v = interp1(C(ie, P.ix0, iz), C(ie, P.ix1, iz), P.tx);
            break;
        }
        default:
            v = comp(ex, ez, k, pol, P.what);
        }
        out[s * P.nout + k] = v;
    }
}

}  // namespace

extern "C" int srwx_launch_extract(const float2* ex, const float2* ez, const srwx_extract& P, float* out,
                                   hipStream_t st)
{
    if (P.nout <= 0) return 0;
    hipLaunchKernelGGL(k_extract, dim3((unsigned)((P.nout + 255) / 256)), dim3(256), 0, st, ex, ez, P, out);
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

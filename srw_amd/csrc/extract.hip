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
// Also: the phase (double output), the intensity integrated over photon
// energy, the single-electron flux vs photon energy and the running-average /
// sum modes of arMeth (srradmnp.cpp:83-143, 729-945; gmmeth.h:171-203).
// The arithmetic repeats the reference's exactly (float component, float
// differences scaled by double interpolation weights, one rounding to float)
// with contraction off, so the result is bit-identical to a CPU evaluation.
#include "srw_prop.h"

namespace {

#pragma clang fp contract(off)

// srradmnp.h:434-454 FormalPhase: argument of re + i im, the ratio in float.
__device__ __forceinline__ double formal_phase(float re, float im)
{
    const double HalfPi = 1.5707963267949, Pi = 3.1415926535898;
    if (re != 0.f) {
        double a = atan((double)(im / re));
        if (im <= 0.f) return re < 0.f ? a - Pi : a;
        return re < 0.f ? a + Pi : a;
    }
    if (im == 0.f) return 0.;
    return im > 0.f ? HalfPi : -HalfPi;
}

// srradmnp.h:367-430 IntensityComponent; what (the reference's internal
// Int_or_ReE): 0 / 1 intensity, 2 phase, 3 Re E, other values Im E.
__device__ __forceinline__ float comp(const float2* ex, const float2* ez, long long o, int pol, int what)
{
    float xr = 0.f, xi = 0.f, zr = 0.f, zi = 0.f;   // a missing component reads as zero
    if (ex) { float2 v = ex[o]; xr = v.x; xi = v.y; }
    if (ez) { float2 v = ez[o]; zr = v.x; zi = v.y; }
    const bool inten = (what == 0 || what == 1);
    const bool ph = (what == 2);
    switch (pol) {
    case 0: return inten ? xr * xr + xi * xi : (ph ? (float)formal_phase(xr, xi) : (what == 3 ? xr : xi));
    case 1: return inten ? zr * zr + zi * zi : (ph ? (float)formal_phase(zr, zi) : (what == 3 ? zr : zi));
    case 2: { float a = xr + zr, b = xi + zi;
        return (float)(inten ? 0.5 * (double)(a * a + b * b) : (ph ? formal_phase(a, b) : 0.70710678 * (double)(what == 3 ? a : b))); }
    case 3: { float a = xr - zr, b = xi - zi;
        return (float)(inten ? 0.5 * (double)(a * a + b * b) : (ph ? formal_phase(a, b) : 0.70710678 * (double)(what == 3 ? a : b))); }
    case 5: { float a = xr - zi, b = xi + zr;
        return (float)(inten ? 0.5 * (double)(a * a + b * b) : (ph ? formal_phase(a, b) : 0.70710678 * (double)(what == 3 ? a : b))); }
    case 4: { float a = xr + zi, b = xi - zr;
        return (float)(inten ? 0.5 * (double)(a * a + b * b) : (ph ? formal_phase(a, b) : 0.70710678 * (double)(what == 3 ? a : b))); }
    case -1: return xr * xr + xi * xi + zr * zr + zi * zi;
    case -2: return xr * xr + xi * xi - (zr * zr + zi * zi);
    case -3: return (float)(2. * (double)(xr * zr + xi * zi));
    case -4: return (float)(2. * (double)(xr * zi - xi * zr));
    default: return inten ? xr * xr + xi * xi + zr * zr + zi * zi
                          : (ph ? (float)formal_phase(xr, xi) : (what == 3 ? xr : xi));
    }
}

// srradmnp.h:345-363; the phase is taken at the first point, not interpolated
__device__ __forceinline__ float interp1(float a, float b, double t, int what)
{
    return what == 2 ? a : (float)((double)(b - a) * t + (double)a);
}

__device__ __forceinline__ float interp2(float i00, float i10, float i01, float i11, double t1, double t2, int what)
{
    if (what == 2) return i00;
    const double t12 = t1 * t2;
    return (float)((double)(i00 - i01 - i10 + i11) * t12 + (double)(i10 - i00) * t1 + (double)(i01 - i00) * t2 +
                   (double)i00);
}

// gmmeth.h:171-203 Integ1D_FuncDefByArray with its loop bounds as written
// (the Simpson sums start at the second interior pair); f(i) gives the i-th
// sample as a double.
template <class F>
__device__ double integ1d(F f, long long np, double step)
{
    if (np < 2 || step == 0) return 0;
    if (np == 2) return 0.5 * step * (f(0) + f(1));
    const bool even = (np == ((np >> 1) << 1));
    const long long nsim = even ? np - 1 : np;
    double s1 = 0, s2 = 0;
    long long t = 1;
    for (long long i = 1; i < ((nsim - 3) >> 1); i++) { s1 += f(t); s2 += f(t + 1); t += 2; }
    s1 += f(t); t++;
    double res = (step / 3.) * (f(0) + 4. * s1 + 2. * s2 + f(t));
    if (even) res += 0.5 * step * (f(nsim - 1) + f(nsim));
    return res;
}

__global__ void __launch_bounds__(256) k_extract(const float2* __restrict__ ex, const float2* __restrict__ ez,
                                                 srwx_extract P, void* __restrict__ outv)
{
    const long long k = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= P.nout) return;
    const long long ne = P.ne, nx = P.nx;
    auto at = [&](long long ie, long long ix, long long iz) { return (iz * nx + ix) * ne + ie; };
    const int npol = P.pol == -5 ? 4 : 1;
    const int w = P.what;
    for (int s = 0; s < npol; s++) {
        const int pol = P.pol == -5 ? -1 - s : P.pol;
        auto C = [&](long long ie, long long ix, long long iz) { return comp(ex, ez, at(ie, ix, iz), pol, w); };
        double v;
        if (P.intOverE) {
            // intensity integrated over photon energy at each output point
            auto f = [&](long long ie) -> double {
                if (P.dep == 1) return interp1(C(ie, k, P.iz0), C(ie, k, P.iz1), P.tz, 0);
                if (P.dep == 2) return interp1(C(ie, P.ix0, k), C(ie, P.ix1, k), P.tx, 0);
                const long long iz = k / nx, ix = k - iz * nx;
                return C(ie, ix, iz);
            };
            v = integ1d(f, ne, P.eStep);
        } else {
            switch (P.dep) {
            case 0:
                v = interp2(C(k, P.ix0, P.iz0), C(k, P.ix1, P.iz0), C(k, P.ix0, P.iz1), C(k, P.ix1, P.iz1), P.tx, P.tz, w);
                break;
            case 1:
                v = interp2(C(P.ie0, k, P.iz0), C(P.ie1, k, P.iz0), C(P.ie0, k, P.iz1), C(P.ie1, k, P.iz1), P.te, P.tz, w);
                break;
            case 2:
                v = interp2(C(P.ie0, P.ix0, k), C(P.ie1, P.ix0, k), C(P.ie0, P.ix1, k), C(P.ie1, P.ix1, k), P.te, P.tx, w);
                break;
            case 3: {
                const long long iz = k / nx, ix = k - iz * nx;
                v = interp1(C(P.ie0, ix, iz), C(P.ie1, ix, iz), P.te, w);
                break;
            }
            case 4: {
                const long long ix = k / ne, ie = k - ix * ne;
                v = interp1(C(ie, ix, P.iz0), C(ie, ix, P.iz1), P.tz, w);
                break;
            }
            case 5: {
                const long long iz = k / ne, ie = k - iz * ne;
                v = interp1(C(ie, P.ix0, iz), C(ie, P.ix1, iz), P.tx, w);
                break;
            }
            default:
                v = comp(ex, ez, k, pol, w);
            }
        }
        // srradmnp.cpp:742-752, 869-939: overwrite, running average, or sum
        const long long o = s * P.nout + k;
        if (P.dbl) {
            double* out = (double*)outv;
            out[o] = P.iter == 0. ? v : (P.iter > 0. ? (out[o] * P.iter + v) * P.invIter : out[o] + v);
        } else {
            float* out = (float*)outv;
            out[o] = P.iter == 0. ? (float)v
                                  : (P.iter > 0. ? (float)(((double)out[o] * P.iter + v) * P.invIter) : out[o] + (float)v);
        }
    }
}

// Single-electron flux vs photon energy (srradmnp.cpp:83-143), in the
// reference's summation order: k_flux_rows forms each row's sum over x
// (edge points halved), k_flux_energies adds the rows (edge rows halved).
__global__ void __launch_bounds__(256) k_flux_rows(const float2* __restrict__ ex, const float2* __restrict__ ez,
                                                   srwx_extract P, double* __restrict__ rows)
{
    const long long k = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= P.ne * P.nz) return;
    const long long ie = k % P.ne, iz = k / P.ne;
    double sx = 0.;
    for (long long ix = 0; ix < P.nx; ix++) {
        double c = comp(ex, ez, (iz * P.nx + ix) * P.ne + ie, P.pol, 0);
        if (ix == 0 || ix == P.nx - 1) c *= 0.5;
        sx += c;
    }
    if (iz == 0 || iz == P.nz - 1) sx *= 0.5;
    rows[k] = sx;
}

__global__ void __launch_bounds__(256) k_flux_energies(const double* __restrict__ rows, srwx_extract P,
                                                       float* __restrict__ out)
{
    const long long ie = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (ie >= P.ne) return;
    double sum = 0.;
    for (long long iz = 0; iz < P.nz; iz++) sum += rows[iz * P.ne + ie];
    out[ie] = (float)(sum * P.xStep * P.zStep * 1E+06);
}

}  // namespace

extern "C" int srwx_launch_extract(const float2* ex, const float2* ez, const srwx_extract& P, void* out,
                                   double* rows, hipStream_t st)
{
    if (P.nout <= 0) return 0;
    if (P.what == 4) {
        long long n = P.ne * P.nz;
        hipLaunchKernelGGL(k_flux_rows, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, ex, ez, P, rows);
        hipLaunchKernelGGL(k_flux_energies, dim3((unsigned)((P.ne + 255) / 256)), dim3(256), 0, st, rows, P,
                           (float*)out);
    } else {
        hipLaunchKernelGGL(k_extract, dim3((unsigned)((P.nout + 255) / 256)), dim3(256), 0, st, ex, ez, P, out);
    }
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

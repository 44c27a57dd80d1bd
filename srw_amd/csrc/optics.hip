// optics.hip - pointwise optical elements of a propagation container (thin
// lens, rectangular / circular aperture and obstacle) on the slice-major
// device field, for gfx950.
//
// Reference algorithms (paths relative to the reference source root):
//   thin lens            cpp/src/core/sroptfoc.h:227-242 (RadPointModifier)
//   apertures/obstacles  cpp/src/core/sroptapt.h:150-166, 244-258, 330-343, 425-440
//   float cos/sin        cpp/src/core/sroptelm.h:856-872 (CosAndSin)
//
// One thread per transverse point of one photon energy; the thread updates
// both field components (slices 2 ie and 2 ie + 1).  HBM-bound: 32 bytes read
// and written per point (apertures skip the write inside the opening).
#include "srw_prop.h"

namespace {

// sroptelm.h:856-872: reduction to [-pi/2, pi/2] with the reference's
// constants, Taylor series to x^10 / x^11, rounded to float.
__device__ __forceinline__ void cos_sin_opt(double x, float& c, float& s)
{
    const double HalfPI = 1.5707963267949, PI = 3.141592653590, TwoPI = 6.2831853071796;
    const double ThreePIdTwo = 4.7123889803847, One_dTwoPI = 0.1591549430919;
    if (x < -1.E+08 || x > 1.E+08) { c = (float)cos(x); s = (float)sin(x); return; }
    x -= TwoPI * (double)(long long)(x * One_dTwoPI);
    if (x < 0.) x += TwoPI;
    bool neg = false;
    if (x > ThreePIdTwo) x -= TwoPI;
    else if (x > HalfPI) { x -= PI; neg = true; }
    const double q = x * x;
    float cc = (float)(1. + q * (-0.5 + q * (0.041666666666667 + q * (-0.0013888888888889 +
                       q * (0.000024801587301587 + q * -2.755731922E-07)))));
    float ss = (float)(x * (1. + q * (-0.16666666666667 + q * (0.0083333333333333 + q * (-0.0001984126984127 +
                       q * (2.755731922E-06 + q * -2.505210839E-08))))));
    c = neg ? -cc : cc;
    s = neg ? -ss : ss;
}

__global__ void __launch_bounds__(256) k_point_elem(float2* d, long long nxy, int ne, srwo_elem E)
{
    const long long p = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    const int ie = blockIdx.y;
    if (p >= nxy || ie >= ne) return;
    const long long iz = p / E.nx, ix = p - iz * E.nx;
    const double xr = E.xStart + ix * E.xStep - E.cx, zr = E.zStart + iz * E.zStep - E.cy;
    float2* a = d + (2LL * ie) * nxy + p;
    float2* b = a + nxy;
    if (E.kind == 0) {
        const double e = E.eStart + ie * E.eStep;
        const double ph = -(e * 2.533865612E+06) * (xr * xr / E.p0 + zr * zr / E.p1);
        float c, s;
        cos_sin_opt(ph, c, s);
        float2 u = *a, v = *b;
        *a = make_float2(u.x * c - u.y * s, u.x * s + u.y * c);
        *b = make_float2(v.x * c - v.y * s, v.x * s + v.y * c);
        return;
    }
    bool inside;
    if (E.kind <= 2) {
        const double hx = 0.5 * E.p0 + 1.E-10, hz = 0.5 * E.p1 + 1.E-10;
        inside = xr >= -hx && xr <= hx && zr >= -hz && zr <= hz;
    } else {
        inside = xr * xr + zr * zr <= 0.25 * E.p0 * E.p0;
    }
    const bool zero = (E.kind == 1 || E.kind == 3) ? !inside : inside;
    if (zero) { *a = make_float2(0.f, 0.f); *b = make_float2(0.f, 0.f); }
}

}  // namespace

extern "C" int srwo_launch_point_elem(float2* d, long long nxy, int ne, const srwo_elem& E, hipStream_t st)
{
    if (nxy <= 0 || ne <= 0 || ne > 65535) return -1;
    dim3 grid((unsigned)((nxy + 255) / 256), (unsigned)ne);
    hipLaunchKernelGGL(k_point_elem, grid, dim3(256), 0, st, d, nxy, ne, E);
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

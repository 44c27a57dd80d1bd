// srw_optics.h - device helpers shared by the optics kernels.
#pragma once
#include <hip/hip_runtime.h>
#include <math.h>

// float sin/cos of the reference's optics code: reduction into [-pi/2, pi/2]
// with its rounded constants, then Taylor polynomials to x^10 / x^11
// (sroptelm.h:105-115, 856-872).  Phase treatments use it so that removing
// and adding the same phase round-trips like the reference's.
__device__ __forceinline__ void cos_sin_optics(double x, float& c, float& s)
{
    const double HalfPI = 1.5707963267949, PI = 3.141592653590, TwoPI = 6.2831853071796;
    const double ThreePIdTwo = 4.7123889803847, One_dTwoPI = 0.1591549430919;
    if (x < -1.E+08 || x > 1.E+08) { c = (float)cos(x); s = (float)sin(x); return; }
    x -= TwoPI * (double)(long long)(x * One_dTwoPI);
    if (x < 0.) x += TwoPI;
    bool neg = false;
    if (x > ThreePIdTwo) x -= TwoPI;
    else if (x > HalfPI) { x -= PI; neg = true; }
    double q = x * x;
    float cc = (float)(1. + q * (-0.5 + q * (0.041666666666667 + q * (-0.0013888888888889 +
                       q * (0.000024801587301587 + q * -2.755731922E-07)))));
    float ss = (float)(x * (1. + q * (-0.16666666666667 + q * (0.0083333333333333 + q * (-0.0001984126984127 +
                       q * (2.755731922E-06 + q * -2.505210839E-08))))));
    c = neg ? -cc : cc;
    s = neg ? -ss : ss;
}

// Multiplies a stored float value by (c + i s), rounding back to float.
__device__ __forceinline__ float2 rotate_f(float2 a, float c, float s)
{
    return make_float2((float)((double)a.x * c - (double)a.y * s), (float)((double)a.x * s + (double)a.y * c));
}

// pi*1e6/1.23984186: pi/lambda per eV of the phase treatment (sroptelm.cpp:4780-4782)
#define SRW_QUAD_CONST (3.14159265358979 * 1.E+06 / 1.23984186)

// srw_sr_dev.h - device helpers shared by the SR integral kernels (sr.hip,
// sr_auto.hip): trajectory interpolation, the phase constant and the
// terminating terms of the radiation integral.
#pragma once
#include "srw_dev.h"
#include <math.h>

// pi/lambda [1/m] for photon energy e [eV] with SRW's SR-core constants
// (srradint.cpp:48-57).
#define SRWD_KPH(e) (3.141592653590 * 1.E+06 * 0.80654658 * (e))

// ------------------------------------------------------------------------
// Interpolation structure (FP64)
// ------------------------------------------------------------------------
__device__ __forceinline__ void coefs5(double h, const double* f, double* c)
{   // srtrjdat.h:1037-1056
    double hi = 1. / h, hi2 = hi * hi;
    double d1 = 0.016666666666667 * hi, d2 = 0.041666666666667 * hi2;
    double d3 = d2 * hi, d4 = d3 * hi, d5 = 0.2 * d4 * hi;
    double f0 = f[0], f1 = f[1], f2 = f[2], f3 = f[3], f4 = f[4], f5 = f[5];
    c[0] = f2;
    c[1] = (3 * f0 - 30 * f1 - 20 * f2 + 60 * f3 - 15 * f4 + 2 * f5) * d1;
    c[2] = -(f0 - 16 * f1 + 30 * f2 - 16 * f3 + f4) * d2;
    c[3] = -(f0 + f1 - 10 * f2 + 14 * f3 - 7 * f4 + f5) * d3;
    c[4] = (f0 - 4 * (f1 + f3) + 6 * f2 + f4) * d4;
    c[5] = (-f0 + 5 * (f1 - f4) + 10 * (f3 - f2) + f5) * d5;
}

__device__ __forceinline__ long long offset5(long long i, long long np)
{
    if (i < 2) return 0;
    if (i < np - 3) return i - 2;
    if (i < np - 2) return i - 3;
    return i - 4;
}

// Evaluates coordinate, angle, integral of squared angle and field of one
// plane at s (srtrjdat.cpp:1185-1248 with byDiv, srtrjdat.h:431-472 without).
struct trj_pt { double crd, bt, ib, b, dbds; };

__device__ __forceinline__ trj_pt eval_plane(const srwd_trj_info& I, const double* crd, const double* ib,
                                             double bmult, double s, bool byDiv)
{
    long long np = I.np;
    long long idx = byDiv ? (long long)((s - I.sStart) / I.sStep) : (long long)((s - I.sStart) * I.invStep);
    if (idx >= np - 1) idx = np - 2;
    if (idx < 0) idx = 0;
    double r = s - (I.sStart + I.sStep * idx);
    if (idx < 2) r -= (2 - idx) * I.sStep;
    else if (idx < np - 3) ;
    else if (idx < np - 2) r += I.sStep;
    else r += 2 * I.sStep;
    long long off = offset5(idx, np);
    double c[6], q[6];
    coefs5(I.sStep, crd + off, c);
    coefs5(I.sStep, ib + off, q);
    double bt[5] = {c[1], 2. * c[2], 3. * c[3], 4. * c[4], 5. * c[5]};
    double bb[4] = {bt[1] * bmult, 2. * bt[2] * bmult, 3. * bt[3] * bmult, 4. * bt[4] * bmult};
    trj_pt P;
    P.ib = q[0] + r * (q[1] + r * (q[2] + r * (q[3] + r * (q[4] + r * q[5]))));
    P.crd = c[0] + r * (c[1] + r * (c[2] + r * (c[3] + r * (c[4] + r * c[5]))));
    P.bt = bt[0] + r * (bt[1] + r * (bt[2] + r * (bt[3] + r * bt[4])));
    P.b = bb[0] + r * (bb[1] + r * (bb[2] + r * bb[3]));
    P.dbds = bb[1] + r * (bb[2] * 2 + r * (bb[3] * 3));
    return P;
}

// ------------------------------------------------------------------------
// End terms (FP64)
// ------------------------------------------------------------------------
__device__ __forceinline__ void cos_sin_ref(double x, double& c, double& s)
{
    s = sin(x); c = cos(x);
}

struct cplx { double re, im; };
__device__ __forceinline__ cplx cmul(cplx a, cplx b) { return {a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re}; }

// srradint.cpp:398-871, coordinate representation, 3-term far-field expansion.
// Returns false when the expansion is rejected (residual zero, no derivative).
static __device__ bool residual(const srwd_trj_info& I, const double* ev, double s, double xo, double yo, double zo,
                         double eph, double actNorm, cplx* R, cplx* D)
{
    double dBzds = ev[0], Bz = ev[1], Btx = ev[2], X = ev[3], IBx = ev[4];
    double dBxds = ev[5], Bx = ev[6], Btz = ev[7], Z = ev[8], IBz = ev[9];
    double kph = SRWD_KPH(eph);
    double inv = 1. / (yo - s), dx = xo - X, dz = zo - Z;
    double Ph = kph * (s * I.gammaEm2 + IBx + IBz + (dx * dx + dz * dz) * inv);
    double cph, sph;
    cos_sin_ref(Ph, cph, sph);
    cplx eip = {cph, sph};
    double bxn = Btx - dx * inv, bzn = Btz - dz * inv;
    double dPhds = kph * (I.gammaEm2 + bxn * bxn + bzn * bzn);
    double odp = 1. / dPhds;
    double Ax = bxn * inv, Az = bzn * inv;
    double cBxBz = I.betaNorm * Bz, cBzBx = -I.betaNorm * Bx;
    double cxA = cBxBz + Ax, czA = cBzBx + Az, tk = 2. * kph;
    double d2 = tk * (bxn * cxA + bzn * czA);
    double dAxds = (2. * Ax + cBxBz) * inv, dAzds = (2. * Az + cBzBx) * inv;
    double q = d2 * odp;
    double t2x = (dAxds - Ax * q) * odp, t2z = (dAzds - Az * q) * odp;
    double cdBz = I.betaNorm * dBzds, cdBx = -I.betaNorm * dBxds;
    double d3 = tk * (cxA * cxA + czA * czA + bxn * (inv * (2. * bxn * inv + cBxBz) + cdBz) +
                      bzn * (inv * (2. * bzn * inv + cBzBx) + cdBx));
    double d2Ax = inv * (3. * dAxds + cdBz), d2Az = inv * (3. * dAzds + cdBx);
    double odp2 = odp * odp, qq = d2 * d2 * odp2;
    double t3x = (-d2Ax + (3. * dAxds * d2 + Ax * d3) * odp - 3. * Ax * qq) * odp2;
    double t3z = (-d2Az + (3. * dAzds * d2 + Az * d3) * odp - 3. * Az * qq) * odp2;
    double t1x = Ax;
    bool ok2 = ((t2x == 0.) && (t1x == 0.)) || ((t2x != 0.) && (fabs(t2x) < 0.7 * fabs(t1x)));
    bool ok3 = ok2 && (((t3x == 0.) && (t2x == 0.)) || ((t3x != 0.) && (fabs(t3x) < 0.7 * fabs(t2x))));
    if (!ok2) return false;
    cplx pX = {Ax, t2x}, pZ = {Az, t2z};
    if (ok3) { pX.re += t3x; pZ.re += t3z; }
    // multiply by i/dPhds then by exp(i Ph)
    cplx mX = {-pX.im * odp, pX.re * odp}, mZ = {-pZ.im * odp, pZ.re * odp};
    cplx ex = cmul(mX, eip), ez = cmul(mZ, eip);
    R[0] = {actNorm * ex.re, actNorm * ex.im};
    R[1] = {actNorm * ez.re, actNorm * ez.im};
    D[0] = cmul({dAxds, Ax * dPhds}, eip);
    D[1] = cmul({dAzds, Az * dPhds}, eip);
    return true;
}


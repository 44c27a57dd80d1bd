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
// srradint.h:867-880 with the constants of srradint.cpp:48-60.
__device__ __forceinline__ void cos_sin_taylor(double x, double& c, double& s)
{
    const double PI = 3.141592653590, TwoPI = 2. * PI, ThreePIdTwo = 1.5 * PI, HalfPI = 0.5 * PI;
    const double One_dTwoPI = 1. / TwoPI;
    const double a2c = -0.5, a4c = 1. / 24., a6c = -1. / 720., a8c = 1. / 40320., a10c = -1. / 3628800.;
    const double a3s = -1. / 6., a5s = 1. / 120., a7s = -1. / 5040., a9s = 1. / 362880., a11s = -1. / 39916800.;
    x -= TwoPI * (double)(int)(x * One_dTwoPI);
    if (x < 0.) x += TwoPI;
    bool chs = false;
    if (x > ThreePIdTwo) x -= TwoPI;
    else if (x > HalfPI) { x -= PI; chs = true; }
    double x2 = x * x;
    c = 1. + x2 * (a2c + x2 * (a4c + x2 * (a6c + x2 * (a8c + x2 * a10c))));
    s = x * (1. + x2 * (a3s + x2 * (a5s + x2 * (a7s + x2 * (a9s + x2 * a11s)))));
    if (chs) { c = -c; s = -s; }
}

struct cplx { double re, im; };
__device__ __forceinline__ cplx cmul(cplx a, cplx b) { return {a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re}; }

// Radiation integral along the field-free straight line through the edge
// point sc (srradint.cpp:1052-1201), from sSt to sFi, added with sign -1 to R
// (the left edge, sc = sFi): composite rule of 5 points, then level doubling
// until a level whose midpoint phases advance by at most 5*pi/4 changes |E|^2
// by at most prec*|E|^2; dd is the difference of the integrand's derivative
// between the ends.  Capped at 22 levels (the reference has no cap).
struct free_line { double sc, btx, x0, btz, z0, ib0, xo, yo, zo, kph, gEm2; };

__device__ __forceinline__ void free_term(const free_line& L, double s, double& c, double& sn, double& Ax,
                                          double& Az, double& Ph)
{
    double inv = 1. / (L.yo - s);
    double dx = L.xo - (L.x0 + L.btx * (s - L.sc)), dz = L.zo - (L.z0 + L.btz * (s - L.sc));
    double Nx = dx * inv, Nz = dz * inv;
    Ph = L.kph * (s * L.gEm2 + dx * Nx + dz * Nz + (L.btx * L.btx + L.btz * L.btz) * (s - L.sc) + L.ib0);
    Ax = (L.btx - Nx) * inv;
    Az = (L.btz - Nz) * inv;
    cos_sin_taylor(Ph, c, sn);
}

static __device__ void left_pre_resid(const free_line& L, double sSt, double sFi, const cplx* dd, double actNorm,
                                      double prec, cplx* R)
{
    const double wfe = 0.46666666666667, wf1 = 1.0666666666667, wf2 = 0.93333333333333, wd = 0.066666666666667;
    const double fivePIdFour = 1.25 * 3.141592653590;
    long long npl = 5;
    double step = (sFi - sSt) / (npl - 1);
    double c, sn, Ax, Az, Ph;
    double fe[4], s1[4] = {0, 0, 0, 0}, s2[4] = {0, 0, 0, 0};
    free_term(L, sSt, c, sn, Ax, Az, Ph);
    fe[0] = Ax * c; fe[1] = Ax * sn; fe[2] = Az * c; fe[3] = Az * sn;
    double phInit = Ph, s = sSt + step;
    free_term(L, s, c, sn, Ax, Az, Ph);
    s1[0] += Ax * c; s1[1] += Ax * sn; s1[2] += Az * c; s1[3] += Az * sn; s += step;
    free_term(L, s, c, sn, Ax, Az, Ph);
    s2[0] += Ax * c; s2[1] += Ax * sn; s2[2] += Az * c; s2[3] += Az * sn; s += step;
    free_term(L, s, c, sn, Ax, Az, Ph);
    s1[0] += Ax * c; s1[1] += Ax * sn; s1[2] += Az * c; s1[3] += Az * sn; s += step;
    free_term(L, s, c, sn, Ax, Az, Ph);
    fe[0] += Ax * c; fe[1] += Ax * sn; fe[2] += Az * c; fe[3] += Az * sn;
    for (int m = 0; m < 4; m++) fe[m] *= wfe;
    double wdd[4] = {wd * dd[0].re, wd * dd[0].im, wd * dd[1].re, wd * dd[1].im};
    double r0[4] = {R[0].re, R[0].im, R[1].re, R[1].im}, out[4];
    double an = -actNorm * step, sq = 0.;
    for (int m = 0; m < 4; m++) {
        out[m] = r0[m] + an * (fe[m] + wf1 * s1[m] + wf2 * s2[m] + step * wdd[m]);
        sq += out[m] * out[m];
    }
    npl--;
    for (int lev = 0; lev < 22; lev++) {
        for (int m = 0; m < 4; m++) { s2[m] += s1[m]; s1[m] = 0.; }
        bool last = true;
        double phPrev = phInit, half = 0.5 * step;
        s = sSt + half;
        for (long long i = 0; i < npl; i++) {
            free_term(L, s, c, sn, Ax, Az, Ph);
            s1[0] += Ax * c; s1[1] += Ax * sn; s1[2] += Az * c; s1[3] += Az * sn;
            s += step;
            if (Ph - phPrev > fivePIdFour) last = false;
            phPrev = Ph;
        }
        double ah = -actNorm * half, lsq = 0.;
        for (int m = 0; m < 4; m++) {
            out[m] = r0[m] + ah * (fe[m] + wf1 * s1[m] + wf2 * s2[m] + half * wdd[m]);
            lsq += out[m] * out[m];
        }
        bool notYet = last ? (fabs(lsq - sq) > prec * lsq) : true;
        sq = lsq;
        step = half; npl *= 2;
        if (!notYet) break;
    }
    R[0] = {out[0], out[1]};
    R[1] = {out[2], out[3]};
}

// Far-field quantities at an integration edge (srradint.cpp:432-503): the
// phase derivative, the amplitudes and their derivatives, the second term of
// the expansion and its acceptance test |t2x| < 0.7 |t1x| (:507-510).
struct edge_geo {
    double kph, inv, dx, dz, bxn, bzn, dPhds, odp, Ax, Az, cBxBz, cBzBx, cxA, czA, tk, d2, dAxds, dAzds, q, t2x, t2z;
    bool ok2;
};

__device__ __forceinline__ void edge_geometry(const srwd_trj_info& I, const double* ev, double s, double xo,
                                              double yo, double zo, double eph, edge_geo& g)
{
    double Bz = ev[1], Btx = ev[2], X = ev[3];
    double Bx = ev[6], Btz = ev[7], Z = ev[8];
    g.kph = SRWD_KPH(eph);
    g.inv = 1. / (yo - s);
    g.dx = xo - X;
    g.dz = zo - Z;
    g.bxn = Btx - g.dx * g.inv;
    g.bzn = Btz - g.dz * g.inv;
    g.dPhds = g.kph * (I.gammaEm2 + g.bxn * g.bxn + g.bzn * g.bzn);
    g.odp = 1. / g.dPhds;
    g.Ax = g.bxn * g.inv;
    g.Az = g.bzn * g.inv;
    g.cBxBz = I.betaNorm * Bz;
    g.cBzBx = -I.betaNorm * Bx;
    g.cxA = g.cBxBz + g.Ax;
    g.czA = g.cBzBx + g.Az;
    g.tk = 2. * g.kph;
    g.d2 = g.tk * (g.bxn * g.cxA + g.bzn * g.czA);
    g.dAxds = (2. * g.Ax + g.cBxBz) * g.inv;
    g.dAzds = (2. * g.Az + g.cBzBx) * g.inv;
    g.q = g.d2 * g.odp;
    g.t2x = (g.dAxds - g.Ax * g.q) * g.odp;
    g.t2z = (g.dAzds - g.Az * g.q) * g.odp;
    const double t1x = g.Ax;
    g.ok2 = ((g.t2x == 0.) && (t1x == 0.)) || ((g.t2x != 0.) && (fabs(g.t2x) < 0.7 * fabs(t1x)));
}

// Phase at the edge with the reference's own sine and cosine (srradint.cpp:449):
// a point whose integral nearly cancels against its end terms depends on them.
__device__ __forceinline__ void edge_phase(const srwd_trj_info& I, const double* ev, double s, const edge_geo& g,
                                           double& c, double& sn)
{
    double Ph = g.kph * (s * I.gammaEm2 + ev[4] + ev[9] + (g.dx * g.dx + g.dz * g.dz) * g.inv);
    cos_sin_taylor(Ph, c, sn);
}

// Derivative of the integrand at the edge (srradint.cpp:841-846).
__device__ __forceinline__ void edge_derivative(const edge_geo& g, double c, double sn, cplx* D)
{
    cplx eip = {c, sn};
    D[0] = cmul({g.dAxds, g.Ax * g.dPhds}, eip);
    D[1] = cmul({g.dAzds, g.Az * g.dPhds}, eip);
}

// Observation coordinates and photon energy of point q (request-local, the
// reference's order: rows, columns, photon energies of the computed region).
__device__ __forceinline__ void req_point(const srwd_req& Rq, const double* coord, long long q, double& xo,
                                          double& zo, double& eph)
{
    int ie = (int)(q % Rq.ne);
    long long r = q / Rq.ne;
    int ix = (int)(r % Rq.nxc), iz = (int)(r / Rq.nxc);
    xo = coord[Rq.xOff + ix];
    zo = coord[Rq.zOff + iz];
    eph = coord[Rq.eOff + ie];
}

// The edge derivative that the reference carries into point p when p's own
// expansion at that edge is rejected: it sets the derivative only where the
// expansion is accepted and keeps the previous point's value otherwise (zero
// before the first accepted point of a call).  So: the derivative of the last
// earlier point with an accepted expansion, found by walking back through p's
// block of SRWD_EDGE_BLOCK points and then from the block table `last`
// (2 values per block: left, right edge).
__device__ __noinline__ void carried_derivative(const srwd_trj_info& I, const double* ev, double s, double yo,
                                                const srwd_req& Rq, const double* coord, const long long* last,
                                                int end, long long p, cplx* D)
{
    const long long b = p / SRWD_EDGE_BLOCK, lo = b * SRWD_EDGE_BLOCK;
    double xo, zo, eph;
    edge_geo g;
    long long q = p - 1;
    for (; q >= lo; q--) {
        req_point(Rq, coord, q, xo, zo, eph);
        edge_geometry(I, ev, s, xo, yo, zo, eph, g);
        if (g.ok2) break;
    }
    if (q < lo) {
        q = b > 0 ? last[2 * (Rq.blk0 + b - 1) + end] : -1;
        if (q < 0) {
            D[0] = {0., 0.};
            D[1] = {0., 0.};
            return;
        }
        req_point(Rq, coord, q, xo, zo, eph);
        edge_geometry(I, ev, s, xo, yo, zo, eph, g);
    }
    double c, sn;
    edge_phase(I, ev, s, g, c, sn);
    edge_derivative(g, c, sn, D);
}

// Result of residual(): the far-field expansion with the edge derivative set,
// the near-field residual (no derivative, no warning), or a rejection (zero
// residual, no derivative; the caller raises the warning).
enum { SRWD_RESID_OK = 1, SRWD_RESID_NEAR = 0, SRWD_RESID_BAD = -1 };

// srradint.cpp:398-871, coordinate representation: 3-term far-field
// expansion.  When the 2-term expansion is rejected at the left edge
// (left = true) and the field there is zero (|B| < 1e-10 T, srtrjdat.h:119),
// the near-field residual of srradint.cpp:589-652, which SRWLIB enables for
// the left edge only (srradint.cpp:3610-3611): the far-field term at a point
// sNew further upstream on the straight line plus the integral from sNew to
// the edge (relative precision nearPrec).
static __device__ int residual(const srwd_trj_info& I, const double* ev, double s, double xo, double yo,
                               double zo, double eph, double actNorm, cplx* R, cplx* D, bool left = false,
                               double nearPrec = 1e-4)
{
    double dBzds = ev[0], Bz = ev[1], Btx = ev[2], X = ev[3], IBx = ev[4];
    double dBxds = ev[5], Bx = ev[6], Btz = ev[7], Z = ev[8], IBz = ev[9];
    edge_geo g;
    edge_geometry(I, ev, s, xo, yo, zo, eph, g);
    double cph, sph;
    edge_phase(I, ev, s, g, cph, sph);
    cplx eip = {cph, sph};
    const double kph = g.kph, dx = g.dx, dz = g.dz, bxn = g.bxn, bzn = g.bzn, dPhds = g.dPhds;
    const double inv = g.inv, odp = g.odp, Ax = g.Ax, Az = g.Az, cBxBz = g.cBxBz, cBzBx = g.cBzBx;
    const double cxA = g.cxA, czA = g.czA, tk = g.tk, d2 = g.d2, dAxds = g.dAxds, dAzds = g.dAzds;
    const double t2x = g.t2x, t2z = g.t2z;
    double cdBz = I.betaNorm * dBzds, cdBx = -I.betaNorm * dBxds;
    double d3 = tk * (cxA * cxA + czA * czA + bxn * (inv * (2. * bxn * inv + cBxBz) + cdBz) +
                      bzn * (inv * (2. * bzn * inv + cBzBx) + cdBx));
    double d2Ax = inv * (3. * dAxds + cdBz), d2Az = inv * (3. * dAzds + cdBx);
    double odp2 = odp * odp, qq = d2 * d2 * odp2;
    double t3x = (-d2Ax + (3. * dAxds * d2 + Ax * d3) * odp - 3. * Ax * qq) * odp2;
    double t3z = (-d2Az + (3. * dAzds * d2 + Az * d3) * odp - 3. * Az * qq) * odp2;
    const bool ok2 = g.ok2;
    bool ok3 = ok2 && (((t3x == 0.) && (t2x == 0.)) || ((t3x != 0.) && (fabs(t3x) < 0.7 * fabs(t2x))));
    if (!ok2) {
        if (!left || !(fabs(Bx) < 1.E-10) || !(fabs(Bz) < 1.E-10)) return SRWD_RESID_BAD;
        // srradint.cpp:591-652
        double gEm2 = I.gammaEm2;
        double sNew1 = s - 2. * (3.141592653590 / (kph * gEm2));
        double sNew2 = 0.5 * (yo - sqrt(yo * yo + 4. * (dx * dx + dz * dz) / (gEm2 + Btx * Btx + Btz * Btz)));
        double sNew = (sNew1 < sNew2) ? sNew1 : sNew2;
        double invN = 1. / (yo - sNew);
        double dxN = xo - (X + Btx * (sNew - s)), dzN = zo - (Z + Btz * (sNew - s));
        double bxN = Btx - dxN * invN, bzN = Btz - dzN * invN;
        double dPhN = kph * (gEm2 + bxN * bxN + bzN * bzN);
        double two2 = 2. * invN * invN;
        double AxN = bxN * invN, AzN = bzN * invN;
        double dAxN = bxN * two2, dAzN = bzN * two2;
        double d2N = tk * (bxN * AxN + bzN * AzN);
        double odN = 1. / dPhN, qN = d2N * odN;
        double PhN = kph * (sNew * gEm2 + (IBx + IBz) + (Btx * Btx + Btz * Btz) * (sNew - s) +
                            (dxN * dxN + dzN * dzN) * invN);
        double cN, sN, c0 = cph, s0 = sph;
        cos_sin_taylor(PhN, cN, sN);
        cplx dd[2];
        dd[0] = {(dAxN * cN - AxN * dPhN * sN) - (dAxds * c0 - Ax * dPhds * s0),
                 (AxN * dPhN * cN + dAxN * sN) - (Ax * dPhds * c0 + dAxds * s0)};
        dd[1] = {(dAzN * cN - AzN * dPhN * sN) - (dAzds * c0 - Az * dPhds * s0),
                 (AzN * dPhN * cN + dAzN * sN) - (Az * dPhds * c0 + dAzds * s0)};
        double pXr = (AxN * qN - dAxN) * odN * odN, pXi = AxN * odN;
        double pZr = (AzN * qN - dAzN) * odN * odN, pZi = AzN * odN;
        if (fabs(pXi) < fabs(pXr) || fabs(pZi) < fabs(pZr)) return SRWD_RESID_BAD;
        double ac = actNorm * cN, as = actNorm * sN;
        R[0] = {pXr * ac - pXi * as, pXi * ac + pXr * as};
        R[1] = {pZr * ac - pZi * as, pZi * ac + pZr * as};
        free_line L = {s, Btx, X, Btz, Z, IBx + IBz, xo, yo, zo, kph, gEm2};
        left_pre_resid(L, sNew, s, dd, actNorm, nearPrec, R);
        return SRWD_RESID_NEAR;
    }
    cplx pX = {Ax, t2x}, pZ = {Az, t2z};
    if (ok3) { pX.re += t3x; pZ.re += t3z; }
    // multiply by i/dPhds then by exp(i Ph)
    cplx mX = {-pX.im * odp, pX.re * odp}, mZ = {-pZ.im * odp, pZ.re * odp};
    cplx ex = cmul(mX, eip), ez = cmul(mZ, eip);
    R[0] = {actNorm * ex.re, actNorm * ex.im};
    R[1] = {actNorm * ez.re, actNorm * ez.im};
    edge_derivative(g, cph, sph, D);
    return SRWD_RESID_OK;
}

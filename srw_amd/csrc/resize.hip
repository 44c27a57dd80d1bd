// resize.hip - resampling of wavefront slices onto a new transverse mesh
// (the interpolation of SRW's RadResizeCore) for gfx950.
//
// Reference algorithms (paths relative to the reference source root):
//   range-only enlargement   cpp/src/core/sroptelm.cpp:2850-2920
//   cell interpolation       cpp/src/core/sroptelm.cpp:3038-3300, sroptelm.h:460-760
//   cell constants           cpp/src/core/srstraux.h:1588-1683
//   phase treatment          cpp/src/core/sroptelm.cpp:4774-4990
//
// One thread per new element (ix, iz, ie), energy fastest so that threads of
// a wave read neighbouring words of the interleaved layout.  Each thread
// loads the 4x4 cell of both components (64 floats, L1/L2 resident across
// neighbouring threads), builds the bicubic forms of Re, Im and of the
// intensity, and rescales the complex value to the interpolated intensity.
// The cell sums keep the reference's float arithmetic (products and sums of
// floats, then double), so values match to rounding.
#include "srw_optics.h"
#include "srw_resize.h"

namespace {

// Element order of a launch over (ix, iz, ie): x fastest when the layout's
// x stride is 1 (slice-major fields), energy fastest otherwise (SRW's
// order), so that consecutive threads touch consecutive addresses.
__device__ __forceinline__ void split_index(long long i, const srwr_layout& l, int nx, int nz, int ne, int& ix,
                                            int& iz, int& ie)
{
    if (l.perX == 1) {
        ix = (int)(i % nx);
        long long q = i / nx;
        iz = (int)(q % nz);
        ie = (int)(q / nz);
    } else {
        ie = (int)(i % ne);
        long long p = i / ne;
        ix = (int)(p % nx);
        iz = (int)(p / nx);
    }
}

__device__ __forceinline__ void phase_cs(const srwr_phase& P, int ix, int iz, int ie, float& c, float& s)
{
    double ePh = P.eStart + P.eStep * ie;
    double ph = 0.;
    if (P.fz) { double z = (P.zStart - P.zc) + iz * P.zStep; ph = P.sign * (SRW_QUAD_CONST / P.Rz) * ePh * z * z; }
    if (P.fx) { double x = (P.xStart - P.xc) + ix * P.xStep; ph += P.sign * (SRW_QUAD_CONST / P.Rx) * ePh * x * x; }
    cos_sin_optics(ph, c, s);
}

__global__ void k_phase(float2* ex, float2* ez, srwr_phase P)
{
    long long n = (long long)P.nx * P.nz * P.ne;
    long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    int ix, iz, ie;
    split_index(i, P.l, P.nx, P.nz, P.ne, ix, iz, ie);
    float c, s;
    phase_cs(P, ix, iz, ie, c, s);
    long long o = ie * P.l.perE + iz * P.l.perZ + ix * P.l.perX;
    if (ex) ex[o] = rotate_f(ex[o], c, s);
    if (ez) ez[o] = rotate_f(ez[o], c, s);
}

// Bicubic coefficients a[4*px + pz] of a 4x4 float cell F[4*iz + ix]
// (sroptelm.h:526-546): sums in float, scaled in double.
__device__ __forceinline__ void bicubic(const float* F, double* a)
{
#pragma clang fp contract(off)
    const float f00 = F[0], f10 = F[1], f20 = F[2], f30 = F[3];
    const float f01 = F[4], f11 = F[5], f21 = F[6], f31 = F[7];
    const float f02 = F[8], f12 = F[9], f22 = F[10], f32 = F[11];
    const float f03 = F[12], f13 = F[13], f23 = F[14], f33 = F[15];
    const double c6 = 0.1666666667, c36 = 0.027777777778, c12 = 0.083333333333;
    a[0] = f11;
    a[1] = (-2 * f10 - 3 * f11 + 6 * f12 - f13) * c6;
    a[2] = (f10 + f12 - 2 * f11) * 0.5;
    a[3] = (f13 - f10 + 3 * (f11 - f12)) * c6;
    a[4] = (-2 * f01 - 3 * f11 + 6 * f21 - f31) * c6;
    a[5] = (4 * f00 + 6 * (f01 + f10 - f23 - f32) - 12 * (f02 + f20) + 2 * (f03 + f30) + 9 * f11 - 18 * (f12 + f21) +
            3 * (f13 + f31) + 36 * f22 + f33) * c36;
    a[6] = (-2 * (f00 + f02 - f31) + 4 * f01 - 3 * (f10 + f12) + 6 * (f11 + f20 + f22) - 12 * f21 - f30 - f32) * c12;
    a[7] = (2 * (f00 - f03) + 6 * (-f01 + f02 - f20 + f23) + 3 * (f10 - f13 - f31 + f32) + 9 * (f12 - f11) +
            18 * (f21 - f22) + f30 - f33) * c36;
    a[8] = (f01 + f21 - 2 * f11) * 0.5;
    a[9] = (2 * (-f00 + f13 - f20) - 3 * (f21 + f01) + 6 * (f02 + f11 + f22) + 4 * f10 - 12 * f12 - f23 - f03) * c12;
    a[10] = (f00 + f02 + f22 + f20 - 2 * (f01 + f10 + f12 + f21) + 4 * f11) * 0.25;
    a[11] = (-f00 + f03 - f20 + f23 + 3 * (f01 - f02 + f21 - f22) + 2 * (f10 - f13) + 6 * (f12 - f11)) * c12;
    a[12] = (f31 - f01 + 3 * (f11 - f21)) * 0.16666666667;
    a[13] = (2 * (f00 - f30) + 3 * (f01 - f13 + f23 - f31) + 6 * (-f02 - f10 + f20 + f32) + 9 * (f21 - f11) +
             18 * (f12 - f22) + f03 - f33) * c36;
    a[14] = (f30 + f32 - f00 - f02 + 2 * (f01 - f31) + 3 * (f10 + f12 - f20 - f22) + 6 * (f21 - f11)) * c12;
    a[15] = (f00 - f03 - f30 + f33 + 3 * (-f01 + f02 - f10 + f13 + f20 - f23 + f31 - f32) +
             9 * (f11 - f12 - f21 + f22)) * c36;
}

__device__ __forceinline__ double bicubic_eval(const double* a, double x, double z)
{   // sroptelm.h:594-606 term order
#pragma clang fp contract(off)
    double xE2 = x * x, xz = x * z, zE2 = z * z;
    double xE3 = xE2 * x, xE2z = xE2 * z, xzE2 = x * zE2, zE3 = zE2 * z, xE2zE2 = xE2 * zE2;
    double xE3z = xE3 * z, xE3zE2 = xE3 * zE2, xE3zE3 = xE3 * zE3, xE2zE3 = xE2 * zE3, xzE3 = x * zE3;
    return a[15] * xE3zE3 + a[14] * xE3zE2 + a[13] * xE3z + a[12] * xE3 + a[11] * xE2zE3 + a[10] * xE2zE2 +
           a[9] * xE2z + a[8] * xE2 + a[7] * xzE3 + a[6] * xzE2 + a[5] * xz + a[4] * x + a[3] * zE3 + a[2] * zE2 +
           a[1] * z + a[0];
}

__device__ __forceinline__ double bilinear_eval(const float* F, int ix, int iz, double x, double z)
{   // sroptelm.h:563-569, 650-657
#pragma clang fp contract(off)
    float f00 = F[4 * iz + ix], f10 = F[4 * iz + ix + 1], f01 = F[4 * iz + 4 + ix], f11 = F[4 * iz + 4 + ix + 1];
    double a0 = f00, a10 = 1. * (f10 - f00), a01 = 1. * (f01 - f00), a11 = 1. * (f00 - f01 - f10 + f11);
    return a11 * (x * z) + a10 * x + a01 * z + a0;
}

// Intensity of the 16 cell values, scaled to an average of at most 1, and the
// scale (sroptelm.cpp:3150-3200); low: the cell holds an exact zero, which
// selects the bilinear form (sroptelm.h:650-657).
__device__ __forceinline__ void cell_intensity(const float* re, const float* im, float* In, float& norm, bool& low)
{
    for (int k = 0; k < 16; k++) In[k] = __fadd_rn(__fmul_rn(re[k], re[k]), __fmul_rn(im[k], im[k]));
    float sum = In[0];
    for (int k = 1; k < 16; k++) sum = __fadd_rn(sum, In[k]);
    float avg = (float)(0.0625 * (double)sum);
    norm = 1.f;
    if (fabsf(avg) > 1.f) {
        float a = (float)(1. / avg);
        for (int k = 0; k < 16; k++) In[k] = __fmul_rn(In[k], a);
        norm = avg;
    }
    low = false;
    for (int k = 0; k < 16; k++) low = low || (re[k] == 0.f) || (im[k] == 0.f);
}

__device__ __forceinline__ float2 bilinear_point(const float* re, const float* im, const float* In, int ixRel,
                                                 int izRel, double xr, double zr, float& FI)
{
    int ix = ixRel < 0 ? 0 : (ixRel > 2 ? 2 : ixRel);
    int iz = izRel < 0 ? 0 : (izRel > 2 ? 2 : izRel);
    float F0 = (float)bilinear_eval(re, ix, iz, xr, zr);
    float F1 = (float)bilinear_eval(im, ix, iz, xr, zr);
    double b = bilinear_eval(In, ix, iz, xr, zr);
    FI = (float)((b > 0.) ? b : 0.);
    return make_float2(F0, F1);
}

__device__ __forceinline__ float2 bicubic_point(const double* aRe, const double* aIm, const double* aIn, double xr,
                                                double zr, float& FI)
{
    float F0 = (float)bicubic_eval(aRe, xr, zr);
    float F1 = (float)bicubic_eval(aIm, xr, zr);
    double b = bicubic_eval(aIn, xr, zr);
    FI = (float)((b > 0.) ? b : 0.);
    return make_float2(F0, F1);
}

// Rescales the interpolated complex value to the interpolated intensity
// (sroptelm.h:688-697).
__device__ __forceinline__ float2 finish_point(float2 F, float FI, float norm, bool improve)
{
    float F0 = F.x, F1 = F.y;
    if (!improve) return F;
    FI = __fmul_rn(FI, norm);
    float appI = __fadd_rn(__fmul_rn(F0, F0), __fmul_rn(F1, F1));
    if (appI != 0.f) {
        float f = (float)sqrt((double)__fdiv_rn(FI, appI));
        F0 = __fmul_rn(F0, f); F1 = __fmul_rn(F1, f);
    }
    return make_float2(F0, F1);
}

// One component at relative cell position (xr, zr); re/im: the 16 cell values.
__device__ __forceinline__ float2 resample_point(const float* re, const float* im, int ixRel, int izRel, double xr,
                                                 double zr, bool improve)
{
    float In[16], norm, FI;
    bool low;
    cell_intensity(re, im, In, norm, low);
    float2 F;
    if (!low) {
        double aRe[16], aIm[16], aIn[16];
        bicubic(re, aRe); bicubic(im, aIm); bicubic(In, aIn);
        F = bicubic_point(aRe, aIm, aIn, xr, zr, FI);
    } else F = bilinear_point(re, im, In, ixRel, izRel, xr, zr, FI);
    return finish_point(F, FI, norm, improve);
}

// Interval of the old mesh that holds abs, the first node of its 4-point
// stencil and the relative coordinate in steps (sroptelm.cpp:3094-3140).
__device__ __forceinline__ void cell_of(double abs, double start, double step, int n, int& c, int& st, double& rel)
{
    double is = 1. / step;
    c = (int)((abs - start) * is + 1.E-06);
    rel = abs - (start + c * step);
    if (c == n - 1) { st = c - 3; rel += 2. * step; }
    else if (c == n - 2) { st = c - 2; rel += step; }
    else if (c == 0) { st = c; rel -= step; }
    else st = c - 1;
    rel *= is;
    // a full 4x4 cell inside the old mesh (the host requires nx, nz >= 4)
    st = st < 0 ? 0 : (st > n - 4 ? n - 4 : st);
}

// Whether new point (ix, iz) at (xAbs, zAbs) takes a value from the old mesh
// (otherwise it stays zero, srradstr.cpp:4726, 4761-4775).
__device__ __forceinline__ bool filled_point(const srwr_params& P, int ix, int iz, double xAbs, double zAbs)
{
    if (ix < P.ix0 || ix > P.ix1 || iz < P.iz0 || iz > P.iz1) return false;
    if (P.zeroOutside && (zAbs < P.zMin - P.zTol || zAbs > P.zMax + P.zTol || xAbs < P.xMin - P.xTol ||
                          xAbs > P.xMax + P.xTol))
        return false;
    return true;
}

__global__ void __launch_bounds__(256) k_resize(const float2* __restrict__ oex, const float2* __restrict__ oez,
                                                float2* nex, float2* nez, srwr_params P)
{
    long long n = (long long)P.nxn * P.nzn * P.ne;
    long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    int ix, iz, ie;
    split_index(i, P.ln, P.nxn, P.nzn, P.ne, ix, iz, ie);
    long long on = ie * P.ln.perE + iz * P.ln.perZ + ix * P.ln.perX;
    float2 vx = make_float2(0.f, 0.f), vz = make_float2(0.f, 0.f);
    double zAbs = P.zStartN + iz * P.zStepN, xAbs = P.xStartN + ix * P.xStepN;
    if (filled_point(P, ix, iz, xAbs, zAbs)) {
        if (P.onlyLarger) {
            long izo = (long)((zAbs - P.zStartO) * (1. / P.zStepO) + 1.E-08);
            long ixo = (long)((xAbs - P.xStartO) * (1. / P.xStepO) + 1.E-08);
            izo = izo < 0 ? 0 : (izo > P.nzo - 1 ? P.nzo - 1 : izo);
            ixo = ixo < 0 ? 0 : (ixo > P.nxo - 1 ? P.nxo - 1 : ixo);
            long long oo = ie * P.lo.perE + izo * P.lo.perZ + ixo * P.lo.perX;
            if (oex) vx = oex[oo];
            if (oez) vz = oez[oo];
        } else {
            if (P.cellMode) return;   // written by k_resize_cells
            int izc, izSt, ixc, ixSt;
            double zRel, xRel;
            cell_of(zAbs, P.zStartO, P.zStepO, P.nzo, izc, izSt, zRel);
            cell_of(xAbs, P.xStartO, P.xStepO, P.nxo, ixc, ixSt, xRel);
            long long base = ie * P.lo.perE + izSt * P.lo.perZ + ixSt * P.lo.perX;
            float re[16], im[16];
            if (oex) {
                for (int j = 0; j < 4; j++)
                    for (int k = 0; k < 4; k++) {
                        float2 v = oex[base + j * P.lo.perZ + k * P.lo.perX];
                        re[4 * j + k] = v.x; im[4 * j + k] = v.y;
                    }
                vx = resample_point(re, im, ixc - ixSt, izc - izSt, xRel, zRel, P.improve != 0);
            }
            if (oez) {
                for (int j = 0; j < 4; j++)
                    for (int k = 0; k < 4; k++) {
                        float2 v = oez[base + j * P.lo.perZ + k * P.lo.perX];
                        re[4 * j + k] = v.x; im[4 * j + k] = v.y;
                    }
                vz = resample_point(re, im, ixc - ixSt, izc - izSt, xRel, zRel, P.improve != 0);
            }
        }
    }
    bool zero = vx.x == 0.f && vx.y == 0.f && vz.x == 0.f && vz.y == 0.f;
    if (P.hasPost && !zero) {   // the same rounding as k_phase applied to the stored values
        float c, s;
        phase_cs(P.post, ix, iz, ie, c, s);
        vx = rotate_f(vx, c, s);
        vz = rotate_f(vz, c, s);
    }
    if (nex) nex[on] = vx;
    if (nez) nez[on] = vz;
}

// Interpolated points, one thread per (old interval cx, old interval cz,
// energy): the bicubic forms of the cell are built once and evaluated at every
// new point whose interval is (cx, cz), with the per-point arithmetic of
// k_resize (which writes all other points).  Candidate new indices come from
// inverting the interval rule with one point of margin; each is kept only
// when its own interval computation gives (cx, cz).
__global__ void __launch_bounds__(256) k_resize_cells(const float2* __restrict__ oex, const float2* __restrict__ oez,
                                                      float2* nex, float2* nez, srwr_params P, int cx0, int ncx,
                                                      int cz0, int ncz)
{
    long long n = (long long)ncx * ncz * P.ne;
    long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    int cx, cz, ie;
    if (P.lo.perX == 1) { cx = (int)(i % ncx); long long q = i / ncx; cz = (int)(q % ncz); ie = (int)(q / ncz); }
    else { ie = (int)(i % P.ne); long long q = i / P.ne; cx = (int)(q % ncx); cz = (int)(q / ncx); }
    cx += cx0; cz += cz0;
    auto cand = [](int c, double startO, double stepO, double startN, double stepN, int lo, int hi, int& a, int& b) {
        double l = startO + c * stepO;
        a = (int)floor((l - startN) / stepN) - 1;
        b = (int)ceil((l + stepO - startN) / stepN) + 1;
        a = a < lo ? lo : a;
        b = b > hi ? hi : b;
    };
    int ixA, ixB, izA, izB;
    cand(cx, P.xStartO, P.xStepO, P.xStartN, P.xStepN, P.ix0, P.ix1, ixA, ixB);
    cand(cz, P.zStartO, P.zStepO, P.zStartN, P.zStepN, P.iz0, P.iz1, izA, izB);
    if (ixA > ixB || izA > izB) return;
    // the new indices of this interval pair (contiguous: the interval is
    // monotone in the index) and the cell they share
    int ixSt = -1, izSt = -1, ixF = 0, ixL = -1, izF = 0, izL = -1;
    for (int iz = izA; iz <= izB; iz++) {
        int c, st; double r;
        cell_of(P.zStartN + iz * P.zStepN, P.zStartO, P.zStepO, P.nzo, c, st, r);
        if (c == cz) { if (izSt < 0) { izSt = st; izF = iz; } izL = iz; }
    }
    for (int ix = ixA; ix <= ixB; ix++) {
        int c, st; double r;
        cell_of(P.xStartN + ix * P.xStepN, P.xStartO, P.xStepO, P.nxo, c, st, r);
        if (c == cx) { if (ixSt < 0) { ixSt = st; ixF = ix; } ixL = ix; }
    }
    if (ixSt < 0 || izSt < 0) return;
    const long long base = ie * P.lo.perE + izSt * P.lo.perZ + ixSt * P.lo.perX;
    for (int comp = 0; comp < 2; comp++) {
        const float2* o = comp == 0 ? oex : oez;
        float2* d = comp == 0 ? nex : nez;
        if (o == nullptr || d == nullptr) continue;
        float re[16], im[16], In[16], norm;
        bool low;
        for (int j = 0; j < 4; j++)
            for (int k = 0; k < 4; k++) {
                float2 v = o[base + j * P.lo.perZ + k * P.lo.perX];
                re[4 * j + k] = v.x; im[4 * j + k] = v.y;
            }
        cell_intensity(re, im, In, norm, low);
        // every new point of the cell, its value from `value` (the forms, or
        // the bilinear fallback); the two cases in separate loops so that the
        // cell values are dead once the forms are built
        auto points = [&](auto value) {
            for (int iz = izF; iz <= izL; iz++) {
                double zAbs = P.zStartN + iz * P.zStepN;
                int izc, st; double zRel;
                cell_of(zAbs, P.zStartO, P.zStepO, P.nzo, izc, st, zRel);
                for (int ix = ixF; ix <= ixL; ix++) {
                    double xAbs = P.xStartN + ix * P.xStepN;
                    int ixc, st2; double xRel;
                    cell_of(xAbs, P.xStartO, P.xStepO, P.nxo, ixc, st2, xRel);
                    if (!filled_point(P, ix, iz, xAbs, zAbs)) continue;
                    float FI;
                    float2 F = value(ixc - ixSt, izc - izSt, xRel, zRel, FI);
                    F = finish_point(F, FI, norm, P.improve != 0);
                    if (P.hasPost) {
                        float c, s;
                        phase_cs(P.post, ix, iz, ie, c, s);
                        F = rotate_f(F, c, s);
                    }
                    d[ie * P.ln.perE + iz * P.ln.perZ + ix * P.ln.perX] = F;
                }
            }
        };
        if (low) {
            points([&](int ixr, int izr, double xr, double zr, float& FI) {
                return bilinear_point(re, im, In, ixr, izr, xr, zr, FI);
            });
        } else {
            double aRe[16], aIm[16], aIn[16];
            bicubic(re, aRe); bicubic(im, aIm); bicubic(In, aIn);
            points([&](int, int, double xr, double zr, float& FI) { return bicubic_point(aRe, aIm, aIn, xr, zr, FI); });
        }
    }
}

}  // namespace

extern "C" {
int srwr_launch_phase(float2* ex, float2* ez, const srwr_phase& P, hipStream_t st)
{
    long long n = (long long)P.nx * P.nz * P.ne;
    if (n <= 0) return 0;
    hipLaunchKernelGGL(k_phase, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, ex, ez, P);
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

int srwr_launch_resize(const float2* oex, const float2* oez, float2* nex, float2* nez, const srwr_params& P,
                       hipStream_t st)
{
    long long n = (long long)P.nxn * P.nzn * P.ne;
    if (n <= 0) return 0;
    if (!P.onlyLarger && (P.nxo < 4 || P.nzo < 4)) return -1;
    hipLaunchKernelGGL(k_resize, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, oex, oez, nex, nez, P);
    if (P.cellMode && !P.onlyLarger) {
        // interval range of the filled new points (monotone in the index), one interval of margin
        auto interval = [](double abs, double start, double step) { return (int)((abs - start) * (1. / step) + 1.E-06); };
        int cx0 = interval(P.xStartN + P.ix0 * P.xStepN, P.xStartO, P.xStepO) - 1;
        int cx1 = interval(P.xStartN + P.ix1 * P.xStepN, P.xStartO, P.xStepO) + 1;
        int cz0 = interval(P.zStartN + P.iz0 * P.zStepN, P.zStartO, P.zStepO) - 1;
        int cz1 = interval(P.zStartN + P.iz1 * P.zStepN, P.zStartO, P.zStepO) + 1;
        if (cx1 >= cx0 && cz1 >= cz0) {
            int ncx = cx1 - cx0 + 1, ncz = cz1 - cz0 + 1;
            long long m = (long long)ncx * ncz * P.ne;
            hipLaunchKernelGGL(k_resize_cells, dim3((unsigned)((m + 255) / 256)), dim3(256), 0, st, oex, oez, nex, nez,
                               P, cx0, ncx, cz0, ncz);
        }
    }
    return hipGetLastError() == hipSuccess ? 0 : -1;
}
}

// srw_fft_ops.h - pointwise operations of the drift and of the SRW-convention
// FFT, shared by the power-of-2 passes (fft.hip) and the mixed-radix passes
// (fft_gen.hip), so both compute the same values.
#pragma once
#include "srw_fft.h"

__device__ __forceinline__ float2 cmulf(float2 a, float2 b)
{
    return make_float2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}

__device__ __forceinline__ float2 cis_rev(double rev)
{   // exp(i 2pi rev) with the argument reduced in FP64
    float r = (float)(rev - rint(rev));
    return make_float2(__builtin_amdgcn_cosf(r), __builtin_amdgcn_sinf(r));
}

// Input of the forward transform of element (ix, iy) of an nx x ny slice:
// preOp 1 removes the quadratic phase (sroptdrf.h:812-821), preOp 2 applies
// the linear pre-shift of a non-centred input (gmfft.h:250-276); then the
// sign (-1)^(ix+iy) of the centred transform.
__device__ __forceinline__ float2 srwf_pre(const srwf_slice& S, int preOp, int ix, int iy, int nx, int ny, float2 v)
{
    if (preOp == 1) {
        double rx = S.xStart + ix * S.xStep - S.xc, rz = S.zStart + iy * S.zStep - S.zc;
        v = cmulf(v, cis_rev(S.pre_rev * (S.invRx * rx * rx + S.invRz * rz * rz)));
    } else if (preOp == 2) {
        v = cmulf(v, cis_rev(S.sh0x * ((ix - (nx >> 1)) * S.xStep) + S.sh0y * ((iy - (ny >> 1)) * S.zStep)));
    }
    if ((ix + iy) & 1) { v.x = -v.x; v.y = -v.y; }
    return v;
}

// Angular-space kernel at natural-order DFT indices (kx, ky), with the
// normalisation: op 0 angular representation (sroptdrf.h:615-644), op 1
// analytic quadratic-phase treatment (sroptdrf.h:835-845).
__device__ __forceinline__ float2 srwf_mid(const srwf_slice& S, int op, int kx, int ky, int NX, int NY, float2 v)
{
    double qx = (kx - (NX >> 1)) * S.qStepX, qz = (ky - (NY >> 1)) * S.qStepZ;
    double rev;
    if (op == 0) {
        double q2 = qx * qx + qz * qz, cq = S.c1 * q2;
        rev = S.c0_rev * q2 * (1. + cq + 2. * cq * cq) + S.path_rev;
    } else {
        rev = S.c0x_rev * qx * qx + S.c0z_rev * qz * qz + S.phT_rev;
    }
    float2 w = cis_rev(rev);
    w.x *= S.midScale; w.y *= S.midScale;
    return cmulf(v, w);
}

// Output multiplier of the inverse transform at (ix, iy): the sign
// (-1)^(ix+iy) and, for postOp >= 1, the new quadratic phase of the analytic
// mode (sroptdrf.h:846-860).
__device__ __forceinline__ float2 srwf_post(const srwf_slice& S, int postOp, int ix, int iy)
{
    float2 m = make_float2(((ix + iy) & 1) ? -1.f : 1.f, 0.f);
    if (postOp >= 1) {
        double rx = S.nxStart + ix * S.nxStep - S.xc, rz = S.nzStart + iy * S.nzStep - S.zc;
        float2 w = cis_rev(S.post_rev * (S.invRxL * rx * rx + S.invRzL * rz * rz) + S.path3_rev);
        m = make_float2(m.x * w.x, m.x * w.y);
    }
    return m;
}

// Standalone centred transform value at natural-order indices (kx, ky): the
// sign (-1)^(kx+ky+NX/2+NY/2), the step product and the post-shift
// (gmfft.cpp:579-615, gmfft.h:420-472).
__device__ __forceinline__ float2 srwf_std_out(const srwf_slice& S, int kx, int ky, int NX, int NY, float2 v)
{
    float sg = ((kx + ky + (NX >> 1) + (NY >> 1)) & 1) ? -S.midScale : S.midScale;
    v.x *= sg; v.y *= sg;
    if (S.postShift) {
        double rev = S.sh1x * ((kx - (NX >> 1)) * S.qStepX) + S.sh1y * ((ky - (NY >> 1)) * S.qStepZ);
        v = cmulf(v, cis_rev(rev));
    }
    return v;
}

// srw_fft.h - per-slice parameters and launch wrappers of the FFT / drift passes.
#pragma once
#include <hip/hip_runtime.h>

// Parameters of one N x N complex slice (one field component at one photon
// energy).  Phases are in revolutions; see fft.hip for the pass structure.
struct srwf_slice {
    double xStart, xStep, zStart, zStep;        // input coordinate mesh
    double nxStart, nxStep, nzStart, nzStep;    // output coordinate mesh (mode 3 rescaled)
    double xc, zc;                              // wavefront centre
    double invRx, invRz, invRxL, invRzL;        // 1/R before and 1/(R+L) after the drift
    double pre_rev, post_rev, path3_rev;        // -pi/lambda, +pi/lambda (per 2pi), absolute-path term
    double qStepX, qStepZ;                      // angular-representation steps 1/(N dx)
    double c0_rev, c1, path_rev;                // mode 0 kernel
    double c0x_rev, c0z_rev, phT_rev;           // mode 3 kernel
    double sh0x, sh0y, sh1x, sh1y;              // linear-phase shifts of the standalone FFT
    float midScale;
    int postShift;
    int ny;
    int pad;
};

extern "C" {
int srwf_rows_per_wg(int logn);
int srwf_cols_per_wg(int logn);
int srwf_launch_rows_fwd(float2* data, long long sliceStride, int nSlices, const srwf_slice* dSp, const float2* tw,
                         int logn, int ny, int preOp, int inv, hipStream_t st);
int srwf_launch_cols(float2* data, long long sliceStride, int nSlices, const srwf_slice* dSp, const float2* twy,
                     int logny, int lognx, int op, int inv, float2* out, hipStream_t st);
int srwf_launch_rows_inv(float2* data, long long sliceStride, int nSlices, const srwf_slice* dSp, const float2* tw,
                         int logn, int ny, int postOp, float* inten, long long intenStride, hipStream_t st);
int srwf_launch_gather(const float2* src, float2* dst, int ne, long long nxy, int scatter, hipStream_t st);
// Both field components between SRW's [iy][ix][ie] order (ex, ez) and the
// slice-major d[(2*ie + c)*nxy + p]: toSlices = 1 reads ex/ez and writes d,
// 0 reads d and writes ex/ez.
int srwf_launch_slices(float2* ex, float2* ez, float2* d, int ne, long long nxy, int toSlices, hipStream_t st);
int srwf_init_attributes();
}

// srw_resize.h - resampling of a wavefront onto a new transverse mesh.
#pragma once
#include <hip/hip_runtime.h>

// Element (ix, iz, ie) of a component is at p[ie*perE + iz*perZ + ix*perX]
// (float2 units).
struct srwr_layout {
    long long perX, perZ, perE;
};

// Quadratic phase exp(sign * i pi/lambda ((x-xc)^2/Rx + (z-zc)^2/Rz)) applied
// in place with float storage (sign -1 removes, +1 adds).
struct srwr_phase {
    srwr_layout l;
    int nx, nz, ne;
    double xStart, xStep, zStart, zStep, eStart, eStep;
    double Rx, Rz, xc, zc;
    int fx, fz, sign;
};

struct srwr_params {
    srwr_layout lo, ln;                     // old / new layouts
    int nxo, nzo, nxn, nzn, ne;
    double xStartO, xStepO, zStartO, zStepO;
    double xStartN, xStepN, zStartN, zStepN;
    double eStart, eStep;
    int ix0, ix1, iz0, iz1;                 // filled range of the new mesh
    int onlyLarger;                         // copy without interpolation
    int hasPost;                            // apply `post` to each new value before it is stored
    srwr_phase post;                        // on the new mesh and layout
    int improve;                            // rescale Re, Im to the interpolated intensity
    int zeroOutside;                        // zero new points outside [min - tol, max + tol] of the old mesh
    double xMin, xMax, zMin, zMax, xTol, zTol;
    int cellMode;                           // interpolate per old cell (k_resize_cells) instead of per new point
};

extern "C" {
int srwr_launch_phase(float2* ex, float2* ez, const srwr_phase& P, hipStream_t st);
int srwr_launch_resize(const float2* oex, const float2* oez, float2* nex, float2* nez, const srwr_params& P,
                       hipStream_t st);
}

// srw_prop.h - host helpers of the drift propagation shared by the C ABI files.
#pragma once
#include "srw_fft.h"

namespace srwamd {

// Drift buffer variables of the analytic quadratic-phase mode
// (reference sroptdrf.cpp:889-1081).
struct DriftGeom {
    double invRx, invRz, Lx, Lz, invRxL, invRzL, kx, kxc, kz, kzc, mult, phT;
};

const float2* twiddles(int logn);
int fft_ready();
int drift_mode_from_prop(const double* pp, int& mode, int& sub);
void mode3_setup(double Rx, double Rz, double L, DriftGeom& g);
void drift_slice(srwf_slice& S, int mode, const DriftGeom& g, double L, int treat, double e, long nx, long nz,
                 double xStart, double xStep, double zStart, double zStep, double xc, double zc);
int drift_passes(float2* d, long long sliceStride, int nSlices, const srwf_slice* dSp, int lognx, int lognz,
                 int mode, int postOp, float* inten, long long intenStride, hipStream_t st);

}  // namespace srwamd

// One extraction request of srwlCalcIntFromElecField: component what
// (0 intensity, 3 Re E, 6 Im E) of polarisation pol (-5: the four Stokes
// parameters one after another) vs dependence dep, with the (i0, i1, t)
// interpolation of every coordinate that is not varied.
struct srwx_extract {
    int pol, what, dep, dbl;      // what: the reference's internal Int_or_ReE; dbl: double output (phase)
    int intOverE, pad;            // integrate the intensity over photon energy at each output point
    long long ne, nx, nz, nout;
    long long ie0, ie1, ix0, ix1, iz0, iz1;
    double te, tx, tz;
    double eStep, xStep, zStep;
    double iter, invIter;         // arMeth: 0 overwrite, > 0 running average over iter values, < 0 sum
};

// Pointwise element of a container on the slice-major field (slice 2 ie +
// component holds [iz][ix]): kind 0 thin lens (p0, p1 = Fx, Fy), 1 / 2
// rectangular aperture / obstacle (p0, p1 = Dx, Dy), 3 / 4 circular aperture /
// obstacle (p0 = diameter); (cx, cy) the element's centre.
struct srwo_elem {
    int kind, pad;
    long long nx;
    double p0, p1, cx, cy;
    double xStart, xStep, zStart, zStep, eStart, eStep;
};

extern "C" {
int srwo_launch_point_elem(float2* d, long long nxy, int ne, const srwo_elem& E, hipStream_t st);
// rows: ne * nz doubles of scratch for the flux (what == 4)
int srwx_launch_extract(const float2* ex, const float2* ez, const srwx_extract& P, void* out, double* rows,
                        hipStream_t st);
int srwp_launch_intensity(const float2* ex, const float2* ez, int ne, int ie0, int ie1, float t, int pol,
                          long long nxy, float* out, hipStream_t st);
}

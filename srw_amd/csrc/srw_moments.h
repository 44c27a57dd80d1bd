// srw_moments.h - wavefront statistical moments on the device.
#pragma once
#include <hip/hip_runtime.h>

// One photon-energy slice of a wavefront.  Element (ix, iz) of a component
// is at ex[iz*perZ + ix*perX] (float2 units); either pointer may be null.
struct srwm_mslice {
    const float2* ex;
    const float2* ez;
    long long perX, perZ;
    int nx, nz;
    double xStart, xStep, zStart, zStep, ePh;
    double Rx, Rz, xc, zc;    // radius and centre of the quadratic phase removed before the sums
    int fx, fz;               // remove the x / z part of that phase
};

#define SRWM_NSUM 22

extern "C" {
// Device work space (bytes) of srwm_launch_moments for these sizes.
long long srwm_work_bytes(int nSlices, int nxMax, int nzMax, int lanesOverSlices);
// For each slice: the 22 moment sums (sums[s*24 + k]) and the index limits
// of the 90 % power rectangle (lims[s*4 + k]).  nxMax / nzMax bound every
// slice's nx / nz.  lanesOverSlices selects the mapping for slices s whose
// data lie at base + s (SRW's energy-fastest order); otherwise each slice's
// rows should be contiguous in x.
int srwm_launch_moments(const srwm_mslice* dSlices, int nSlices, int nxMax, int nzMax, int lanesOverSlices,
                        void* dWork, double* dSums, int* dLims, hipStream_t st);
}

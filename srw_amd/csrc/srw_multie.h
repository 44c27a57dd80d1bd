// srw_multie.h - partially coherent (multi-electron) accumulation kernels.
#pragma once
#include <hip/hip_runtime.h>

// Propagated mesh of one electron.
struct srwm_emesh {
    double xStart, xStep, zStart, zStep;
    int nx, nz;
};

// Result mesh.
struct srwm_rmesh {
    double xStart, xStep, zStart, zStep;
    int nx, nz;
};

extern "C" {
// Srep (optional): electrons j of the batch in groups of perRep also add
// their weighted intensity into Srep[(j / perRep) * nr + point] (per-replicate
// sums of the randomized quasi-Monte Carlo sampler).
int srwm_launch_accum(const float* I, long long iStride, const srwm_emesh* em, const float* w, int nb,
                      srwm_rmesh rm, double* S1, double* S2, double* norm2, double* Srep, int perRep,
                      hipStream_t st);
int srwm_launch_stokes_out(const double* S1, double scale, long long n, float* out, hipStream_t st);
}

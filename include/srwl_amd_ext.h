/*
 * srwl_amd_ext.h - instrumentation and test hooks of libsrw_amd.so.
 *
 * None of these has a counterpart in the reference API.  They let a caller
 * time the device kernels, keep results resident in device memory between
 * calls, and check individual device components (the electron sampler and
 * the proposal network) against a CPU implementation.
 */
#ifndef SRW_AMD_EXT_H
#define SRW_AMD_EXT_H

#include "srwlib.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Enables (on != 0) HIP-event timing of the SR-integral and FFT kernels on
 * the library's stream.  Returns 0 or an SRWL error code. */
SRWL_EXPORT int srwlAmdSetTiming(int on);

/* Copies up to n (<= 16) statistics into out: [0] last SR-kernel ms,
 * [1] last integrand evaluations, [2] last SR work blocks, [3] last s-range
 * split, [4] last FFT-pass ms, [5] last FFT-pass bytes, [6] sum SR ms,
 * [7] sum SR evaluations, [8] SR calls, [9] sum FFT ms, [10] sum FFT bytes,
 * [11] FFT calls (sums since srwlAmdResetStats), [12] last resampling ms,
 * [13] last resampling bytes, [14] device ms of the last propagation
 * (after the upload, before the download), [15] points of the last SR pass
 * whose FP32 sums failed the condition check and were recomputed in FP64. */
SRWL_EXPORT int srwlAmdGetStats(double* out, int n);
SRWL_EXPORT int srwlAmdResetStats(void);

/* on != 0: srwlCalcElecFieldSR leaves the field in device memory and skips
 * the copy into pWfr->arEx/arEy (throughput measurement of the device path). */
SRWL_EXPORT int srwlAmdSetKeepOnDevice(int on);

/* Forces the number of longitudinal partitions of the SR integral
 * (0: automatic). */
SRWL_EXPORT int srwlAmdSetSplit(int nSplit);

/* Forces the number of row bands in which a single request's field is
 * computed and streamed into the caller's host arrays (0: automatic, at most
 * 32; SRWL_AMD_ERR_UNSUPPORTED above). */
SRWL_EXPORT int srwlAmdSetBands(int nBands);

/* on != 0 (default): a single request whose field takes the banded download
 * and has energy tiles is computed in one pass whose blocks finish their own
 * points and count finished row bands, each band downloaded as soon as it is
 * complete; 0: row bands computed one after another. */
SRWL_EXPORT int srwlAmdSetStreamRows(int on);

/* 1 when the last single-request call took the streamed-rows download to the
 * end (no fall back to a plain download), 0 otherwise. */
SRWL_EXPORT int srwlAmdLastStreamed(void);

/* on != 0: integrate trajectories with the serial RK4 kernel (two threads per
 * trajectory) instead of the parallel-in-s solver; both give the serial RK4
 * recurrence, the switch exists to compare them. */
SRWL_EXPORT int srwlAmdSetTrajSerial(int on);

/* Tolerance of the condition check of the FP32 SR sums (method 0): a point
 * whose estimated rounding error exceeds tol times the magnitude of its field
 * (both components) is recomputed in FP64.  Default 1e-4; 0 turns the check
 * off. */
SRWL_EXPORT int srwlAmdSetCondTol(double tol);

/* e = 8 (default) or 4: requests with at least e photon energies integrate
 * them e at a time, advancing each point's phasor from one energy to the
 * next; 0: one energy per workgroup (for comparison). */
SRWL_EXPORT int srwlAmdSetEnergyTiles(int e);

/* Final level (1..20) of every computed point of the last method-1 field
 * computation (precPar[0] = 1), in the reference's point order: rows, then
 * columns, then photon energies of the computed region, request by request.
 * Copies up to cap values into out (when given); returns the count. */
SRWL_EXPORT long long srwlAmdLastAutoLevels(int* out, long long cap);

/* on != 0: resample wavefronts with one thread per new point (each point
 * builds its own interpolation forms) instead of one thread per old-mesh cell
 * (forms built once per cell); both give identical values, the switch exists
 * to compare them. */
SRWL_EXPORT int srwlAmdSetResizePerPoint(int on);

/* FFT tile shapes (rows per workgroup and threads of the forward row pass,
 * columns per workgroup and threads of the column pass, rows per workgroup
 * and threads of the inverse row pass, 1 when the column pass takes two
 * adjacent column blocks per workgroup).  srwlAmdSetFftTiles(t7) makes every
 * pass use t7 (a shape that does not fit falls back to the default),
 * srwlAmdSetFftTiles(NULL) returns to the autotuner; srwlAmdSetTuning(0)
 * selects the fixed default shapes instead of the autotuner.
 * srwlAmdGetFftTiles copies the shapes the autotuner chose for a drift of
 * 2^lognx x 2^lognz slices (mode 0 or 3, postOp 0 field, 1 field with the new
 * quadratic phase, 2 intensity) and returns 1, or returns 0 when that shape
 * has not been tuned yet.  Every shape gives the same values. */
SRWL_EXPORT int srwlAmdSetFftTiles(const int* t7);
SRWL_EXPORT int srwlAmdSetTuning(int on);
SRWL_EXPORT int srwlAmdGetFftTiles(int lognx, int lognz, int mode, int postOp, int* out7);

/* Six standard normals of electron i of the counter-based sampler used by
 * srwlEmitPropMultiE (Philox4x32-10 + Box-Muller, see srw_rng.h). */
SRWL_EXPORT int srwlAmdNormals(unsigned long long seed, long long i, double* g6);

/* Initial parameters (11368 floats) of the proposal network for a seed;
 * returns the parameter count. */
SRWL_EXPORT int srwlAmdProposalInit(float* P, unsigned long long seed);

/* Device forward pass, mixture-loss gradient and backward pass of the
 * proposal network: P (11368 parameters), in6 (beam descriptor), g (nb x 6
 * standardized samples), wh (nb weights), alpha (defensive weight).
 * Outputs G (11368 gradient values), out104 (network output), loss. */
SRWL_EXPORT int srwlAmdProposalGrad(const float* P, const float* in6, const float* g, const float* wh,
                                    int nb, float alpha, float* G, float* out104, float* loss);

#ifdef __cplusplus
}
#endif
#endif

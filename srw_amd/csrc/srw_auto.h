// srw_auto.h - internal C ABI of the adaptive SR integral (SRW method 1,
// cpp/src/core/srradint.cpp:1731-2046) between auto_api.cpp and sr_auto.hip.
#pragma once
#include "srw_dev.h"

// Deepest level evaluated.  The reference keeps doubling until a level holds
// more than 8e8 points (level 29, srradint.cpp:1734, 2025-2031); here a point
// that has not converged after level 20 (2^21 new points) ends the call with
// CAN_NOT_COMPUTE_RADIATION_INTEGRAL.
#define SRWD_AUTO_MAX_LEVEL 20
// Levels up to 11 take tabulated trajectory data (srradint.h:1092-1098 with
// MaxNumPoToSave = 10000); deeper levels evaluate it per point with the
// multiplying index rule (srtrjdat.h:431-472).  Levels up to this one are
// tabulated on the device; deeper ones are evaluated in the kernel.
#define SRWD_AUTO_REF_TAB_LEVEL 11
#define SRWD_AUTO_DEV_TAB_LEVEL 13

// Level L >= 1 holds 2^(L+1) points; level 0 holds 5.  Tables of consecutive
// levels are concatenated: level L starts at entry srwd_auto_off(L).
__host__ __device__ inline long long srwd_auto_off(int L) { return L == 0 ? 0 : (1LL << (L + 1)) + 1; }
__host__ __device__ inline long long srwd_auto_count(int L) { return L == 0 ? 5 : (1LL << (L + 1)); }

// Trajectory data at one level point (FP64, 64 bytes): 1/(yObs - s), X, Z,
// angles, s/gamma^2 + integrals of the squared angles, and s.
struct srwd_auto_tab {
    double inv, X, Z, Btx, Btz, Q, s, pad;
};

// One request: its computed region (rows nzc, columns nxc, energies ne, in
// the reference's order: rows outermost, energies fastest) and its output.
struct srwd_auto_job {
    double xStart, xStep, zStart, zStep, eStart, eStep;
    double normConst;     // sqrt(alpha * I * 1e9 / e)
    double relPrec;
    int nx, nz, nxc, nzc, ne, symX, symZ, termin;
    long long outOffset;  // float2 offset of (0, 0, 0) in the output arrays
    const srwd_req* req;  // the request (device): observation coordinates of the reference's loop
    const double* coord;  // coordinate table (device)
    const long long* last;   // edge-derivative state table (device; nullptr without terminating terms)
};

extern "C" {
// Level tables 0..SRWD_AUTO_DEV_TAB_LEVEL of one trajectory.
int srwd_launch_auto_tables(const srwd_trj_info* dInfo, const double* dX, const double* dY, const double* dIBx,
                            const double* dIBz, srwd_auto_tab* dTab, hipStream_t st);
// Runs points p0 + q of a request through the levels, q in [0, n) or q =
// idx[k] for k < nIdx when idx is given.  thr (per q, optional): absolute
// threshold of the convergence test when >= 0, relative test otherwise.
// Per level L it stores |E|^2 in S[L*n + q] and the field (Ex re, im, Ez re,
// im) in F[L*n + q]; bit L of guard[q] is set when the level's phase steps
// stay within pi.  lvl[q] and done[q] receive the final level.  err is set to 1 when a
// point given its threshold (thr) needs more than SRWD_AUTO_MAX_LEVEL levels,
// warn when a terminating term was rejected.
int srwd_launch_auto_points(const srwd_auto_job& J, const srwd_trj_info* dInfo, const srwd_edge* dEdge,
                            const srwd_auto_tab* dTab, const double* dX, const double* dY, const double* dIBx,
                            const double* dIBz, long long p0, long long n, const int* idx, int nIdx,
                            const double* thr, double* S, float4* F, unsigned* guard, int* lvl, int* done,
                            int* err, int* warn, hipStream_t st);
// Exclusive running maximum of the final |E|^2 (S at level lvl[q]) in point
// order, started at M0[0]; Mtot[0] receives the maximum over all n points and
// M0.  Then every point's final level is recomputed with that maximum: the
// first level L >= 1 (up to done[q]) with its guard bit set and
// |S_L - S_{L-1}| within relPrec*M (relPrec*S_L while M is 0).  Changed levels are written to lvl
// and counted in cnt[0]; points whose stored levels do not reach the test are
// appended to ext (cnt[1]) with thr[q] = relPrec*M.
int srwd_launch_auto_scan(long long n, const double* S, int* lvl, const unsigned* guard, const int* done,
                          double relPrec, const double* M0, double* Mex, double* Mtot, double* thr, int* ext,
                          int* cnt, hipStream_t st);
// Writes the final fields (with the mirrored parts of a symmetric mesh).
int srwd_launch_auto_out(const srwd_auto_job& J, long long p0, long long n, const float4* F, const int* lvl,
                         float* dEx, float* dEz, hipStream_t st);
}

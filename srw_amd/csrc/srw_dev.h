// srw_dev.h - device-side data structures and the internal C ABI between the
// host library (srwlib_api.cpp) and the HIP kernels.  Plain structs and
// pointers only; every launch wrapper takes an explicit stream.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#define SRWD_MAX_HARM 8
#define SRWD_MAX_UND 4

// Periodic undulator in a container (reference srmagfld.h:401-502,
// srmagfld.cpp:128-174).  Element frames are translations only.  Per-harmonic
// constants are precomputed on the host from the reference's formulas.
struct srwd_harm {
    double kw;        // 2pi/(per/n) with the reference's 2pi = 6.2831853072
    double kwpi;      // kw/pi, for cospi/sinpi
    double shift;     // phase offset along z: zrr = zr + shift
    double Bm0;       // K/(per*93.372904175766)
    double halfThis;  // half length of the periodic part (+per/4 if symmetric)
    double p, c4;     // termination exponent and polynomial coefficients
    double sL, sR;    // termination sign factors
    int isX, sym;
};

struct srwd_und {
    int nh;
    int pad;
    srwd_harm h[SRWD_MAX_HARM];
    double halfTotTerm, cx, cy, cz;
};

struct srwd_field {
    int n;
    srwd_und u[SRWD_MAX_UND];
};

// One electron's trajectory job: RK4 over np points in ct in [ctStart, ctEnd]
// (reference srgtrjdt.cpp:24-365).
struct srwd_traj_job {
    double init[5];      // x, x', y, y', z at ct = 0
    double gammaInt;     // internal gamma (srebmdat.h:136-141)
    double mult;         // nq*e/(m c gamma)
    double ctStart, ctEnd;
    long long np;
    int field;           // index into the field array
    int pad;
};

// Interpolation / integration setup of one trajectory (srtrjdat.cpp:2006-2088,
// srradint.cpp:3579-3614).
struct srwd_trj_info {
    long long np;            // trajectory points
    double sStart, sStep, invStep;
    double gammaEm2, betaNorm;
    double x0, dxds0, z0, dzds0, s0;
    int horNZ, verNZ;
    double sIntegStart, sIntegFin, h;   // integration range and step
    long long nInt;          // integration points (odd)
    double yObs;             // longitudinal observation position
};

// Per integration point, energy independent (FP32), 32 bytes: phase increments
// relative to the chunk reference (read with scalar loads), the weighted
// 1/R^2 b and b times the projected positions (staged in LDS per chunk).
struct srwd_rec {
    float dInv, dU, dV, b, bPx, bPz, pad0, pad1;
};

// Per chunk reference (FP64).
struct srwd_chunk {
    double inv, U, V, pad;
};

// Trajectory values with field at an integration edge:
// [0..4] dBzds, Bz, Btx, X, IBx ; [5..9] dBxds, Bx, Btz, Z, IBz
struct srwd_edge {
    double v[2][10];
};

// A field computation item: one (trajectory, photon energy) pair on a mesh.
struct srwd_item {
    int traj;                 // trajectory index (tables, info, edges)
    int arow;                 // row of the per-(trajectory, energy) phase table
    double eph;               // photon energy [eV]
    double xStart, xStep, zStart, zStep;
    int nx, nz;               // full mesh
    int nxc, nzc;             // computed region (symmetry)
    int iz0;                  // first mesh row of this item (items may cover a band of the computed rows)
    int symX, symZ;
    int ne, ie;               // output layout: energies per point and this energy's index
    long long outOffset;      // float2 offset of (iz=0, ix=0, ie) in the output arrays
    long long pointBase;      // first point of this item in the partial-sum workspace
    int blockBase;            // first 256-point block of this item
    int npts;                 // nxc * nzc
    double actNorm;           // normalisation constant * eph * 0.80654658e-3
    int drow;                 // phase-table row of the energy step (energy tiles), -1 if none
    int tiled;                // 1: computed by the energy-tile kernel (its FP32 error model differs)
    double de;                // energy step of the request's mesh
};

#define SRWD_CHUNK 256
// Photon energies per workgroup of the energy-tile SR kernel (4 or 8): the
// phasor of energy k is advanced from that of energy k - 1 by one complex
// multiply.
#define SRWD_TILE_E 8

// One request of a field computation, for the parts that follow the
// reference's point loop (rows, then columns, then photon energies of the
// computed region, srradint.cpp:342-388): its observation coordinates as that
// loop forms them, running sums of the steps (x[nxc], z[nzc] and e[ne] at
// xOff, zOff, eOff of the coordinate table), and its blocks of 256 points in
// the edge-derivative state table.
struct srwd_req {
    int nxc, nzc, ne, pad;
    int xOff, zOff, eOff, pad2;
    long long blk0;           // first block of the request in the state table
    long long nPts;           // nzc * nxc * ne
};
#define SRWD_EDGE_BLOCK 256
// FP64 values per integration point for the FP64 recomputation: X, Z, X',
// Z', s/gamma^2 + the integrals of the squared angles, 1/(yObs - s).
#define SRWD_T64 6

extern "C" {
// Launch wrappers implemented in sr.hip; all asynchronous on stream st.
// Iterate-buffer length per component of the parallel-in-s RK4 for
// trajectories of up to ld points (512 threads per trajectory).
inline long long srwd_pic_cap(long long ld) { return ((ld - 1 + 511) / 512) * 512 + 1; }

// dScratch (20*srwd_pic_cap(ld) doubles per trajectory) selects the
// parallel-in-s solver; nullptr runs the serial kernel.
// singleHarmonic: every field of dFields is one undulator with one harmonic.
int srwd_launch_traj(const srwd_traj_job* dJobs, const srwd_field* dFields, int nTraj, double* dX,
                     double* dXp, double* dY, double* dYp, double* dZ, long long ld, double* dScratch,
                     int singleHarmonic, hipStream_t st);
int srwd_launch_nz(srwd_trj_info* dInfo, int nTraj, const double* dX, const double* dXp, const double* dY,
                   const double* dYp, long long ldT, hipStream_t st);
int srwd_launch_tables(const srwd_trj_info* dInfo, int nTraj, long long maxInt, const double* dX,
                       const double* dXp, const double* dY, const double* dYp, long long ldT,
                       double* dIBx, double* dIBz, srwd_rec* dRec, srwd_chunk* dChunk, double* dQ,
                       long long ldI, long long ldC, srwd_edge* dEdge, double* dT64, hipStream_t st);
int srwd_launch_rx(const srwd_trj_info* dInfo, int nTraj, long long maxNp, const double* dX, const double* dY,
                   const double* dIBx, const double* dIBz, long long ldT, double* dSamp, const double* dMesh,
                   double* dRx, hipStream_t st);
int srwd_launch_phase_table(const double* dQ, long long ldI, const int* dRowTraj, const double* dRowE,
                            const long long* dRowN, int nRows, long long maxInt, float* dA, hipStream_t st);
// A store of the FP64 pass recorded for the host (offset in float2 units of
// both components, the values stored).
struct srwd_patch {
    long long off;
    float2 ex, ez;
};

// Finishing inside the energy-tile kernel (one pass without s-range split):
// the arguments of the finalize pass; after its stores, block i writes
// `epoch` to blockDone[i] (host-coherent memory, system scope).
struct srwd_fuse {
    const srwd_edge* edges;
    int termin;
    float2 *Ex, *Ez;
    int* warn;
    const double* cond;
    double condTol;
    int* fpList;
    int* fpCount;
    const srwd_req* reqs;
    const double* coord;
    const long long* last;
    unsigned* blockDone;     // per launched block
    unsigned epoch;
};

// SR integral over a list of 256-point blocks: dBlocks[i] = {item, block of
// the item}.  tiled = 0: one photon energy per block; tiled = 4 or 8: the
// block's item and the tiled - 1 items after it (consecutive photon energies
// of one request and band, with the same points).
int srwd_launch_sr(const srwd_item* dItems, const int2* dBlocks, int nBlocks, int tiled, const srwd_trj_info* dInfo,
                   const srwd_rec* dRec, const srwd_chunk* dChunk, const float* dA, long long ldI,
                   long long ldC, int nSplit, long long nPointsPad, double* dPart, hipStream_t st);
// The energy-tile kernel (tiled = 4 or 8) with the finalize pass done by each
// block after its sums (nSplit must be 1; dPart holds the sums as in the
// two-pass form), and the per-block completion flags of F.
int srwd_launch_sr_fused(const srwd_item* dItems, const int2* dBlocks, int nBlocks, int tiled,
                         const srwd_trj_info* dInfo, const srwd_rec* dRec, const srwd_chunk* dChunk, const float* dA,
                         long long ldI, long long ldC, long long nPointsPad, double* dPart, const srwd_fuse& F,
                         hipStream_t st);
// Edge-derivative state of the reference's point loop: per block of
// SRWD_EDGE_BLOCK points of each request (dBlkReq: its request) and edge, the
// last point at or before the block's end whose far-field expansion at that
// edge is accepted (request-local index; -1: none).  dLast holds 2 values per
// block.
int srwd_launch_edge_state(const srwd_req* dReq, int nReq, const int* dBlkReq, long long nBlk,
                           const srwd_trj_info* dInfo, const srwd_edge* dEdge, const double* dCoord, int termin,
                           long long* dLast, hipStream_t st);
// Entries (first item, 256-point block in the item) of groups of g items
// (g consecutive photon energies of one request, g divides 256; 1 for single
// items).  dCond (5 doubles per trajectory, from srwd_launch_cond_sums)
// enables the condition check of the FP32 sums (estimated error <= condTol
// of the point's field magnitude): points that fail it are appended to
// dFpList (item's band-relative block * 256 + point in the block;
// *dFpCount entries, zeroed by the caller) instead of being stored, for
// srwd_launch_sr_fp64.  dCond == nullptr stores every point.
int srwd_launch_finalize(const srwd_item* dItems, const int2* dEntries, int nEntries, int g,
                         const srwd_trj_info* dInfo, const srwd_edge* dEdge, int termin, const double* dPart,
                         int nSplit, long long nPointsPad, float* dEx, float* dEz, int* dWarn, const double* dCond,
                         double condTol, int* dFpList, int* dFpCount, const srwd_req* dReq, const double* dCoord,
                         const long long* dLast, hipStream_t st);
int srwd_launch_cond_sums(const srwd_trj_info* dInfo, int nTraj, const srwd_rec* dRec, long long ldI, double* dCond,
                          hipStream_t st);
// FP64 recomputation and store of the listed points (block indices relative
// to dBlockItem, as passed to srwd_launch_finalize).
// dT64: FP64 trajectory values (X, Z, X', Z', integral of the squared angles)
// per integration point, SRWD_T64 doubles, as written by srwd_launch_tables.
// dSums: optional scratch of 4 doubles per list entry (the longest list the
// call can produce); with it a long list's sums run in a kernel of their own.
// dPatch (optional, capacity patchCap): every store is also recorded there
// (*dPatchCount counts them, zeroed by the caller, and may exceed the
// capacity: then the records past it are dropped).
int srwd_launch_sr_fp64(const srwd_item* dItems, const int* dBlockItem, const srwd_trj_info* dInfo,
                        const srwd_edge* dEdge, int termin, const double* dT64, long long ldI, float* dEx,
                        float* dEz, int* dWarn, const int* dFpList, const int* dFpCount, const srwd_req* dReq,
                        const double* dCoord, const long long* dLast, double* dSums, srwd_patch* dPatch,
                        int* dPatchCount, int patchCap, hipStream_t st);
}

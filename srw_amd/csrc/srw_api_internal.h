// srw_api_internal.h - host-side state shared by the C ABI implementation files.
#pragma once
#include "../../include/srwlib.h"
#include "../../include/srwl_amd_ext.h"
#include "srw_dev.h"
#include "srw_moments.h"
#include "srw_resize.h"
#include "srw_fft.h"
#include <vector>

namespace srwamd {

// Grow-only device buffer.  While an Arena is active (the no-allocation
// entry point), get() hands out pieces of the arena instead of allocating.
struct DBuf {
    void* p = nullptr;
    size_t n = 0;
    void* get(size_t bytes);
    void release();
};

// Bump allocator over caller-provided device memory: each DBuf asking within
// one arena session gets its own 256-byte aligned piece, kept for the
// session (asked again, a piece at least as large is returned as is).
struct Arena {
    char* base = nullptr;
    size_t cap = 0, used = 0;
    bool overflow = false;
    struct Piece { const DBuf* buf; void* p; size_t bytes; };
    Piece pieces[64];             // one per buffer of the call (fixed: the arena allocates nothing)
    int nPieces = 0;
    void* get(const DBuf* b, size_t bytes);
};
// Most row bands of one request (the FP64 counts of the bands are read back
// through a pinned array of 64).
#define SRWL_AMD_MAX_BANDS 32
// FP64 store records a streamed-rows call patches into the caller's arrays
// (more: the whole field is downloaded again), and the longest the host
// waits for the blocks' flags before it falls back to a plain download.
#define SRWL_AMD_PATCH_CAP (1 << 18)
#define SRWL_AMD_STREAM_WAIT_S 60
extern Arena* g_arena;   // non-null while a no-allocation call runs

struct Buffers {
    DBuf X, Xp, Y, Yp, IBx, IBz, Samp, Jobs, Fields, Info;
    DBuf Rec, Chunk, Q, Edge, Mesh, Rx, Items, BlockItem, RowTraj, RowE, RowN, A, Part, Warn;
    DBuf Cond, FpList, FpCount;   // condition check of the FP32 sums and its FP64 recomputation list
    DBuf T64;                     // FP64 trajectory values per integration point (FP64 recomputation)
    DBuf Req, Coord, BlkReq, Last;   // requests, observation coordinates, edge-derivative state
    DBuf SrBlocks;                // SR launch lists {item, block}
    DBuf TileBands, Patch, PatchCount;   // streamed rows: row bands per tile block, FP64 store records
    DBuf Ex, Ez;                  // field outputs
    DBuf W0, W1, W2, W3;          // propagation / accumulation work buffers
    DBuf Stokes, Draws, Wt, Net, Grad, Rep;
    DBuf Pic;                     // parallel-in-s RK4 iterates
    DBuf MomSl, MomWork, MomSums, MomLims;   // wavefront moments
    DBuf RsO, RsN;                // resampling
    DBuf AutoTab, AutoFlags, AutoS, AutoF, AutoWork;   // adaptive SR integral (method 1)
    DBuf Tune;                    // scratch slices of the FFT tile autotuner
    DBuf GenT;                    // transposed slices of the mixed-radix FFT passes
};

struct Context {
    bool ready = false;
    int device = -1;
    hipStream_t stream = nullptr;
    Buffers buf;
    bool timing = false;
    hipEvent_t ev0 = nullptr, ev1 = nullptr, ev2 = nullptr, ev3 = nullptr;
    float lastSrMs = 0;
    long long lastSrEvals = 0;
    int lastSrBlocks = 0, lastSrSplit = 0;
    int forceSplit = 0;
    int forceBands = 0;           // row bands of a single request's streamed download (0: automatic)
    bool streamRows = true;       // one pass counting finished row bands instead of banded passes
    unsigned* blockDone = nullptr; // host-coherent completion flags of a streamed-rows pass, per tile block
    size_t blockDoneCap = 0;
    unsigned streamEpoch = 0;
    hipEvent_t evTiles = nullptr;  // the streamed-rows tile kernel has finished
    char* patchHost = nullptr;    // pinned: FP64 store count, then the store records
    bool lastStreamFallback = false;
    bool lastStreamed = false;    // the last single-request call streamed rows without falling back
    bool trajSerial = false;      // test switch: serial RK4 kernel instead of the parallel one
    bool keepOnDevice = false;    // benchmark switch: leave results in HBM
    bool resizePerPoint = false;  // test switch: resample per new point instead of per old cell
    bool tuneTiles = true;        // FFT tile shapes from the autotuner (else the fixed defaults)
    bool tilesForced = false;     // test switch: every FFT pass uses forcedTiles
    srwf_tiles forcedTiles = {0, 0, 0, 0, 0, 0, 0};
    double sumSrMs = 0, sumSrEvals = 0;
    int nSrCalls = 0;
    double lastFftMs = 0, lastFftBytes = 0;
    double sumFftMs = 0, sumFftBytes = 0;
    int nFftCalls = 0;
    double lastRsMs = 0, lastRsBytes = 0;   // last resampling of a propagation
    double lastPropDevMs = 0;               // device part of the last propagation (after upload, before download)
    int* warnHost = nullptr;                // pinned: termination warning of the last SR pass
    // Streamed download of field results: a second stream copies finished
    // row bands into pinned staging while the next band is computed; the
    // calling thread moves them on into the caller's arrays.
    hipStream_t copyStream = nullptr;
    hipStream_t stream2 = nullptr;          // second compute stream: alternate row bands overlap their tails
    hipEvent_t evFork = nullptr, evJoin = nullptr;
    char* pin = nullptr;                    // pinned staging (grow-only)
    size_t pinBytes = 0;
    char* up = nullptr;           // pinned staging of setup uploads (bump allocated per SR call)
    size_t upBytes = 0, upUsed = 0;
    char* momPin = nullptr;       // pinned slice table and sums of queued moments
    size_t momPinBytes = 0;
    std::vector<hipEvent_t> bandEv;         // per band: computed, downloaded, SR start, SR end
    int nTimePairs = 0;                     // timed SR launches of the last pass (bandEv[4b+2], [4b+3])
    int* fpCountHost = nullptr;             // pinned: points recomputed in FP64, per band (at most 64 bands)
    int fpBands = 0;
    long long lastFp64Points = 0;           // points of the last SR pass recomputed in FP64
    double condTol = 1e-4;                  // condition check of the FP32 sums (0: off)
    int tileE = SRWD_TILE_E;                // energies per energy-tile SR workgroup (0: single-energy kernel only)
    long long pendingEvals = 0;             // integrand evaluations of the last SR pass
    std::vector<int> autoLevels;            // final levels of the last method-1 pass, in point order
    int autoSweeps = 0;                     // fixed-point sweeps of the last method-1 pass (largest chunk)
};

extern Context g_ctx;
extern int (*g_wfrModif)(int, SRWLWfr*, char);
extern char* (*g_allocArray)(char, long long);
extern int (*g_progress)(double);

int ensure_device();

// One single-electron SR request.
struct SrJob {
    SRWLParticle part;
    double Iavg = 0;
    int field = 0;
    long long np = 0;
    double ctStart = 0, ctEnd = 0;
    const SRWLPrtTrj* userTrj = nullptr;
    SRWLRadMesh mesh;
    double step = 0, sLim0 = 0, sLim1 = 0;
    int meth = 0;                 // precPar[0]: 0 manual, 1 adaptive
    double relPrec = 0;           // precPar[1] of method 1
};

// Computed region of a request's mesh: the whole mesh, or the half / quarter
// that the symmetry shortcut computes (srradint.cpp:319-345, 3231-3260).
struct SrRegion {
    double xStep = 0, zStep = 0;
    int symX = 0, symZ = 0, nxc = 0, nzc = 0;
};
SrRegion sr_region(const SRWLRadMesh& m, const srwd_trj_info& I);

struct SrResult {
    double Robs = 0, RobsErr = 0, xc = 0, zc = 0;
    int warn = 0;
};

int field_from_c(const SRWLMagFldC* c, srwd_field& f, double& gsStart, double& gsEnd);
int setup_job_from_field(SrJob& J, const SRWLParticle& p, const SRWLMagFldC* mag, const double* precPar,
                         std::vector<srwd_field>& fields);
// Host destination of a single request's field, filled band by band while
// later bands are computed (see sr_run).
struct HostOut {
    float* ex = nullptr;
    float* ez = nullptr;
    // more device work to queue behind the field before the host starts
    // moving bands (optional; gets the request's results so far)
    int (*queueMore)(void* ctx, const std::vector<SrResult>& res) = nullptr;
    void* queueCtx = nullptr;
};
int sr_run(std::vector<SrJob>& jobs, const std::vector<srwd_field>& fields, float* dEx, float* dEz,
           const std::vector<long long>& outOff, std::vector<SrResult>& res, int termin,
           bool defer = false, bool geomOnly = false, const HostOut* hostOut = nullptr);
// Pinned staging of at least `bytes` (grow-only); nullptr on failure.
char* pinned_staging(size_t bytes);
// Copies device memory to host memory through the pinned staging on the copy
// stream, in pieces, after the work already queued on the compute stream.
int download(void* dst, const void* src, size_t bytes);
// Host memcpy split over a pool of threads (large copies out of pinned
// memory): one range, or several ranges as one job.
struct CopySeg {
    void* dst;
    const void* src;
    size_t n;
};
void par_memcpy(void* dst, const void* src, size_t bytes);
void par_memcpy(const CopySeg* seg, int nSeg);
// Mesh size of one axis for the automatic mesh of precPar[6] > 0.
long sr_auto_mesh_size(double start, double fin, long n, double c, double R, double lambda, double f);
// Completes a deferred sr_run once the stream has been synchronised.
void sr_complete(std::vector<SrResult>& res);
// Method 1 for every job, after sr_run has set up trajectories, edges and
// wavefront radii (auto_api.cpp); synchronous.
int sr_auto_run(const std::vector<SrJob>& jobs, const std::vector<srwd_trj_info>& info, const srwd_trj_info* dInfo,
                const double* dX, const double* dY, const double* dIBx, const double* dIBz, long long ldT,
                const srwd_edge* dEdge, int termin, float* dEx, float* dEz, const std::vector<long long>& outOff,
                std::vector<SrResult>& res, const srwd_req* dReq, const double* dCoord, const long long* dLast);

// Wavefront moments (moments_api.cpp).  Component element (ix, iz, ie) of a
// device wavefront is at base[ie*perE + iz*perZ + ix*perX] (float2 units).
// srradstr.cpp:2610-2674: whether the reference would subtract a linear phase
// term before a lens centred at (newXc, newZc) (reported as unsupported).
bool lin_phase_would_subtract(const SRWLWfr& w, double newXc, double newZc);

int wfr_moments(const SRWLWfr& w, const float2* dEx, const float2* dEz, long long perE, long long perX,
                long long perZ, double* momX, double* momZ);
int device_moments(const std::vector<srwm_mslice>& sl, std::vector<double>& sums, std::vector<int>& lims,
                   bool lanesOverSlices = false);
void moments_from_sums(const double* S, const srwm_mslice& s, double* mom);
// The moments of wfr_moments split in two: kernels and read-back queued on
// the library stream (into pinned memory, so the host is not held), then the
// moments formed on the host once the stream has been synchronised.
struct MomQueued {
    std::vector<srwm_mslice> sl;
    double* sums = nullptr;   // pinned, 24 per slice (the library's; valid until the next queue)
    int n = 0;
    bool hasX = false, hasZ = false;
};
int wfr_moments_queue(const SRWLWfr& w, const float2* dEx, const float2* dEz, long long perE, long long perX,
                      long long perZ, MomQueued& q);
void wfr_moments_collect(const MomQueued& q, double* momX, double* momZ);
srwm_mslice wfr_slice(const SRWLWfr& w, const float2* dEx, const float2* dEz, long long perE, long long perX,
                      long long perZ, long ie);
int drift_radius(const SRWLWfr& w, int sub, double L, double* momX, double* momZ, int (*recompute)(void*),
                 void* ctx, double& Rx, double& Rz);
void drift_moments(const SRWLWfr& w, double L, double* mom);
void drift_4x4(double* P, double L);

// Resampling (resize_api.cpp).
struct ResizeAxis {
    long n = 0, i0 = 0, i1 = 0;     // new size and the index range filled from the old mesh
    double start = 0, step = 0;
};
struct ResizePlan {
    double pxm = 1, pxd = 1, pzm = 1, pzd = 1;
    ResizeAxis x, z;
    int onlyLarger = 0, fx = 0, fz = 0;
    int improve = 1, zeroOutside = 0;
};
long next_fft_size(long n);
int resize_plan(const SRWLWfr& w, const double* par, int npar, ResizePlan& P);
int device_resize(const SRWLWfr& w, const ResizePlan& P, float2* oX, float2* oZ, const srwr_layout& lo, float2* nX,
                  float2* nZ, const srwr_layout& ln, hipStream_t st);
int realloc_wfr_arrays(SRWLWfr& w, const ResizePlan& P);

}  // namespace srwamd

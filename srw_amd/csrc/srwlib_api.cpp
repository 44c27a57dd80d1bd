// srwlib_api.cpp - host implementation of the SRWLIB C ABI (include/srwlib.h)
// over the HIP kernels of this library.
//
// Every computation runs on the GPU; there is no CPU fallback.  When no
// device is usable the entry points return SRWL_AMD_ERR_GPU.
//
// Reference call semantics followed here:
//   srwlCalcElecFieldSR    cpp/src/lib/srwlib.cpp:403-540
//   srwlCalcPartTraj       cpp/src/lib/srwlib.cpp:328-373
//   srwlUtiGetErrText etc. cpp/src/lib/srwlib.cpp:78-140
#include "srw_api_internal.h"
#include "srw_fft.h"

#include <emmintrin.h>
#include <unistd.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <condition_variable>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

namespace srwamd {

Context g_ctx;

int (*g_wfrModif)(int, SRWLWfr*, char) = nullptr;
char* (*g_allocArray)(char, long long) = nullptr;
int (*g_progress)(double) = nullptr;

int ensure_device()
{
    if (g_ctx.ready) return 0;
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n < 1) return SRWL_AMD_ERR_GPU;
    int dev = g_ctx.device < 0 ? 0 : g_ctx.device;
    if (dev >= n) return SRWL_AMD_ERR_GPU;
    if (hipSetDevice(dev) != hipSuccess) return SRWL_AMD_ERR_GPU;
    if (hipStreamCreateWithFlags(&g_ctx.stream, hipStreamNonBlocking) != hipSuccess) return SRWL_AMD_ERR_GPU;
    if (hipStreamCreateWithFlags(&g_ctx.copyStream, hipStreamNonBlocking) != hipSuccess) return SRWL_AMD_ERR_GPU;
    if (hipStreamCreateWithFlags(&g_ctx.stream2, hipStreamNonBlocking) != hipSuccess) return SRWL_AMD_ERR_GPU;
    if (hipEventCreateWithFlags(&g_ctx.evFork, hipEventDisableTiming) != hipSuccess) return SRWL_AMD_ERR_GPU;
    if (hipEventCreateWithFlags(&g_ctx.evJoin, hipEventDisableTiming) != hipSuccess) return SRWL_AMD_ERR_GPU;
    if (hipEventCreateWithFlags(&g_ctx.evTiles, hipEventDisableTiming) != hipSuccess) return SRWL_AMD_ERR_GPU;
    if (hipHostMalloc((void**)&g_ctx.warnHost, sizeof(int), hipHostMallocDefault) != hipSuccess) return SRWL_AMD_ERR_GPU;
    if (hipHostMalloc((void**)&g_ctx.fpCountHost, sizeof(int) * 64, hipHostMallocDefault) != hipSuccess)
        return SRWL_AMD_ERR_GPU;
    *g_ctx.warnHost = 0;
    g_ctx.device = dev;
    g_ctx.ready = true;
    return 0;
}

Arena* g_arena = nullptr;

void* Arena::get(const DBuf* b, size_t bytes)
{
    int k = 0;
    while (k < nPieces && pieces[k].buf != b) k++;
    if (k < nPieces && pieces[k].bytes >= bytes) return pieces[k].p;
    // new buffer, or larger now: a new piece (the old one is left unused)
    size_t need = (bytes + 255) & ~(size_t)255;
    if (used + need > cap || (k == nPieces && nPieces == 64)) { overflow = true; return nullptr; }
    void* p = base + used;
    used += need;
    if (k == nPieces) nPieces++;
    pieces[k] = {b, p, bytes};
    return p;
}

void* DBuf::get(size_t bytes)
{
    if (bytes == 0) bytes = 256;
    if (g_arena) return g_arena->get(this, bytes);
    if (bytes > n) {
        if (p) (void)hipFree(p);
        p = nullptr;
        n = 0;
        if (hipMalloc(&p, bytes) != hipSuccess) { p = nullptr; return nullptr; }
        n = bytes;
    }
    return p;
}

void DBuf::release()
{
    if (p) (void)hipFree(p);
    p = nullptr;
    n = 0;
}

// srmagcnt.cpp:20-140 / srmagfld.cpp:128-174: container of periodic undulators.
int field_from_c(const SRWLMagFldC* c, srwd_field& f, double& gsStart, double& gsEnd)
{
    if (c == nullptr || c->arMagFld == nullptr || c->arMagFldTypes == nullptr || c->nElem <= 0)
        return SRWL_NO_FUNC_ARG_DATA;
    if (c->nElem > SRWD_MAX_UND) return SRWL_AMD_ERR_UNSUPPORTED;
    bool cen = c->arXc && c->arYc && c->arZc;
    bool ax = c->arVx && c->arVy && c->arVz;
    std::memset(&f, 0, sizeof(f));
    f.n = c->nElem;
    gsStart = 1.e+23; gsEnd = -1.e+23;
    for (int i = 0; i < c->nElem; i++) {
        if (c->arMagFldTypes[i] != 'u') return SRWL_AMD_ERR_UNSUPPORTED;
        if (ax && (c->arVx[i] != 0. || c->arVy[i] != 0. || c->arVz[i] != 1.)) return SRWL_AMD_ERR_UNSUPPORTED;
        if (c->arAng && c->arAng[i] != 0.) return SRWL_AMD_ERR_UNSUPPORTED;
        const SRWLMagFldU* U = (const SRWLMagFldU*)c->arMagFld[i];
        if (U == nullptr) return SRWL_NO_FUNC_ARG_DATA;
        if (U->per <= 0) return SRWL_ERR_INCORRECT_MAG_FIELD_PER;
        if (U->nPer <= 0) return SRWL_ERR_INCORRECT_PER_MAG_FIELD_LENGTH;
        if (U->nHarm > SRWD_MAX_HARM || U->nHarm < 0) return SRWL_AMD_ERR_UNSUPPORTED;
        srwd_und& u = f.u[i];
        const double pi = 3.1415926535898, twoPi = 6.2831853072, piE2 = pi * pi;
        const double factK2B = 93.372904175766;
        double per = U->per;
        double totLen = (2 * U->nPer) * U->per * 0.5;
        double halfTot = 0.5 * totLen;
        u.halfTotTerm = halfTot + 4 * per;
        u.cx = cen ? c->arXc[i] : 0.; u.cy = cen ? c->arYc[i] : 0.; u.cz = cen ? c->arZc[i] : 0.;
        u.nh = U->nHarm;
        for (int k = 0; k < U->nHarm; k++) {
            const SRWLMagFldH& H = U->arHarm[k];
            if (H.n <= 0) return SRWL_AMD_ERR_UNSUPPORTED;
            srwd_harm& h = u.h[k];
            h.isX = (H.h_or_v == 'h' || H.h_or_v == 'H') ? 1 : 0;
            h.sym = H.s;
            double K = 93.37290417576577 * per * H.B;
            double ph = H.ph;
            long long n2pi = (long long)(std::fabs(ph) / twoPi);
            if (ph > 0) ph -= twoPi * n2pi; else if (ph < 0) ph += twoPi * n2pi;
            double perH = per / H.n;
            h.kw = twoPi / perH;
            h.kwpi = h.kw / 3.141592653589793;
            h.shift = ph * perH / twoPi;
            h.Bm0 = K / (per * factK2B);
            h.halfThis = halfTot + ((H.s == 1) ? 0.25 * per : 0.);
            double perE2 = perH * perH;
            h.p = 2. * piE2 / (3. * perE2);
            h.c4 = 4. * piE2 / (9. * perE2);
            double halfPerH = 0.5 * perH;
            double ztL = -h.halfThis + 0.5 * halfPerH, ztR = h.halfThis - 0.5 * halfPerH;
            double ttL = (H.s == 1) ? std::cos(h.kw * ztL) : std::sin(h.kw * ztL);
            double ttR = (H.s == 1) ? std::cos(h.kw * ztR) : std::sin(h.kw * ztR);
            h.sL = (ttL < 0) ? -1. : 1.;
            h.sR = (ttR > 0) ? -1. : 1.;
        }
        // Longitudinal limits of the element as the container reports them:
        // srTMagFieldPeriodic stores cz -/+ halfTotTerm (srmagfld.cpp:148-149)
        // and GetMagnFieldLongLim translates them by the element's centre
        // once more (srmagelem.h:79-90), so the default trajectory range of
        // an element centred at cz is 2 cz -/+ halfTotTerm.  The field itself
        // is placed at cz.
        gsStart = std::min(gsStart, 2. * u.cz - u.halfTotTerm);
        gsEnd = std::max(gsEnd, 2. * u.cz + u.halfTotTerm);
    }
    return 0;
}

static inline double gamma_internal(const SRWLParticle& p)
{   // srebmdat.h:136-141 with srwlib.cpp:331-333
    return p.gamma * p.relE0 * 0.51099890221e-03 * 1000. / 0.511003414;
}

#define HIPCHK(x) do { if ((x) != hipSuccess) return SRWL_AMD_ERR_GPU; } while (0)

// Moments of a device field in SRW's [iy][ix][ie] order: energy ie, point
// (ix, iy) at ie + ne*(ix + nx*iy).
static int field_moments(SRWLWfr& w, const float2* dEx, const float2* dEz)
{
    long long ne = w.mesh.ne, nx = w.mesh.nx;
    return wfr_moments(w, dEx, dEz, 1, ne, nx * ne, w.arMomX, w.arMomY);
}

// Host-to-device copy of setup data through a pinned staging area (bump
// allocated per sr_run call), so that the copy is queued without the host
// waiting for the stream, as a copy out of pageable memory can.  In a
// no-allocation call the data are copied straight from the caller's memory.
static int h2d(void* dst, const void* src, size_t bytes, hipStream_t st)
{
    if (bytes == 0) return 0;
    if (g_arena) return hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, st) == hipSuccess ? 0 : SRWL_AMD_ERR_GPU;
    size_t at = (g_ctx.upUsed + 255) & ~(size_t)255;
    if (at + bytes > g_ctx.upBytes) {
        // a larger area: the current one may still feed queued copies
        if (hipStreamSynchronize(st) != hipSuccess) return SRWL_AMD_ERR_GPU;
        if (g_ctx.up) (void)hipHostFree(g_ctx.up);
        g_ctx.up = nullptr;
        g_ctx.upBytes = 0;
        size_t want = std::max<size_t>({(size_t)1 << 20, 2 * (at + bytes), bytes});
        if (hipHostMalloc((void**)&g_ctx.up, want, hipHostMallocDefault) != hipSuccess) {
            g_ctx.up = nullptr;
            return SRWL_ERR_MEMORY_ALLOCATION_FAILURE;
        }
        g_ctx.upBytes = want;
        at = 0;
    }
    std::memcpy(g_ctx.up + at, src, bytes);
    g_ctx.upUsed = at + bytes;
    return hipMemcpyAsync(dst, g_ctx.up + at, bytes, hipMemcpyHostToDevice, st) == hipSuccess ? 0 : SRWL_AMD_ERR_GPU;
}

// Runs the single-electron SR computation for a set of requests.  Outputs
// stay on the device: request j writes ne*nx*ny complex values of Ex and Ez
// starting at float2 offset outOff[j] of dEx / dEz.
int sr_run(std::vector<SrJob>& jobs, const std::vector<srwd_field>& fields, float* dEx, float* dEz,
           const std::vector<long long>& outOff, std::vector<SrResult>& res, int termin, bool defer, bool geomOnly,
           const HostOut* hostOut)
{
    int rc = ensure_device();
    if (rc) return rc;
    hipStream_t st = g_ctx.stream;
    int nT = (int)jobs.size();
    if (nT == 0) return 0;
    res.assign(nT, SrResult());
    g_ctx.lastStreamed = false;
    if (g_ctx.upUsed) {   // the staging area is reused once its copies are done
        HIPCHK(hipStreamSynchronize(st));
        g_ctx.upUsed = 0;
    }
    Buffers& B = g_ctx.buf;
    // --- trajectories -----------------------------------------------------
    long long ldT = 0;
    std::vector<srwd_traj_job> tj(nT);
    std::vector<srwd_trj_info> info(nT);
    for (int j = 0; j < nT; j++) {
        SrJob& J = jobs[j];
        const SRWLParticle& p = J.part;
        double g = gamma_internal(p);
        if (J.np < 6) return SRWL_INCORRECT_PARAM_FOR_SR_COMP;
        srwd_traj_job& T = tj[j];
        T.init[0] = p.x; T.init[1] = p.xp; T.init[2] = p.y; T.init[3] = p.yp; T.init[4] = p.z;
        T.gammaInt = g;
        T.mult = p.nq * 1.602176462e-19 / (9.10938188e-31 * 2.99792458e+08 * g);
        T.ctStart = J.ctStart; T.ctEnd = J.ctEnd; T.np = J.np; T.field = J.field;
        ldT = std::max(ldT, J.np);
        srwd_trj_info& I = info[j];
        std::memset(&I, 0, sizeof(I));
        I.np = J.np;
        I.sStart = J.ctStart + p.z;
        I.sStep = (J.ctEnd - J.ctStart) / (J.np - 1);
        I.invStep = 1. / I.sStep;
        I.gammaEm2 = 1. / g / g;
        I.betaNorm = -1.60217646263E-19 / (g * 9.1093818872E-31 * 2.99792458E+08);
        I.x0 = p.x; I.dxds0 = p.xp; I.z0 = p.y; I.dzds0 = p.yp; I.s0 = p.z;
        // srradint.cpp:3579-3614
        I.sIntegStart = I.sStart;
        I.sIntegFin = I.sStart + I.sStep * (I.np - 1);
        if (J.sLim0 < J.sLim1) {
            if (J.sLim0 > I.sIntegStart && J.sLim0 < I.sIntegFin) I.sIntegStart = J.sLim0;
            if (J.sLim1 > I.sIntegStart && J.sLim1 < I.sIntegFin) I.sIntegFin = J.sLim1;
        }
        if (J.meth == 1) {
            // adaptive method: no manual integration points (the manual
            // tables below get a 3-point placeholder)
            if (!(J.relPrec > 0)) return SRWL_INCORRECT_PARAM_FOR_SR_COMP;
            I.nInt = 3;
        } else {
            // srradint.h:1043-1058, 814-830
            if (!(J.step > 0)) return SRWL_INCORRECT_PARAM_FOR_SR_COMP;
            int b0 = (int)((I.sIntegFin - I.sIntegStart) / J.step) >> 1;
            I.nInt = ((long long)b0 << 1) + 1;
            if (I.nInt < 3) return SRWL_INCORRECT_PARAM_FOR_SR_COMP;
        }
        I.h = (I.sIntegFin - I.sIntegStart) / (I.nInt - 1);
        I.yObs = J.mesh.zStart;
        // srradint.cpp:92-123
        if (g <= 1.) return SRWL_ERR_ELEC_BEAM_NOT_ULTRARELATIVISTIC;
        if (I.s0 < I.sStart || I.s0 > I.sStart + I.sStep * (I.np - 1)) return SRWL_ERR_S0_OUT_OF_FIELD_LIMITS;
        if (I.yObs < I.sIntegFin) return SRWL_ERR_OBS_WITHIN_INTEGR_LIMITS;
        if ((I.yObs - I.sIntegFin) < 3. * 1.239842E-06 / J.mesh.eStart) return SRWL_ERR_WAVELENGTH_COMPARABLE_WITH_DISTANCE;
        if (g <= 10.) res[j].warn = SRWL_WARN_ELEC_BEAM_NOT_ULTRARELATIVISTIC;
    }
    size_t trjBytes = sizeof(double) * (size_t)ldT * nT;
    double* dX = (double*)B.X.get(trjBytes); double* dXp = (double*)B.Xp.get(trjBytes);
    double* dY = (double*)B.Y.get(trjBytes); double* dYp = (double*)B.Yp.get(trjBytes);
    double* dIBx = (double*)B.IBx.get(trjBytes); double* dIBz = (double*)B.IBz.get(trjBytes);
    double* dSamp = (double*)B.Samp.get(4 * trjBytes);
    srwd_traj_job* dJobs = (srwd_traj_job*)B.Jobs.get(sizeof(srwd_traj_job) * nT);
    srwd_field* dFields = (srwd_field*)B.Fields.get(sizeof(srwd_field) * std::max<size_t>(fields.size(), 1));
    srwd_trj_info* dInfo = (srwd_trj_info*)B.Info.get(sizeof(srwd_trj_info) * nT);
    if (!dX || !dXp || !dY || !dYp || !dIBx || !dIBz || !dSamp || !dJobs || !dFields || !dInfo)
        return SRWL_ERR_MEMORY_ALLOCATION_FAILURE;
    if (int e_ = h2d(dJobs, tj.data(), sizeof(srwd_traj_job) * nT, st)) return e_;
    if (!fields.empty())
        if (int e_ = h2d(dFields, fields.data(), sizeof(srwd_field) * fields.size(), st)) return e_;
    if (int e_ = h2d(dInfo, info.data(), sizeof(srwd_trj_info) * nT, st)) return e_;
    bool anyRK = false;
    for (int j = 0; j < nT; j++) {
        if (jobs[j].userTrj) {
            const SRWLPrtTrj* u = jobs[j].userTrj;
            size_t b = sizeof(double) * u->np;
            if (int e_ = h2d(dX + j * ldT, u->arX, b, st)) return e_;
            if (int e_ = h2d(dXp + j * ldT, u->arXp, b, st)) return e_;
            if (int e_ = h2d(dY + j * ldT, u->arY, b, st)) return e_;
            if (int e_ = h2d(dYp + j * ldT, u->arYp, b, st)) return e_;
        } else anyRK = true;
    }
    if (anyRK) {
        // user-supplied trajectories are re-uploaded after the RK pass overwrote them
        double* dPic = nullptr;
        if (!g_ctx.trajSerial) {
            dPic = (double*)B.Pic.get(sizeof(double) * 20 * (size_t)srwd_pic_cap(ldT) * nT);
            if (!dPic) return SRWL_ERR_MEMORY_ALLOCATION_FAILURE;
        }
        int one = 1;
        for (auto& f : fields) one = one && (f.n == 1 && f.u[0].nh == 1);
        if (srwd_launch_traj(dJobs, dFields, nT, dX, dXp, dY, dYp, nullptr, ldT, dPic, one, st)) return SRWL_AMD_ERR_GPU;
        for (int j = 0; j < nT; j++) {
            if (!jobs[j].userTrj) continue;
            const SRWLPrtTrj* u = jobs[j].userTrj;
            size_t b = sizeof(double) * u->np;
            if (int e_ = h2d(dX + j * ldT, u->arX, b, st)) return e_;
            if (int e_ = h2d(dXp + j * ldT, u->arXp, b, st)) return e_;
            if (int e_ = h2d(dY + j * ldT, u->arY, b, st)) return e_;
            if (int e_ = h2d(dYp + j * ldT, u->arYp, b, st)) return e_;
        }
    }
    if (srwd_launch_nz(dInfo, nT, dX, dXp, dY, dYp, ldT, st)) return SRWL_AMD_ERR_GPU;
    // --- tables ------------------------------------------------------------
    long long maxInt = 0;
    for (auto& I : info) maxInt = std::max(maxInt, I.nInt);
    long long ldI = maxInt;
    long long ldC = (maxInt + SRWD_CHUNK - 1) / SRWD_CHUNK;
    srwd_rec* dRec = (srwd_rec*)B.Rec.get(sizeof(srwd_rec) * (size_t)ldI * nT);
    srwd_chunk* dChunk = (srwd_chunk*)B.Chunk.get(sizeof(srwd_chunk) * (size_t)ldC * nT);
    double* dQ = (double*)B.Q.get(sizeof(double) * (size_t)ldI * nT);
    srwd_edge* dEdge = (srwd_edge*)B.Edge.get(sizeof(srwd_edge) * nT);
    double* dMesh = (double*)B.Mesh.get(sizeof(double) * 4 * nT);
    double* dRx = (double*)B.Rx.get(sizeof(double) * 4 * nT);
    double* dT64 = g_ctx.condTol > 0 ? (double*)B.T64.get(sizeof(double) * SRWD_T64 * (size_t)ldI * nT) : nullptr;
    if (!dRec || !dChunk || !dQ || !dEdge || !dMesh || !dRx || (g_ctx.condTol > 0 && !dT64))
        return SRWL_ERR_MEMORY_ALLOCATION_FAILURE;
    if (srwd_launch_tables(dInfo, nT, maxInt, dX, dXp, dY, dYp, ldT, dIBx, dIBz, dRec, dChunk, dQ, ldI, ldC,
                           dEdge, dT64, st))
        return SRWL_AMD_ERR_GPU;
    std::vector<double> meshLim(4 * nT);
    for (int j = 0; j < nT; j++) {
        const SRWLRadMesh& m = jobs[j].mesh;
        double xStep = m.nx > 1 ? (m.xFin - m.xStart) / (m.nx - 1) : 0;
        double zStep = m.ny > 1 ? (m.yFin - m.yStart) / (m.ny - 1) : 0;
        meshLim[4 * j] = m.xStart; meshLim[4 * j + 1] = m.xStart + (m.nx - 1) * xStep;
        meshLim[4 * j + 2] = m.yStart; meshLim[4 * j + 3] = m.yStart + (m.ny - 1) * zStep;
    }
    if (int e_ = h2d(dMesh, meshLim.data(), sizeof(double) * 4 * nT, st)) return e_;
    if (srwd_launch_rx(dInfo, nT, ldT, dX, dY, dIBx, dIBz, ldT, dSamp, dMesh, dRx, st)) return SRWL_AMD_ERR_GPU;
    std::vector<double> rx(4 * nT);
    HIPCHK(hipMemcpyAsync(rx.data(), dRx, sizeof(double) * 4 * nT, hipMemcpyDeviceToHost, st));
    HIPCHK(hipMemcpyAsync(info.data(), dInfo, sizeof(srwd_trj_info) * nT, hipMemcpyDeviceToHost, st));
    HIPCHK(hipStreamSynchronize(st));
    for (int j = 0; j < nT; j++) {
        res[j].Robs = rx[4 * j]; res[j].RobsErr = rx[4 * j + 1];
        res[j].xc = rx[4 * j + 2]; res[j].zc = rx[4 * j + 3];
    }
    if (geomOnly) return 0;
    // --- requests in the reference's point order ----------------------------
    // Observation coordinates as the reference's loop forms them (running
    // sums of the steps, srradint.cpp:342-388) and the edge-derivative state.
    std::vector<srwd_req> reqs(nT);
    std::vector<double> coord;
    std::vector<int> blkReq;
    for (int j = 0; j < nT; j++) {
        const SRWLRadMesh& m = jobs[j].mesh;
        SrRegion G = sr_region(m, info[j]);
        srwd_req& R = reqs[j];
        std::memset(&R, 0, sizeof(R));
        R.nxc = G.nxc; R.nzc = G.nzc; R.ne = (int)m.ne;
        R.nPts = (long long)G.nzc * G.nxc * m.ne;
        R.blk0 = (long long)blkReq.size();
        R.xOff = (int)coord.size();
        double v = m.xStart;
        for (int ix = 0; ix < G.nxc; ix++) { coord.push_back(v); v += G.xStep; }
        R.zOff = (int)coord.size();
        v = m.yStart;
        for (int iz = 0; iz < G.nzc; iz++) { coord.push_back(v); v += G.zStep; }
        R.eOff = (int)coord.size();
        double eStep = m.ne > 1 ? (m.eFin - m.eStart) / (m.ne - 1) : 0;
        v = m.eStart;
        for (int ie = 0; ie < (int)m.ne; ie++) { coord.push_back(v); v += eStep; }
        if (termin > 0) {
            long long nb = (R.nPts + SRWD_EDGE_BLOCK - 1) / SRWD_EDGE_BLOCK;
            blkReq.insert(blkReq.end(), (size_t)nb, j);
        }
    }
    srwd_req* dReq = (srwd_req*)B.Req.get(sizeof(srwd_req) * nT);
    double* dCoord = (double*)B.Coord.get(sizeof(double) * std::max<size_t>(1, coord.size()));
    long long nEdgeBlk = (long long)blkReq.size();
    int* dBlkReq = nEdgeBlk ? (int*)B.BlkReq.get(sizeof(int) * (size_t)nEdgeBlk) : nullptr;
    long long* dLast = nEdgeBlk ? (long long*)B.Last.get(sizeof(long long) * 2 * (size_t)nEdgeBlk) : nullptr;
    if (!dReq || !dCoord || (nEdgeBlk && (!dBlkReq || !dLast))) return SRWL_ERR_MEMORY_ALLOCATION_FAILURE;
    if (int e_ = h2d(dReq, reqs.data(), sizeof(srwd_req) * nT, st)) return e_;
    if (!coord.empty())
        if (int e_ = h2d(dCoord, coord.data(), sizeof(double) * coord.size(), st)) return e_;
    if (nEdgeBlk) {
        if (int e_ = h2d(dBlkReq, blkReq.data(), sizeof(int) * (size_t)nEdgeBlk, st)) return e_;
        if (srwd_launch_edge_state(dReq, nT, dBlkReq, nEdgeBlk, dInfo, dEdge, dCoord, termin, dLast, st))
            return SRWL_AMD_ERR_GPU;
    }
    if (jobs[0].meth == 1) {
        for (auto& J : jobs) if (J.meth != 1) return SRWL_AMD_ERR_BATCH;
        if (g_ctx.timing) HIPCHK(hipEventRecord(g_ctx.ev0, st));
        rc = sr_auto_run(jobs, info, dInfo, dX, dY, dIBx, dIBz, ldT, dEdge, termin, dEx, dEz, outOff, res, dReq, dCoord,
                         dLast);
        if (rc) return rc;
        if (g_ctx.timing) HIPCHK(hipEventRecord(g_ctx.ev1, st));
        if (hostOut) {
            const SRWLRadMesh& m = jobs[0].mesh;
            size_t n = sizeof(float) * 2 * (size_t)m.ne * m.nx * m.ny;
            rc = download(hostOut->ex, dEx, n);
            if (!rc) rc = download(hostOut->ez, dEz, n);
            if (rc) return rc;
        }
        long long ev = 0;
        for (int L : g_ctx.autoLevels) ev += 4LL * (1LL << L) + 1;
        g_ctx.pendingEvals = ev;
        *g_ctx.warnHost = 0;
        if (defer) return 0;
        HIPCHK(hipStreamSynchronize(st));
        sr_complete(res);
        return 0;
    }
    // --- items -------------------------------------------------------------
    // One item per (trajectory, photon energy) and band of computed mesh rows.
    // A single request whose field goes to host arrays is cut into row bands:
    // band b's download (rows b and their mirror images) overlaps the
    // computation of band b + 1.
    std::vector<int> rowTraj;
    std::vector<double> rowE;
    std::vector<long long> rowN;
    const double normBase = 1. / 137.0360411;
    std::vector<int> rowBase(nT);
    for (int j = 0; j < nT; j++) {
        rowBase[j] = (int)rowTraj.size();
        const SRWLRadMesh& m = jobs[j].mesh;
        for (int ie = 0; ie < (int)m.ne; ie++) {
            rowTraj.push_back(j); rowE.push_back(coord[reqs[j].eOff + ie]); rowN.push_back(info[j].nInt);
        }
    }
    // energy tiles: a request with at least SRWD_TILE_E photon energies gets a
    // phase-table row for its energy step, and its energies in groups of
    // SRWD_TILE_E go to the energy-tile kernel
    std::vector<int> dRowOf(nT, -1);
    std::vector<double> eStepOf(nT, 0.);
    for (int j = 0; j < nT; j++) {
        const SRWLRadMesh& m = jobs[j].mesh;
        if (g_ctx.tileE < 2 || m.ne < g_ctx.tileE) continue;
        eStepOf[j] = (m.eFin - m.eStart) / (m.ne - 1);
        dRowOf[j] = (int)rowTraj.size();
        rowTraj.push_back(j); rowE.push_back(eStepOf[j]); rowN.push_back(info[j].nInt);
    }
    int nBands = 1;
    if (hostOut && nT == 1 && !g_arena) {
        SrRegion G = sr_region(jobs[0].mesh, info[0]);
        const SRWLRadMesh& m = jobs[0].mesh;
        double outBytes = 16. * m.ne * m.nx * m.ny;
        if (outBytes >= 16. * (1 << 20)) nBands = std::min(8, G.nzc);
        if (g_ctx.forceBands > 0) nBands = std::min(g_ctx.forceBands, G.nzc);
    }
    // Streamed rows: instead of computing row bands one after another, one
    // pass whose energy-tile blocks (in row order) finish their own points
    // and then set a completion flag in host-coherent memory; the host
    // downloads each band as soon as the flags of its blocks are set, and patches
    // the points the FP64 pass recomputes afterwards.  Taken when the banded
    // download would be, the request has energy tiles, and one pass without
    // an s-range split fills the GPU.
    int nDl = 0;
    if (nBands > 1 && g_ctx.streamRows && dRowOf[0] >= 0 && g_ctx.forceSplit <= 1) {
        SrRegion G = sr_region(jobs[0].mesh, info[0]);
        const long long ne = jobs[0].mesh.ne, blocks = ((long long)G.nxc * G.nzc + 255) / 256;
        if ((ne / g_ctx.tileE + ne % g_ctx.tileE) * blocks >= 1024) {
            // bands only set the copies' granularity here: more of them keep
            // the copy after the last flag short
            nDl = g_ctx.forceBands > 0 ? nBands : std::min(16, G.nzc);
            nBands = 1;
        }
    }
    std::vector<srwd_item> items;
    std::vector<int> blockItem;
    std::vector<int> bandBlock0(nBands + 1, 0), bandRow0(nBands + 1, 0);
    // SR launch lists: {item, block of the item}; tile blocks name the first
    // item of a group of SRWD_TILE_E energies
    std::vector<int2> tileBlocks, restBlocks;
    std::vector<int> bandTile0(nBands + 1, 0), bandRest0(nBands + 1, 0);
    for (int b = 0; b < nBands; b++) {
        bandBlock0[b] = (int)blockItem.size();
        bandTile0[b] = (int)tileBlocks.size();
        bandRest0[b] = (int)restBlocks.size();
        for (int j = 0; j < nT; j++) {
            SrJob& J = jobs[j];
            const srwd_trj_info& I = info[j];
            const SRWLRadMesh& m = J.mesh;
            int nx = (int)m.nx, nz = (int)m.ny, ne = (int)m.ne;
            SrRegion G = sr_region(m, I);
            // bands shrink towards the end (row fraction 1 - (1 - b/n)^1.5 before
            // band b): the last band's download, which nothing overlaps, is short
            auto cut = [&](int q) {
                return q >= nBands ? G.nzc
                                   : (int)std::lround(G.nzc * (1. - std::pow(1. - (double)q / nBands, 1.5)));
            };
            int r0 = cut(b), r1 = cut(b + 1);
            if (nBands > 1) { bandRow0[b] = r0; bandRow0[b + 1] = r1; }
            if (r1 <= r0) continue;
            double normConst = std::sqrt(normBase * J.Iavg * 1.E+09 / 1.602189246E-19);
            const int tE = g_ctx.tileE;
            const int nTiled = dRowOf[j] >= 0 ? (ne / tE) * tE : 0;
            for (int ie = 0; ie < ne; ie++) {
                srwd_item it;
                std::memset(&it, 0, sizeof(it));
                it.traj = j;
                it.eph = coord[reqs[j].eOff + ie];
                it.arow = rowBase[j] + ie;
                it.xStart = m.xStart; it.xStep = G.xStep; it.zStart = m.yStart; it.zStep = G.zStep;
                it.nx = nx; it.nz = nz; it.nxc = G.nxc; it.nzc = G.nzc;
                it.iz0 = r0;
                it.symX = G.symX; it.symZ = G.symZ;
                it.ne = ne; it.ie = ie;
                it.outOffset = outOff[j] + ie;
                it.npts = G.nxc * (r1 - r0);
                it.blockBase = (int)blockItem.size() - bandBlock0[b];   // relative to the band's first block
                it.pointBase = (long long)blockItem.size() * 256;
                it.actNorm = normConst * it.eph * 0.80654658E-03;
                it.drow = dRowOf[j];
                it.de = eStepOf[j];
                it.tiled = ie < nTiled ? 1 : 0;
                int nb = (it.npts + 255) / 256;
                const int me = (int)items.size();
                for (int q = 0; q < nb; q++) blockItem.push_back(me);
                if (!it.tiled) for (int q = 0; q < nb; q++) restBlocks.push_back(make_int2(me, q));
                else if (ie % tE == 0) for (int q = 0; q < nb; q++) tileBlocks.push_back(make_int2(me, q));
                items.push_back(it);
            }
        }
    }
    bandBlock0[nBands] = (int)blockItem.size();
    bandTile0[nBands] = (int)tileBlocks.size();
    bandRest0[nBands] = (int)restBlocks.size();
    std::vector<int> dlRow0, dlLast;
    if (nDl) {
        // the tile blocks in row order (every energy group of a 256-point
        // block before the next block), so that rows are finished in order
        std::stable_sort(tileBlocks.begin(), tileBlocks.end(), [](int2 a, int2 b) { return a.y < b.y; });
        SrRegion G = sr_region(jobs[0].mesh, info[0]);
        dlRow0.resize(nDl + 1);
        for (int q = 0; q <= nDl; q++)
            dlRow0[q] = q >= nDl ? G.nzc : (int)std::lround(G.nzc * (1. - std::pow(1. - (double)q / nDl, 1.5)));
        auto bandOf = [&](int r) {
            int b = 0;
            while (b + 1 < nDl && dlRow0[b + 1] <= r) b++;
            return b;
        };
        // last tile block (in launch order) whose points touch each band
        dlLast.assign(nDl, -1);
        for (int i = 0; i < (int)tileBlocks.size(); i++) {
            const srwd_item& it = items[tileBlocks[i].x];
            const int p0 = tileBlocks[i].y * 256, p1 = std::min(p0 + 255, it.npts - 1);
            for (int q = bandOf(it.iz0 + p0 / it.nxc); q <= bandOf(it.iz0 + p1 / it.nxc); q++) dlLast[q] = i;
        }
        for (int q = 1; q < nDl; q++) dlLast[q] = std::max(dlLast[q], dlLast[q - 1]);
        bandRow0 = dlRow0;
    }
    int nItems = (int)items.size();
    int nRows = (int)rowTraj.size();
    int nBlocks = (int)blockItem.size();
    if (nBlocks == 0) return 0;
    long long nPointsPad = (long long)nBlocks * 256;
    // split the integration range when a band's mesh alone does not fill the GPU
    long long minChunks = ldC;
    for (auto& I : info) minChunks = std::min(minChunks, (I.nInt + SRWD_CHUNK - 1) / SRWD_CHUNK);
    // workgroups of a band: every block of the single-energy items, one per
    // group of SRWD_TILE_E energies for the tiles (the tile kernel's
    // workgroups each do SRWD_TILE_E times the work, so fewer fill the GPU)
    const int nWG = (int)(tileBlocks.size() + restBlocks.size());
    const int bandWG = (nWG + nBands - 1) / nBands;
    const int wgTarget = tileBlocks.empty() ? 4096 : 1024;
    int nSplit = (int)std::min<long long>(std::max<long long>(1, (wgTarget + bandWG - 1) / bandWG), minChunks);
    if (g_ctx.forceSplit > 0) nSplit = (int)std::min<long long>(g_ctx.forceSplit, minChunks);
    srwd_item* dItems = (srwd_item*)B.Items.get(sizeof(srwd_item) * nItems);
    int* dBlockItem = (int*)B.BlockItem.get(sizeof(int) * nBlocks);
    int2* dSrBlocks = (int2*)B.SrBlocks.get(sizeof(int2) * (size_t)std::max(1, nWG));
    int* dRowTraj = (int*)B.RowTraj.get(sizeof(int) * nRows);
    double* dRowE = (double*)B.RowE.get(sizeof(double) * nRows);
    long long* dRowN = (long long*)B.RowN.get(sizeof(long long) * nRows);
    float* dA = (float*)B.A.get(sizeof(float) * (size_t)ldI * nRows);
    double* dPart = (double*)B.Part.get(sizeof(double) * 4 * (size_t)nSplit * nPointsPad);
    int* dWarn = (int*)B.Warn.get(sizeof(int));
    // condition check of the FP32 sums and the FP64 recomputation list (one band at a time)
    int maxBandBlocks = 0;
    for (int b = 0; b < nBands; b++) maxBandBlocks = std::max(maxBandBlocks, bandBlock0[b + 1] - bandBlock0[b]);
    double* dCond = (double*)B.Cond.get(sizeof(double) * 5 * nT);
    // one list per compute stream: consecutive bands run on alternate streams
    const int nLists = nBands > 1 ? 2 : 1;
    int* dFpList = (int*)B.FpList.get(sizeof(int) * 256 * (size_t)maxBandBlocks * nLists);
    int* dFpCount = (int*)B.FpCount.get(sizeof(int) * (size_t)nBands);
    if (!dItems || !dBlockItem || !dSrBlocks || !dRowTraj || !dRowE || !dRowN || !dA || !dPart || !dWarn || !dCond || !dFpList ||
        !dFpCount)
        return SRWL_ERR_MEMORY_ALLOCATION_FAILURE;
    if (int e_ = h2d(dItems, items.data(), sizeof(srwd_item) * nItems, st)) return e_;
    if (int e_ = h2d(dBlockItem, blockItem.data(), sizeof(int) * nBlocks, st)) return e_;
    int2* dTileBlocks = dSrBlocks;
    int2* dRestBlocks = dSrBlocks + tileBlocks.size();
    if (!tileBlocks.empty())
        if (int e_ = h2d(dTileBlocks, tileBlocks.data(), sizeof(int2) * tileBlocks.size(), st)) return e_;
    if (!restBlocks.empty())
        if (int e_ = h2d(dRestBlocks, restBlocks.data(), sizeof(int2) * restBlocks.size(), st)) return e_;
    if (int e_ = h2d(dRowTraj, rowTraj.data(), sizeof(int) * nRows, st)) return e_;
    if (int e_ = h2d(dRowE, rowE.data(), sizeof(double) * nRows, st)) return e_;
    if (int e_ = h2d(dRowN, rowN.data(), sizeof(long long) * nRows, st)) return e_;
    HIPCHK(hipMemsetAsync(dWarn, 0, sizeof(int), st));
    if (srwd_launch_phase_table(dQ, ldI, dRowTraj, dRowE, dRowN, nRows, maxInt, dA, st)) return SRWL_AMD_ERR_GPU;
    if (srwd_launch_cond_sums(dInfo, nT, dRec, ldI, dCond, st)) return SRWL_AMD_ERR_GPU;
    HIPCHK(hipMemsetAsync(dFpCount, 0, sizeof(int) * (size_t)nBands, st));
    g_ctx.lastSrBlocks = nBlocks;
    g_ctx.lastSrSplit = nSplit;
    // events of a banded pass: per band a timing pair around the SR launch,
    // "computed" and "downloaded"; a single band is timed with ev0 / ev1
    if (nBands > 1) {
        size_t needEv = 4 * (size_t)nBands;
        while (g_ctx.bandEv.size() < needEv) {
            hipEvent_t e;
            HIPCHK(hipEventCreate(&e));
            g_ctx.bandEv.push_back(e);
        }
    }
    g_ctx.nTimePairs = (g_ctx.timing && nBands > 1) ? nBands : 0;
    // host side of the streamed download
    const SRWLRadMesh& m0 = jobs[0].mesh;
    size_t rowBytes = sizeof(float) * 2 * (size_t)m0.nx * m0.ne;
    size_t compBytes = rowBytes * m0.ny;
    char* pin = nullptr;
    if (nBands > 1) {
        pin = pinned_staging(2 * compBytes);
        if (!pin) return SRWL_ERR_MEMORY_ALLOCATION_FAILURE;
    }
    auto rows_of_band = [&](int b, int* ranges) {   // [a, b) of computed rows and of their mirror images
        int nz = (int)m0.ny, hz = nz >> 1;
        ranges[0] = bandRow0[b]; ranges[1] = bandRow0[b + 1];
        ranges[2] = ranges[3] = 0;
        if (items[0].symZ) {
            int hi = std::min(bandRow0[b + 1], hz);
            if (bandRow0[b] < hi) { ranges[2] = nz - hi; ranges[3] = nz - bandRow0[b]; }
        }
    };
    auto host_copy_band = [&](int b) {
        int r[4];
        rows_of_band(b, r);
        CopySeg seg[4];
        int nSeg = 0;
        for (int k = 0; k < 4; k += 2) {
            if (r[k + 1] <= r[k]) continue;
            size_t o = rowBytes * r[k], n = rowBytes * (r[k + 1] - r[k]);
            seg[nSeg++] = CopySeg{(char*)hostOut->ex + o, pin + o, n};
            seg[nSeg++] = CopySeg{(char*)hostOut->ez + o, pin + compBytes + o, n};
        }
        par_memcpy(seg, nSeg);
    };
    if (nDl) {
        // --- streamed rows (see above) ---------------------------------------
        pin = pinned_staging(2 * compBytes);
        if (!pin) return SRWL_ERR_MEMORY_ALLOCATION_FAILURE;
        const int nTile = (int)tileBlocks.size(), nRest = (int)restBlocks.size();
        if ((size_t)nTile > g_ctx.blockDoneCap) {
            if (g_ctx.blockDone) (void)hipHostFree(g_ctx.blockDone);
            g_ctx.blockDone = nullptr;
            g_ctx.blockDoneCap = 0;
            if (hipHostMalloc((void**)&g_ctx.blockDone, sizeof(unsigned) * (size_t)nTile, hipHostMallocCoherent) !=
                hipSuccess) {
                g_ctx.blockDone = nullptr;
                return SRWL_ERR_MEMORY_ALLOCATION_FAILURE;
            }
            std::memset(g_ctx.blockDone, 0, sizeof(unsigned) * (size_t)nTile);
            g_ctx.blockDoneCap = nTile;
        }
        // flags of this call carry its own epoch (never 0), so they need no clearing
        const unsigned epoch = ++g_ctx.streamEpoch ? g_ctx.streamEpoch : ++g_ctx.streamEpoch;
        if (!g_ctx.patchHost &&
            hipHostMalloc((void**)&g_ctx.patchHost, sizeof(srwd_patch) * SRWL_AMD_PATCH_CAP + 256,
                          hipHostMallocDefault) != hipSuccess) {
            g_ctx.patchHost = nullptr;
            return SRWL_ERR_MEMORY_ALLOCATION_FAILURE;
        }
        srwd_patch* dPatch = (srwd_patch*)B.Patch.get(sizeof(srwd_patch) * SRWL_AMD_PATCH_CAP);
        int* dPatchCount = (int*)B.PatchCount.get(sizeof(int));
        if (!dPatch || !dPatchCount) return SRWL_ERR_MEMORY_ALLOCATION_FAILURE;
        HIPCHK(hipMemsetAsync(dPatchCount, 0, sizeof(int), st));
        volatile unsigned* done = g_ctx.blockDone;
        while (g_ctx.bandEv.size() < 4 * (size_t)nDl) {
            hipEvent_t e;
            HIPCHK(hipEventCreate(&e));
            g_ctx.bandEv.push_back(e);
        }
        const double* cnd = g_ctx.condTol > 0 ? dCond : nullptr;
        // the single-energy items (energies left over from the groups) on the
        // second stream, finished there by the two-pass form
        HIPCHK(hipEventRecord(g_ctx.evFork, st));
        HIPCHK(hipStreamWaitEvent(g_ctx.stream2, g_ctx.evFork, 0));
        if (srwd_launch_sr(dItems, dRestBlocks, nRest, 0, dInfo, dRec, dChunk, dA, ldI, ldC, 1, nPointsPad, dPart,
                           g_ctx.stream2) ||
            srwd_launch_finalize(dItems, dRestBlocks, nRest, 1, dInfo, dEdge, termin, dPart, 1, nPointsPad, dEx, dEz,
                                 dWarn, cnd, g_ctx.condTol, dFpList, dFpCount, dReq, dCoord, dLast, g_ctx.stream2))
            return SRWL_AMD_ERR_GPU;
        HIPCHK(hipEventRecord(g_ctx.evJoin, g_ctx.stream2));
        srwd_fuse F{};
        F.edges = dEdge; F.termin = termin; F.Ex = (float2*)dEx; F.Ez = (float2*)dEz; F.warn = dWarn;
        F.cond = cnd; F.condTol = g_ctx.condTol; F.fpList = dFpList; F.fpCount = dFpCount;
        F.reqs = dReq; F.coord = dCoord; F.last = dLast; F.blockDone = g_ctx.blockDone; F.epoch = epoch;
        if (g_ctx.timing) HIPCHK(hipEventRecord(g_ctx.ev0, st));
        if (srwd_launch_sr_fused(dItems, dTileBlocks, nTile, g_ctx.tileE, dInfo, dRec, dChunk, dA, ldI, ldC,
                                 nPointsPad, dPart, F, st))
            return SRWL_AMD_ERR_GPU;
        if (g_ctx.timing) HIPCHK(hipEventRecord(g_ctx.ev1, st));
        HIPCHK(hipEventRecord(g_ctx.evTiles, st));
        HIPCHK(hipStreamWaitEvent(st, g_ctx.evJoin, 0));
        // FP64 recomputation, every store also recorded for the host
        if (g_ctx.condTol > 0 &&
            srwd_launch_sr_fp64(dItems, dBlockItem, dInfo, dEdge, termin, dT64, ldI, dEx, dEz, dWarn, dFpList,
                                dFpCount, dReq, dCoord, dLast, dPart, dPatch, dPatchCount, SRWL_AMD_PATCH_CAP, st))
            return SRWL_AMD_ERR_GPU;
        HIPCHK(hipMemcpyAsync(g_ctx.warnHost, dWarn, sizeof(int), hipMemcpyDeviceToHost, st));
        HIPCHK(hipMemcpyAsync(g_ctx.fpCountHost, dFpCount, sizeof(int), hipMemcpyDeviceToHost, st));
        HIPCHK(hipMemcpyAsync(g_ctx.patchHost, dPatchCount, sizeof(int), hipMemcpyDeviceToHost, st));
        g_ctx.fpBands = 1;
        long long evs = 0;
        for (auto& it : items) evs += (long long)it.npts * info[it.traj].nInt;
        g_ctx.pendingEvals = evs;
        if (hostOut->queueMore) {
            rc = hostOut->queueMore(hostOut->queueCtx, res);
            if (rc) return rc;
        }
        // Each band's copy is queued as soon as the flags of its blocks are
        // set, whatever the host copies' progress (the staging holds the
        // whole field), and the bands are copied on into the caller's arrays
        // in order as their copies land.  A block's flag is visible once its
        // stores are; blocks finish nearly in launch order, so a cursor walks
        // the flags in that order.  Once the kernel is known to have
        // finished, a flag still missing means the flags do not reach the
        // host: fall back.
        const auto t0 = std::chrono::steady_clock::now();
        bool late = false;
        int cursor = 0, nextDl = 0, nextCopy = 0;
        for (unsigned spin = 0; nextCopy < nDl && !late;) {
            while (cursor < nTile && done[cursor] == epoch) cursor++;
            bool progress = false;
            while (nextDl < nDl && cursor > dlLast[nextDl]) {
                const int b = nextDl++;
                if (b == 0) HIPCHK(hipStreamWaitEvent(g_ctx.copyStream, g_ctx.evJoin, 0));
                int r[4];
                rows_of_band(b, r);
                for (int k = 0; k < 4; k += 2) {
                    if (r[k + 1] <= r[k]) continue;
                    size_t o = rowBytes * r[k], n = rowBytes * (r[k + 1] - r[k]);
                    HIPCHK(hipMemcpyAsync(pin + o, (char*)dEx + o, n, hipMemcpyDeviceToHost, g_ctx.copyStream));
                    HIPCHK(hipMemcpyAsync(pin + compBytes + o, (char*)dEz + o, n, hipMemcpyDeviceToHost,
                                          g_ctx.copyStream));
                }
                HIPCHK(hipEventRecord(g_ctx.bandEv[4 * b + 1], g_ctx.copyStream));
                progress = true;
            }
            if (nextCopy < nextDl && hipEventQuery(g_ctx.bandEv[4 * nextCopy + 1]) == hipSuccess) {
                host_copy_band(nextCopy++);
                progress = true;
            }
            if (progress || (++spin & 1023)) continue;
            const bool timedOut =
                std::chrono::steady_clock::now() - t0 > std::chrono::seconds(SRWL_AMD_STREAM_WAIT_S);
            if (nextDl < nDl) {
                const bool finished = hipEventQuery(g_ctx.evTiles) == hipSuccess;
                if ((finished && done[cursor] != epoch) || timedOut) late = true;
            } else if (timedOut) {
                late = true;
            }
        }
        (void)hipGetLastError();   // the queries' "not ready"
        HIPCHK(hipStreamSynchronize(st));
        const int nPatch = *(volatile int*)g_ctx.patchHost;
        g_ctx.lastStreamFallback = late || nPatch > SRWL_AMD_PATCH_CAP;
        g_ctx.lastStreamed = !g_ctx.lastStreamFallback;
        if (g_ctx.lastStreamFallback) {
            // flags not all seen, or more FP64 stores than the patch area
            // holds: the whole field, now final, the plain way
            HIPCHK(hipStreamSynchronize(g_ctx.copyStream));
            rc = download(hostOut->ex, dEx, compBytes);
            if (!rc) rc = download(hostOut->ez, dEz, compBytes);
            if (rc) return rc;
        } else {
            // the FP64-recomputed values over what the bands carried
            srwd_patch* P = (srwd_patch*)(g_ctx.patchHost + 256);
            if (nPatch > 0)
                HIPCHK(hipMemcpy(P, dPatch, sizeof(srwd_patch) * (size_t)nPatch, hipMemcpyDeviceToHost));
            float2* hx = (float2*)hostOut->ex;
            float2* hz = (float2*)hostOut->ez;
            for (int k = 0; k < nPatch; k++) { hx[P[k].off] = P[k].ex; hz[P[k].off] = P[k].ez; }
        }
        if (defer) return 0;
        HIPCHK(hipStreamSynchronize(st));
        sr_complete(res);
        return 0;
    }
    if (nBands > 1) {
        // bands alternate between the two compute streams, so that a band's
        // last blocks share the GPU with the next band's first ones
        HIPCHK(hipEventRecord(g_ctx.evFork, st));
        HIPCHK(hipStreamWaitEvent(g_ctx.stream2, g_ctx.evFork, 0));
    }
    for (int b = 0; b < nBands; b++) {
        int b0 = bandBlock0[b], nb = bandBlock0[b + 1] - b0;
        hipStream_t sb = (nBands > 1 && (b & 1)) ? g_ctx.stream2 : st;
        int* fpList = dFpList + (size_t)256 * maxBandBlocks * (b & (nLists - 1));
        if (nb > 0) {
            if (g_ctx.timing) HIPCHK(hipEventRecord(nBands > 1 ? g_ctx.bandEv[4 * b + 2] : g_ctx.ev0, sb));
            const int nTile = bandTile0[b + 1] - bandTile0[b], nRest = bandRest0[b + 1] - bandRest0[b];
            // an unbanded pass runs the few single-energy blocks on the second
            // stream, beside the energy tiles, instead of after them
            const bool side = nBands == 1 && nTile > 0 && nRest > 0;
            if (side) {
                HIPCHK(hipEventRecord(g_ctx.evFork, sb));
                HIPCHK(hipStreamWaitEvent(g_ctx.stream2, g_ctx.evFork, 0));
            }
            if (srwd_launch_sr(dItems, dRestBlocks + bandRest0[b], nRest, 0, dInfo, dRec, dChunk, dA, ldI, ldC, nSplit,
                               nPointsPad, dPart, side ? g_ctx.stream2 : sb) ||
                srwd_launch_sr(dItems, dTileBlocks + bandTile0[b], nTile, g_ctx.tileE, dInfo, dRec, dChunk, dA, ldI,
                               ldC, nSplit, nPointsPad, dPart, sb))
                return SRWL_AMD_ERR_GPU;
            if (side) {
                HIPCHK(hipEventRecord(g_ctx.evJoin, g_ctx.stream2));
                HIPCHK(hipStreamWaitEvent(sb, g_ctx.evJoin, 0));
            }
            if (g_ctx.timing) HIPCHK(hipEventRecord(nBands > 1 ? g_ctx.bandEv[4 * b + 3] : g_ctx.ev1, sb));
            // energy tiles in groups (runs of consecutive photon energies stored together), then the rest
            const double* cnd = g_ctx.condTol > 0 ? dCond : nullptr;
            if (srwd_launch_finalize(dItems, dTileBlocks + bandTile0[b], nTile, g_ctx.tileE > 0 ? g_ctx.tileE : 1,
                                     dInfo, dEdge, termin, dPart, nSplit, nPointsPad, dEx, dEz, dWarn, cnd,
                                     g_ctx.condTol, fpList, dFpCount + b, dReq, dCoord, dLast, sb) ||
                srwd_launch_finalize(dItems, dRestBlocks + bandRest0[b], nRest, 1, dInfo, dEdge, termin, dPart,
                                     nSplit, nPointsPad, dEx, dEz, dWarn, cnd, g_ctx.condTol, fpList, dFpCount + b,
                                     dReq, dCoord, dLast, sb))
                return SRWL_AMD_ERR_GPU;
            // a single band's partial sums are consumed by now: their buffer
            // (at least 4 doubles per point) holds the FP64 pass's sums
            if (g_ctx.condTol > 0 && srwd_launch_sr_fp64(dItems, dBlockItem + b0, dInfo, dEdge, termin, dT64, ldI, dEx,
                                                         dEz, dWarn, fpList, dFpCount + b, dReq, dCoord, dLast,
                                                         nBands == 1 ? dPart : nullptr, nullptr, nullptr, 0, sb))
                return SRWL_AMD_ERR_GPU;
        }
        if (nBands > 1) {
            // download band b on the copy stream once computed
            HIPCHK(hipEventRecord(g_ctx.bandEv[4 * b], sb));
            HIPCHK(hipStreamWaitEvent(g_ctx.copyStream, g_ctx.bandEv[4 * b], 0));
            int r[4];
            rows_of_band(b, r);
            for (int k = 0; k < 4; k += 2) {
                if (r[k + 1] <= r[k]) continue;
                size_t o = rowBytes * r[k], n = rowBytes * (r[k + 1] - r[k]);
                HIPCHK(hipMemcpyAsync(pin + o, (char*)dEx + o, n, hipMemcpyDeviceToHost, g_ctx.copyStream));
                HIPCHK(hipMemcpyAsync(pin + compBytes + o, (char*)dEz + o, n, hipMemcpyDeviceToHost,
                                      g_ctx.copyStream));
            }
            HIPCHK(hipEventRecord(g_ctx.bandEv[4 * b + 1], g_ctx.copyStream));
        }
    }
    if (nBands > 1) {
        HIPCHK(hipEventRecord(g_ctx.evJoin, g_ctx.stream2));
        HIPCHK(hipStreamWaitEvent(st, g_ctx.evJoin, 0));
    }
    // the warning flag and the FP64 point counts are read back with the caller's final synchronisation
    HIPCHK(hipMemcpyAsync(g_ctx.warnHost, dWarn, sizeof(int), hipMemcpyDeviceToHost, st));
    HIPCHK(hipMemcpyAsync(g_ctx.fpCountHost, dFpCount, sizeof(int) * (size_t)nBands, hipMemcpyDeviceToHost, st));
    g_ctx.fpBands = nBands;
    long long ev = 0;
    for (auto& it : items) ev += (long long)it.npts * info[it.traj].nInt;
    g_ctx.pendingEvals = ev;
    if (nBands > 1) {
        // work that follows the field on the device is queued before the host
        // starts waiting for bands
        if (hostOut && hostOut->queueMore) {
            rc = hostOut->queueMore(hostOut->queueCtx, res);
            if (rc) return rc;
        }
        // every band is queued; the host moves each downloaded band on into the
        // caller's arrays while the later ones are still computed
        for (int b = 0; b < nBands; b++) {
            HIPCHK(hipEventSynchronize(g_ctx.bandEv[4 * b + 1]));
            host_copy_band(b);
        }
    } else if (hostOut && g_arena) {
        // no pinned staging in a no-allocation call: plain copies, finished
        // by the caller's synchronisation
        HIPCHK(hipMemcpyAsync(hostOut->ex, dEx, compBytes, hipMemcpyDeviceToHost, st));
        HIPCHK(hipMemcpyAsync(hostOut->ez, dEz, compBytes, hipMemcpyDeviceToHost, st));
    } else if (hostOut) {
        size_t n = compBytes;
        rc = download(hostOut->ex, dEx, n);
        if (!rc) rc = download(hostOut->ez, dEz, n);
        if (rc) return rc;
    }
    if (defer) return 0;
    HIPCHK(hipStreamSynchronize(st));
    sr_complete(res);
    return 0;
}

char* pinned_staging(size_t bytes)
{
    if (bytes > g_ctx.pinBytes) {
        if (g_ctx.pin) (void)hipHostFree(g_ctx.pin);
        g_ctx.pin = nullptr;
        g_ctx.pinBytes = 0;
        if (hipHostMalloc((void**)&g_ctx.pin, bytes, hipHostMallocDefault) != hipSuccess) {
            g_ctx.pin = nullptr;
            return nullptr;
        }
        g_ctx.pinBytes = bytes;
    }
    return g_ctx.pin;
}

namespace {
// A copy whose stores bypass the caches (16-byte streaming stores once the
// destination is aligned): the caller's arrays are written once and not read
// back here, so reading each destination line before writing it (what an
// ordinary store does) only costs bandwidth.
void copy_stream(void* dst, const void* src, size_t n)
{
    char* d = (char*)dst;
    const char* s = (const char*)src;
    const size_t head = std::min(n, (size_t)((16 - ((uintptr_t)d & 15)) & 15));
    std::memcpy(d, s, head);
    d += head; s += head; n -= head;
    size_t k = 0;
    for (; k + 64 <= n; k += 64) {
        const __m128i a = _mm_loadu_si128((const __m128i*)(s + k));
        const __m128i b = _mm_loadu_si128((const __m128i*)(s + k + 16));
        const __m128i c = _mm_loadu_si128((const __m128i*)(s + k + 32));
        const __m128i e = _mm_loadu_si128((const __m128i*)(s + k + 48));
        _mm_stream_si128((__m128i*)(d + k), a);
        _mm_stream_si128((__m128i*)(d + k + 16), b);
        _mm_stream_si128((__m128i*)(d + k + 32), c);
        _mm_stream_si128((__m128i*)(d + k + 48), e);
    }
    std::memcpy(d + k, s + k, n - k);
    _mm_sfence();
}

// Host copies out of the pinned staging: a pool of worker threads started on
// first use and kept for the life of the process (starting threads per copy
// costs more than a small band's copy).  A job is a list of ranges cut into
// 1 MiB pieces that the workers and the calling thread take in turn.  A forked
// child, which has none of the workers, copies on its own thread.
class CopyPool {
public:
    explicit CopyPool(unsigned nWorkers) : pid_(getpid())
    {
        for (unsigned i = 0; i < nWorkers; i++) std::thread([this] { worker(); }).detach();
        workers_ = nWorkers;
    }
    void copy(const CopySeg* seg, int nSeg)
    {
        size_t total = 0;
        for (int k = 0; k < nSeg; k++) total += seg[k].n;
        if (workers_ == 0 || total <= kPiece || getpid() != pid_) {
            for (int k = 0; k < nSeg; k++) copy_stream(seg[k].dst, seg[k].src, seg[k].n);
            return;
        }
        std::lock_guard<std::mutex> one(job_);   // one job at a time
        pieces_.clear();
        for (int k = 0; k < nSeg; k++)
            for (size_t o = 0; o < seg[k].n; o += kPiece)
                pieces_.push_back(CopySeg{(char*)seg[k].dst + o, (const char*)seg[k].src + o,
                                          std::min(kPiece, seg[k].n - o)});
        next_.store(0);
        {
            std::lock_guard<std::mutex> l(m_);
            pending_ = workers_;
            gen_++;
        }
        cv_.notify_all();
        run();
        std::unique_lock<std::mutex> l(m_);
        cvDone_.wait(l, [&] { return pending_ == 0; });
    }

private:
    static constexpr size_t kPiece = 1 << 20;
    void run()
    {
        for (size_t k; (k = next_.fetch_add(1)) < pieces_.size();)
            copy_stream(pieces_[k].dst, pieces_[k].src, pieces_[k].n);
    }
    void worker()
    {
        unsigned seen = 0;
        for (;;) {
            {
                std::unique_lock<std::mutex> l(m_);
                cv_.wait(l, [&] { return gen_ != seen; });
                seen = gen_;
            }
            run();
            std::lock_guard<std::mutex> l(m_);
            if (--pending_ == 0) cvDone_.notify_one();
        }
    }
    pid_t pid_;
    unsigned workers_ = 0;
    std::mutex job_, m_;
    std::condition_variable cv_, cvDone_;
    std::vector<CopySeg> pieces_;
    std::atomic<size_t> next_{0};
    unsigned pending_ = 0, gen_ = 0;
};

CopyPool& copy_pool()
{
    // never destroyed: the workers wait on it until the process ends
    static CopyPool* pool =
        new CopyPool(std::min(7u, std::max(1u, std::thread::hardware_concurrency()) - 1));
    return *pool;
}
}  // namespace

void par_memcpy(const CopySeg* seg, int nSeg) { copy_pool().copy(seg, nSeg); }

void par_memcpy(void* dst, const void* src, size_t bytes)
{
    CopySeg s{dst, src, bytes};
    par_memcpy(&s, 1);
}

int download(void* dst, const void* src, size_t bytes)
{
    // pieces of up to 32 MiB through the pinned staging: the DMA of piece k + 1
    // overlaps the host copy of piece k
    const size_t piece = 32 << 20;
    size_t np = (bytes + piece - 1) / piece;
    char* pin = pinned_staging(std::min(bytes, 2 * piece));
    if (!pin) return SRWL_ERR_MEMORY_ALLOCATION_FAILURE;
    while (g_ctx.bandEv.size() < 4) {
        hipEvent_t e;
        HIPCHK(hipEventCreate(&e));
        g_ctx.bandEv.push_back(e);
    }
    HIPCHK(hipEventRecord(g_ctx.bandEv[0], g_ctx.stream));
    HIPCHK(hipStreamWaitEvent(g_ctx.copyStream, g_ctx.bandEv[0], 0));
    for (size_t k = 0; k < np; k++) {
        size_t o = k * piece, n = std::min(piece, bytes - o);
        char* stage = pin + (k & 1) * piece;
        if (k >= 2) HIPCHK(hipEventSynchronize(g_ctx.bandEv[2 + (k & 1)]));   // stage free again
        HIPCHK(hipMemcpyAsync(stage, (const char*)src + o, n, hipMemcpyDeviceToHost, g_ctx.copyStream));
        HIPCHK(hipEventRecord(g_ctx.bandEv[2 + (k & 1)], g_ctx.copyStream));
        if (k >= 1) {
            size_t op = (k - 1) * piece, npv = std::min(piece, bytes - op);
            HIPCHK(hipEventSynchronize(g_ctx.bandEv[2 + ((k - 1) & 1)]));
            par_memcpy((char*)dst + op, pin + ((k - 1) & 1) * piece, npv);
        }
    }
    if (np > 0) {
        size_t k = np - 1, o = k * piece, n = std::min(piece, bytes - o);
        HIPCHK(hipEventSynchronize(g_ctx.bandEv[2 + (k & 1)]));
        par_memcpy((char*)dst + o, pin + (k & 1) * piece, n);
    }
    return 0;
}

SrRegion sr_region(const SRWLRadMesh& m, const srwd_trj_info& I)
{
    SrRegion G;
    int nx = (int)m.nx, nz = (int)m.ny;
    G.xStep = nx > 1 ? (m.xFin - m.xStart) / (nx - 1) : 0;
    G.zStep = nz > 1 ? (m.yFin - m.yStart) / (nz - 1) : 0;
    double xS = m.xStart, xE = m.xStart + (nx - 1) * G.xStep;
    double zS = m.yStart, zE = m.yStart + (nz - 1) * G.zStep;
    // srradint.cpp:3231-3260
    if (!I.verNZ && nx > 1) {
        double tol = G.xStep * 0.01;
        G.symX = (std::fabs(I.dxds0 * (I.yObs - I.s0)) < tol) && (std::fabs(0.5 * (xS + xE) - I.x0) < tol);
    }
    if (!I.horNZ && nz > 1) {
        double tol = G.zStep * 0.01;
        G.symZ = (std::fabs(I.dzds0 * (I.yObs - I.s0)) < tol) && (std::fabs(0.5 * (zS + zE) - I.z0) < tol);
    }
    G.nzc = nz; G.nxc = nx;
    // srradint.cpp:331-345: rows / columns beyond the centre are mirrored
    double z = zS;
    for (int iz = 0; iz < nz; iz++) { if (G.symZ && (z - I.z0) > G.zStep * 0.001) { G.nzc = iz; break; } z += G.zStep; }
    double x = xS;
    for (int ix = 0; ix < nx; ix++) { if (G.symX && (x - I.x0) > G.xStep * 0.001) { G.nxc = ix; break; } x += G.xStep; }
    return G;
}

// srradstr.cpp:1063-1119 CheckNxNzForSR, coordinate representation: the
// number of points along one mesh axis [start, fin] that samples the
// wavefront's curvature phase pi (x - c)^2 / (lambda R) at the mesh edges with
// at most pi / (1.2 f) per step; even, a fast FFT size, at least 8.
long sr_auto_mesh_size(double start, double fin, long n, double c, double R, double lambda, double f)
{
    double step = n > 1 ? (fin - start) / (n - 1) : 0.;
    double halfLamR = 0.5 * lambda * R;
    double sRel = start - c, eRel = start + (n - 1) * step - c;
    double d0 = std::fabs(halfLamR / sRel), d1 = std::fabs(halfLamR / eRel);
    double d = std::min(d0, d1) / (f * 1.2);
    long m = (long)(std::fabs(eRel - sRel) / d) + 1;
    if (((m >> 1) << 1) != m) m++;
    m = next_fft_size(m);
    return std::max(m, 8L);
}

// After the stream has been synchronised: timing statistics and the
// warning of the last sr_run.
void sr_complete(std::vector<SrResult>& res)
{
    if (g_ctx.timing) {
        float ms = 0;
        if (g_ctx.nTimePairs > 0) {
            for (int b = 0; b < g_ctx.nTimePairs; b++) {
                float t = 0;
                (void)hipEventElapsedTime(&t, g_ctx.bandEv[4 * b + 2], g_ctx.bandEv[4 * b + 3]);
                ms += t;
            }
            g_ctx.nTimePairs = 0;
        } else {
            (void)hipEventElapsedTime(&ms, g_ctx.ev0, g_ctx.ev1);
        }
        g_ctx.lastSrMs = ms;
        g_ctx.lastSrEvals = g_ctx.pendingEvals;
        g_ctx.sumSrMs += ms;
        g_ctx.sumSrEvals += (double)g_ctx.pendingEvals;
        g_ctx.nSrCalls += 1;
    }
    if (*g_ctx.warnHost)
        for (auto& r : res) if (!r.warn) r.warn = SRWL_WARN_BAD_RAD_INT_TERMINATIONS;
    long long fp = 0;
    for (int b = 0; b < g_ctx.fpBands; b++) fp += g_ctx.fpCountHost[b];
    g_ctx.lastFp64Points = fp;
    g_ctx.fpBands = 0;
}

// srwlib.cpp:403-452 + 328-373: trajectory range for a request without a
// tabulated trajectory.
int setup_job_from_field(SrJob& J, const SRWLParticle& p, const SRWLMagFldC* mag, const double* precPar,
                         std::vector<srwd_field>& fields)
{
    srwd_field f;
    double gs0, gs1;
    int rc = field_from_c(mag, f, gs0, gs1);
    if (rc) return rc;
    J.field = (int)fields.size();
    fields.push_back(f);
    J.part = p;
    J.np = precPar ? (long long)(int)precPar[4] : 100000;
    double a = precPar ? precPar[2] : 0., b = precPar ? precPar[3] : 0.;
    double sSt = a - p.z, sEn = b - p.z;
    if (sSt >= sEn) {
        sSt = gs0; sEn = gs1;
        if (sSt > p.z) sSt = p.z;
        if (sEn < p.z) sEn = p.z;
        sSt -= p.z; sEn -= p.z;
    }
    J.ctStart = sSt; J.ctEnd = sEn;
    return 0;
}

}  // namespace srwamd

using namespace srwamd;

// ---------------------------------------------------------------------------
// Exported C ABI
// ---------------------------------------------------------------------------
extern "C" {

SRWL_EXPORT void srwlUtiSetWfrModifFunc(int (*pExtFunc)(int action, SRWLWfr* pWfrIn, char pol))
{
    g_wfrModif = pExtFunc;
}

SRWL_EXPORT void srwlUtiSetAllocArrayFunc(char* (*pExtFunc)(char type, long long len))
{
    g_allocArray = pExtFunc;
}

SRWL_EXPORT void srwlUtiSetProgrIndFunc(int (*pExtFunc)(double curVal))
{
    g_progress = pExtFunc;
}

SRWL_EXPORT int srwlUtiVerNo(char* verNoStr, int code)
{
    if (verNoStr == nullptr) return 0;
    std::strcpy(verNoStr, code == 2 ? "0.10" : "4.10");
    return 0;
}

SRWL_EXPORT int srwlUtiGetErrText(char* t, int erNo)
{
    if (t == nullptr) return 0;
    const char* s = "Unknown error or warning";
    switch (erNo) {
    case SRWL_ERR_MEMORY_ALLOCATION_FAILURE: s = "Memory allocation failure"; break;
    case SRWL_ERR_OBS_WITHIN_INTEGR_LIMITS: s = "Longitudinal position of observation point is within radiation integration limits"; break;
    case SRWL_ERR_ELEC_BEAM_NOT_ULTRARELATIVISTIC: s = "Electron beam is not ultra-relativistic"; break;
    case SRWL_ERR_S0_OUT_OF_FIELD_LIMITS: s = "Initial longitudinal position of electron is out of the magnetic field definition limits"; break;
    case SRWL_ERR_WAVELENGTH_COMPARABLE_WITH_DISTANCE: s = "Radiation wavelength is comparable with the observation distance"; break;
    case SRWL_ERR_IN_FFT: s = "Error in FFT (mesh sizes must be even and supported)"; break;
    case SRWL_ERR_INCORRECT_MAG_FIELD_PER: s = "Incorrect magnetic field period"; break;
    case SRWL_ERR_INCORRECT_PER_MAG_FIELD_LENGTH: s = "Incorrect number of periods of magnetic field"; break;
    case SRWL_NO_FUNC_ARG_DATA: s = "No data for function argument"; break;
    case SRWL_INCORRECT_TRJ_STRUCT: s = "Incorrect trajectory structure"; break;
    case SRWL_INCORRECT_PARAM_FOR_TRJ_COMP: s = "Incorrect parameters for trajectory calculation"; break;
    case SRWL_INCORRECT_PARAM_FOR_SR_COMP: s = "Incorrect parameters for SR calculation"; break;
    case SRWL_INCORRECT_PARAM_FOR_INT_EXTR: s = "Incorrect parameters for intensity extraction"; break;
    case SRWL_INCORRECT_WFR_STRUCT: s = "Incorrect wavefront structure"; break;
    case SRWL_WFR_EXT_MODIF_FAILED: s = "Wavefront modification by the calling application failed"; break;
    case SRWL_INCORRECT_PARAM_FOR_RESIZE: s = "Incorrect parameters for wavefront resizing"; break;
    case SRWL_RESIZE_NOT_SUPPORTED_WITH_SUCH_PARAMS:
        s = "Resizing of the wavefront is not supported with these parameters"; break;
    case SRWL_INCORRECT_PARAM_FOR_WFR_PROP: s = "Incorrect parameters for wavefront propagation"; break;
    case SRWL_WFR_EXT_FUNC_NOT_DEFINED: s = "Wavefront modification function of the calling application is not defined"; break;
    case SRWL_INCORRECT_PARAM_FOR_FFT: s = "Incorrect parameters for FFT"; break;
    case SRWL_AMD_ERR_GPU: s = "GPU (HIP) error or no usable GPU"; break;
    case SRWL_AMD_ERR_UNSUPPORTED: s = "Parameter combination not supported by this build"; break;
    case SRWL_AMD_ERR_BATCH: s = "Inconsistent batch request"; break;
    case SRWL_AMD_ERR_NOALLOC_INIT:
        s = "srwlAmdNoAllocInit must be called on this device before srwlCalcElecFieldSRNoAlloc"; break;
    case SRWL_AMD_ERR_NOALLOC_SIZE:
        s = "Request exceeds the fixed scratch of srwlCalcElecFieldSRNoAlloc (method 0, at most 65536 field "
            "points ne*nx*ny, at most 20001 trajectory points, no automatic mesh)"; break;
    case SRWL_WARN_FLUX_ONLY_VS_PHOTON_ENERGY: s = "Flux can only be extracted vs photon energy"; break;
    case SRWL_WARN_BAD_RAD_INT_TERMINATIONS: s = "Terminating terms of the radiation integral were not computed"; break;
    case SRWL_WARN_ELEC_BEAM_NOT_ULTRARELATIVISTIC: s = "Electron beam may be not ultra-relativistic"; break;
    default: break;
    }
    std::strcpy(t, s);
    return 0;
}

SRWL_EXPORT int srwlUtiGPUProc(int op, double* arParGPU)
{
    if (op == 1) {
        if (arParGPU && arParGPU[0] > 0 && !g_ctx.ready) g_ctx.device = (int)arParGPU[1];
        return ensure_device();
    }
    if (op == 0) return 0;
    return SRWL_AMD_ERR_UNSUPPORTED;
}

SRWL_EXPORT int srwlCalcPartTraj(SRWLPrtTrj* pTrj, SRWLMagFldC* pMagFld, double* precPar)
{
    (void)precPar;
    if (pTrj == nullptr || pMagFld == nullptr) return SRWL_NO_FUNC_ARG_DATA;
    if (!pTrj->arX || !pTrj->arXp || !pTrj->arY || !pTrj->arYp || pTrj->np <= 0) return SRWL_INCORRECT_TRJ_STRUCT;
    int rc = ensure_device();
    if (rc) return rc;
    srwd_field f;
    double gs0, gs1;
    rc = field_from_c(pMagFld, f, gs0, gs1);
    if (rc) return rc;
    SRWLParticle& p = pTrj->partInitCond;
    if (pTrj->ctStart >= pTrj->ctEnd) {
        double a = gs0, b = gs1;
        if (a > p.z) a = p.z;
        if (b < p.z) b = p.z;
        pTrj->ctStart = a - p.z; pTrj->ctEnd = b - p.z;
    }
    srwd_traj_job J;
    std::memset(&J, 0, sizeof(J));
    double g = gamma_internal(p);
    J.init[0] = p.x; J.init[1] = p.xp; J.init[2] = p.y; J.init[3] = p.yp; J.init[4] = p.z;
    J.gammaInt = g;
    J.mult = p.nq * 1.602176462e-19 / (9.10938188e-31 * 2.99792458e+08 * g);
    J.ctStart = pTrj->ctStart; J.ctEnd = pTrj->ctEnd; J.np = pTrj->np; J.field = 0;
    hipStream_t st = g_ctx.stream;
    Buffers& B = g_ctx.buf;
    size_t b = sizeof(double) * pTrj->np;
    double* dX = (double*)B.X.get(b); double* dXp = (double*)B.Xp.get(b);
    double* dY = (double*)B.Y.get(b); double* dYp = (double*)B.Yp.get(b); double* dZ = (double*)B.IBx.get(b);
    srwd_traj_job* dJ = (srwd_traj_job*)B.Jobs.get(sizeof(J));
    srwd_field* dF = (srwd_field*)B.Fields.get(sizeof(f));
    if (!dX || !dXp || !dY || !dYp || !dZ || !dJ || !dF) return SRWL_ERR_MEMORY_ALLOCATION_FAILURE;
    HIPCHK(hipMemcpyAsync(dJ, &J, sizeof(J), hipMemcpyHostToDevice, st));
    HIPCHK(hipMemcpyAsync(dF, &f, sizeof(f), hipMemcpyHostToDevice, st));
    double* dPic = nullptr;
    if (!g_ctx.trajSerial) {
        dPic = (double*)B.Pic.get(sizeof(double) * 20 * (size_t)srwd_pic_cap(pTrj->np));
        if (!dPic) return SRWL_ERR_MEMORY_ALLOCATION_FAILURE;
    }
    int one = (f.n == 1 && f.u[0].nh == 1);
    if (srwd_launch_traj(dJ, dF, 1, dX, dXp, dY, dYp, dZ, pTrj->np, dPic, one, st)) return SRWL_AMD_ERR_GPU;
    HIPCHK(hipMemcpyAsync(pTrj->arX, dX, b, hipMemcpyDeviceToHost, st));
    HIPCHK(hipMemcpyAsync(pTrj->arXp, dXp, b, hipMemcpyDeviceToHost, st));
    HIPCHK(hipMemcpyAsync(pTrj->arY, dY, b, hipMemcpyDeviceToHost, st));
    HIPCHK(hipMemcpyAsync(pTrj->arYp, dYp, b, hipMemcpyDeviceToHost, st));
    if (pTrj->arZ) HIPCHK(hipMemcpyAsync(pTrj->arZ, dZ, b, hipMemcpyDeviceToHost, st));
    HIPCHK(hipStreamSynchronize(st));
    if (pTrj->arZp) {
        // srgtrjdt.cpp:150: longitudinal relative velocity from the transverse ones
        double gEm2 = 1. / g / g;
        for (long long i = 0; i < pTrj->np; i++) {
            double x = -(gEm2 + pTrj->arXp[i] * pTrj->arXp[i] + pTrj->arYp[i] * pTrj->arYp[i]);
            pTrj->arZp[i] = (std::fabs(x) > 0.01) ? std::sqrt(1. + x)
                : 1. + x * (0.5 + x * (-0.125 + x * (0.0625 + x * (-0.0390625 + x * (0.02734375 +
                  x * (-0.0205078125 + x * 0.01611328125))))));
        }
    }
    return 0;
}

SRWL_EXPORT int srwlCalcElecFieldSR(SRWLWfr* pWfr, SRWLPrtTrj* pTrj, SRWLMagFldC* pMagFld, double* precPar,
                                    int nPrecPar)
{
    if (pWfr == nullptr || precPar == nullptr) return SRWL_INCORRECT_PARAM_FOR_SR_COMP;
    bool trjDef = !(pTrj == nullptr || ((!pTrj->arX || !pTrj->arXp) && (!pTrj->arY || !pTrj->arYp)) || pTrj->np <= 0);
    bool fldDef = !(pMagFld == nullptr || pMagFld->arMagFld == nullptr || pMagFld->arMagFldTypes == nullptr ||
                    pMagFld->nElem <= 0);
    if (!trjDef && !fldDef) return SRWL_INCORRECT_PARAM_FOR_SR_COMP;
    int meth = (int)precPar[0];
    if (meth != 0 && meth != 1) return SRWL_AMD_ERR_UNSUPPORTED;      // method 2: not in this build
    double sampFact = nPrecPar > 6 ? precPar[6] : 0.;
    if (pWfr->arEx == nullptr || pWfr->arEy == nullptr) return SRWL_INCORRECT_WFR_STRUCT;
    if (pWfr->mesh.nx <= 0 || pWfr->mesh.ny <= 0 || pWfr->mesh.ne <= 0) return SRWL_INCORRECT_WFR_STRUCT;
    std::vector<SrJob> jobs(1);
    std::vector<srwd_field> fields;
    SrJob& J = jobs[0];
    if (trjDef) {
        if (!pTrj->arX || !pTrj->arXp || !pTrj->arY || !pTrj->arYp) return SRWL_AMD_ERR_UNSUPPORTED;
        pWfr->partBeam.partStatMom1 = pTrj->partInitCond;
        J.part = pTrj->partInitCond;
        J.np = pTrj->np;
        J.ctStart = pTrj->ctStart; J.ctEnd = pTrj->ctEnd;
        J.userTrj = pTrj;
        J.field = 0;
    } else {
        int rc = setup_job_from_field(J, pWfr->partBeam.partStatMom1, pMagFld, precPar, fields);
        if (rc) return rc;
    }
    J.Iavg = pWfr->partBeam.Iavg;
    J.mesh = pWfr->mesh;
    J.meth = meth;
    J.step = precPar[1];
    J.relPrec = precPar[1];
    J.sLim0 = precPar[2]; J.sLim1 = precPar[3];
    int termin = (nPrecPar <= 0 || nPrecPar > 5) ? (int)(char)precPar[5] : 1;
    int rc = ensure_device();
    if (rc) return rc;
    std::vector<long long> off(1, 0);
    std::vector<SrResult> res;
    if (sampFact > 0.) {
        // srradstr.cpp:324-375: mesh sizes from the wavefront radius and
        // centre, the wavefront reallocated by the caller when they change
        rc = sr_run(jobs, fields, nullptr, nullptr, off, res, termin, false, true);
        if (rc) return rc;
        const SRWLRadMesh& m = pWfr->mesh;
        double lambda = 1.239842E-06 / m.eStart;
        long nxN = sr_auto_mesh_size(m.xStart, m.xFin, m.nx, res[0].xc, res[0].Robs, lambda, sampFact);
        long nzN = sr_auto_mesh_size(m.yStart, m.yFin, m.ny, res[0].zc, res[0].Robs, lambda, sampFact);
        if (nxN != m.nx || nzN != m.ny) {
            pWfr->mesh.nx = nxN; pWfr->mesh.ny = nzN;
            if (g_wfrModif == nullptr) return SRWL_WFR_EXT_FUNC_NOT_DEFINED;
            if ((*g_wfrModif)(2, pWfr, 0)) return SRWL_WFR_EXT_MODIF_FAILED;
            if (pWfr->arEx == nullptr || pWfr->arEy == nullptr) return SRWL_WFR_EXT_MODIF_FAILED;
            J.mesh = pWfr->mesh;
        }
    }
    long long n = (long long)J.mesh.ne * J.mesh.nx * J.mesh.ny;
    float* dEx = (float*)g_ctx.buf.Ex.get(sizeof(float) * 2 * n);
    float* dEz = (float*)g_ctx.buf.Ez.get(sizeof(float) * 2 * n);
    if (!dEx || !dEz) return SRWL_ERR_MEMORY_ALLOCATION_FAILURE;
    // the field reaches the caller's arrays band by band while the rest is
    // computed; the moments (srradint.cpp:3570-3574: every energy slice, then
    // the mean energy) are queued behind the field before the host copies
    // begin, so that they run while the last bands travel
    struct MomCtx {
        SRWLWfr* w;
        const float2 *ex, *ez;
        MomQueued q;
        bool queued = false;
    } mc{pWfr, (const float2*)dEx, (const float2*)dEz};
    HostOut ho;
    ho.ex = (float*)pWfr->arEx; ho.ez = (float*)pWfr->arEy;
    if ((pWfr->arMomX || pWfr->arMomY) && !g_arena) {
        ho.queueMore = [](void* p, const std::vector<SrResult>& r) -> int {
            MomCtx& m = *(MomCtx*)p;
            SRWLWfr w = *m.w;   // the radius and centre the moments use
            w.Rx = w.Ry = r[0].Robs;
            w.dRx = w.dRy = r[0].RobsErr;
            w.xc = r[0].xc; w.yc = r[0].zc;
            int rc = wfr_moments_queue(w, m.ex, m.ez, 1, w.mesh.ne, (long long)w.mesh.nx * w.mesh.ne, m.q);
            m.queued = rc == 0;
            return rc;
        };
        ho.queueCtx = &mc;
    }
    rc = sr_run(jobs, fields, dEx, dEz, off, res, termin, true, false, g_ctx.keepOnDevice ? nullptr : &ho);
    if (rc) return rc;
    pWfr->Rx = pWfr->Ry = res[0].Robs;
    pWfr->dRx = pWfr->dRy = res[0].RobsErr;
    pWfr->xc = res[0].xc; pWfr->yc = res[0].zc;
    if ((pWfr->arMomX || pWfr->arMomY) && !mc.queued) {
        rc = field_moments(*pWfr, (const float2*)dEx, (const float2*)dEz);
        if (rc) return rc;
    }
    HIPCHK(hipStreamSynchronize(g_ctx.stream));
    if (mc.queued) wfr_moments_collect(mc.q, pWfr->arMomX, pWfr->arMomY);
    sr_complete(res);
    pWfr->avgPhotEn = 0.5 * (pWfr->mesh.eStart + pWfr->mesh.eFin);
    pWfr->presCA = 0; pWfr->presFT = 0;
    return res[0].warn;
}

SRWL_EXPORT int srwlCalcElecFieldSRBatch(SRWLWfr** arWfr, SRWLMagFldC** arMagFld, int nWfr, double* precPar,
                                         int nPrecPar)
{
    if (arWfr == nullptr || arMagFld == nullptr || precPar == nullptr || nWfr <= 0) return SRWL_AMD_ERR_BATCH;
    int meth = (int)precPar[0];
    if (meth != 0 && meth != 1) return SRWL_AMD_ERR_UNSUPPORTED;
    if (nPrecPar > 6 && precPar[6] > 0) return SRWL_AMD_ERR_UNSUPPORTED;
    std::vector<SrJob> jobs(nWfr);
    std::vector<srwd_field> fields;
    std::vector<long long> off(nWfr);
    long long tot = 0;
    for (int i = 0; i < nWfr; i++) {
        SRWLWfr* w = arWfr[i];
        if (w == nullptr || w->arEx == nullptr || w->arEy == nullptr) return SRWL_INCORRECT_WFR_STRUCT;
        int rc = setup_job_from_field(jobs[i], w->partBeam.partStatMom1, arMagFld[i], precPar, fields);
        if (rc) return rc;
        jobs[i].Iavg = w->partBeam.Iavg;
        jobs[i].mesh = w->mesh;
        jobs[i].meth = meth;
        jobs[i].step = precPar[1];
        jobs[i].relPrec = precPar[1];
        jobs[i].sLim0 = precPar[2]; jobs[i].sLim1 = precPar[3];
        off[i] = tot;
        tot += (long long)w->mesh.ne * w->mesh.nx * w->mesh.ny;
    }
    int termin = (nPrecPar <= 0 || nPrecPar > 5) ? (int)(char)precPar[5] : 1;
    int rc = ensure_device();
    if (rc) return rc;
    float* dEx = (float*)g_ctx.buf.Ex.get(sizeof(float) * 2 * tot);
    float* dEz = (float*)g_ctx.buf.Ez.get(sizeof(float) * 2 * tot);
    if (!dEx || !dEz) return SRWL_ERR_MEMORY_ALLOCATION_FAILURE;
    std::vector<SrResult> res;
    rc = sr_run(jobs, fields, dEx, dEz, off, res, termin, true);
    if (rc) return rc;
    int warn = 0;
    std::vector<srwm_mslice> momSl;
    std::vector<std::pair<int, int>> momOwner;   // (request, energy) of each slice
    long long nAdj = 0;
    for (int i = 0; i < nWfr; i++) {
        SRWLWfr* w = arWfr[i];
        long long n = (long long)w->mesh.ne * w->mesh.nx * w->mesh.ny;
        if (!g_ctx.keepOnDevice) {
            HIPCHK(hipMemcpyAsync(w->arEx, dEx + 2 * off[i], sizeof(float) * 2 * n, hipMemcpyDeviceToHost,
                                  g_ctx.stream));
            HIPCHK(hipMemcpyAsync(w->arEy, dEz + 2 * off[i], sizeof(float) * 2 * n, hipMemcpyDeviceToHost,
                                  g_ctx.stream));
        }
        w->Rx = w->Ry = res[i].Robs; w->dRx = w->dRy = res[i].RobsErr;
        w->xc = res[i].xc; w->yc = res[i].zc;
        if ((w->arMomX || w->arMomY) && (w->mesh.nx > 1 || w->mesh.ny > 1)) {
            long long ne = w->mesh.ne, nx = w->mesh.nx;
            for (long ie = 0; ie < ne; ie++) {
                momSl.push_back(wfr_slice(*w, (const float2*)dEx + off[i], (const float2*)dEz + off[i], 1, ne, nx * ne,
                                          ie));
                momOwner.push_back({i, (int)ie});
                if (ne > 1) nAdj++;
            }
        }
        w->avgPhotEn = 0.5 * (w->mesh.eStart + w->mesh.eFin);
        w->presCA = 0; w->presFT = 0;
    }
    if (!momSl.empty()) {
        // all requests' slices in one device pass
        std::vector<double> sums;
        std::vector<int> lims;
        rc = device_moments(momSl, sums, lims, 2 * nAdj > (long long)momSl.size());
        if (rc) return rc;
        for (size_t k = 0; k < momSl.size(); k++) {
            SRWLWfr* w = arWfr[momOwner[k].first];
            int ie = momOwner[k].second;
            if (w->arMomX) moments_from_sums(&sums[24 * k], momSl[k], w->arMomX + 11 * ie);
            if (w->arMomY) moments_from_sums(&sums[24 * k + 11], momSl[k], w->arMomY + 11 * ie);
        }
    }
    HIPCHK(hipStreamSynchronize(g_ctx.stream));
    sr_complete(res);
    for (auto& r : res) if (r.warn) warn = r.warn;
    return warn;
}

// Timing/introspection hooks used by the benchmark (no counterpart in the reference).
SRWL_EXPORT int srwlAmdSetTiming(int on)
{
    int rc = ensure_device();
    if (rc) return rc;
    if (on && !g_ctx.timing) {
        HIPCHK(hipEventCreate(&g_ctx.ev0));
        HIPCHK(hipEventCreate(&g_ctx.ev1));
        HIPCHK(hipEventCreate(&g_ctx.ev2));
        HIPCHK(hipEventCreate(&g_ctx.ev3));
    }
    g_ctx.timing = on != 0;
    return 0;
}

SRWL_EXPORT int srwlAmdGetStats(double* out, int n)
{
    if (out == nullptr) return 0;
    double v[16] = {g_ctx.lastSrMs, (double)g_ctx.lastSrEvals, (double)g_ctx.lastSrBlocks, (double)g_ctx.lastSrSplit,
                    g_ctx.lastFftMs, g_ctx.lastFftBytes, g_ctx.sumSrMs, g_ctx.sumSrEvals, (double)g_ctx.nSrCalls,
                    g_ctx.sumFftMs, g_ctx.sumFftBytes, (double)g_ctx.nFftCalls,
                    g_ctx.lastRsMs, g_ctx.lastRsBytes, g_ctx.lastPropDevMs, (double)g_ctx.lastFp64Points};
    for (int i = 0; i < n && i < 16; i++) out[i] = v[i];
    return 0;
}

SRWL_EXPORT int srwlAmdResetStats(void)
{
    g_ctx.sumSrMs = g_ctx.sumSrEvals = 0; g_ctx.nSrCalls = 0;
    g_ctx.sumFftMs = g_ctx.sumFftBytes = 0; g_ctx.nFftCalls = 0;
    return 0;
}

SRWL_EXPORT int srwlAmdSetKeepOnDevice(int on)
{
    g_ctx.keepOnDevice = on != 0;
    return 0;
}

SRWL_EXPORT int srwlAmdSetSplit(int nSplit)
{
    g_ctx.forceSplit = nSplit;
    return 0;
}

SRWL_EXPORT int srwlAmdSetStreamRows(int on)
{
    g_ctx.streamRows = on != 0;
    return 0;
}

SRWL_EXPORT int srwlAmdLastStreamed()
{
    return g_ctx.lastStreamed ? 1 : 0;
}

SRWL_EXPORT int srwlAmdSetBands(int nBands)
{
    if (nBands < 0 || nBands > SRWL_AMD_MAX_BANDS) return SRWL_AMD_ERR_UNSUPPORTED;
    g_ctx.forceBands = nBands;
    return 0;
}

SRWL_EXPORT int srwlAmdSetTrajSerial(int on)
{
    g_ctx.trajSerial = on != 0;
    return 0;
}

SRWL_EXPORT int srwlAmdSetEnergyTiles(int e)
{
    if (e != 0 && e != 4 && e != 8) return SRWL_AMD_ERR_UNSUPPORTED;
    g_ctx.tileE = e;
    return 0;
}

SRWL_EXPORT int srwlAmdSetCondTol(double tol)
{
    if (!(tol >= 0)) return SRWL_AMD_ERR_UNSUPPORTED;
    g_ctx.condTol = tol;
    return 0;
}

SRWL_EXPORT int srwlAmdSetResizePerPoint(int on)
{
    g_ctx.resizePerPoint = on != 0;
    return 0;
}

}  // extern "C"

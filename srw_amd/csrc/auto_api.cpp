// auto_api.cpp - host side of the adaptive SR integral (SRW method 1,
// cpp/src/core/srradint.cpp:1731-2046 called from ComputeTotalRadDistrDirectOut,
// srradint.cpp:289-394).
//
// Per request: level tables of its trajectory, then the points of its
// computed region in chunks of at most kChunk, in the reference's order.  The
// running maximum of |E|^2 carries from chunk to chunk; it restarts with
// every request, as every srwlCalcElecFieldSR call of the reference does.
#include "srw_api_internal.h"
#include "srw_auto.h"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

namespace srwamd {

#define HIPCHK(x) do { if ((x) != hipSuccess) return SRWL_AMD_ERR_GPU; } while (0)

static const long long kChunk = 1LL << 20;
// fixed-point sweeps per chunk before giving up (each one settles at least
// the first unsettled point; in practice two or three suffice)
static const int kMaxSweeps = 4096;

int sr_auto_run(const std::vector<SrJob>& jobs, const std::vector<srwd_trj_info>& info, const srwd_trj_info* dInfo,
                const double* dX, const double* dY, const double* dIBx, const double* dIBz, long long ldT,
                const srwd_edge* dEdge, int termin, float* dEx, float* dEz, const std::vector<long long>& outOff,
                std::vector<SrResult>& res, const srwd_req* dReq, const double* dCoord, const long long* dLast)
{
    hipStream_t st = g_ctx.stream;
    Buffers& B = g_ctx.buf;
    const long long nTab = srwd_auto_off(SRWD_AUTO_DEV_TAB_LEVEL + 1);
    srwd_auto_tab* dTab = (srwd_auto_tab*)B.AutoTab.get(sizeof(srwd_auto_tab) * nTab);
    int* dFlags = (int*)B.AutoFlags.get(sizeof(int) * 8 + 2 * sizeof(double));
    if (!dTab || !dFlags) return SRWL_ERR_MEMORY_ALLOCATION_FAILURE;
    int* dCnt = dFlags;            // [0] changed, [1] extended
    int* dErr = dFlags + 2;
    int* dWarn = dFlags + 3;
    double* dM = (double*)(dFlags + 4);   // [0] carried maximum, [1] maximum after the chunk
    g_ctx.autoSweeps = 0;
    long long allPts = 0;
    for (size_t j = 0; j < jobs.size(); j++) {
        SrRegion G = sr_region(jobs[j].mesh, info[j]);
        allPts += (long long)G.nzc * G.nxc * jobs[j].mesh.ne;
    }
    g_ctx.autoLevels.assign((size_t)allPts, 0);
    long long lvBase = 0;
    const double normBase = 1. / 137.0360411;
    for (size_t j = 0; j < jobs.size(); j++) {
        const SrJob& R = jobs[j];
        const srwd_trj_info& I = info[j];
        const SRWLRadMesh& m = R.mesh;
        SrRegion G = sr_region(m, I);
        srwd_auto_job J;
        std::memset(&J, 0, sizeof(J));
        J.xStart = m.xStart; J.xStep = G.xStep; J.zStart = m.yStart; J.zStep = G.zStep;
        J.eStart = m.eStart; J.eStep = m.ne > 1 ? (m.eFin - m.eStart) / (m.ne - 1) : 0;
        J.normConst = std::sqrt(normBase * R.Iavg * 1.E+09 / 1.602189246E-19);
        J.relPrec = R.relPrec;
        J.nx = (int)m.nx; J.nz = (int)m.ny; J.nxc = G.nxc; J.nzc = G.nzc; J.ne = (int)m.ne;
        J.symX = G.symX; J.symZ = G.symZ; J.termin = termin;
        J.outOffset = outOff[j];
        J.req = dReq + j;
        J.coord = dCoord;
        J.last = dLast;
        const srwd_trj_info* dI = dInfo + j;
        const double *x = dX + j * ldT, *y = dY + j * ldT, *ibx = dIBx + j * ldT, *ibz = dIBz + j * ldT;
        if (srwd_launch_auto_tables(dI, x, y, ibx, ibz, dTab, st)) return SRWL_AMD_ERR_GPU;
        HIPCHK(hipMemsetAsync(dFlags, 0, sizeof(int) * 8 + 2 * sizeof(double), st));
        long long total = (long long)G.nzc * G.nxc * m.ne;
        for (long long p0 = 0; p0 < total; p0 += kChunk) {
            long long n = std::min(kChunk, total - p0);
            const size_t nl = SRWD_AUTO_MAX_LEVEL + 1;
            double* S = (double*)B.AutoS.get(sizeof(double) * nl * n);
            float4* F = (float4*)B.AutoF.get(sizeof(float4) * nl * n);
            char* w = (char*)B.AutoWork.get((sizeof(unsigned) + 3 * sizeof(int) + 2 * sizeof(double)) * n);
            if (!S || !F || !w) return SRWL_ERR_MEMORY_ALLOCATION_FAILURE;
            double* Mex = (double*)w;
            double* thr = Mex + n;
            unsigned* guard = (unsigned*)(thr + n);
            int* lvl = (int*)(guard + n);
            int* done = lvl + n;
            int* ext = done + n;
            if (srwd_launch_auto_points(J, dI, dEdge + j, dTab, x, y, ibx, ibz, p0, n, nullptr, 0, nullptr, S, F, guard,
                                        lvl, done, dErr, dWarn, st))
                return SRWL_AMD_ERR_GPU;
            int sweeps = 0;
            for (;;) {
                int h[2];
                HIPCHK(hipMemsetAsync(dCnt, 0, 2 * sizeof(int), st));
                if (srwd_launch_auto_scan(n, S, lvl, guard, done, J.relPrec, dM, Mex, dM + 1, thr, ext, dCnt, st))
                    return SRWL_AMD_ERR_GPU;
                HIPCHK(hipMemcpyAsync(h, dCnt, 2 * sizeof(int), hipMemcpyDeviceToHost, st));
                int e = 0;
                HIPCHK(hipMemcpyAsync(&e, dErr, sizeof(int), hipMemcpyDeviceToHost, st));
                HIPCHK(hipStreamSynchronize(st));
                if (e) return SRWL_ERR_CAN_NOT_COMPUTE_RADIATION_INTEGRAL;
                sweeps++;
                if (h[0] == 0 && h[1] == 0) break;
                if (sweeps >= kMaxSweeps) return SRWL_ERR_CAN_NOT_COMPUTE_RADIATION_INTEGRAL;
                if (h[1] > 0 &&
                    srwd_launch_auto_points(J, dI, dEdge + j, dTab, x, y, ibx, ibz, p0, n, ext, h[1], thr, S, F,
                                            guard, lvl, done, dErr, dWarn, st))
                    return SRWL_AMD_ERR_GPU;
            }
            g_ctx.autoSweeps = std::max(g_ctx.autoSweeps, sweeps);
            if (srwd_launch_auto_out(J, p0, n, F, lvl, dEx, dEz, st)) return SRWL_AMD_ERR_GPU;
            HIPCHK(hipMemcpyAsync(dM, dM + 1, sizeof(double), hipMemcpyDeviceToDevice, st));
            HIPCHK(hipMemcpyAsync(g_ctx.autoLevels.data() + lvBase + p0, lvl, sizeof(int) * n, hipMemcpyDeviceToHost,
                                  st));
        }
        lvBase += total;
        int wv = 0;
        HIPCHK(hipMemcpyAsync(&wv, dWarn, sizeof(int), hipMemcpyDeviceToHost, st));
        HIPCHK(hipStreamSynchronize(st));
        if (wv && !res[j].warn) res[j].warn = SRWL_WARN_BAD_RAD_INT_TERMINATIONS;
    }
    return 0;
}

}  // namespace srwamd

extern "C" {

// Final levels of the points of the last method-1 computation, in the
// reference's point order (no counterpart in the reference; for tests).
SRWL_EXPORT long long srwlAmdLastAutoLevels(int* out, long long cap)
{
    long long n = (long long)srwamd::g_ctx.autoLevels.size();
    if (out) std::memcpy(out, srwamd::g_ctx.autoLevels.data(), sizeof(int) * (size_t)std::min(n, cap));
    return n;
}

}  // extern "C"

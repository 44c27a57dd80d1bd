// moments_api.cpp - host side of the wavefront moments and of the state a
// drift updates besides the field (moments, radius estimate, 4x4 matrix).
//
// Reference semantics (paths relative to the reference source root):
//   moments from the sums       cpp/src/core/sroptelm.cpp:1948-2078
//   central moments             cpp/src/core/srstraux.h:348-372
//   radius for arProp[3] = 1    cpp/src/core/sroptdrf.cpp:1129-1370
//   radius for arProp[3] = 2    cpp/src/core/sroptdrf.cpp:1085-1116
//   drift moment transport      cpp/src/core/sroptdrf.cpp:21-110, sroptdrf.h:967-1027
//   4x4 propagation matrix      cpp/src/core/sroptdrf.h:1002-1012, sroptelm.cpp:5594-5640
#include "srw_api_internal.h"
#include "srw_moments.h"

#include <cmath>
#include <cstring>
#include <vector>

namespace srwamd {

// Queues the moment sums of the slices on the library stream: slice table
// in, sums and limits out through the host arrays given (pinned memory lets
// the caller go on before the stream has reached them).
static int queue_sums(const srwm_mslice* sl, int n, int nxMax, int nzMax, bool lanesOverSlices, double* sums,
                      int* lims)
{
    hipStream_t st = g_ctx.stream;
    Buffers& B = g_ctx.buf;
    srwm_mslice* dSl = (srwm_mslice*)B.MomSl.get(sizeof(srwm_mslice) * n);
    void* dWork = B.MomWork.get((size_t)srwm_work_bytes(n, nxMax, nzMax, lanesOverSlices ? 1 : 0));
    double* dSums = (double*)B.MomSums.get(sizeof(double) * 24 * (size_t)n);
    int* dLims = (int*)B.MomLims.get(sizeof(int) * 4 * (size_t)n);
    if (!dSl || !dWork || !dSums || !dLims) return SRWL_ERR_MEMORY_ALLOCATION_FAILURE;
    if (hipMemcpyAsync(dSl, sl, sizeof(srwm_mslice) * n, hipMemcpyHostToDevice, st) != hipSuccess)
        return SRWL_AMD_ERR_GPU;
    if (srwm_launch_moments(dSl, n, nxMax, nzMax, lanesOverSlices ? 1 : 0, dWork, dSums, dLims, st))
        return SRWL_AMD_ERR_GPU;
    if (hipMemcpyAsync(sums, dSums, sizeof(double) * 24 * n, hipMemcpyDeviceToHost, st) != hipSuccess)
        return SRWL_AMD_ERR_GPU;
    if (lims && hipMemcpyAsync(lims, dLims, sizeof(int) * 4 * n, hipMemcpyDeviceToHost, st) != hipSuccess)
        return SRWL_AMD_ERR_GPU;
    return 0;
}

int device_moments(const std::vector<srwm_mslice>& sl, std::vector<double>& sums, std::vector<int>& lims,
                   bool lanesOverSlices)
{
    int n = (int)sl.size();
    sums.assign((size_t)n * 24, 0.);
    lims.assign((size_t)n * 4, 0);
    if (n == 0) return 0;
    int rc = ensure_device();
    if (rc) return rc;
    int nxMax = 1, nzMax = 1;
    for (auto& s : sl) { nxMax = std::max(nxMax, s.nx); nzMax = std::max(nzMax, s.nz); }
    rc = queue_sums(sl.data(), n, nxMax, nzMax, lanesOverSlices, sums.data(), lims.data());
    if (rc) return rc;
    if (hipStreamSynchronize(g_ctx.stream) != hipSuccess) return SRWL_AMD_ERR_GPU;
    return 0;
}

int wfr_moments_queue(const SRWLWfr& w, const float2* dEx, const float2* dEz, long long perE, long long perX,
                      long long perZ, MomQueued& q)
{
    q.n = 0;
    if (w.mesh.nx <= 1 && w.mesh.ny <= 1) return 0;
    int rc = ensure_device();
    if (rc) return rc;
    int n = (int)w.mesh.ne;
    // pinned: slice table, then 24 sums per slice
    size_t slBytes = ((sizeof(srwm_mslice) * n + 255) / 256) * 256;
    size_t bytes = slBytes + sizeof(double) * 24 * n;
    if (bytes > g_ctx.momPinBytes) {
        if (g_ctx.momPin) (void)hipHostFree(g_ctx.momPin);
        g_ctx.momPin = nullptr;
        g_ctx.momPinBytes = 0;
        if (hipHostMalloc((void**)&g_ctx.momPin, bytes, hipHostMallocDefault) != hipSuccess) {
            g_ctx.momPin = nullptr;
            return SRWL_ERR_MEMORY_ALLOCATION_FAILURE;
        }
        g_ctx.momPinBytes = bytes;
    }
    srwm_mslice* sl = (srwm_mslice*)g_ctx.momPin;
    q.sums = (double*)(g_ctx.momPin + slBytes);
    int nxMax = 1, nzMax = 1;
    for (long ie = 0; ie < n; ie++) {
        sl[ie] = wfr_slice(w, dEx, dEz, perE, perX, perZ, ie);
        nxMax = std::max(nxMax, sl[ie].nx);
        nzMax = std::max(nzMax, sl[ie].nz);
    }
    q.sl.assign(sl, sl + n);
    rc = queue_sums(sl, n, nxMax, nzMax, perE == 1 && n > 1, q.sums, nullptr);
    if (rc) return rc;
    q.n = n;
    q.hasX = dEx != nullptr;
    q.hasZ = dEz != nullptr;
    return 0;
}

void wfr_moments_collect(const MomQueued& q, double* momX, double* momZ)
{
    for (int i = 0; i < q.n; i++) {
        if (momX && q.hasX) moments_from_sums(&q.sums[24 * i], q.sl[i], momX + 11 * i);
        if (momZ && q.hasZ) moments_from_sums(&q.sums[24 * i + 11], q.sl[i], momZ + 11 * i);
    }
}

// 11 moments of one component from its 11 sums; values pass through float as
// in the reference (sroptelm.cpp:1948-2078, coordinate representation).
void moments_from_sums(const double* S, const srwm_mslice& s, double* mom)
{
    const double TwoPi = 3.141592653590 * 2., FourPi = TwoPi * 2.;
    double lam2 = 2. * (1.239842E-06 / (FourPi * s.ePh)), lam2e2 = lam2 * lam2;
    bool scX = s.nx > 1, scZ = s.nz > 1;
    if (S[0] == 0.) { for (int k = 0; k < 11; k++) mom[k] = 0.; return; }
    double inv = 1. / S[0];
    double iX = 1. / s.xStep, iZ = 1. / s.zStep;
    mom[0] = (float)((scX && scZ) ? S[0] * s.xStep * s.zStep * 1.E+06 : 0.);
    mom[1] = (float)(scX ? S[1] * inv : s.xStart);
    mom[2] = (float)(scX ? S[2] * (inv * lam2 * iX) : 0.);
    mom[3] = (float)(scZ ? S[3] * inv : s.zStart);
    mom[4] = (float)(scZ ? S[4] * (inv * lam2 * iZ) : 0.);
    mom[5] = (float)(scX ? S[5] * inv : s.xStart * s.xStart);
    mom[6] = (float)(scX ? S[6] * (inv * lam2 * iX) : 0.);
    mom[7] = (float)(scX ? S[7] * (inv * lam2e2 * (iX * iX)) : 0.);
    mom[8] = (float)(scZ ? S[8] * inv : s.zStart * s.zStart);
    mom[9] = (float)(scZ ? S[9] * (inv * lam2 * iZ) : 0.);
    mom[10] = (float)(scZ ? S[10] * (inv * lam2e2 * (iZ * iZ)) : 0.);
}

// Slice descriptor of energy ie of a wavefront whose component element
// (ix, iz, ie) is at base[ie*perE + iz*perZ + ix*perX] (float2 units).
srwm_mslice wfr_slice(const SRWLWfr& w, const float2* dEx, const float2* dEz, long long perE, long long perX,
                      long long perZ, long ie)
{
    srwm_mslice s;
    std::memset(&s, 0, sizeof(s));
    long nx = w.mesh.nx, nz = w.mesh.ny, ne = w.mesh.ne;
    s.ex = dEx ? dEx + ie * perE : nullptr;
    s.ez = dEz ? dEz + ie * perE : nullptr;
    s.perX = perX; s.perZ = perZ;
    s.nx = (int)nx; s.nz = (int)nz;
    s.xStart = w.mesh.xStart; s.zStart = w.mesh.yStart;
    s.xStep = nx > 1 ? (w.mesh.xFin - w.mesh.xStart) / (nx - 1) : 0.;
    s.zStep = nz > 1 ? (w.mesh.yFin - w.mesh.yStart) / (nz - 1) : 0.;
    double eStep = ne > 1 ? (w.mesh.eFin - w.mesh.eStart) / (ne - 1) : 0.;
    s.ePh = w.mesh.eStart + eStep * ie;
    s.Rx = w.Rx; s.Rz = w.Ry; s.xc = w.xc; s.zc = w.yc;
    // sroptelm.cpp:4692-4770: the radius is trusted when its error is below 40 %
    s.fx = std::fabs(w.dRx) < 0.4 * std::fabs(w.Rx);
    s.fz = std::fabs(w.dRy) < 0.4 * std::fabs(w.Ry);
    return s;
}

// Moments of every photon energy of a device-resident wavefront into
// momX / momZ (11 per energy; either may be null).
int wfr_moments(const SRWLWfr& w, const float2* dEx, const float2* dEz, long long perE, long long perX,
                long long perZ, double* momX, double* momZ)
{
    if (w.mesh.nx <= 1 && w.mesh.ny <= 1) return 0;
    std::vector<srwm_mslice> sl;
    for (long ie = 0; ie < w.mesh.ne; ie++) sl.push_back(wfr_slice(w, dEx, dEz, perE, perX, perZ, ie));
    std::vector<double> sums;
    std::vector<int> lims;
    // energy-fastest order: the slices' data are adjacent, one lane per slice
    int rc = device_moments(sl, sums, lims, perE == 1 && w.mesh.ne > 1);
    if (rc) return rc;
    for (size_t i = 0; i < sl.size(); i++) {
        if (momX && dEx) moments_from_sums(&sums[24 * i], sl[i], momX + 11 * i);
        if (momZ && dEz) moments_from_sums(&sums[24 * i + 11], sl[i], momZ + 11 * i);
    }
    return 0;
}

// srstraux.h:348-372; c = {Mxx, Mxpxp, Mzz, Mzpzp, sqrt|Mxpxp|, sqrt|Mzpzp|}
static bool central_moments(const double* m, double* c)
{
    double Mxx = m[5] - m[1] * m[1], Mxpxp = m[7] - m[2] * m[2];
    double Mzz = m[8] - m[3] * m[3], Mzpzp = m[10] - m[4] * m[4];
    bool ok = !(Mxx < 0 || Mxpxp < 0 || Mzz < 0 || Mzpzp < 0);
    if (Mxx < 0.1 * m[5]) Mxx = 0.1 * m[5];
    if (Mxpxp < 0.1 * m[7]) Mxpxp = 0.1 * m[7];
    if (Mzz < 0.1 * m[8]) Mzz = 0.1 * m[8];
    if (Mzpzp < 0.1 * m[10]) Mzpzp = 0.1 * m[10];
    c[0] = Mxx; c[1] = Mxpxp; c[2] = Mzz; c[3] = Mzpzp;
    c[4] = std::sqrt(std::fabs(Mxpxp)); c[5] = std::sqrt(std::fabs(Mzpzp));
    return ok;
}

static double sub1_axis(double s1, double L, double smin)
{
    double Rmin = 2. * smin, s2 = s1 + L, eff;
    if (std::fabs(s1) < smin) {
        if (std::fabs(s2) < smin) {
            if (std::fabs(s1) < std::fabs(s2)) eff = (L > 0) ? Rmin : -Rmin;
            else eff = (L > 0) ? -Rmin - L : Rmin - L;
        } else eff = (L > 0) ? Rmin : -Rmin;
    } else {
        eff = s1 + smin * smin / s1;
        double minR = 0.007 * std::fabs(eff);
        if (Rmin < minR) Rmin = minR;
        double effL = eff + L;
        if (std::fabs(effL) < Rmin) eff = (effL < 0) ? -Rmin - L : Rmin - L;
    }
    return eff;
}

// Central photon energy index (sroptdrf.cpp:1139-1151).
static long center_energy(const SRWLWfr& w, double& photEn0)
{
    long ne = w.mesh.ne;
    double eStep = ne > 1 ? (w.mesh.eFin - w.mesh.eStart) / (ne - 1) : 0.;
    photEn0 = w.mesh.eStart;
    long ieCen = 0;
    if (ne > 1 && eStep > 0) {
        photEn0 = w.avgPhotEn;
        double d = (photEn0 - w.mesh.eStart) / eStep;
        ieCen = (long)d;
        if ((d - ieCen) > 0.5) ieCen++;
        if (ieCen < 0) ieCen = 0; else if (ieCen >= ne) ieCen = ne - 1;
    }
    return ieCen;
}

// Radius used by the analytic quadratic-phase drift.  sub = 2: the radius
// kept 3 errors away from 0 and -L; sub = 1: from the angular spread of the
// moments at the central energy.  recompute() refreshes momX/momZ from the
// current field when they are empty or unusable.
int drift_radius(const SRWLWfr& w, int sub, double L, double* momX, double* momZ, int (*recompute)(void*),
                 void* ctx, double& Rx, double& Rz)
{
    if (sub == 2) {
        Rx = w.Rx; Rz = w.Ry;
        double mx = 3. * w.dRx, mz = 3. * w.dRy;
        double RxL = Rx + L;
        if (std::fabs(Rx) < mx) Rx = mx * (Rx >= 0 ? 1. : -1.);
        else if (std::fabs(RxL) < mx) Rx = mx * (RxL >= 0 ? 1. : -1.) - L;
        double RzL = Rz + L;
        if (std::fabs(Rz) < mz) Rz = mz * (Rz >= 0 ? 1. : -1.);
        else if (std::fabs(RzL) < mz) Rz = mz * (RzL >= 0 ? 1. : -1.) - L;
        return 0;
    }
    const double infLarge = 1E+23, coefAngRange = 0.2, fourPi = 4. * 3.14159265359;
    double photEn0;
    long ieCen = center_energy(w, photEn0);
    double* mx = momX + 11 * ieCen;
    double* mz = momZ + 11 * ieCen;
    if (mx[0] == 0 && mz[0] == 0) { int rc = recompute(ctx); if (rc) return rc; }
    double* use = (mx[0] >= mz[0]) ? mx : mz;
    double c[6];
    bool ok = central_moments(use, c);
    if (!ok || c[4] == 0 || c[5] == 0) {
        int rc = recompute(ctx);
        if (rc) return rc;
        central_moments(use, c);
    }
    if (c[4] == 0 || c[5] == 0) {
        Rx = (w.Rx == 0) ? infLarge : w.Rx;
        Rz = (w.Ry == 0) ? infLarge : w.Ry;
        return 0;
    }
    double lam = 1.239842E-06 / photEn0;
    double ax = coefAngRange * c[4], az = coefAngRange * c[5];
    Rx = sub1_axis(w.Rx, L, lam / (fourPi * ax * ax));
    Rz = sub1_axis(w.Ry, L, lam / (fourPi * az * az));
    return 0;
}

// Moment transport through a drift of length L on the (already updated)
// output mesh, with the diffraction-limited spot floor.
void drift_moments(const SRWLWfr& w, double L, double* mom)
{
    long nx = w.mesh.nx, nz = w.mesh.ny, ne = w.mesh.ne;
    double xStep = nx > 1 ? (w.mesh.xFin - w.mesh.xStart) / (nx - 1) : 0.;
    double zStep = nz > 1 ? (w.mesh.yFin - w.mesh.yStart) / (nz - 1) : 0.;
    double eStep = ne > 1 ? (w.mesh.eFin - w.mesh.eStart) / (ne - 1) : 0.;
    for (long ie = 0; ie < ne; ie++) {
        double* M = mom + 11 * ie;
        if (M[0] == 0.) continue;
        double X = M[1] + L * M[2], Z = M[3] + L * M[4];
        double XX = M[5] + 2. * L * M[6] + L * L * M[7], XXP = M[6] + L * M[7];
        double ZZ = M[8] + 2. * L * M[9] + L * L * M[10], ZZP = M[9] + L * M[10];
        M[1] = X; M[3] = Z; M[5] = XX; M[6] = XXP; M[8] = ZZ; M[9] = ZZP;
        double lam = 1.239842e-06 / (w.mesh.eStart + ie * eStep);
        double sx = lam * L / ((nx - 1) * xStep), sz = lam * L / ((nz - 1) * zStep);
        if (M[5] - M[1] * M[1] < sx * sx) M[5] = sx * sx + M[1] * M[1];
        if (M[8] - M[3] * M[3] < sz * sz) M[8] = sz * sz + M[3] * M[3];
    }
}

// M <- D M, V <- D V for the drift matrix D (arElecPropMatr: M row-major, then V).
void drift_4x4(double* P, double L)
{
    const double D[16] = {1., L, 0., 0., 0., 1., 0., 0., 0., 0., 1., L, 0., 0., 0., 1.};
    double R[16], V[4];
    for (int i = 0; i < 4; i++) {
        V[i] = 0.;
        for (int j = 0; j < 4; j++) {
            double r = 0.;
            for (int k = 0; k < 4; k++) r += D[4 * i + k] * P[4 * k + j];
            R[4 * i + j] = r;
            V[i] += D[4 * i + j] * P[16 + j];
        }
    }
    std::memcpy(P, R, sizeof(R));
    for (int i = 0; i < 4; i++) P[16 + i] = V[i];
}

}  // namespace srwamd

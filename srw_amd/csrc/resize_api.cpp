// resize_api.cpp - host side of the wavefront resampling (srwlResizeElecField
// and the resize instructions of a propagation container).
//
// Reference semantics (paths relative to the reference source root):
//   C entry point          cpp/src/lib/srwlib.cpp:860-900
//   new mesh               cpp/src/core/sroptelm.cpp:2242-2500 (centre 0.5, no FFT-side resizing)
//   good FFT sizes         cpp/src/ext/genmath/gmfft.cpp:26-225
//   array reallocation     cpp/src/core/srradstr.cpp:760-790 (wavefront-modification callback)
#include "srw_api_internal.h"
#include "srw_resize.h"

#include <cmath>
#include <cstring>

namespace srwamd {

// Smallest size >= n of the reference's table of good FFT sizes: the even
// numbers 2^a 3^b 5^c 7^d 11^e 13^f with e, f <= 1 (the table, up to 100000).
long next_fft_size(long n)
{
    if (n < 4) return 4;
    for (long m = n;; m++) {
        if (m & 1) continue;
        long r = m;
        int e11 = 0, e13 = 0;
        while (r % 2 == 0) r /= 2;
        while (r % 3 == 0) r /= 3;
        while (r % 5 == 0) r /= 5;
        while (r % 7 == 0) r /= 7;
        while (r % 11 == 0) { r /= 11; e11++; }
        while (r % 13 == 0) { r /= 13; e13++; }
        if (r == 1 && e11 <= 1 && e13 <= 1) return m;
    }
}

// One axis of the new mesh: size, start, step and the filled index range.
void resize_axis(long nOld, double startOld, double stepOld, double pm, double pd, ResizeAxis& a)
{
    double pTot = pm * pd;
    long nNew = nOld;
    if (pTot != 1.) {
        double d = pTot * nNew;
        long l = (long)d;
        nNew = ((d - l) >= 0.5) ? l + 1 : l;
        nNew = next_fft_size(nNew);
    }
    double magOld = (nOld - 1) * stepOld;
    if (pm > 1.) {
        long inner = (pd != 1.) ? (long)(nNew / pm) : nOld;
        long outer = nNew - inner;
        long left = outer >> 1;
        if ((left << 1) != outer) left++;
        a.step = magOld / (inner - 1);
        a.start = startOld - left * a.step;
        a.i0 = left; a.i1 = left + inner - 1;
    } else {
        double endOld = startOld + (nOld - 1) * stepOld;
        double mid = 0.5 * (startOld + endOld);
        double magNew = magOld * pm;
        a.step = magNew / (nNew - 1);
        if (pd == 1) { a.step = stepOld; magNew = (nNew - 1) * a.step; }
        a.start = mid - 0.5 * magNew;
        a.i0 = 0; a.i1 = nNew - 1;
    }
    a.n = nNew;
}

// Validates par (= [fft side, pxm, pxd, pzm, pzd, rel. centre x, rel. centre z])
// and fills the new mesh; returns 0, an error, or -1 when nothing changes.
int resize_plan(const SRWLWfr& w, const double* par, int npar, ResizePlan& P)
{
    if (npar < 5) return SRWL_INCORRECT_PARAM_FOR_RESIZE;
    P.pxm = par[1]; P.pxd = par[2]; P.pzm = par[3]; P.pzd = par[4];
    if (P.pxm == 1. && P.pxd == 1. && P.pzm == 1. && P.pzd == 1.) return -1;   // sroptelm.cpp:2244
    if (par[0] != 0) return SRWL_AMD_ERR_UNSUPPORTED;                         // resizing on the FFT side
    if (npar > 5 && std::fabs(par[5] - 0.5) > 1.e-06) return SRWL_AMD_ERR_UNSUPPORTED;   // off-centre resizing
    if (npar > 6 && std::fabs(par[6] - 0.5) > 1.e-06) return SRWL_AMD_ERR_UNSUPPORTED;
    if (!(P.pxm > 0) || !(P.pxd > 0) || !(P.pzm > 0) || !(P.pzd > 0)) return SRWL_INCORRECT_PARAM_FOR_RESIZE;
    long nx = w.mesh.nx, nz = w.mesh.ny;
    if (nx < 2 || nz < 2) return SRWL_INCORRECT_PARAM_FOR_RESIZE;
    double xStep = (w.mesh.xFin - w.mesh.xStart) / (nx - 1), zStep = (w.mesh.yFin - w.mesh.yStart) / (nz - 1);
    resize_axis(nx, w.mesh.xStart, xStep, P.pxm, P.pxd, P.x);
    resize_axis(nz, w.mesh.yStart, zStep, P.pzm, P.pzd, P.z);
    // sroptelm.cpp:3040-3046: pure range enlargement at the same step copies points
    P.onlyLarger = (P.pxd == 1. && P.pzd == 1. && P.pxm >= 1. && P.pzm >= 1.) &&
                   (std::fabs(xStep - P.x.step) < 1.e-05 && std::fabs(zStep - P.z.step) < 1.e-05);
    if (!P.onlyLarger && (nx < 4 || nz < 4)) return SRWL_INCORRECT_PARAM_FOR_RESIZE;
    // sroptelm.cpp:3051-3060 (WaveFrontTermCanBeTreated): quadratic phase removed around the interpolation
    P.fx = std::fabs(w.dRx) < 0.4 * std::fabs(w.Rx);
    P.fz = std::fabs(w.dRy) < 0.4 * std::fabs(w.Ry);
    if (P.onlyLarger) P.fx = P.fz = 0;
    P.improve = 1;
    P.zeroOutside = 0;
    return 0;
}

static srwr_phase phase_desc(const SRWLWfr& w, long nx, long nz, double xStart, double xStep, double zStart,
                             double zStep, const srwr_layout& l, int fx, int fz, int sign)
{
    srwr_phase p;
    std::memset(&p, 0, sizeof(p));
    p.l = l;
    p.nx = (int)nx; p.nz = (int)nz; p.ne = (int)w.mesh.ne;
    p.xStart = xStart; p.xStep = xStep; p.zStart = zStart; p.zStep = zStep;
    p.eStart = w.mesh.eStart;
    p.eStep = w.mesh.ne > 1 ? (w.mesh.eFin - w.mesh.eStart) / (w.mesh.ne - 1) : 0.;
    p.Rx = w.Rx; p.Rz = w.Ry; p.xc = w.xc; p.zc = w.yc;
    p.fx = fx; p.fz = fz; p.sign = sign;
    return p;
}

// Resamples a device wavefront (old arrays modified in place by the phase
// treatment) into the new arrays.  w holds the old mesh on entry.
int device_resize(const SRWLWfr& w, const ResizePlan& P, float2* oX, float2* oZ, const srwr_layout& lo, float2* nX,
                  float2* nZ, const srwr_layout& ln, hipStream_t st)
{
    long nx = w.mesh.nx, nz = w.mesh.ny;
    double xStep = (w.mesh.xFin - w.mesh.xStart) / (nx - 1), zStep = (w.mesh.yFin - w.mesh.yStart) / (nz - 1);
    if (P.fx || P.fz) {
        srwr_phase ph = phase_desc(w, nx, nz, w.mesh.xStart, xStep, w.mesh.yStart, zStep, lo, P.fx, P.fz, -1);
        if (srwr_launch_phase(oX, oZ, ph, st)) return SRWL_AMD_ERR_GPU;
    }
    srwr_params R;
    std::memset(&R, 0, sizeof(R));
    R.lo = lo; R.ln = ln;
    R.nxo = (int)nx; R.nzo = (int)nz; R.nxn = (int)P.x.n; R.nzn = (int)P.z.n; R.ne = (int)w.mesh.ne;
    R.xStartO = w.mesh.xStart; R.xStepO = xStep; R.zStartO = w.mesh.yStart; R.zStepO = zStep;
    R.xStartN = P.x.start; R.xStepN = P.x.step; R.zStartN = P.z.start; R.zStepN = P.z.step;
    R.eStart = w.mesh.eStart;
    R.eStep = w.mesh.ne > 1 ? (w.mesh.eFin - w.mesh.eStart) / (w.mesh.ne - 1) : 0.;
    R.ix0 = (int)P.x.i0; R.ix1 = (int)P.x.i1; R.iz0 = (int)P.z.i0; R.iz1 = (int)P.z.i1;
    R.onlyLarger = P.onlyLarger;
    R.improve = P.improve;
    R.zeroOutside = P.zeroOutside;
    R.xMin = w.mesh.xStart; R.xMax = w.mesh.xFin; R.zMin = w.mesh.yStart; R.zMax = w.mesh.yFin;
    R.xTol = 1.e-06 * P.x.step; R.zTol = 1.e-06 * P.z.step;
    // the phase is added back to each new value before it is stored
    R.hasPost = (P.fx || P.fz) ? 1 : 0;
    if (R.hasPost) R.post = phase_desc(w, P.x.n, P.z.n, P.x.start, P.x.step, P.z.start, P.z.step, ln, P.fx, P.fz, 1);
    R.cellMode = g_ctx.resizePerPoint ? 0 : 1;
    if (srwr_launch_resize(oX, oZ, nX, nZ, R, st)) return SRWL_AMD_ERR_GPU;
    return 0;
}

// Writes the new mesh into the caller's structure and has the caller
// reallocate its field arrays through the registered callback.
int realloc_wfr_arrays(SRWLWfr& w, const ResizePlan& P)
{
    w.mesh.nx = P.x.n; w.mesh.ny = P.z.n;
    w.mesh.xStart = P.x.start; w.mesh.xFin = P.x.start + (P.x.n - 1) * P.x.step;
    w.mesh.yStart = P.z.start; w.mesh.yFin = P.z.start + (P.z.n - 1) * P.z.step;
    if (g_wfrModif == nullptr) return SRWL_WFR_EXT_FUNC_NOT_DEFINED;
    if ((*g_wfrModif)(2, &w, 0)) return SRWL_WFR_EXT_MODIF_FAILED;
    if (w.arEx == nullptr || w.arEy == nullptr) return SRWL_WFR_EXT_MODIF_FAILED;
    return 0;
}

}  // namespace srwamd

using namespace srwamd;

#define HIPCHK(x) do { if ((x) != hipSuccess) return SRWL_AMD_ERR_GPU; } while (0)

extern "C" {

// reference srwlib.h:775 (srwlib.cpp:860-900), coordinate resizing
SRWL_EXPORT int srwlResizeElecField(SRWLWfr* pWfr, char type, double* par)
{
    if (pWfr == nullptr || par == nullptr) return SRWL_INCORRECT_PARAM_FOR_RESIZE;
    bool isCorA = (type == 'c') || (type == 'C') || (type == 'a') || (type == 'A');
    bool isForT = (type == 'f') || (type == 'F') || (type == 't') || (type == 'T');
    if (!(isCorA || isForT)) return SRWL_INCORRECT_PARAM_FOR_RESIZE;
    if (isForT) return SRWL_AMD_ERR_UNSUPPORTED;            // resizing vs photon energy
    if (pWfr->presCA != 0) return SRWL_AMD_ERR_UNSUPPORTED;  // angular representation
    if (pWfr->arEx == nullptr || pWfr->arEy == nullptr) return SRWL_INCORRECT_WFR_STRUCT;
    ResizePlan P;
    int rc = resize_plan(*pWfr, par, 7, P);
    if (rc == -1) return 0;
    if (rc) return rc;
    rc = ensure_device();
    if (rc) return rc;
    hipStream_t st = g_ctx.stream;
    long long ne = pWfr->mesh.ne;
    long long nOld = ne * pWfr->mesh.nx * pWfr->mesh.ny, nNew = ne * P.x.n * P.z.n;
    float2* dO = (float2*)g_ctx.buf.RsO.get(sizeof(float2) * 2 * nOld);
    float2* dN = (float2*)g_ctx.buf.RsN.get(sizeof(float2) * 2 * nNew);
    if (!dO || !dN) return SRWL_ERR_MEMORY_ALLOCATION_FAILURE;
    HIPCHK(hipMemcpyAsync(dO, pWfr->arEx, sizeof(float2) * nOld, hipMemcpyHostToDevice, st));
    HIPCHK(hipMemcpyAsync(dO + nOld, pWfr->arEy, sizeof(float2) * nOld, hipMemcpyHostToDevice, st));
    srwr_layout lo = {ne, ne * pWfr->mesh.nx, 1};
    srwr_layout ln = {ne, ne * P.x.n, 1};
    rc = device_resize(*pWfr, P, dO, dO + nOld, lo, dN, dN + nNew, ln, st);
    if (rc) return rc;
    HIPCHK(hipStreamSynchronize(st));
    rc = realloc_wfr_arrays(*pWfr, P);
    if (rc) return rc;
    HIPCHK(hipMemcpyAsync(pWfr->arEx, dN, sizeof(float2) * nNew, hipMemcpyDeviceToHost, st));
    HIPCHK(hipMemcpyAsync(pWfr->arEy, dN + nNew, sizeof(float2) * nNew, hipMemcpyDeviceToHost, st));
    HIPCHK(hipStreamSynchronize(st));
    return 0;
}

// reference srwlib.h:787 (srwlib.cpp:914-931, srradstr.cpp:4218-4502 and
// ResizeCoreXZ :4504-4904): resampling onto the given mesh.  par = [FFT
// side, allow the quadratic-phase treatment, allow the Re/Im correction from
// the intensity], as the reference's Python binding fills it ({0, 1, 1}).
SRWL_EXPORT int srwlResizeElecFieldMesh(SRWLWfr* pWfr, SRWLRadMesh* pMesh, double* par)
{
    if (pWfr == nullptr || pMesh == nullptr || par == nullptr) return SRWL_INCORRECT_PARAM_FOR_RESIZE;
    if (pWfr->presCA != 0 || pWfr->presFT != 0) return SRWL_AMD_ERR_UNSUPPORTED;
    const SRWLRadMesh& mo = pWfr->mesh;
    const SRWLRadMesh& mn = *pMesh;
    long ne = mo.ne, nx = mo.nx, nz = mo.ny;
    double eStep = ne > 1 ? (mo.eFin - mo.eStart) / (ne - 1) : 0.;
    double xStep = nx > 1 ? (mo.xFin - mo.xStart) / (nx - 1) : 0.;
    double zStep = nz > 1 ? (mo.yFin - mo.yStart) / (nz - 1) : 0.;
    double eStepN = mn.ne <= 1 ? 0. : (mn.eFin - mn.eStart) / (mn.ne - 1);
    double xStepN = mn.nx <= 1 ? 0. : (mn.xFin - mn.xStart) / (mn.nx - 1);
    double zStepN = mn.ny <= 1 ? 0. : (mn.yFin - mn.yStart) / (mn.ny - 1);
    const double relTol = 1.e-06;
    bool reqE = !(((mn.ne == ne) && std::fabs(mn.eStart - mo.eStart) <= relTol * eStepN &&
                   std::fabs(eStepN - eStep) <= relTol * eStepN) || ne <= 1);
    bool reqX = !(((mn.nx == nx) && std::fabs(mn.xStart - mo.xStart) <= relTol * xStepN &&
                   std::fabs(xStepN - xStep) <= relTol * xStepN) || nx <= 1);
    bool reqZ = !(((mn.ny == nz) && std::fabs(mn.yStart - mo.yStart) <= relTol * zStepN &&
                   std::fabs(zStepN - zStep) <= relTol * zStepN) || nz <= 1);
    if (!(reqE || reqX || reqZ)) return 0;
    if (reqE && (reqX || reqZ)) return SRWL_RESIZE_NOT_SUPPORTED_WITH_SUCH_PARAMS;
    if (reqE) return SRWL_AMD_ERR_UNSUPPORTED;      // resampling vs photon energy
    if (par[0] != 0.) return SRWL_AMD_ERR_UNSUPPORTED;   // resampling on the FFT side
    if (pWfr->arEx == nullptr || pWfr->arEy == nullptr) return SRWL_INCORRECT_WFR_STRUCT;
    if (nx < 4 || nz < 4) return SRWL_INCORRECT_PARAM_FOR_RESIZE;   // a 4x4 cell per new point
    if (mn.nx < 1 || mn.ny < 1) return SRWL_INCORRECT_PARAM_FOR_RESIZE;
    ResizePlan P;
    P.x.n = mn.nx; P.x.start = mn.xStart; P.x.step = xStepN;
    P.z.n = mn.ny; P.z.start = mn.yStart; P.z.step = zStepN;
    // filled index range: the new points inside the old mesh (srradstr.cpp:4581-4608)
    double xFinO = mo.xStart + xStep * (nx - 1), zFinO = mo.yStart + zStep * (nz - 1);
    long ix0 = 0, ix1 = mn.nx - 1, iz0 = 0, iz1 = mn.ny - 1;
    if (mn.xStart < mo.xStart) {
        ix0 = (long)((mo.xStart - mn.xStart) / xStepN - 1.e-13);
        if (mn.xStart + (ix0 + 0.1) * xStepN < mo.xStart) ix0++;
        if (ix0 < 0) ix0 = 0;
    }
    if (xFinO < mn.xFin) {
        ix1 = (long)((xFinO - mn.xStart) / xStepN + 1.e-13);
        if (ix1 >= mn.nx) ix1 = mn.nx - 1;
    }
    if (mn.yStart < mo.yStart) {
        iz0 = (long)((mo.yStart - mn.yStart) / zStepN - 1.e-13);
        if (mn.yStart + (iz0 + 0.1) * zStepN < mo.yStart) iz0++;
        if (iz0 < 0) iz0 = 0;
    }
    if (zFinO < mn.yFin) {
        iz1 = (long)((zFinO - mn.yStart) / zStepN + 1.e-13);
        if (iz1 >= mn.ny) iz1 = mn.ny - 1;
    }
    // meshes that do not intersect give a zero field (srradstr.cpp:4382-4425, 4610-4631)
    auto outside = [](double a0, double a1, double b0, double b1) {
        return (a0 < b0 && a1 < b0 && a0 < b1 && a1 < b1) || (a0 > b0 && a1 > b0 && a0 > b1 && a1 > b1);
    };
    bool disjoint = outside(mn.xStart, mn.xFin, mo.xStart, xFinO) || outside(mn.yStart, mn.yFin, mo.yStart, zFinO);
    if (disjoint) { ix0 = 1; ix1 = 0; iz0 = 1; iz1 = 0; }   // empty range: all zero
    P.x.i0 = ix0; P.x.i1 = ix1; P.z.i0 = iz0; P.z.i1 = iz1;
    P.onlyLarger = 0;
    P.improve = par[2] != 0. ? 1 : 0;
    P.zeroOutside = 1;
    // srradstr.h:786-811 (coordinate representation): radius error below 20 %
    bool allowQuad = par[1] != 0.;
    P.fx = allowQuad && std::fabs(pWfr->dRx) < 0.2 * std::fabs(pWfr->Rx);
    P.fz = allowQuad && std::fabs(pWfr->dRy) < 0.2 * std::fabs(pWfr->Ry);
    int rc = ensure_device();
    if (rc) return rc;
    hipStream_t st = g_ctx.stream;
    long long nOld = (long long)ne * nx * nz, nNew = (long long)ne * mn.nx * mn.ny;
    float2* dO = (float2*)g_ctx.buf.RsO.get(sizeof(float2) * 2 * nOld);
    float2* dN = (float2*)g_ctx.buf.RsN.get(sizeof(float2) * 2 * nNew);
    if (!dO || !dN) return SRWL_ERR_MEMORY_ALLOCATION_FAILURE;
    HIPCHK(hipMemcpyAsync(dO, pWfr->arEx, sizeof(float2) * nOld, hipMemcpyHostToDevice, st));
    HIPCHK(hipMemcpyAsync(dO + nOld, pWfr->arEy, sizeof(float2) * nOld, hipMemcpyHostToDevice, st));
    srwr_layout lo = {ne, ne * nx, 1};
    srwr_layout ln = {ne, ne * (long long)mn.nx, 1};
    rc = device_resize(*pWfr, P, dO, dO + nOld, lo, dN, dN + nNew, ln, st);
    if (rc) return rc;
    HIPCHK(hipStreamSynchronize(st));
    rc = realloc_wfr_arrays(*pWfr, P);
    if (rc) return rc;
    HIPCHK(hipMemcpyAsync(pWfr->arEx, dN, sizeof(float2) * nNew, hipMemcpyDeviceToHost, st));
    HIPCHK(hipMemcpyAsync(pWfr->arEy, dN + nNew, sizeof(float2) * nNew, hipMemcpyDeviceToHost, st));
    HIPCHK(hipStreamSynchronize(st));
    return 0;
}

}  // extern "C"

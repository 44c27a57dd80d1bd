// multie_api.cpp - partially coherent emission and propagation: batches of
// macro-electrons through the SR integral, the drift passes and the
// re-gridded intensity accumulation, all on the device.
//
// Reference behaviour (paths relative to the reference source root):
//   beam sampling        env/python/srwpy/srwlib.py:9000-9029, 9793-9803
//   per-electron loop    env/python/srwpy/srwlib.py:9747-10559
//   averaging            env/python/srwpy/srwlib.py:1083-1240 (avg_update_interp, order 1)
// The reference draws normals with Python's random.gauss; this build uses a
// counter-based Philox4x32-10 stream (one counter block per electron), so a
// run is reproducible and independent of the batch size.
#include "srw_api_internal.h"
#include "srw_multie.h"
#include "srw_prop.h"
#include "../../include/srw_rng.h"

#include <chrono>
#include <cmath>
#include <cstring>
#include <vector>

namespace srwamd {

// srwlib.py:9000-9029: Gaussian beam parameters of the x, y phase planes.
struct BeamSampler {
    double x0, xp0, y0, yp0, gamma0, relEnSpr;
    double SigPX, SigQX, AX, GX, SigPY, SigQY, AY, GY;
    bool init(const SRWLPartBeam& b)
    {
        x0 = b.partStatMom1.x; xp0 = b.partStatMom1.xp;
        y0 = b.partStatMom1.y; yp0 = b.partStatMom1.yp;
        gamma0 = b.partStatMom1.gamma;
        const double* m = b.arStatMom2;
        relEnSpr = std::sqrt(m[10]);
        double dx = m[0] * m[2] - m[1] * m[1], dy = m[3] * m[5] - m[4] * m[4];
        if (!(dx > 0) || !(dy > 0)) return false;
        double mx = 0.5 / dx;
        double BX = m[0] * mx; GX = m[2] * mx; AX = m[1] * mx;
        SigPX = 1. / std::sqrt(2. * GX);
        SigQX = std::sqrt(GX / (2. * (BX * GX - AX * AX)));
        double my = 0.5 / dy;
        double BY = m[3] * my; GY = m[5] * my; AY = m[4] * my;
        SigPY = 1. / std::sqrt(2. * GY);
        SigQY = std::sqrt(GY / (2. * (BY * GY - AY * AY)));
        return true;
    }
    // srwlib.py:9793-9803
    void map(const double* g, SRWLParticle& p) const
    {
        double pxp = SigQX * g[0];
        double px = SigPX * g[1] + AX * pxp / GX;
        double pyp = SigQY * g[2];
        double py = SigPY * g[3] + AY * pyp / GY;
        p.x = x0 + px; p.xp = xp0 + pxp;
        p.y = y0 + py; p.yp = yp0 + pyp;
        p.gamma = gamma0 * (1. + relEnSpr * g[4]);
    }
};

}  // namespace srwamd

using namespace srwamd;

#define HIPCHK(x) do { if ((x) != hipSuccess) return SRWL_AMD_ERR_GPU; } while (0)

extern "C" {

// Draws the standard-normal vector of electron i (test hook of the sampler).
SRWL_EXPORT int srwlAmdNormals(unsigned long long seed, long long i, double* g6)
{
    srw_normals6(seed, (unsigned long long)i, g6);
    return 0;
}

SRWL_EXPORT int srwlEmitPropMultiE(SRWLStokes* pStokes, SRWLWfr* pWfr0, SRWLMagFldC* pMagFld, SRWLOptC* pOpt,
                                   double* precPar, int nPrecPar, long long nMacroElec, double* arPar,
                                   double* arOut)
{
    if (!pStokes || !pWfr0 || !pMagFld || !precPar || nMacroElec <= 0) return SRWL_AMD_ERR_BATCH;
    if (pStokes->arS0 == nullptr || pStokes->mesh.nx < 2 || pStokes->mesh.ny < 2) return SRWL_AMD_ERR_BATCH;
    if ((int)precPar[0] != 0 || (nPrecPar > 6 && precPar[6] > 0)) return SRWL_AMD_ERR_UNSUPPORTED;
    if (pWfr0->mesh.ne != 1) return SRWL_AMD_ERR_UNSUPPORTED;
    long nx = pWfr0->mesh.nx, nz = pWfr0->mesh.ny;
    int lx = 0, lz = 0;
    while ((1L << lx) < nx) lx++;
    while ((1L << lz) < nz) lz++;
    if ((1L << lx) != nx || (1L << lz) != nz || lx < 4 || lz < 4 || lx > 11 || lz > 11) return SRWL_ERR_IN_FFT;
    int nDrift = pOpt ? pOpt->nElem : 0;
    std::vector<int> modes(nDrift);
    for (int i = 0; i < nDrift; i++) {
        if (std::strcmp(pOpt->arOptTypes[i], "drift") != 0) return SRWL_AMD_ERR_UNSUPPORTED;
        const double* pp = (pOpt->arProp && i < pOpt->nProp) ? pOpt->arProp[i] : nullptr;
        int rc = drift_mode_from_prop(pp, modes[i]);
        if (rc) return rc;
    }
    BeamSampler bs;
    if (!bs.init(pWfr0->partBeam)) return SRWL_AMD_ERR_BATCH;
    unsigned long long seed = arPar ? (unsigned long long)arPar[0] : 1ull;
    long long B = (arPar && arPar[1] > 0) ? (long long)arPar[1] : 256;
    if (B > nMacroElec) B = nMacroElec;
    double target = (arPar && arPar[3] > 0) ? arPar[3] : 0.;
    int rc = fft_ready();
    if (rc) return rc;
    hipStream_t st = g_ctx.stream;
    long long nxy = (long long)nx * nz;
    srwm_rmesh rm;
    rm.nx = (int)pStokes->mesh.nx; rm.nz = (int)pStokes->mesh.ny;
    rm.xStart = pStokes->mesh.xStart; rm.xStep = (pStokes->mesh.xFin - pStokes->mesh.xStart) / (rm.nx - 1);
    rm.zStart = pStokes->mesh.yStart; rm.zStep = (pStokes->mesh.yFin - pStokes->mesh.yStart) / (rm.nz - 1);
    long long nr = (long long)rm.nx * rm.nz;
    Buffers& Bf = g_ctx.buf;
    float* dF = (float*)Bf.Ex.get(sizeof(float2) * 2 * nxy * B);
    float* dI = (float*)Bf.W3.get(sizeof(float) * nxy * B);
    double* dS = (double*)Bf.Stokes.get(sizeof(double) * 2 * nr);
    srwm_emesh* dEm = (srwm_emesh*)Bf.W1.get(sizeof(srwm_emesh) * B);
    float* dW = (float*)Bf.Wt.get(sizeof(float) * B);
    srwf_slice* dSp = (srwf_slice*)Bf.W2.get(sizeof(srwf_slice) * 2 * B);
    double* dN2 = (double*)Bf.Grad.get(sizeof(double) * B);
    if (!dF || !dI || !dS || !dEm || !dW || !dSp || !dN2) return SRWL_ERR_MEMORY_ALLOCATION_FAILURE;
    HIPCHK(hipMemsetAsync(dS, 0, sizeof(double) * 2 * nr, st));
    std::vector<srwd_field> fields(1);
    double gs0, gs1;
    rc = field_from_c(pMagFld, fields[0], gs0, gs1);
    if (rc) return rc;
    int termin = (nPrecPar <= 0 || nPrecPar > 5) ? (int)(char)precPar[5] : 1;
    double xStep0 = (pWfr0->mesh.xFin - pWfr0->mesh.xStart) / (nx - 1);
    double zStep0 = (pWfr0->mesh.yFin - pWfr0->mesh.yStart) / (nz - 1);
    auto t0 = std::chrono::steady_clock::now();
    long long done = 0;
    double sumW = 0, sumW2 = 0;
    double errEst = -1;
    std::vector<double> S1h, S2h;
    while (done < nMacroElec) {
        long long nb = std::min<long long>(B, nMacroElec - done);
        std::vector<SrJob> jobs(nb);
        std::vector<long long> off(nb);
        std::vector<float> w(nb, 1.f);
        for (long long j = 0; j < nb; j++) {
            double g[6];
            srw_normals6(seed, (unsigned long long)(done + j), g);
            SRWLParticle p = pWfr0->partBeam.partStatMom1;
            bs.map(g, p);
            std::vector<srwd_field> dummy;
            SrJob& J = jobs[j];
            J.part = p;
            J.field = 0;
            J.np = (long long)(int)precPar[4];
            double a = precPar[2], b = precPar[3];
            double sSt = a - p.z, sEn = b - p.z;
            if (sSt >= sEn) {
                sSt = gs0; sEn = gs1;
                if (sSt > p.z) sSt = p.z;
                if (sEn < p.z) sEn = p.z;
                sSt -= p.z; sEn -= p.z;
            }
            J.ctStart = sSt; J.ctEnd = sEn;
            J.Iavg = pWfr0->partBeam.Iavg;
            J.mesh = pWfr0->mesh;
            J.step = precPar[1];
            J.sLim0 = precPar[2]; J.sLim1 = precPar[3];
            off[j] = 2 * j * nxy;
            sumW += w[j]; sumW2 += (double)w[j] * w[j];
        }
        std::vector<SrResult> res;
        rc = sr_run(jobs, fields, dF, dF + 2 * nxy, off, res, termin);
        if (rc > 0) return rc;
        // drift(s): per electron wavefront radius, centre and mesh
        std::vector<srwm_emesh> em(nb);
        std::vector<double> Rx(nb), Rz(nb);
        for (long long j = 0; j < nb; j++) {
            em[j].xStart = pWfr0->mesh.xStart; em[j].xStep = xStep0;
            em[j].zStart = pWfr0->mesh.yStart; em[j].zStep = zStep0;
            em[j].nx = (int)nx; em[j].nz = (int)nz;
            Rx[j] = Rz[j] = res[j].Robs;
        }
        std::vector<srwf_slice> sp(2 * nb);
        for (int d = 0; d < nDrift; d++) {
            const SRWLOptD* D = (const SRWLOptD*)pOpt->arOpt[d];
            for (long long j = 0; j < nb; j++) {
                DriftGeom g;
                std::memset(&g, 0, sizeof(g));
                if (modes[d] == 3) {
                    mode3_setup(Rx[j], Rz[j], res[j].RobsErr, res[j].RobsErr, D->L, g);
                    if (g.kx < 0 || g.kz < 0) return SRWL_AMD_ERR_UNSUPPORTED;
                }
                drift_slice(sp[2 * j], modes[d], g, D->L, D->treat, pWfr0->mesh.eStart, nx, nz, em[j].xStart,
                            em[j].xStep, em[j].zStart, em[j].zStep, res[j].xc, res[j].zc);
                sp[2 * j + 1] = sp[2 * j];
                if (modes[d] == 3) {
                    em[j].xStart = sp[2 * j].nxStart; em[j].xStep = sp[2 * j].nxStep;
                    em[j].zStart = sp[2 * j].nzStart; em[j].zStep = sp[2 * j].nzStep;
                }
                Rx[j] += D->L; Rz[j] += D->L;
            }
            HIPCHK(hipMemcpyAsync(dSp, sp.data(), sizeof(srwf_slice) * 2 * nb, hipMemcpyHostToDevice, st));
            bool last = (d == nDrift - 1);
            rc = drift_passes((float2*)dF, nxy, (int)(2 * nb), dSp, lx, lz, modes[d], last ? 2 : 1, dI, nxy, st);
            if (rc) return rc;
        }
        if (nDrift == 0) {
            // intensity of the emitted field itself
            for (long long j = 0; j < nb; j++)
                if (srwp_launch_intensity((const float2*)(dF + 2 * off[j]), (const float2*)(dF + 2 * off[j] + 2 * nxy),
                                          1, 0, 0, 0.f, 6, nxy, dI + j * nxy, st))
                    return SRWL_AMD_ERR_GPU;
        }
        HIPCHK(hipMemcpyAsync(dEm, em.data(), sizeof(srwm_emesh) * nb, hipMemcpyHostToDevice, st));
        HIPCHK(hipMemcpyAsync(dW, w.data(), sizeof(float) * nb, hipMemcpyHostToDevice, st));
        if (srwm_launch_accum(dI, nxy, dEm, dW, (int)nb, rm, dS, dS + nr, nullptr, st)) return SRWL_AMD_ERR_GPU;
        done += nb;
        if (target > 0 || done == nMacroElec) {
            S1h.resize(nr); S2h.resize(nr);
            HIPCHK(hipMemcpyAsync(S1h.data(), dS, sizeof(double) * nr, hipMemcpyDeviceToHost, st));
            HIPCHK(hipMemcpyAsync(S2h.data(), dS + nr, sizeof(double) * nr, hipMemcpyDeviceToHost, st));
            HIPCHK(hipStreamSynchronize(st));
            double a1 = 0, a2 = 0;
            for (long long r = 0; r < nr; r++) { a1 += S1h[r] * S1h[r]; a2 += S2h[r]; }
            double N = (double)done;
            errEst = (done > 1 && a1 > 0) ? std::sqrt(std::max(0., N * (a2 - a1 / N) / ((N - 1.) * a1))) : -1;
            if (target > 0 && errEst > 0 && errEst <= target) break;
        }
    }
    double scale = 1. / (double)done;
    if (srwm_launch_stokes_out(dS, scale, nr, (float*)dI, st)) return SRWL_AMD_ERR_GPU;
    HIPCHK(hipMemcpyAsync(pStokes->arS0, dI, sizeof(float) * nr, hipMemcpyDeviceToHost, st));
    HIPCHK(hipStreamSynchronize(st));
    double secs = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    if (arOut) {
        arOut[0] = (double)done;
        arOut[1] = sumW2 > 0 ? sumW * sumW / sumW2 : 0;
        arOut[2] = sumW;
        arOut[3] = secs;
        arOut[4] = errEst;
        arOut[5] = (double)B;
        arOut[6] = 0;
        arOut[7] = 0;
    }
    return 0;
}

}  // extern "C"

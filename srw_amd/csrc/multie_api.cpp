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
#include "srw_moments.h"
#include "srw_prop.h"
#include "srw_learn.h"
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

// Learned proposal state: network parameters and Adam moments on the device,
// the current mixture on the host (for sampling and p/q weights).
struct Proposal {
    float* P = nullptr; float* m = nullptr; float* v = nullptr; float* G = nullptr;
    float* act = nullptr; float* gout = nullptr; float* loss = nullptr; float* in = nullptr;
    float* gs = nullptr; float* wh = nullptr;
    int step = 0;
    double alpha = 0.25;
    double logit[SRWL_K], mu[SRWL_K][SRWL_D], ls[SRWL_K][SRWL_D];
};

// Initial parameters: scaled normal hidden weights and small output weights,
// output biases zero, so that every component starts as N(0, I) and the
// initial proposal is p itself; training then moves it toward q*.
void proposal_init_params(std::vector<float>& P, unsigned long long seed)
{
    P.assign(SRWL_NPARAM, 0.f);
    auto fill = [&](int off, int n, double sc, unsigned long long sub) {
        for (int i = 0; i < n; i += 6) {
            double g[6];
            srw_normals6(seed ^ (0x5151ull * sub), (unsigned long long)i, g);
            for (int k = 0; k < 6 && i + k < n; k++) P[off + i + k] = (float)(sc * g[k]);
        }
    };
    fill(SRWL_OFF_W1, SRWL_H * SRWL_IN, 1. / std::sqrt((double)SRWL_IN), 1);
    fill(SRWL_OFF_W2, SRWL_H * SRWL_H, 1. / std::sqrt((double)SRWL_H), 2);
    fill(SRWL_OFF_W3, SRWL_OUT * SRWL_H, 0.01 / std::sqrt((double)SRWL_H), 3);
}

void proposal_set(Proposal& Q, const float* o)
{
    for (int k = 0; k < SRWL_K; k++) {
        Q.logit[k] = o[k];
        for (int d = 0; d < SRWL_D; d++) {
            Q.mu[k][d] = o[SRWL_K + k * SRWL_D + d];
            double v = o[SRWL_K + SRWL_K * SRWL_D + k * SRWL_D + d];
            Q.ls[k][d] = std::min(std::max(v, -3.), 1.);
        }
    }
}

// log p(g) and log q(g) of the defensive mixture (same formulas as
// k_gmm_grad) over the first nd coordinates (a mixture of diagonal Gaussians
// marginalised over the others)
void proposal_logpq(const Proposal& Q, const double* g, double& lp, double& lq, int nd = SRWL_D)
{
    const double L2P = 1.8378770664093453;
    lp = -0.5 * nd * L2P;
    for (int d = 0; d < nd; d++) lp -= 0.5 * g[d] * g[d];
    double lmax = -1e300;
    for (int k = 0; k < SRWL_K; k++) lmax = std::max(lmax, Q.logit[k]);
    double se = 0;
    for (int k = 0; k < SRWL_K; k++) se += std::exp(Q.logit[k] - lmax);
    double logZ = lmax + std::log(se);
    double lN[SRWL_K], mm = -1e300;
    for (int k = 0; k < SRWL_K; k++) {
        double s = Q.logit[k] - logZ;
        for (int d = 0; d < nd; d++) {
            double z = (g[d] - Q.mu[k][d]) * std::exp(-Q.ls[k][d]);
            s += -0.5 * z * z - Q.ls[k][d] - 0.5 * L2P;
        }
        lN[k] = s;
        mm = std::max(mm, s);
    }
    double ms = 0;
    for (int k = 0; k < SRWL_K; k++) ms += std::exp(lN[k] - mm);
    double lm = mm + std::log(ms);
    double top = std::max(lm, lp);
    lq = top + std::log(Q.alpha * std::exp(lp - top) + (1. - Q.alpha) * std::exp(lm - top));
}

// Draw of electron i from q: the base normals of plain sampling, a mixture
// choice from a third Philox block.
void proposal_draw(const Proposal& Q, unsigned long long seed, unsigned long long i, double* g)
{
    double z[6];
    srw_normals6(seed, i, z);
    uint32_t ctr[4] = {(uint32_t)i, 2u, (uint32_t)(i >> 32), 0u}, u[4];
    srw_philox4x32_10(ctr, (uint32_t)seed, (uint32_t)(seed >> 32), u);
    double a = srw_u01(u[0]);
    if (a < Q.alpha) { for (int d = 0; d < 6; d++) g[d] = z[d]; return; }
    double b = srw_u01(u[1]);
    double lmax = -1e300;
    for (int k = 0; k < SRWL_K; k++) lmax = std::max(lmax, Q.logit[k]);
    double pk[SRWL_K], se = 0;
    for (int k = 0; k < SRWL_K; k++) { pk[k] = std::exp(Q.logit[k] - lmax); se += pk[k]; }
    int kk = SRWL_K - 1;
    double cum = 0;
    for (int k = 0; k < SRWL_K; k++) { cum += pk[k] / se; if (b < cum) { kk = k; break; } }
    for (int d = 0; d < SRWL_D; d++) g[d] = Q.mu[kk][d] + std::exp(Q.ls[kk][d]) * z[d];
}

// The single Gaussian the quasi-Monte Carlo path draws from: the mixture's
// Gaussians moment-matched per coordinate, widths held at >= 1 (the beam
// distribution's), so that p/q stays bounded without a defensive part.  An
// affine map of the normals keeps the integrand as smooth in the Sobol'
// point as under p, which a mixture's component choice would not.
void proposal_single(const Proposal& Q, double* m, double* sd)
{
    double lmax = -1e300, se = 0., pk[SRWL_K];
    for (int k = 0; k < SRWL_K; k++) lmax = std::max(lmax, Q.logit[k]);
    for (int k = 0; k < SRWL_K; k++) { pk[k] = std::exp(Q.logit[k] - lmax); se += pk[k]; }
    for (int d = 0; d < 5; d++) {
        double a = 0., b = 0.;
        for (int k = 0; k < SRWL_K; k++) {
            const double w = pk[k] / se, s2 = std::exp(2. * Q.ls[k][d]);
            a += w * Q.mu[k][d];
            b += w * (s2 + Q.mu[k][d] * Q.mu[k][d]);
        }
        m[d] = a;
        sd[d] = std::max(1., std::sqrt(std::max(b - a * a, 0.)));
    }
}

// log p(g) - log q(g) for the single Gaussian of proposal_single (5-D).
double single_logw(const double* m, const double* sd, const double* g)
{
    double r = 0.;
    for (int d = 0; d < 5; d++) {
        const double z = (g[d] - m[d]) / sd[d];
        r += -0.5 * g[d] * g[d] + 0.5 * z * z + std::log(sd[d]);
    }
    return r;
}

}  // namespace srwamd

using namespace srwamd;

#define HIPCHK(x) do { if ((x) != hipSuccess) return SRWL_AMD_ERR_GPU; } while (0)

extern "C" {

// Forward + mixture gradient + backward of the proposal network for given
// parameters, descriptor, samples g (nb x 6) and weights (test hook used to
// check the hand-written kernels against a CPU implementation).
SRWL_EXPORT int srwlAmdProposalGrad(const float* P, const float* in6, const float* g, const float* wh, int nb,
                                    float alpha, float* G, float* out104, float* loss)
{
    int rc = ensure_device();
    if (rc) return rc;
    hipStream_t st = g_ctx.stream;
    size_t n = 2 * SRWL_NPARAM + SRWL_NACT + SRWL_OUT + 2 + 8 + 7 * (size_t)nb;
    float* b = (float*)g_ctx.buf.Grad.get(sizeof(float) * n);
    if (!b) return SRWL_ERR_MEMORY_ALLOCATION_FAILURE;
    float *dP = b, *dG = dP + SRWL_NPARAM, *dAct = dG + SRWL_NPARAM, *dGo = dAct + SRWL_NACT, *dL = dGo + SRWL_OUT;
    float *dIn = dL + 2, *dg = dIn + 8, *dw = dg + 6 * nb;
    HIPCHK(hipMemcpyAsync(dP, P, sizeof(float) * SRWL_NPARAM, hipMemcpyHostToDevice, st));
    HIPCHK(hipMemcpyAsync(dIn, in6, sizeof(float) * 6, hipMemcpyHostToDevice, st));
    HIPCHK(hipMemcpyAsync(dg, g, sizeof(float) * 6 * nb, hipMemcpyHostToDevice, st));
    HIPCHK(hipMemcpyAsync(dw, wh, sizeof(float) * nb, hipMemcpyHostToDevice, st));
    HIPCHK(hipMemsetAsync(dGo, 0, sizeof(float) * (SRWL_OUT + 2), st));
    if (srwl_launch_mlp_forward(dP, dIn, dAct, st)) return SRWL_AMD_ERR_GPU;
    if (srwl_launch_gmm_grad(dAct, dg, dw, nb, alpha, dGo, dL, st)) return SRWL_AMD_ERR_GPU;
    if (srwl_launch_mlp_backward(dP, dAct, dGo, dG, st)) return SRWL_AMD_ERR_GPU;
    HIPCHK(hipMemcpyAsync(G, dG, sizeof(float) * SRWL_NPARAM, hipMemcpyDeviceToHost, st));
    HIPCHK(hipMemcpyAsync(out104, dAct + SRWL_ACT_OUT, sizeof(float) * SRWL_OUT, hipMemcpyDeviceToHost, st));
    HIPCHK(hipMemcpyAsync(loss, dL, sizeof(float), hipMemcpyDeviceToHost, st));
    HIPCHK(hipStreamSynchronize(st));
    return 0;
}

SRWL_EXPORT int srwlAmdProposalInit(float* P, unsigned long long seed)
{
    std::vector<float> v;
    proposal_init_params(v, seed);
    std::memcpy(P, v.data(), sizeof(float) * SRWL_NPARAM);
    return SRWL_NPARAM;
}

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
    std::vector<int> modes(nDrift), subs(nDrift);
    std::vector<char> isDrift(nDrift);
    for (int i = 0; i < nDrift; i++) {
        const char* ty = pOpt->arOptTypes[i];
        isDrift[i] = std::strcmp(ty, "drift") == 0;
        if (std::strcmp(ty, "aperture") == 0 || std::strcmp(ty, "obstacle") == 0) {
            const SRWLOptA* A = (const SRWLOptA*)pOpt->arOpt[i];
            const char ao = A->ap_or_ob;
            if ((A->shape != 'r' && A->shape != 'c') || (ao != 'a' && ao != 'A' && ao != 'o' && ao != 'O'))
                return SRWL_AMD_ERR_UNSUPPORTED;
        } else if (!isDrift[i] && std::strcmp(ty, "lens") != 0) {
            return SRWL_AMD_ERR_UNSUPPORTED;
        }
        const double* pp = (pOpt->arProp && i < pOpt->nProp) ? pOpt->arProp[i] : nullptr;
        int rc = drift_mode_from_prop(pp, modes[i], subs[i]);
        if (rc) return rc;
        // resampling instructions (arProp[5..8] != 1, sroptcnt.cpp:314-327) are
        // not applied per electron here: refuse them instead of ignoring them
        const int npp = (pOpt->arPropN != nullptr) ? (int)(unsigned char)pOpt->arPropN[i] : 9;
        if (pp && npp > 8 && (pp[5] != 1. || pp[6] != 1. || pp[7] != 1. || pp[8] != 1.))
            return SRWL_AMD_ERR_UNSUPPORTED;
    }
    // a trailing post-resize instruction (sroptcnt.cpp:408-423) likewise
    if (pOpt && pOpt->arProp && pOpt->nProp > pOpt->nElem) {
        const double* pp = pOpt->arProp[pOpt->nElem];
        const int npp = (pOpt->arPropN != nullptr) ? (int)(unsigned char)pOpt->arPropN[pOpt->nElem] : 9;
        if (pp && npp > 8 && (pp[5] != 1. || pp[6] != 1. || pp[7] != 1. || pp[8] != 1.))
            return SRWL_AMD_ERR_UNSUPPORTED;
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
    int useQ = (arPar && arPar[2] > 0) ? 1 : 0;
    // arPar[6] > 0: randomized quasi-Monte Carlo draws (digitally shifted
    // Sobol' points, arPar[7] independent replicates, 16 by default) instead
    // of pseudo-random ones; the error estimate is the spread of the
    // replicate averages.  Batches hold whole groups of one point per
    // replicate.
    const int rqmc = (arPar && arPar[6] > 0) ? 1 : 0;
    const int R = rqmc ? ((arPar && arPar[7] > 1) ? (int)arPar[7] : 16) : 1;
    if (rqmc) {
        if (nMacroElec < R) return SRWL_AMD_ERR_BATCH;
        B = std::max<long long>(R, (B / R) * R);
    }
    uint32_t sobolV[SRW_SOBOL_DIMS][32];
    if (rqmc) srw_sobol_directions(sobolV);
    double* dSrep = nullptr;
    if (rqmc) {
        dSrep = (double*)Bf.Rep.get(sizeof(double) * R * nr);
        if (!dSrep) return SRWL_ERR_MEMORY_ALLOCATION_FAILURE;
        HIPCHK(hipMemsetAsync(dSrep, 0, sizeof(double) * R * nr, st));
    }
    std::vector<double> Srh;
    // arPar[5] > 0: also estimate the gain of the variance-optimal proposal
    // q* ~ p ||I|| for this estimator from the electrons drawn (plain runs)
    int diag = (arPar && arPar[5] > 0 && !useQ) ? 1 : 0;
    double sumN2 = 0., sumN = 0.;
    int trainSteps = (arPar && arPar[4] > 0) ? (int)arPar[4] : 32;
    // training set: every electron drawn so far (up to kReplay, oldest
    // dropped first) with its standardized phase-space point, its weight p/q
    // at the time it was drawn and its intensity norm
    const long long kReplay = 8192;
    std::vector<float> repG;
    std::vector<double> repW, repC;
    Proposal Q;
    std::vector<float> hOut(SRWL_OUT);
    double lastLoss = 0;
    // The proposal is used only where the electrons drawn so far predict a
    // gain: the variance of w I under the current q over that of I under p,
    // both estimated from the replay set, must fall below kUseRatio;
    // otherwise the batch is drawn from p itself, plain sampling being the
    // fallback.
    const double kUseRatio = 0.8;
    double lastRatio = -1.;
    double sgM[5] = {0, 0, 0, 0, 0}, sgS[5] = {1, 1, 1, 1, 1};   // single Gaussian of the RQMC path
    long long mixBatches = 0;
    if (useQ) {
        float* base = (float*)Bf.Net.get(sizeof(float) * (4 * SRWL_NPARAM + SRWL_NACT + SRWL_OUT + 8 + 7 * kReplay + 64));
        if (!base) return SRWL_ERR_MEMORY_ALLOCATION_FAILURE;
        Q.P = base; Q.m = Q.P + SRWL_NPARAM; Q.v = Q.m + SRWL_NPARAM; Q.G = Q.v + SRWL_NPARAM;
        Q.act = Q.G + SRWL_NPARAM; Q.gout = Q.act + SRWL_NACT; Q.loss = Q.gout + SRWL_OUT; Q.in = Q.loss + 2;
        Q.gs = Q.in + 8; Q.wh = Q.gs + 6 * kReplay;
        std::vector<float> P0;
        proposal_init_params(P0, 0x6c65617264ull);
        HIPCHK(hipMemcpyAsync(Q.P, P0.data(), sizeof(float) * SRWL_NPARAM, hipMemcpyHostToDevice, st));
        HIPCHK(hipMemsetAsync(Q.m, 0, sizeof(float) * 2 * SRWL_NPARAM, st));
        // beam / beamline descriptor: log-scaled rms sizes, divergences, energy spread, total drift
        const double* m2 = pWfr0->partBeam.arStatMom2;
        double Ltot = 0;
        for (int d = 0; d < nDrift; d++)
            if (isDrift[d]) Ltot += ((const SRWLOptD*)pOpt->arOpt[d])->L;
        float in[8] = {(float)std::log(std::sqrt(m2[0]) / 1e-5 + 1e-30), (float)std::log(std::sqrt(m2[2]) / 1e-6 + 1e-30),
                       (float)std::log(std::sqrt(m2[3]) / 1e-5 + 1e-30), (float)std::log(std::sqrt(m2[5]) / 1e-6 + 1e-30),
                       (float)std::log(std::sqrt(m2[10]) / 1e-3 + 1e-30), (float)(Ltot / 10.), 0.f, 0.f};
        HIPCHK(hipMemcpyAsync(Q.in, in, sizeof(float) * 8, hipMemcpyHostToDevice, st));
    }
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
    double errEst = -1, lastA1 = 0;
    std::vector<double> S1h, S2h;
    while (done < nMacroElec) {
        long long nb = std::min<long long>(B, nMacroElec - done);
        if (rqmc) nb = (nb / R) * R;
        if (nb <= 0) break;
        const long long perRep = nb / R;
        std::vector<SrJob> jobs(nb);
        std::vector<long long> off(nb);
        std::vector<float> w(nb, 1.f);
        std::vector<float> gsamp(6 * nb);
        if (useQ) {
            if (srwl_launch_mlp_forward(Q.P, Q.in, Q.act, st)) return SRWL_AMD_ERR_GPU;
            HIPCHK(hipMemcpyAsync(hOut.data(), Q.act + SRWL_ACT_OUT, sizeof(float) * SRWL_OUT, hipMemcpyDeviceToHost, st));
            HIPCHK(hipStreamSynchronize(st));
            proposal_set(Q, hOut.data());
            // the first batch, before any training, is drawn from p itself;
            // later ones from the mixture only where it predicts a gain
            Q.alpha = 0.25;
            if (rqmc) proposal_single(Q, sgM, sgS);
            bool use = false;
            if (Q.step > 0 && !repW.empty() && done > 0) {
                const double m2 = lastA1 / ((double)done * (double)done);
                double a = 0., b = 0.;
                for (size_t j = 0; j < repW.size(); j++) {
                    double gg[SRWL_D], lw;
                    for (int d = 0; d < SRWL_D; d++) gg[d] = repG[6 * j + d];
                    if (rqmc) lw = single_logw(sgM, sgS, gg);
                    else {
                        double lp, lq;
                        proposal_logpq(Q, gg, lp, lq);
                        lw = lp - lq;
                    }
                    const double n2 = repC[j] * repC[j];
                    a += repW[j] * std::exp(lw) * n2;
                    b += repW[j] * n2;
                }
                const double nrep = (double)repW.size();
                const double vq = a / nrep - m2, vp = b / nrep - m2;
                lastRatio = vp > 0. ? vq / vp : -1.;
                use = lastRatio >= 0. && lastRatio < kUseRatio;
            }
            if (!use) {
                Q.alpha = 1.0;
                for (int d = 0; d < 5; d++) { sgM[d] = 0.; sgS[d] = 1.; }
            } else mixBatches++;
        }
        for (long long j = 0; j < nb; j++) {
            double g[6];
            if (useQ && rqmc) {
                // slot j: replicate j / perRep, its point done / R + j % perRep
                const unsigned r = (unsigned)(j / perRep);
                const unsigned long long k = (unsigned long long)(done / R + j % perRep);
                for (int d = 0; d < 5; d++)
                    g[d] = sgM[d] + sgS[d] * srw_norm_inv(srw_u01(srw_sobol_word(sobolV, d, k) ^ srw_rqmc_shift(seed, r, d)));
                g[5] = 0.;
                w[j] = (float)std::exp(single_logw(sgM, sgS, g));
                for (int d = 0; d < 6; d++) gsamp[6 * j + d] = (float)g[d];
            } else if (useQ) {
                proposal_draw(Q, seed, (unsigned long long)(done + j), g);
                double lp, lq;
                proposal_logpq(Q, g, lp, lq);
                w[j] = (float)std::exp(lp - lq);
                for (int d = 0; d < 6; d++) gsamp[6 * j + d] = (float)g[d];
            } else if (rqmc) {
                // slot j: replicate j / perRep, its point done / R + j % perRep
                srw_rqmc_normals6(sobolV, seed, (unsigned)(j / perRep), (unsigned long long)(done / R + j % perRep), g);
            } else srw_normals6(seed, (unsigned long long)(done + j), g);
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
        // per-electron wavefront state seen by the drifts (ne = 1)
        std::vector<SRWLWfr> ws(nb);
        for (long long j = 0; j < nb; j++) {
            SRWLWfr& q = ws[j];
            std::memset(&q, 0, sizeof(q));
            q.mesh = pWfr0->mesh;
            q.Rx = q.Ry = Rx[j]; q.dRx = q.dRy = res[j].RobsErr;
            q.xc = res[j].xc; q.yc = res[j].zc;
            q.avgPhotEn = pWfr0->mesh.eStart;
        }
        // moments of each electron's field, formed when a drift first needs them
        // and then carried through the drifts like the reference does
        std::vector<double> mom;
        bool momValid = false;
        auto emesh_to_wfr = [&](long long j) {
            SRWLWfr& q = ws[j];
            q.mesh.xStart = em[j].xStart; q.mesh.xFin = em[j].xStart + (nx - 1) * em[j].xStep;
            q.mesh.yStart = em[j].zStart; q.mesh.yFin = em[j].zStart + (nz - 1) * em[j].zStep;
        };
        struct OneCtx { SRWLWfr* w; const float2* d; long long nxy, nx; double *mx, *mz; };
        auto recompute = [](void* c) -> int {
            OneCtx* m = (OneCtx*)c;
            return wfr_moments(*m->w, m->d, m->d + m->nxy, 0, 1, m->nx, m->mx, m->mz);
        };
        for (int d = 0; d < nDrift; d++) {
            if (!isDrift[d]) {
                // thin lens / aperture / obstacle on each electron's own mesh
                // (sroptfoc.h:44-70, 227-242; sroptapt.h:150-166, 244-258, 330-343,
                // 425-440).  The oracle's per-electron wavefronts carry no
                // moments, so a later drift that needs them computes them from
                // the field again.
                const char* ty = pOpt->arOptTypes[d];
                const bool lens = std::strcmp(ty, "lens") == 0;
                const SRWLOptL* L = (const SRWLOptL*)pOpt->arOpt[d];
                const SRWLOptA* A = (const SRWLOptA*)pOpt->arOpt[d];
                for (long long j = 0; j < nb; j++) {
                    srwo_elem E{};
                    E.nx = nx; E.xStart = em[j].xStart; E.xStep = em[j].xStep; E.zStart = em[j].zStart;
                    E.zStep = em[j].zStep; E.eStart = pWfr0->mesh.eStart; E.eStep = 0.;
                    if (lens) {
                        emesh_to_wfr(j);
                        ws[j].xc = res[j].xc; ws[j].yc = res[j].zc;
                        if (lin_phase_would_subtract(ws[j], L->x, L->y)) return SRWL_AMD_ERR_UNSUPPORTED;
                        E.kind = 0; E.p0 = L->Fx; E.p1 = L->Fy; E.cx = L->x; E.cy = L->y;
                    } else {
                        const bool circ = A->shape == 'c', ob = A->ap_or_ob == 'o' || A->ap_or_ob == 'O';
                        E.kind = circ ? (ob ? 4 : 3) : (ob ? 2 : 1);
                        E.p0 = A->Dx; E.p1 = A->Dy; E.cx = A->x; E.cy = A->y;
                    }
                    if (srwo_launch_point_elem((float2*)dF + 2 * j * nxy, nxy, 1, E, st)) return SRWL_AMD_ERR_GPU;
                    if (lens) {
                        const double difX = (L->Fx == ws[j].Rx) ? 1E-23 : L->Fx - ws[j].Rx;
                        const double difZ = (L->Fy == ws[j].Ry) ? 1E-23 : L->Fy - ws[j].Ry;
                        const double mX = L->Fx / difX, mZ = L->Fy / difZ;
                        ws[j].Rx *= mX; ws[j].Ry *= mZ;
                        ws[j].dRx *= mX * mX; ws[j].dRy *= mZ * mZ;
                        res[j].xc = L->x - (L->x - res[j].xc) * mX;
                        res[j].zc = L->y - (L->y - res[j].zc) * mZ;
                        ws[j].xc = res[j].xc; ws[j].yc = res[j].zc;
                        Rx[j] = ws[j].Rx; Rz[j] = ws[j].Ry;
                    }
                    if (ws[j].Rx != 0 && ws[j].dRx == 0) ws[j].dRx = std::fabs(0.1 * ws[j].Rx);
                    if (ws[j].Ry != 0 && ws[j].dRy == 0) ws[j].dRy = std::fabs(0.1 * ws[j].Ry);
                }
                momValid = false;
                continue;
            }
            const SRWLOptD* D = (const SRWLOptD*)pOpt->arOpt[d];
            if (modes[d] == 3 && subs[d] == 1 && !momValid) {
                mom.assign(22 * (size_t)nb, 0.);
                std::vector<srwm_mslice> sl(nb);
                for (long long j = 0; j < nb; j++) {
                    emesh_to_wfr(j);
                    // electron j's Ex is slice 2j of dF, its Ez slice 2j + 1, each [iz][ix]
                    const float2* base = (const float2*)dF + 2 * j * nxy;
                    sl[j] = wfr_slice(ws[j], base, base + nxy, 0, 1, nx, 0);
                }
                std::vector<double> sums;
                std::vector<int> lims;
                rc = device_moments(sl, sums, lims);
                if (rc) return rc;
                for (long long j = 0; j < nb; j++) {
                    moments_from_sums(&sums[24 * j], sl[j], &mom[22 * j]);
                    moments_from_sums(&sums[24 * j + 11], sl[j], &mom[22 * j + 11]);
                }
                momValid = true;
            }
            for (long long j = 0; j < nb; j++) {
                DriftGeom g;
                std::memset(&g, 0, sizeof(g));
                if (modes[d] == 3) {
                    emesh_to_wfr(j);
                    double Rxe, Rze;
                    const float2* base = (const float2*)dF + 2 * j * nxy;
                    OneCtx oc;
                    oc.w = &ws[j]; oc.d = base; oc.nxy = nxy; oc.nx = nx;
                    oc.mx = momValid ? &mom[22 * j] : nullptr; oc.mz = momValid ? &mom[22 * j + 11] : nullptr;
                    rc = drift_radius(ws[j], subs[d], D->L, oc.mx, oc.mz, recompute, &oc, Rxe, Rze);
                    if (rc) return rc;
                    mode3_setup(Rxe, Rze, D->L, g);
                    if (g.kx < 0 || g.kz < 0) return SRWL_AMD_ERR_UNSUPPORTED;
                }
                drift_slice(sp[2 * j], modes[d], g, D->L, D->treat, pWfr0->mesh.eStart, nx, nz, em[j].xStart,
                            em[j].xStep, em[j].zStart, em[j].zStep, res[j].xc, res[j].zc);
                sp[2 * j + 1] = sp[2 * j];
                if (modes[d] == 3) {
                    em[j].xStart = sp[2 * j].nxStart; em[j].xStep = sp[2 * j].nxStep;
                    em[j].zStart = sp[2 * j].nzStart; em[j].zStep = sp[2 * j].nzStep;
                }
                emesh_to_wfr(j);
                if (momValid) {
                    drift_moments(ws[j], D->L, &mom[22 * j]);
                    drift_moments(ws[j], D->L, &mom[22 * j + 11]);
                }
                Rx[j] += D->L; Rz[j] += D->L;
                ws[j].Rx += D->L; ws[j].Ry += D->L;
            }
            HIPCHK(hipMemcpyAsync(dSp, sp.data(), sizeof(srwf_slice) * 2 * nb, hipMemcpyHostToDevice, st));
            bool last = (d == nDrift - 1);
            rc = drift_passes((float2*)dF, nxy, (int)(2 * nb), dSp, lx, lz, modes[d], last ? 2 : 1, dI, nxy, st);
            if (rc) return rc;
        }
        if (nDrift == 0 || !isDrift[nDrift - 1]) {
            // intensity of the emitted field itself, or after a final lens / aperture
            for (long long j = 0; j < nb; j++) {
                const float2* ex = nDrift == 0 ? (const float2*)(dF + 2 * off[j]) : (const float2*)dF + 2 * j * nxy;
                if (srwp_launch_intensity(ex, ex + nxy, 1, 0, 0, 0.f, 6, nxy, dI + j * nxy, st))
                    return SRWL_AMD_ERR_GPU;
            }
        }
        HIPCHK(hipMemcpyAsync(dEm, em.data(), sizeof(srwm_emesh) * nb, hipMemcpyHostToDevice, st));
        HIPCHK(hipMemcpyAsync(dW, w.data(), sizeof(float) * nb, hipMemcpyHostToDevice, st));
        if (useQ || diag) HIPCHK(hipMemsetAsync(dN2, 0, sizeof(double) * nb, st));
        if (srwm_launch_accum(dI, nxy, dEm, dW, (int)nb, rm, dS, dS + nr, (useQ || diag) ? dN2 : nullptr, dSrep,
                              (int)perRep, st))
            return SRWL_AMD_ERR_GPU;
        if (diag) {
            std::vector<double> n2(nb);
            HIPCHK(hipMemcpyAsync(n2.data(), dN2, sizeof(double) * nb, hipMemcpyDeviceToHost, st));
            HIPCHK(hipStreamSynchronize(st));
            for (long long j = 0; j < nb; j++) { sumN2 += n2[j]; sumN += std::sqrt(std::max(n2[j], 0.)); }
        }
        if (useQ) {
            // online training on this batch: target q* ~ p * ||I||
            std::vector<double> n2(nb);
            HIPCHK(hipMemcpyAsync(n2.data(), dN2, sizeof(double) * nb, hipMemcpyDeviceToHost, st));
            HIPCHK(hipStreamSynchronize(st));
            for (long long j = 0; j < nb; j++) {
                repG.insert(repG.end(), gsamp.begin() + 6 * j, gsamp.begin() + 6 * j + 6);
                repW.push_back(w[j]);
                repC.push_back(std::sqrt(std::max(n2[j], 0.)));
            }
            if ((long long)repW.size() > kReplay) {
                size_t drop = repW.size() - (size_t)kReplay;
                repG.erase(repG.begin(), repG.begin() + 6 * drop);
                repW.erase(repW.begin(), repW.begin() + drop);
                repC.erase(repC.begin(), repC.begin() + drop);
            }
            // weighted maximum likelihood toward q* ~ p ||I||: weight of a
            // sample drawn from q_j is p/q_j * ||I|| (normalized)
            int nr = (int)repW.size();
            std::vector<float> wh(nr);
            double tot = 0;
            for (int j = 0; j < nr; j++) tot += repW[j] * repC[j];
            for (int j = 0; j < nr; j++) wh[j] = tot > 0 ? (float)(repW[j] * repC[j] / tot) : 0.f;
            HIPCHK(hipMemcpyAsync(Q.gs, repG.data(), sizeof(float) * 6 * nr, hipMemcpyHostToDevice, st));
            HIPCHK(hipMemcpyAsync(Q.wh, wh.data(), sizeof(float) * nr, hipMemcpyHostToDevice, st));
            const float trainAlpha = 0.25f;
            for (int e = 0; e < trainSteps; e++) {
                if (srwl_launch_mlp_forward(Q.P, Q.in, Q.act, st)) return SRWL_AMD_ERR_GPU;
                HIPCHK(hipMemsetAsync(Q.gout, 0, sizeof(float) * (SRWL_OUT + 2), st));
                if (srwl_launch_gmm_grad(Q.act, Q.gs, Q.wh, nr, trainAlpha, Q.gout, Q.loss, st))
                    return SRWL_AMD_ERR_GPU;
                if (srwl_launch_mlp_backward(Q.P, Q.act, Q.gout, Q.G, st)) return SRWL_AMD_ERR_GPU;
                Q.step++;
                if (srwl_launch_adam(Q.P, Q.G, Q.m, Q.v, SRWL_NPARAM, 3e-3f, Q.step, st)) return SRWL_AMD_ERR_GPU;
            }
            float lossv = 0;
            HIPCHK(hipMemcpyAsync(&lossv, Q.loss, sizeof(float), hipMemcpyDeviceToHost, st));
            HIPCHK(hipStreamSynchronize(st));
            lastLoss = lossv;
        }
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
            if (rqmc) {
                // variance of the mean from the R replicate means (each of done / R points)
                Srh.resize((size_t)R * nr);
                HIPCHK(hipMemcpyAsync(Srh.data(), dSrep, sizeof(double) * R * nr, hipMemcpyDeviceToHost, st));
                HIPCHK(hipStreamSynchronize(st));
                double nPer = N / R, v = 0.;
                for (long long q = 0; q < nr; q++) {
                    double m = S1h[q] / N, d2 = 0.;
                    for (int k = 0; k < R; k++) { double d = Srh[(size_t)k * nr + q] / nPer - m; d2 += d * d; }
                    v += d2 / ((double)R * (R - 1));
                }
                errEst = a1 > 0 ? std::sqrt(v * N * N / a1) : -1;
            }
            lastA1 = a1;
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
        arOut[6] = lastLoss;
        arOut[7] = useQ ? Q.step : 0;
        if (rqmc) arOut[9] = R;
        if (useQ) {
            arOut[10] = (double)mixBatches;   // batches drawn from the mixture (the others from p)
            arOut[11] = lastRatio;            // last predicted variance ratio q vs p
        }
        if (diag) {
            // N Var = E||I||^2 - ||mu||^2 for p; (E||I||)^2 - ||mu||^2 for q* ~ p ||I||
            double N = (double)done, m2 = lastA1 / (N * N);
            double vp = sumN2 / N - m2, vq = (sumN / N) * (sumN / N) - m2;
            arOut[8] = (vq > 0) ? vp / vq : 0.;
        }
    }
    return 0;
}

}  // extern "C"

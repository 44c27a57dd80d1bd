// propag_api.cpp - host side of the drift propagation, the standalone SRW
// FFT and intensity extraction (C ABI in include/srwlib.h).
//
// Reference call semantics:
//   srwlPropagElecField      cpp/src/lib/srwlib.cpp:1023-1074, sroptcnt.cpp:87-425
//   drift buffer variables   cpp/src/core/sroptdrf.cpp:889-1116
//   srwlUtiFFT               cpp/src/lib/srwlib.cpp:1090-1204, gmfft.cpp:295-893
//   srwlCalcIntFromElecField cpp/src/lib/srwlib.cpp:757-856, srradmnp.cpp:35-1180
#include "srw_api_internal.h"
#include "srw_fft.h"
#include "srw_prop.h"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <functional>
#include <map>
#include <vector>

namespace srwamd {

static const double kTwoPi = 6.283185307179586;
static const double kPiSrw = 3.1415926536;   // sroptdrf.h:619, 812

static int log2_exact(long n)
{
    if (n <= 0 || (n & (n - 1))) return -1;
    int l = 0;
    while ((1L << l) < n) l++;
    return l;
}

// Twiddle tables tw[k] = exp(-i 2pi k / N), k < N/2, cached per size.
const float2* twiddles(int logn)
{
    static std::map<int, DBuf> cache;
    DBuf& b = cache[logn];
    if (b.p) return (const float2*)b.p;
    int n = 1 << logn;
    std::vector<float2> h(n / 2);
    for (int k = 0; k < n / 2; k++) {
        double a = -kTwoPi * k / n;
        h[k] = make_float2((float)std::cos(a), (float)std::sin(a));
    }
    void* p = b.get(sizeof(float2) * h.size());
    if (!p) return nullptr;
    if (hipMemcpy(p, h.data(), sizeof(float2) * h.size(), hipMemcpyHostToDevice) != hipSuccess) return nullptr;
    return (const float2*)p;
}

// Twiddle tables exp(-i 2pi m / n), m < n, of the mixed-radix passes, cached per size.
const float2* gen_twiddles(int n)
{
    static std::map<int, DBuf> cache;
    DBuf& b = cache[n];
    if (b.p) return (const float2*)b.p;
    std::vector<float2> h(n);
    for (int m = 0; m < n; m++) {
        double a = -kTwoPi * m / n;
        h[m] = make_float2((float)std::cos(a), (float)std::sin(a));
    }
    void* p = b.get(sizeof(float2) * h.size());
    if (!p) return nullptr;
    if (hipMemcpy(p, h.data(), sizeof(float2) * h.size(), hipMemcpyHostToDevice) != hipSuccess) return nullptr;
    return (const float2*)p;
}

int fft_ready()
{
    static bool done = false;
    int rc = ensure_device();
    if (rc) return rc;
    if (!done) {
        if (srwf_init_attributes() || srwf_gen_init_attributes()) return SRWL_AMD_ERR_GPU;
        done = true;
    }
    return 0;
}

// Mixed-radix drift passes over nSlices nx x nz slices in place (sizes that
// are not both powers of two); tmp holds nSlices transposed slices.
int drift_passes_gen(float2* d, float2* tmp, long long sliceStride, int nSlices, const srwf_slice* dSp, int nx, int nz,
                     int mode, int postOp, float* inten, long long intenStride, hipStream_t st)
{
    srwf_gplan px, pz;
    if (srwf_gplan_make(nx, &px) || srwf_gplan_make(nz, &pz)) return SRWL_ERR_IN_FFT;
    const float2* twx = gen_twiddles(nx);
    const float2* twz = gen_twiddles(nz);
    if (!twx || !twz) return SRWL_AMD_ERR_GPU;
    int pop = postOp;
    if (pop == 1 && mode != 3) pop = 0;
    const int pre = mode == 3 ? 1 : 0, col = mode == 3 ? 1 : 0;
    if (srwf_launch_grows(d, sliceStride, nSlices, dSp, twx, px, nz, 0, pre, 0, 0, nullptr, 0, st)) return SRWL_AMD_ERR_GPU;
    if (srwf_launch_gtranspose(d, tmp, nz, nx, sliceStride, nSlices, st)) return SRWL_AMD_ERR_GPU;
    if (srwf_launch_gcols(tmp, sliceStride, nSlices, dSp, twz, pz, nx, col, 0, nullptr, st)) return SRWL_AMD_ERR_GPU;
    if (srwf_launch_gtranspose(tmp, d, nx, nz, sliceStride, nSlices, st)) return SRWL_AMD_ERR_GPU;
    if (srwf_launch_grows(d, sliceStride, nSlices, dSp, twx, px, nz, 1, 0, 0, pop, inten, intenStride, st))
        return SRWL_AMD_ERR_GPU;
    return 0;
}

// Sizes of the power-of-2 passes (16..2048 both ways); otherwise the
// mixed-radix passes take even sizes 8..2048 with factors 2, 3, 5, 7, 11, 13.
static bool pow2_sizes(long nx, long nz, int& lx, int& lz)
{
    lx = log2_exact(nx); lz = log2_exact(nz);
    return lx >= 4 && lz >= 4 && lx <= 11 && lz <= 11;
}

static bool gen_sizes(long nx, long nz)
{
    srwf_gplan a;
    return srwf_gplan_make((int)nx, &a) == 0 && srwf_gplan_make((int)nz, &a) == 0;
}

// Drift element parameters (sroptcnt.cpp:183-211: arProp layout)
// mode: 0 angular representation, 3 analytic quadratic phase (sroptdrf.h:223-283);
// sub: the radius estimate of the analytic mode (arProp[3] = 1 or 2).
int drift_mode_from_prop(const double* pp, int& mode, int& sub)
{
    mode = 0; sub = 0;
    if (pp == nullptr) return 0;
    if (pp[0] != 0 || pp[1] != 0) return SRWL_AMD_ERR_UNSUPPORTED;    // automatic resizing
    int at = (int)(char)pp[3];
    if (at == 0) mode = 0;
    else if (at == 1 || at == 2) { mode = 3; sub = at; }
    else return SRWL_AMD_ERR_UNSUPPORTED;   // 3-5: waist modes
    return 0;
}

// sroptdrf.cpp:889-1081 for the radii Rx, Rz chosen by drift_radius().
void mode3_setup(double Rx, double Rz, double L, DriftGeom& g)
{
    const double inf = 1E+23;
    g.invRx = 0; g.invRz = 0; g.Lx = 0; g.Lz = 0; g.invRxL = 1. / L; g.invRzL = 1. / L;
    g.kx = inf; g.kxc = inf; g.kz = inf; g.kzc = inf;
    if (Rx != 0) {
        g.invRx = 1. / Rx; g.kx = (Rx + L) * g.invRx; g.kxc = L / Rx;
        if (-L != Rx) { g.Lx = L / (1. + L * g.invRx); g.invRxL = 1. / (L + Rx); } else { g.Lx = inf; g.invRxL = inf; }
    }
    if (Rz != 0) {
        g.invRz = 1. / Rz; g.kz = (Rz + L) * g.invRz; g.kzc = L / Rz;
        if (-L != Rz) { g.Lz = L / (1. + L * g.invRz); g.invRzL = 1. / (L + Rz); } else { g.Lz = inf; g.invRzL = inf; }
    }
    g.mult = std::sqrt(std::fabs(g.Lx * g.Lz)) / L;
    g.phT = 0.25 * 3.141592653589793 * (2. - (g.Lx < 0 ? -1. : 1.) - (g.Lz < 0 ? -1. : 1.));
}

// Slice parameters of one photon energy for a drift of length L.
void drift_slice(srwf_slice& S, int mode, const DriftGeom& g, double L, int treat, double e, long nx, long nz,
                 double xStart, double xStep, double zStart, double zStep, double xc, double zc)
{
    std::memset(&S, 0, sizeof(S));
    double lam = 1.239842e-06 / e;
    S.xStart = xStart; S.xStep = xStep; S.zStart = zStart; S.zStep = zStep;
    S.qStepX = 1. / (nx * xStep); S.qStepZ = 1. / (nz * zStep);
    S.xc = xc; S.zc = zc;
    S.ny = (int)nz;
    double norm = 1. / ((double)nx * (double)nz);   // dx dz (forward) * dqx dqz (inverse)
    if (mode == 0) {
        S.c0_rev = -kPiSrw * L * lam / kTwoPi;
        S.c1 = 0.25 * lam * lam;
        S.path_rev = (treat == 1) ? (5.067730652e+06) * L * e / kTwoPi : 0.;
        S.midScale = (float)norm;
    } else {
        S.invRx = g.invRx; S.invRz = g.invRz; S.invRxL = g.invRxL; S.invRzL = g.invRzL;
        S.pre_rev = -(kPiSrw / lam) / kTwoPi;
        S.post_rev = (kPiSrw / lam) / kTwoPi;
        S.path3_rev = (treat == 1) ? 2. * (kPiSrw / lam) * L / kTwoPi : 0.;
        S.c0x_rev = -kPiSrw * lam * g.Lx / kTwoPi;
        S.c0z_rev = -kPiSrw * lam * g.Lz / kTwoPi;
        S.phT_rev = g.phT / kTwoPi;
        S.midScale = (float)(norm * g.mult);
        S.nxStart = g.kx * xStart - g.kxc * xc; S.nxStep = xStep * g.kx;
        S.nzStart = g.kz * zStart - g.kzc * zc; S.nzStep = zStep * g.kz;
    }
}

// Autotuner of the FFT tile shapes.  For each (nx, nz, pass operations) the
// first call times every shape that fits, pass by pass, on scratch slices of
// the same size (at most 16, with the caller's slice parameters) and keeps the
// fastest; later calls reuse it.  Shapes change only the work assignment, so
// the values computed do not depend on the choice.
struct TileKey {
    int lx, lz, preOp, colOp, inv, postOp;
    bool operator<(const TileKey& o) const
    {
        int a[6] = {lx, lz, preOp, colOp, inv, postOp}, b[6] = {o.lx, o.lz, o.preOp, o.colOp, o.inv, o.postOp};
        for (int i = 0; i < 6; i++) if (a[i] != b[i]) return a[i] < b[i];
        return false;
    }
};
static std::map<TileKey, srwf_tiles> g_tiles;

static int tune_time(hipStream_t st, hipEvent_t e0, hipEvent_t e1, float& ms, const std::function<int()>& run)
{
    if (run()) return -1;                       // first launch: code load, caches
    if (hipEventRecord(e0, st) != hipSuccess) return -1;
    for (int k = 0; k < 3; k++) if (run()) return -1;
    if (hipEventRecord(e1, st) != hipSuccess) return -1;
    if (hipEventSynchronize(e1) != hipSuccess) return -1;
    if (hipEventElapsedTime(&ms, e0, e1) != hipSuccess) return -1;
    return 0;
}

const srwf_tiles* fft_tiles(int lognx, int lognz, int preOp, int colOp, int inv, int postOp, const srwf_slice* dSp,
                            int nSlices, hipStream_t st)
{
    if (g_ctx.tilesForced) return &g_ctx.forcedTiles;
    if (!g_ctx.tuneTiles) return nullptr;
    TileKey key = {lognx, lognz, preOp, colOp, inv, postOp};
    auto it = g_tiles.find(key);
    if (it != g_tiles.end()) return &it->second;
    srwf_tiles best;
    srwf_default_tiles(lognx, lognz, postOp, &best);
    const long long nxy = 1LL << (lognx + lognz);
    int S = nSlices < 16 ? nSlices : 16;
    if (postOp == 2) S &= ~1;
    const float2* twx = twiddles(lognx);
    const float2* twz = twiddles(lognz);
    // S slices, S output slices of a standalone transform, S/2 intensity slices
    const size_t nScratch = (size_t)3 * S * nxy;
    float2* d = (float2*)g_ctx.buf.Tune.get(sizeof(float2) * nScratch);
    hipEvent_t e0 = nullptr, e1 = nullptr;
    bool ok = S > 0 && d && twx && twz && hipEventCreate(&e0) == hipSuccess && hipEventCreate(&e1) == hipSuccess &&
              hipMemsetAsync(d, 0, sizeof(float2) * nScratch, st) == hipSuccess;
    float2* out = d + (size_t)S * nxy;
    float* inten = (float*)(d + (size_t)2 * S * nxy);
    const int nz = 1 << lognz;
    const int thrs[3] = {256, 512, 1024};
    auto tryPass = [&](int pass) {
        float bestMs = 1e30f;
        srwf_tiles keep = best;
        for (int a = 1; a <= (pass == 1 ? 64 : 32); a <<= 1)
            for (int ti = 0; ti < 3 * (pass == 1 ? 2 : 1); ti++) {
                srwf_tiles c = best;
                if (pass == 0) { c.rpwF = a; c.thrF = thrs[ti]; }
                else if (pass == 1) { c.cpw = a; c.thrC = thrs[ti % 3]; c.pairC = ti / 3; }
                else { c.rpwI = a; c.thrI = thrs[ti]; }
                if (!srwf_tiles_fit(lognx, lognz, postOp, &c)) continue;
                float ms = 0;
                int rc = tune_time(st, e0, e1, ms, [&]() -> int {
                    if (pass == 0) return srwf_launch_rows_fwd(d, nxy, S, dSp, twx, lognx, nz, preOp, inv, &c, st);
                    if (pass == 1) return srwf_launch_cols(d, nxy, S, dSp, twz, lognz, lognx, colOp, inv, out, &c, st);
                    return srwf_launch_rows_inv(d, nxy, S, dSp, twx, lognx, nz, postOp, inten, nxy, &c, st);
                });
                if (rc) { ok = false; return; }
                if (ms < bestMs) { bestMs = ms; keep = c; }
            }
        best = keep;
    };
    if (ok) tryPass(0);
    if (ok) tryPass(1);
    if (ok && postOp >= 0) tryPass(2);
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    if (!ok) {
        (void)hipGetLastError();
        srwf_default_tiles(lognx, lognz, postOp, &best);
    }
    return &(g_tiles[key] = best);
}

// Runs passes A-C over nSlices N x N slices in place.
int drift_passes(float2* d, long long sliceStride, int nSlices, const srwf_slice* dSp, int lognx, int lognz,
                 int mode, int postOp, float* inten, long long intenStride, hipStream_t st)
{
    const float2* twx = twiddles(lognx);
    const float2* twz = twiddles(lognz);
    if (!twx || !twz) return SRWL_AMD_ERR_GPU;
    int nz = 1 << lognz;
    int pop = postOp;
    if (pop == 1 && mode != 3) pop = 0;
    const int pre = mode == 3 ? 1 : 0, col = mode == 3 ? 1 : 0;
    const srwf_tiles* T = fft_tiles(lognx, lognz, pre, col, 0, pop, dSp, nSlices, st);
    if (srwf_launch_rows_fwd(d, sliceStride, nSlices, dSp, twx, lognx, nz, pre, 0, T, st))
        return SRWL_AMD_ERR_GPU;
    if (srwf_launch_cols(d, sliceStride, nSlices, dSp, twz, lognz, lognx, col, 0, nullptr, T, st))
        return SRWL_AMD_ERR_GPU;
    if (srwf_launch_rows_inv(d, sliceStride, nSlices, dSp, twx, lognx, nz, pop, inten, intenStride, T, st))
        return SRWL_AMD_ERR_GPU;
    return 0;
}

// srradstr.cpp:2610-2674: whether the reference would subtract a linear phase
// term before a lens centred at (newXc, newZc); that case is not supported.
bool lin_phase_would_subtract(const SRWLWfr& w, double newXc, double newZc)
{
    const double lambda = 3.1415926535898 / (w.mesh.eStart * 2.53384080189E+06);
    const double xr = w.mesh.xFin - w.mesh.xStart, zr = w.mesh.yFin - w.mesh.yStart;
    bool sx = false, sz = false;
    if (w.Rx != 0)
        sx = std::fabs(w.Rx) > std::fabs(w.dRx) && 0.25 * xr * xr / (lambda * w.Rx) > 10 && std::fabs(newXc - w.xc) / xr > 0.2;
    if (w.Ry != 0)
        sz = std::fabs(w.Ry) > std::fabs(w.dRy) && 0.25 * zr * zr / (lambda * w.Ry) > 10 && std::fabs(newZc - w.yc) / zr > 0.2;
    return sx || sz;
}

}  // namespace srwamd

using namespace srwamd;

#define HIPCHK(x) do { if ((x) != hipSuccess) return SRWL_AMD_ERR_GPU; } while (0)

extern "C" {

// sroptelm.cpp:2094-2180 with the lens matrices ax = [[1, 0], [-1/Fx, 1]],
// az = [[1, 0], [-1/Fy, 1]] (sroptfoc.h:30-42), for every photon energy.
static void lens_moments(long ne, double Fx, double Fz, double* mom)
{
    const double cx = -1. / Fx, cz = -1. / Fz;
    for (long ie = 0; ie < ne; ie++) {
        double* M = mom + 11 * ie;
        const double X = M[1], XP = M[2], Z = M[3], ZP = M[4];
        const double XX = M[5], XXP = M[6], XPXP = M[7], ZZ = M[8], ZZP = M[9], ZPZP = M[10];
        M[2] = cx * X + XP;
        M[4] = cz * Z + ZP;
        M[6] = cx * XX + XXP;
        M[7] = cx * cx * XX + 2. * cx * XXP + XPXP;
        M[9] = cz * ZZ + ZZP;
        M[10] = cz * cz * ZZ + 2. * cz * ZZP + ZPZP;
    }
}

// sroptfoc.h:100-108 and sroptelm.cpp:5594-5640: M <- F M, V <- F V + f.
static void lens_4x4(double* P, const SRWLOptL& L)
{
    const double F[16] = {1., 0., 0., 0., -1. / L.Fx, 1., 0., 0., 0., 0., 1., 0., 0., 0., -1. / L.Fy, 1.};
    const double f[4] = {0., L.x / L.Fx, 0., L.y / L.Fy};
    double R[16], V[4];
    for (int i = 0; i < 4; i++) {
        V[i] = 0.;
        for (int j = 0; j < 4; j++) {
            double r = 0.;
            for (int k = 0; k < 4; k++) r += F[4 * i + k] * P[4 * k + j];
            R[4 * i + j] = r;
            V[i] += F[4 * i + j] * P[16 + j];
        }
    }
    std::memcpy(P, R, sizeof(R));
    for (int i = 0; i < 4; i++) P[16 + i] = V[i] + f[i];
}

// srstraux.h:331-384 (centre and rms size with the 0.1 floor) and the
// aperture's test (sroptapt.h:204-212, 383-394): true when the 1.7-sigma
// extent of either component's moments leaves the opening.
static bool aperture_needs_moments(const double* mx, const double* mz, bool circ, double Dx, double Dz, double cx,
                                   double cy)
{
    const double hx = 0.5 * Dx, hz = circ ? 0.5 * Dx : 0.5 * Dz;
    const double* Ms[2] = {mx, mz};
    for (int k = 0; k < 2; k++) {
        const double* M = Ms[k];
        double mxx = M[5] - M[1] * M[1], mzz = M[8] - M[3] * M[3];
        if (mxx < 0.1 * M[5]) mxx = 0.1 * M[5];
        if (mzz < 0.1 * M[8]) mzz = 0.1 * M[8];
        const double sx = 1.7 * std::sqrt(std::fabs(mxx)), sz = 1.7 * std::sqrt(std::fabs(mzz));
        if (M[1] - sx < cx - hx || M[1] + sx > cx + hx || M[3] - sz < cy - hz || M[3] + sz > cy + hz) return true;
    }
    return false;
}

// Resize instructions of one arProp entry (sroptcnt.cpp:183-211: [5..8] =
// pxm, pxd, pzm, pzd, [9..11] centring); par for resize_plan().
static bool prop_resize_par(const double* pp, int npp, double* par)
{
    if (pp == nullptr) return false;
    par[0] = (npp > 4) ? pp[4] : 0.;
    for (int k = 1; k <= 4; k++) par[k] = pp[4 + k];
    par[5] = par[6] = 0.5;
    if (npp > 9 && pp[9] != 0) par[5] = -1.;   // centre shifts are not supported
    return !(par[1] == 1. && par[2] == 1. && par[3] == 1. && par[4] == 1.);
}

SRWL_EXPORT int srwlPropagElecField(SRWLWfr* pWfr, SRWLOptC* pOpt, int nInt, char** arID, SRWLRadMesh* arIM,
                                    char** arI, double* arParGPU)
{
    (void)arID; (void)arIM; (void)arI; (void)arParGPU;
    if (pWfr == nullptr || pOpt == nullptr) return SRWL_INCORRECT_PARAM_FOR_WFR_PROP;
    if (nInt != 0) return SRWL_AMD_ERR_UNSUPPORTED;
    if (pWfr->presCA != 0 || pWfr->presFT != 0) return SRWL_AMD_ERR_UNSUPPORTED;
    if (pWfr->arEx == nullptr || pWfr->arEy == nullptr) return SRWL_INCORRECT_WFR_STRUCT;
    long nx = pWfr->mesh.nx, nz = pWfr->mesh.ny, ne = pWfr->mesh.ne;
    const long nx0 = nx, nz0 = nz;
    if (nx < 2 || nz < 2 || ne < 1) return SRWL_INCORRECT_WFR_STRUCT;
    auto propN = [&](int i) -> int { return (pOpt->arPropN != nullptr) ? (int)(unsigned char)pOpt->arPropN[i] : 9; };
    for (int i = 0; i < pOpt->nElem; i++) {
        const char* ty = pOpt->arOptTypes[i];
        if (std::strcmp(ty, "aperture") == 0 || std::strcmp(ty, "obstacle") == 0) {
            const SRWLOptA* A = (const SRWLOptA*)pOpt->arOpt[i];
            const char ao = A->ap_or_ob;
            if ((A->shape != 'r' && A->shape != 'c') || (ao != 'a' && ao != 'A' && ao != 'o' && ao != 'O'))
                return SRWL_AMD_ERR_UNSUPPORTED;
        } else if (std::strcmp(ty, "lens") != 0 && std::strcmp(ty, "drift") != 0) {
            return SRWL_AMD_ERR_UNSUPPORTED;
        }
        int mode, sub;
        const double* pp = (pOpt->arProp && i < pOpt->nProp) ? pOpt->arProp[i] : nullptr;
        int rc = drift_mode_from_prop(pp, mode, sub);
        if (rc) return rc;
    }
    int rc = fft_ready();
    if (rc) return rc;
    hipStream_t st = g_ctx.stream;
    long long nxy = (long long)nx * nz;
    long long n = nxy * ne;
    float2* dIn = (float2*)g_ctx.buf.W0.get(sizeof(float2) * 2 * n);
    DBuf* bufCur = &g_ctx.buf.W1;
    DBuf* bufSpare = &g_ctx.buf.RsN;
    float2* dS = (float2*)bufCur->get(sizeof(float2) * 2 * n);
    srwf_slice* dSp = (srwf_slice*)g_ctx.buf.W2.get(sizeof(srwf_slice) * 2 * ne);
    if (!dIn || !dS || !dSp) return SRWL_ERR_MEMORY_ALLOCATION_FAILURE;
    HIPCHK(hipMemcpyAsync(dIn, pWfr->arEx, sizeof(float2) * n, hipMemcpyHostToDevice, st));
    HIPCHK(hipMemcpyAsync(dIn + n, pWfr->arEy, sizeof(float2) * n, hipMemcpyHostToDevice, st));
    // slice-major layout: slice (2*ie + comp) holds [iy][ix]
    if (ne == 1) {
        for (int c = 0; c < 2; c++)
            HIPCHK(hipMemcpyAsync(dS + c * nxy, dIn + c * n, sizeof(float2) * nxy, hipMemcpyDeviceToDevice, st));
    } else {
        if (srwf_launch_slices(dIn, dIn + n, dS, (int)ne, nxy, 1, st)) return SRWL_AMD_ERR_GPU;
    }
    if (g_ctx.timing) HIPCHK(hipEventRecord(g_ctx.ev2, st));
    double xStep = (pWfr->mesh.xFin - pWfr->mesh.xStart) / (nx - 1);
    double zStep = (pWfr->mesh.yFin - pWfr->mesh.yStart) / (nz - 1);
    double eStep = ne > 1 ? (pWfr->mesh.eFin - pWfr->mesh.eStart) / (ne - 1) : 0;
    std::vector<srwf_slice> sp(2 * ne);
    // moments: the caller's arrays, or scratch ones when the structure has none
    std::vector<double> tmpMom;
    double* momX = pWfr->arMomX;
    double* momZ = pWfr->arMomY;
    if (!momX || !momZ) {
        tmpMom.assign(22 * (size_t)ne, 0.);
        if (!momX) momX = tmpMom.data();
        if (!momZ) momZ = tmpMom.data() + 11 * ne;
    }
    struct MomCtx { SRWLWfr* w; const float2* d; long long nxy, nx; double *mx, *mz; } mctx;
    mctx.w = pWfr; mctx.mx = momX; mctx.mz = momZ;
    auto recompute = [](void* c) -> int {
        MomCtx* m = (MomCtx*)c;   // slice-major device field: slice 2*ie + component, [iz][ix]
        return wfr_moments(*m->w, m->d, m->d + m->nxy, 2 * m->nxy, 1, m->nx, m->mx, m->mz);
    };
    // resampling of the device field (sroptcnt.cpp:314-327 before an element, 408-423 after the last)
    auto resize_device = [&](const double* pp, int npp) -> int {
        double par[7];
        if (!prop_resize_par(pp, npp, par)) return 0;
        ResizePlan P;
        int r = resize_plan(*pWfr, par, 7, P);
        if (r == -1) return 0;
        if (r) return r;
        long long nxyN = (long long)P.x.n * P.z.n;
        float2* dN = (float2*)bufSpare->get(sizeof(float2) * 2 * nxyN * ne);
        if (!dN) return SRWL_ERR_MEMORY_ALLOCATION_FAILURE;
        srwr_layout lo = {1, nx, 2 * nxy}, ln = {1, P.x.n, 2 * nxyN};
        if (g_ctx.timing) HIPCHK(hipEventRecord(g_ctx.ev0, st));
        r = device_resize(*pWfr, P, dS, dS + nxy, lo, dN, dN + nxyN, ln, st);
        if (r) return r;
        if (g_ctx.timing) {
            HIPCHK(hipEventRecord(g_ctx.ev1, st));
            HIPCHK(hipEventSynchronize(g_ctx.ev1));
            float ms = 0;
            (void)hipEventElapsedTime(&ms, g_ctx.ev0, g_ctx.ev1);
            g_ctx.lastRsMs = ms;
            // phase removal (read + write old), resampling (read old, write new), phase (read + write new)
            g_ctx.lastRsBytes = 8. * 2. * ne * (2. * nxy + (double)nxy + (double)nxyN + 2. * (double)nxyN);
        }
        std::swap(bufCur, bufSpare);
        dS = dN;
        nx = P.x.n; nz = P.z.n; nxy = nxyN; n = nxy * ne;
        pWfr->mesh.nx = nx; pWfr->mesh.ny = nz;
        pWfr->mesh.xStart = P.x.start; pWfr->mesh.xFin = P.x.start + (nx - 1) * P.x.step;
        pWfr->mesh.yStart = P.z.start; pWfr->mesh.yFin = P.z.start + (nz - 1) * P.z.step;
        xStep = P.x.step; zStep = P.z.step;
        return 0;
    };
    // Thin lens, aperture or obstacle (sroptfoc.h:111-117, 182-195; sroptapt.h:53-62;
    // sroptel2.cpp:41-185): field, moments, wavefront radius, 4x4 matrix.
    auto point_element = [&](const char* ty, const void* el) -> int {
        srwo_elem E{};
        E.nx = nx; E.xStart = pWfr->mesh.xStart; E.xStep = xStep; E.zStart = pWfr->mesh.yStart; E.zStep = zStep;
        E.eStart = pWfr->mesh.eStart; E.eStep = eStep;
        const bool lens = std::strcmp(ty, "lens") == 0;
        const SRWLOptL* L = (const SRWLOptL*)el;
        const SRWLOptA* A = (const SRWLOptA*)el;
        bool circ = false, ob = false;
        if (lens) {
            if (lin_phase_would_subtract(*pWfr, L->x, L->y)) return SRWL_AMD_ERR_UNSUPPORTED;
            E.kind = 0; E.p0 = L->Fx; E.p1 = L->Fy; E.cx = L->x; E.cy = L->y;
        } else {
            circ = A->shape == 'c';
            ob = A->ap_or_ob == 'o' || A->ap_or_ob == 'O';
            E.kind = circ ? (ob ? 4 : 3) : (ob ? 2 : 1);
            E.p0 = A->Dx; E.p1 = A->Dy; E.cx = A->x; E.cy = A->y;
        }
        if (srwo_launch_point_elem(dS, nxy, (int)ne, E, st)) return SRWL_AMD_ERR_GPU;
        if (lens) {
            lens_moments(ne, L->Fx, L->Fy, momX);
            lens_moments(ne, L->Fx, L->Fy, momZ);
            const double difX = (L->Fx == pWfr->Rx) ? 1E-23 : L->Fx - pWfr->Rx;
            const double difZ = (L->Fy == pWfr->Ry) ? 1E-23 : L->Fy - pWfr->Ry;
            const double mX = L->Fx / difX, mZ = L->Fy / difZ;
            pWfr->Rx *= mX; pWfr->Ry *= mZ;
            pWfr->dRx *= mX * mX; pWfr->dRy *= mZ * mZ;
            pWfr->xc = L->x - (L->x - pWfr->xc) * mX;
            pWfr->yc = L->y - (L->y - pWfr->yc) * mZ;
            if (pWfr->arElecPropMatr) lens_4x4(pWfr->arElecPropMatr, *L);
        } else if (pWfr->arMomX && pWfr->arMomY) {
            // sroptapt.cpp:248-300: moments recomputed from the field for the
            // energies whose 1.7-sigma extent leaves the opening (always for obstacles)
            std::vector<char> need(ne);
            bool any = false;
            for (long ie = 0; ie < ne; ie++) {
                need[ie] = ob || aperture_needs_moments(momX + 11 * ie, momZ + 11 * ie, circ, A->Dx, A->Dy, A->x, A->y);
                any = any || need[ie];
            }
            if (any) {
                std::vector<double> nm(22 * (size_t)ne);
                if (wfr_moments(*pWfr, dS, dS + nxy, 2 * nxy, 1, nx, nm.data(), nm.data() + 11 * ne))
                    return SRWL_AMD_ERR_GPU;
                for (long ie = 0; ie < ne; ie++)
                    if (need[ie]) {
                        std::memcpy(momX + 11 * ie, nm.data() + 11 * ie, 11 * sizeof(double));
                        std::memcpy(momZ + 11 * ie, nm.data() + 11 * (ne + ie), 11 * sizeof(double));
                    }
            }
        }
        // sroptel2.cpp:147-148
        if (pWfr->Rx != 0 && pWfr->dRx == 0) pWfr->dRx = std::fabs(0.1 * pWfr->Rx);
        if (pWfr->Ry != 0 && pWfr->dRy == 0) pWfr->dRy = std::fabs(0.1 * pWfr->Ry);
        return 0;
    };
    for (int i = 0; i < pOpt->nElem; i++) {
        const SRWLOptD* d = (const SRWLOptD*)pOpt->arOpt[i];
        const double* pp = (pOpt->arProp && i < pOpt->nProp) ? pOpt->arProp[i] : nullptr;
        rc = resize_device(pp, pp ? propN(i) : 0);
        if (rc) return rc;
        if (std::strcmp(pOpt->arOptTypes[i], "drift") != 0) {
            rc = point_element(pOpt->arOptTypes[i], pOpt->arOpt[i]);
            if (rc) return rc;
            continue;
        }
        int lx, lz;
        const bool p2 = pow2_sizes(nx, nz, lx, lz);
        if (!p2 && !gen_sizes(nx, nz)) return SRWL_ERR_IN_FFT;
        float2* dT = nullptr;
        if (!p2) {
            dT = (float2*)g_ctx.buf.GenT.get(sizeof(float2) * 2 * (size_t)n);
            if (!dT) return SRWL_ERR_MEMORY_ALLOCATION_FAILURE;
        }
        mctx.d = dS; mctx.nxy = nxy; mctx.nx = nx;
        // a wavefront without moment arrays starts every drift from empty
        // scratch moments, so that the radius estimate recomputes them from
        // the current field (e.g. after an aperture clipped it)
        if (!tmpMom.empty()) std::fill(tmpMom.begin(), tmpMom.end(), 0.);
        int mode, sub;
        drift_mode_from_prop(pp, mode, sub);
        DriftGeom g;
        std::memset(&g, 0, sizeof(g));
        if (mode == 3) {
            double Rx, Rz;
            rc = drift_radius(*pWfr, sub, d->L, momX, momZ, recompute, &mctx, Rx, Rz);
            if (rc) return rc;
            mode3_setup(Rx, Rz, d->L, g);
            if (g.kx < 0 || g.kz < 0) return SRWL_AMD_ERR_UNSUPPORTED;   // mirrored output mesh
        }
        for (long ie = 0; ie < ne; ie++) {
            double e = pWfr->mesh.eStart + ie * eStep;
            drift_slice(sp[2 * ie], mode, g, d->L, d->treat, e, nx, nz, pWfr->mesh.xStart, xStep, pWfr->mesh.yStart,
                        zStep, pWfr->xc, pWfr->yc);
            sp[2 * ie + 1] = sp[2 * ie];
        }
        HIPCHK(hipMemcpyAsync(dSp, sp.data(), sizeof(srwf_slice) * 2 * ne, hipMemcpyHostToDevice, st));
        // twiddle tables are built and uploaded on first use of a size: outside the timed passes
        if (p2 && (!twiddles(lx) || !twiddles(lz))) return SRWL_AMD_ERR_GPU;
        if (!p2 && (!gen_twiddles((int)nx) || !gen_twiddles((int)nz))) return SRWL_AMD_ERR_GPU;
        if (g_ctx.timing) HIPCHK(hipEventRecord(g_ctx.ev0, st));
        rc = p2 ? drift_passes(dS, nxy, (int)(2 * ne), dSp, lx, lz, mode, 1, nullptr, 0, st)
                : drift_passes_gen(dS, dT, nxy, (int)(2 * ne), dSp, (int)nx, (int)nz, mode, 1, nullptr, 0, st);
        if (rc) return rc;
        if (g_ctx.timing) {
            HIPCHK(hipEventRecord(g_ctx.ev1, st));
            HIPCHK(hipEventSynchronize(g_ctx.ev1));
            float ms = 0;
            (void)hipEventElapsedTime(&ms, g_ctx.ev0, g_ctx.ev1);
            g_ctx.lastFftMs = ms;
            g_ctx.lastFftBytes = 3. * 2. * 8. * (double)n * 2.;
            g_ctx.sumFftMs += ms; g_ctx.sumFftBytes += g_ctx.lastFftBytes; g_ctx.nFftCalls += 1;
        }
        if (mode == 3) {
            pWfr->mesh.xStart = sp[0].nxStart;
            pWfr->mesh.xFin = sp[0].nxStart + (nx - 1) * sp[0].nxStep;
            pWfr->mesh.yStart = sp[0].nzStart;
            pWfr->mesh.yFin = sp[0].nzStart + (nz - 1) * sp[0].nzStep;
            xStep = sp[0].nxStep; zStep = sp[0].nzStep;
        }
        // sroptdrf.h:181-194: moments, radius and 4x4 matrix follow the field
        drift_moments(*pWfr, d->L, momX);
        drift_moments(*pWfr, d->L, momZ);
        pWfr->Rx += d->L;
        pWfr->Ry += d->L;
        if (pWfr->arElecPropMatr) drift_4x4(pWfr->arElecPropMatr, d->L);
    }
    if (pOpt->arProp && pOpt->nProp > pOpt->nElem) {
        rc = resize_device(pOpt->arProp[pOpt->nElem], propN(pOpt->nElem));
        if (rc) return rc;
    }
    if (ne > 1) {
        dIn = (float2*)g_ctx.buf.W0.get(sizeof(float2) * 2 * n);
        if (!dIn) return SRWL_ERR_MEMORY_ALLOCATION_FAILURE;
        if (srwf_launch_slices(dIn, dIn + n, dS, (int)ne, nxy, 0, st)) return SRWL_AMD_ERR_GPU;
    }
    if (g_ctx.timing) HIPCHK(hipEventRecord(g_ctx.ev3, st));
    if (nx != nx0 || nz != nz0) {
        // the caller's arrays are reallocated for the new mesh (srradstr.cpp:760-790)
        if (g_wfrModif == nullptr) return SRWL_WFR_EXT_FUNC_NOT_DEFINED;
        if ((*g_wfrModif)(2, pWfr, 0)) return SRWL_WFR_EXT_MODIF_FAILED;
        if (pWfr->arEx == nullptr || pWfr->arEy == nullptr) return SRWL_WFR_EXT_MODIF_FAILED;
    }
    if (ne == 1) {
        HIPCHK(hipMemcpyAsync(pWfr->arEx, dS, sizeof(float2) * nxy, hipMemcpyDeviceToHost, st));
        HIPCHK(hipMemcpyAsync(pWfr->arEy, dS + nxy, sizeof(float2) * nxy, hipMemcpyDeviceToHost, st));
    } else {
        HIPCHK(hipMemcpyAsync(pWfr->arEx, dIn, sizeof(float2) * n, hipMemcpyDeviceToHost, st));
        HIPCHK(hipMemcpyAsync(pWfr->arEy, dIn + n, sizeof(float2) * n, hipMemcpyDeviceToHost, st));
    }
    HIPCHK(hipStreamSynchronize(st));
    if (g_ctx.timing) {
        float ms = 0;
        (void)hipEventElapsedTime(&ms, g_ctx.ev2, g_ctx.ev3);
        g_ctx.lastPropDevMs = ms;
    }
    return 0;
}

SRWL_EXPORT int srwlUtiFFT(char* pcData, char typeData, double* arMesh, int nMesh, int dir, double* arParGPU)
{
    (void)arParGPU;
    if (pcData == nullptr || arMesh == nullptr || nMesh < 6) return SRWL_INCORRECT_PARAM_FOR_FFT;
    if (typeData != 'f') return SRWL_AMD_ERR_UNSUPPORTED;
    long nx = (long)arMesh[2], nz = (long)arMesh[5];
    if (nx < 2 || nz < 2 || (nx & 1) || (nz & 1)) return SRWL_INCORRECT_PARAM_FOR_FFT;
    int lx, lz;
    const bool p2 = pow2_sizes(nx, nz, lx, lz);
    if (!p2 && !gen_sizes(nx, nz)) return SRWL_AMD_ERR_UNSUPPORTED;
    int rc = fft_ready();
    if (rc) return rc;
    hipStream_t st = g_ctx.stream;
    long long nxy = (long long)nx * nz;
    float2* d = (float2*)g_ctx.buf.W0.get(sizeof(float2) * nxy);
    float2* o = (float2*)g_ctx.buf.W1.get(sizeof(float2) * nxy);
    srwf_slice* dSp = (srwf_slice*)g_ctx.buf.W2.get(sizeof(srwf_slice));
    if (!d || !o || !dSp) return SRWL_ERR_MEMORY_ALLOCATION_FAILURE;
    // gmfft.h:232-248 SetupLimitsTr and gmfft.cpp:303-317 shift decision
    double xStart = arMesh[0], xStep = arMesh[1], zStart = arMesh[3], zStep = arMesh[4];
    double xStartTr = -0.5 / xStep, zStartTr = -0.5 / zStep;
    double xStepTr = -xStartTr / (nx >> 1), zStepTr = -zStartTr / (nz >> 1);
    double x0 = xStart + 0.5 * nx * xStep, z0 = zStart + 0.5 * nz * zStep;
    bool shX = std::fabs(x0) > 1.E-06 * nx * xStep, shZ = std::fabs(z0) > 1.E-06 * nz * zStep;
    double t0s = dir > 0 ? -1. : 1.;
    srwf_slice S;
    std::memset(&S, 0, sizeof(S));
    S.xStart = xStart; S.xStep = xStep; S.zStart = zStart; S.zStep = zStep;
    S.qStepX = xStepTr; S.qStepZ = zStepTr;
    S.midScale = (float)(xStep * zStep);
    S.postShift = (shX || shZ) ? 1 : 0;
    S.sh1x = shX ? t0s * x0 : 0.;
    S.sh1y = shZ ? t0s * z0 : 0.;
    S.ny = (int)nz;
    HIPCHK(hipMemcpyAsync(dSp, &S, sizeof(S), hipMemcpyHostToDevice, st));
    HIPCHK(hipMemcpyAsync(d, pcData, sizeof(float2) * nxy, hipMemcpyHostToDevice, st));
    int inv = dir > 0 ? 0 : 1;
    if (p2) {
        const float2* twx = twiddles(lx);
        const float2* twz = twiddles(lz);
        if (!twx || !twz) return SRWL_AMD_ERR_GPU;
        const srwf_tiles* T = fft_tiles(lx, lz, 0, 2, inv, -1, dSp, 1, st);
        if (srwf_launch_rows_fwd(d, nxy, 1, dSp, twx, lx, (int)nz, 0, inv, T, st)) return SRWL_AMD_ERR_GPU;
        if (srwf_launch_cols(d, nxy, 1, dSp, twz, lz, lx, 2, inv, o, T, st)) return SRWL_AMD_ERR_GPU;
    } else {
        srwf_gplan px, pz;
        srwf_gplan_make((int)nx, &px);
        srwf_gplan_make((int)nz, &pz);
        const float2* twx = gen_twiddles((int)nx);
        const float2* twz = gen_twiddles((int)nz);
        float2* t = (float2*)g_ctx.buf.GenT.get(sizeof(float2) * nxy);
        if (!twx || !twz || !t) return SRWL_AMD_ERR_GPU;
        if (srwf_launch_grows(d, nxy, 1, dSp, twx, px, (int)nz, 0, 0, inv, 0, nullptr, 0, st)) return SRWL_AMD_ERR_GPU;
        if (srwf_launch_gtranspose(d, t, (int)nz, (int)nx, nxy, 1, st)) return SRWL_AMD_ERR_GPU;
        if (srwf_launch_gcols(t, nxy, 1, dSp, twz, pz, (int)nx, 2, inv, o, st)) return SRWL_AMD_ERR_GPU;
    }
    HIPCHK(hipMemcpyAsync(pcData, o, sizeof(float2) * nxy, hipMemcpyDeviceToHost, st));
    HIPCHK(hipStreamSynchronize(st));
    arMesh[0] = xStartTr; arMesh[1] = xStepTr; arMesh[3] = zStartTr; arMesh[4] = zStepTr;
    return 0;
}

SRWL_EXPORT int srwlAmdSetFftTiles(const int* t7)
{
    g_ctx.tilesForced = t7 != nullptr;
    if (t7) {
        srwf_tiles& T = g_ctx.forcedTiles;
        T.rpwF = t7[0]; T.thrF = t7[1]; T.cpw = t7[2]; T.thrC = t7[3]; T.rpwI = t7[4]; T.thrI = t7[5];
        T.pairC = t7[6];
    }
    return 0;
}

SRWL_EXPORT int srwlAmdSetTuning(int on)
{
    g_ctx.tuneTiles = on != 0;
    return 0;
}

SRWL_EXPORT int srwlAmdGetFftTiles(int lognx, int lognz, int mode, int postOp, int* out7)
{
    int pre = mode == 3 ? 1 : 0;
    if (postOp == 1 && mode != 3) postOp = 0;
    auto it = g_tiles.find(TileKey{lognx, lognz, pre, pre, 0, postOp});
    if (it == g_tiles.end()) return 0;
    const srwf_tiles& T = it->second;
    if (out7) {
        out7[0] = T.rpwF; out7[1] = T.thrF; out7[2] = T.cpw; out7[3] = T.thrC; out7[4] = T.rpwI; out7[5] = T.thrI;
        out7[6] = T.pairC;
    }
    return 1;
}

SRWL_EXPORT int srwlCalcIntFromElecField(char* pInt, SRWLWfr* pWfr, char pol, char intType, char depType, double e,
                                         double x, double y, double* arMeth, void* pFldTrj, double* arParGPU)
{
    (void)arParGPU;
    if (pWfr == nullptr || pInt == nullptr) return SRWL_INCORRECT_PARAM_FOR_INT_EXTR;
    // srwlib.cpp:782: intType -> the reference's internal Int_or_ReE
    static const int conv[] = {0, 1, 4, 5, 2, 3, 6, 7, 8};
    if (intType < 0 || intType > 7 || depType < 0 || depType > 6 || pFldTrj != nullptr) return SRWL_AMD_ERR_UNSUPPORTED;
    const int what = conv[(int)intType];
    // multi-electron intensity and flux (convolution with the electron beam) are not in this build
    if (what == 1 || what == 5) return SRWL_AMD_ERR_UNSUPPORTED;
    if (pol < -5 || pol > 6) return SRWL_AMD_ERR_UNSUPPORTED;
    if (pWfr->arEx == nullptr && pWfr->arEy == nullptr) return SRWL_INCORRECT_WFR_STRUCT;
    const SRWLRadMesh& m = pWfr->mesh;
    if (m.nx <= 0 || m.ny <= 0 || m.ne <= 0) return SRWL_INCORRECT_WFR_STRUCT;
    if (what == 4 && depType != 0) {
        // srradmnp.h:284-288: flux only vs photon energy
        *(float*)pInt = 0.f;
        return SRWL_WARN_FLUX_ONLY_VS_PHOTON_ENERGY;
    }
    int rc = ensure_device();
    if (rc) return rc;
    hipStream_t st = g_ctx.stream;
    srwx_extract P{};
    P.pol = pol; P.what = what; P.dep = depType;
    P.dbl = (what == 2);
    P.ne = m.ne; P.nx = m.nx; P.nz = m.ny;
    P.eStep = m.ne > 1 ? (m.eFin - m.eStart) / (m.ne - 1) : 0.;
    P.xStep = m.nx > 1 ? (m.xFin - m.xStart) / (m.nx - 1) : 0.;
    P.zStep = m.ny > 1 ? (m.yFin - m.yStart) / (m.ny - 1) : 0.;
    // intensity integrated over photon energy: vs x, vs y, vs x and y with
    // ne > 1 (srradmnp.cpp:372, 550, 733); there the components are
    // intensities (srradmnp.cpp:329, 505, 686)
    if (what == 7 && depType >= 1 && depType <= 3) {
        P.intOverE = m.ne > 1;
        P.what = 0;
    }
    // srradmnp.cpp:742-752: arMeth applies to the dependence on x and y
    P.iter = 0.; P.invIter = 1.;
    if (arMeth && depType == 3) {
        if (arMeth[0] == 1) { P.iter = arMeth[1]; P.invIter = 1. / (P.iter + 1.); }
        else if (arMeth[0] == 2) P.iter = -1.;
    }
    // srradmnp.cpp:35-79: cell index and fractional offset of an off-mesh coordinate
    auto coord = [](double arg, double start, double fin, long long n, long long& i0, long long& i1, double& t) {
        if (n <= 1) { i0 = i1 = 0; t = 0.; return; }
        const double step = (fin - start) / (n - 1), inv = 1. / step;
        long long a = (long long)((arg - start) * inv);
        if (a < 0) { a = 0; i1 = 0; }
        else if (a >= n - 1) { a = n - 1; i1 = a; }
        else i1 = a + 1;
        i0 = a;
        t = inv * (arg - a * step - start);
    };
    coord(e, m.eStart, m.eFin, P.ne, P.ie0, P.ie1, P.te);
    coord(x, m.xStart, m.xFin, P.nx, P.ix0, P.ix1, P.tx);
    coord(y, m.yStart, m.yFin, P.nz, P.iz0, P.iz1, P.tz);
    const long long lens[7] = {P.ne, P.nx, P.nz, P.nx * P.nz, P.ne * P.nx, P.ne * P.nz, P.ne * P.nx * P.nz};
    P.nout = lens[(int)depType];
    const int nOut = (pol == -5 && what != 4) ? 4 : 1;
    const size_t esz = P.dbl ? sizeof(double) : sizeof(float);
    const size_t outBytes = esz * nOut * P.nout;
    const long long n = P.ne * P.nx * P.nz;
    float2* d = (float2*)g_ctx.buf.W0.get(sizeof(float2) * 2 * n);
    void* o = g_ctx.buf.W1.get(outBytes);
    double* rows = (double*)g_ctx.buf.W2.get(sizeof(double) * (size_t)(P.ne * P.nz));
    if (!d || !o || !rows) return SRWL_ERR_MEMORY_ALLOCATION_FAILURE;
    float2* dx = pWfr->arEx ? d : nullptr;
    float2* dz = pWfr->arEy ? d + n : nullptr;
    if (dx) HIPCHK(hipMemcpyAsync(dx, pWfr->arEx, sizeof(float2) * n, hipMemcpyHostToDevice, st));
    if (dz) HIPCHK(hipMemcpyAsync(dz, pWfr->arEy, sizeof(float2) * n, hipMemcpyHostToDevice, st));
    // the averaging and sum modes update the caller's previous values
    if (P.iter != 0.) HIPCHK(hipMemcpyAsync(o, pInt, outBytes, hipMemcpyHostToDevice, st));
    if (srwx_launch_extract(dx, dz, P, o, rows, st)) return SRWL_AMD_ERR_GPU;
    HIPCHK(hipMemcpyAsync(pInt, o, what == 4 ? sizeof(float) * P.ne : outBytes, hipMemcpyDeviceToHost, st));
    HIPCHK(hipStreamSynchronize(st));
    return 0;
}

}  // extern "C"

// sr_auto.hip - adaptive near-field SR integral (SRW method 1) for gfx950.
//
// Reference: cpp/src/core/srradint.cpp:1731-2046 (RadIntegrationAuto1),
// srradint.h:988-1036 (FillNextLevel), 867-880 (CosAndSin), srradint.cpp:
// 289-394 (point order, running maximum of |E|^2).
//
// Each observation point runs the reference's level doubling on one
// wavefront (the level's new points split over its 64 lanes), in FP64 with
// the reference's Taylor sine and cosine, so that its convergence decisions
// follow the reference's.  The reference's stopping test becomes
// absolute after the first point of a call: |E|^2 may change by at most
// relPrec times the largest |E|^2 of the points already done.  That makes a
// point's final level depend on every earlier point.  Here every point first
// runs to its own relative test, storing |E|^2 and the field of every level;
// a parallel fixed-point iteration (running maximum in point order, then each
// point's first passing level) reproduces the sequential decisions exactly,
// extending the few points whose threshold is tighter than their own.
#include "srw_sr_dev.h"
#include "srw_auto.h"

// Trajectory data at s (srtrjdat.cpp:1185-1248 when byDiv, srtrjdat.h:431-472
// otherwise), with the straight line of a plane without motion.
__device__ __forceinline__ srwd_auto_tab auto_point(const srwd_trj_info& I, const double* X, const double* Y,
                                                    const double* IBx, const double* IBz, double s, bool byDiv)
{
    srwd_auto_tab t;
    double ibx, ibz;
    if (I.verNZ) { trj_pt p = eval_plane(I, X, IBx, -1. / I.betaNorm, s, byDiv); t.X = p.crd; t.Btx = p.bt; ibx = p.ib; }
    else { double d = s - I.s0; t.Btx = I.dxds0; t.X = I.x0 + I.dxds0 * d; ibx = I.dxds0 * d * I.dxds0; }
    if (I.horNZ) { trj_pt p = eval_plane(I, Y, IBz, 1. / I.betaNorm, s, byDiv); t.Z = p.crd; t.Btz = p.bt; ibz = p.ib; }
    else { double d = s - I.s0; t.Btz = I.dzds0; t.Z = I.z0 + I.dzds0 * d; ibz = I.dzds0 * d * I.dzds0; }
    t.inv = 1. / (I.yObs - s);
    t.Q = s * I.gammaEm2 + (ibx + ibz);
    t.s = s;
    t.pad = 0.;
    return t;
}

// Level point s: level 0 spans [sStart, sEnd] in 4 steps; level L >= 1 is the
// midpoints of level L-1, 2^(L+1) points with spacing (sEnd-sStart)/2^(L+1)
// starting half a spacing in (srradint.cpp:1937-1946).
__device__ __forceinline__ double auto_s(const srwd_trj_info& I, int L, long long i)
{
    double range = I.sIntegFin - I.sIntegStart;
    if (L == 0) return I.sIntegStart + (double)i * (range * 0.25);
    double step = range / (double)(1LL << (L + 1));
    return I.sIntegStart + 0.5 * step + (double)i * step;
}

__global__ void __launch_bounds__(256) k_auto_tables(const srwd_trj_info* __restrict__ info, const double* X,
                                                     const double* Y, const double* IBx, const double* IBz,
                                                     srwd_auto_tab* tab)
{
    long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    const long long total = srwd_auto_off(SRWD_AUTO_DEV_TAB_LEVEL + 1);
    if (e >= total) return;
    const srwd_trj_info I = info[0];
    int L = 0;
    if (e >= 5) L = (63 - __clzll((unsigned long long)(e - 1))) - 1;
    long long i = e - srwd_auto_off(L);
    double s = auto_s(I, L, i);
    tab[e] = auto_point(I, X, Y, IBx, IBz, s, L <= SRWD_AUTO_REF_TAB_LEVEL);
}

// Integrand at a level point (srradint.cpp:1771-1786, 1901-1914).
__device__ __forceinline__ void auto_term(const srwd_auto_tab& t, double xo, double zo, double kph, double& Ph,
                                          double& Ax, double& Az)
{
    double dx = xo - t.X, dz = zo - t.Z;
    double Nx = dx * t.inv, Nz = dz * t.inv;
    Ph = kph * (t.Q + dx * Nx + dz * Nz);
    Ax = (t.Btx - Nx) * t.inv;
    Az = (t.Btz - Nz) * t.inv;
}

__device__ __forceinline__ void auto_point_coords(const srwd_auto_job& J, long long p, double& xo, double& zo,
                                                  double& eph, int& ix, int& iz, int& ie)
{
    ie = (int)(p % J.ne);
    long long r = p / J.ne;
    ix = (int)(r % J.nxc);
    iz = (int)(r / J.nxc);
    req_point(*J.req, J.coord, p, xo, zo, eph);
}

// Sum over the 64 lanes of a wave, the same value on every lane (lane 0's
// sum, broadcast), so that decisions taken from it are uniform.
__device__ __forceinline__ double wave_sum(double v)
{
    for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off);
    return __shfl(v, 0);
}

// One wave per observation point: each level's new points are split into
// contiguous runs over the 64 lanes, the four partial sums are added over the
// wave and every lane takes the (uniform) convergence decision.
__global__ void __launch_bounds__(256) k_auto_points(const srwd_auto_job J, const srwd_trj_info* __restrict__ info,
                                                     const srwd_edge* __restrict__ edge,
                                                     const srwd_auto_tab* __restrict__ tab, const double* X,
                                                     const double* Y, const double* IBx, const double* IBz,
                                                     long long p0, long long n, const int* __restrict__ idx,
                                                     int nIdx, const double* __restrict__ thr, double* S, float4* F,
                                                     unsigned* guard, int* lvl, int* done, int* err, int* warn)
{
    const int lane = threadIdx.x & 63;
    long long k = (long long)blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
    long long q;
    if (idx) { if (k >= nIdx) return; q = idx[k]; }
    else { if (k >= n) return; q = k; }
    const srwd_trj_info I = info[0];
    double xo, zo, eph;
    int ix, iz, ie;
    auto_point_coords(J, p0 + q, xo, zo, eph, ix, iz, ie);
    const double actNorm = J.normConst * eph * 0.80654658E-03;
    const double kph = SRWD_KPH(eph);
    const double thrAbs = thr ? thr[q] : -1.;
    // terminating terms and edge derivatives (srradint.h:345-366), on lane 0
    double o0[4] = {0, 0, 0, 0}, wDD[4] = {0, 0, 0, 0};
    if (J.termin > 0) {
        if (lane == 0) {
            cplx RL[2] = {{0, 0}, {0, 0}}, RR[2] = {{0, 0}, {0, 0}}, DL[2] = {{0, 0}, {0, 0}}, DR[2] = {{0, 0}, {0, 0}};
            bool bad = false;
            const long long p = p0 + q;
            if (J.termin == 1 || J.termin == 2) {
                int r = residual(I, edge->v[0], I.sIntegStart, xo, I.yObs, zo, eph, actNorm, RL, DL, true,
                                 0.01 * J.relPrec);
                bad |= r == SRWD_RESID_BAD;
                if (r != SRWD_RESID_OK)
                    carried_derivative(I, edge->v[0], I.sIntegStart, I.yObs, *J.req, J.coord, J.last, 0, p, DL);
            }
            if (J.termin == 1 || J.termin == 3) {
                int r = residual(I, edge->v[1], I.sIntegFin, xo, I.yObs, zo, eph, actNorm, RR, DR);
                bad |= r == SRWD_RESID_BAD;
                if (r != SRWD_RESID_OK)
                    carried_derivative(I, edge->v[1], I.sIntegFin, I.yObs, *J.req, J.coord, J.last, 1, p, DR);
            }
            if (bad) atomicOr(warn, 1);
            o0[0] = RR[0].re - RL[0].re; o0[1] = RR[0].im - RL[0].im;
            o0[2] = RR[1].re - RL[1].re; o0[3] = RR[1].im - RL[1].im;
            const double wd = 1. / 15.;
            wDD[0] = wd * (DL[0].re - DR[0].re); wDD[1] = wd * (DL[0].im - DR[0].im);
            wDD[2] = wd * (DL[1].re - DR[1].re); wDD[3] = wd * (DL[1].im - DR[1].im);
        }
        for (int m = 0; m < 4; m++) { o0[m] = __shfl(o0[m], 0); wDD[m] = __shfl(wDD[m], 0); }
    }
    const double wfe = 7. / 15., wf1 = 16. / 15., wf2 = 14. / 15.;
    const double PI = 3.141592653590;
    // level 0 on lanes 0..4: ends into wF, odd points into s1, the middle point into s2
    double wF[4], s1[4], s2[4], PhInit;
    {
        double a[3][4] = {{0, 0, 0, 0}, {0, 0, 0, 0}, {0, 0, 0, 0}};
        double ph = 0.;
        if (lane < 5) {
            double Ax, Az, c, sn;
            auto_term(tab[lane], xo, zo, kph, ph, Ax, Az);
            cos_sin_taylor(ph, c, sn);
            int g = (lane == 0 || lane == 4) ? 0 : ((lane & 1) ? 1 : 2);
            a[g][0] = Ax * c; a[g][1] = Ax * sn; a[g][2] = Az * c; a[g][3] = Az * sn;
        }
        PhInit = __shfl(ph, 0);
        for (int m = 0; m < 4; m++) {
            wF[m] = wave_sum(a[0][m]) * wfe;
            s1[m] = wave_sum(a[1][m]);
            s2[m] = wave_sum(a[2][m]);
        }
    }
    double step = (I.sIntegFin - I.sIntegStart) * 0.25;
    double sq = 0.;
    {
        double v[4];
        for (int m = 0; m < 4; m++) {
            v[m] = o0[m] + actNorm * step * (wF[m] + wf1 * s1[m] + wf2 * s2[m] + step * wDD[m]);
            sq += v[m] * v[m];
        }
        if (lane == 0) {
            S[q] = sq;
            F[q] = make_float4((float)v[0], (float)v[1], (float)v[2], (float)v[3]);
        }
    }
    unsigned g = 0;
    int L = 0;
    bool ok = false;
    while (L < SRWD_AUTO_MAX_LEVEL) {
        L++;
        for (int m = 0; m < 4; m++) { s2[m] += s1[m]; s1[m] = 0.; }
        const double half = 0.5 * step;
        const long long np = srwd_auto_count(L);
        // this lane's run of the level's points
        const long long per = (np + 63) >> 6;
        const long long i0 = lane * per, i1 = (i0 + per < np) ? i0 + per : np;
        double part[4] = {0, 0, 0, 0};
        double phFirst = 0., phLast = 0.;
        bool stepsOk = true;
        for (long long i = i0; i < i1; i++) {
            srwd_auto_tab t = (L <= SRWD_AUTO_DEV_TAB_LEVEL) ? tab[srwd_auto_off(L) + i]
                                                              : auto_point(I, X, Y, IBx, IBz, auto_s(I, L, i), false);
            double Ph, Ax, Az, c, sn;
            auto_term(t, xo, zo, kph, Ph, Ax, Az);
            cos_sin_taylor(Ph, c, sn);
            part[0] += Ax * c; part[1] += Ax * sn; part[2] += Az * c; part[3] += Az * sn;
            if (i == i0) phFirst = Ph;
            else if (Ph - phLast > PI) stepsOk = false;
            phLast = Ph;
        }
        // phase step into this run from the previous lane's last point
        // (from the range start for the first point of the level)
        double prevLast = __shfl_up(phLast, 1);
        if (i0 < i1 && phFirst - (lane == 0 ? PhInit : prevLast) > PI) stepsOk = false;
        const bool mayBeLast = __all(stepsOk ? 1 : 0) != 0;
        for (int m = 0; m < 4; m++) s1[m] = wave_sum(part[m]);
        double v[4], locSq = 0.;
        for (int m = 0; m < 4; m++) {
            v[m] = o0[m] + actNorm * half * (wF[m] + wf1 * s1[m] + wf2 * s2[m] + half * wDD[m]);
            locSq += v[m] * v[m];
        }
        if (lane == 0) {
            S[(long long)L * n + q] = locSq;
            F[(long long)L * n + q] = make_float4((float)v[0], (float)v[1], (float)v[2], (float)v[3]);
        }
        if (mayBeLast) g |= 1u << L;
        bool pass = false;
        if (mayBeLast) {
            double test = fabs(locSq - sq);
            pass = !(test > (thrAbs >= 0. ? thrAbs : J.relPrec * locSq));
        }
        sq = locSq;
        step = half;
        if (pass) { ok = true; break; }
    }
    if (lane == 0) {
        guard[q] = g;
        lvl[q] = L;
        done[q] = L;
        // A point that does not settle under its own relative test may still
        // settle under the absolute test the running maximum gives it
        // (k_auto_decide); only a run under its final threshold (thr given)
        // is an error.
        if (!ok && thr) atomicOr(err, 1);
    }
}

// Exclusive running maximum of the final |E|^2 in point order (one block).
#define SRWD_AUTO_SCAN_THREADS 1024
__global__ void __launch_bounds__(SRWD_AUTO_SCAN_THREADS) k_auto_prefix(long long n, const double* __restrict__ S,
                                                                      const int* __restrict__ lvl,
                                                                      const double* __restrict__ M0, double* Mex,
                                                                      double* Mtot)
{
    __shared__ double sm[SRWD_AUTO_SCAN_THREADS];
    const int T = SRWD_AUTO_SCAN_THREADS;
    long long seg = (n + T - 1) / T;
    long long q0 = threadIdx.x * seg, q1 = q0 + seg;
    if (q1 > n) q1 = n;
    double loc = 0.;
    for (long long q = q0; q < q1; q++) {
        double v = S[(long long)lvl[q] * n + q];
        loc = v > loc ? v : loc;
    }
    sm[threadIdx.x] = loc;
    __syncthreads();
    for (int off = 1; off < T; off <<= 1) {
        double o = threadIdx.x >= off ? sm[threadIdx.x - off] : 0.;
        __syncthreads();
        if (o > sm[threadIdx.x]) sm[threadIdx.x] = o;
        __syncthreads();
    }
    double run = M0[0];
    if (threadIdx.x > 0 && sm[threadIdx.x - 1] > run) run = sm[threadIdx.x - 1];
    for (long long q = q0; q < q1; q++) {
        Mex[q] = run;
        double v = S[(long long)lvl[q] * n + q];
        if (v > run) run = v;
    }
    if (threadIdx.x == T - 1) {
        double m = sm[T - 1];
        Mtot[0] = m > M0[0] ? m : M0[0];
    }
}

// Each point's final level under the running maximum (srradint.cpp:2011-2044).
__global__ void __launch_bounds__(256) k_auto_decide(long long n, const double* __restrict__ S, int* lvl,
                                                     const unsigned* __restrict__ guard,
                                                     const int* __restrict__ done, double relPrec,
                                                     const double* __restrict__ Mex, double* thr, int* ext, int* cnt)
{
    long long q = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= n) return;
    const double M = Mex[q];
    const unsigned g = guard[q];
    const int D = done[q];
    int found = -1;
    double prev = S[q];
    for (int L = 1; L <= D; L++) {
        double cur = S[(long long)L * n + q];
        if ((g >> L) & 1u) {
            double lim = M > 0. ? relPrec * M : relPrec * cur;
            if (!(fabs(cur - prev) > lim)) { found = L; break; }
        }
        prev = cur;
    }
    if (found > 0) {
        if (found != lvl[q]) { lvl[q] = found; atomicAdd(&cnt[0], 1); }
    } else {
        thr[q] = M > 0. ? relPrec * M : -1.;
        int k = atomicAdd(&cnt[1], 1);
        ext[k] = (int)q;
    }
}

__global__ void __launch_bounds__(256) k_auto_out(const srwd_auto_job J, long long p0, long long n,
                                                  const float4* __restrict__ F, const int* __restrict__ lvl,
                                                  float2* Ex, float2* Ez)
{
    long long q = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= n) return;
    double xo, zo, eph;
    int ix, iz, ie;
    auto_point_coords(J, p0 + q, xo, zo, eph, ix, iz, ie);
    float4 v = F[(long long)lvl[q] * n + q];
    auto put = [&](int jx, int jz, float sx, float sz) {
        long long o = J.outOffset + ((long long)jz * J.nx + jx) * J.ne + ie;
        Ex[o] = make_float2(sx * v.x, sx * v.y);
        Ez[o] = make_float2(sz * v.z, sz * v.w);
    };
    put(ix, iz, 1.f, 1.f);
    // symmetric parts (srradint.cpp:3264-3350): mirror in z flips Ez, in x flips Ex
    int hz = J.nz >> 1, hx = J.nx >> 1;
    bool mz = J.symZ && iz < hz, mx = J.symX && ix < hx;
    if (mz) put(ix, J.nz - 1 - iz, 1.f, -1.f);
    if (mx) put(J.nx - 1 - ix, iz, -1.f, 1.f);
    if (mz && mx) put(J.nx - 1 - ix, J.nz - 1 - iz, -1.f, -1.f);
}

extern "C" {

int srwd_launch_auto_tables(const srwd_trj_info* dInfo, const double* dX, const double* dY, const double* dIBx,
                            const double* dIBz, srwd_auto_tab* dTab, hipStream_t st)
{
    long long total = srwd_auto_off(SRWD_AUTO_DEV_TAB_LEVEL + 1);
    hipLaunchKernelGGL(k_auto_tables, dim3((unsigned)((total + 255) / 256)), dim3(256), 0, st, dInfo, dX, dY, dIBx,
                       dIBz, dTab);
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

int srwd_launch_auto_points(const srwd_auto_job& J, const srwd_trj_info* dInfo, const srwd_edge* dEdge,
                            const srwd_auto_tab* dTab, const double* dX, const double* dY, const double* dIBx,
                            const double* dIBz, long long p0, long long n, const int* idx, int nIdx,
                            const double* thr, double* S, float4* F, unsigned* guard, int* lvl, int* done, int* err,
                            int* warn, hipStream_t st)
{
    long long m = idx ? nIdx : n;
    if (m <= 0) return 0;
    // one wave per point, four per workgroup
    hipLaunchKernelGGL(k_auto_points, dim3((unsigned)((m + 3) / 4)), dim3(256), 0, st, J, dInfo, dEdge, dTab, dX,
                       dY, dIBx, dIBz, p0, n, idx, nIdx, thr, S, F, guard, lvl, done, err, warn);
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

int srwd_launch_auto_scan(long long n, const double* S, int* lvl, const unsigned* guard, const int* done,
                          double relPrec, const double* M0, double* Mex, double* Mtot, double* thr, int* ext,
                          int* cnt, hipStream_t st)
{
    hipLaunchKernelGGL(k_auto_prefix, dim3(1), dim3(SRWD_AUTO_SCAN_THREADS), 0, st, n, S, lvl, M0, Mex, Mtot);
    hipLaunchKernelGGL(k_auto_decide, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, n, S, lvl, guard, done,
                       relPrec, Mex, thr, ext, cnt);
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

int srwd_launch_auto_out(const srwd_auto_job& J, long long p0, long long n, const float4* F, const int* lvl,
                         float* dEx, float* dEz, hipStream_t st)
{
    hipLaunchKernelGGL(k_auto_out, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, J, p0, n, F, lvl,
                       (float2*)dEx, (float2*)dEz);
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

}  // extern "C"

// sr.hip - FP64 trajectory (RK4), trajectory interpolation tables and the
// near-field SR integral for gfx950.
//
// Reference algorithms (paths relative to the reference source root):
//   field          cpp/src/core/srmagfld.h:401-502
//   RK4 trajectory cpp/src/core/srgtrjdt.cpp:24-365, ext/genmath/gmintrk.h:129-200
//   interpolation  cpp/src/core/srtrjdat.cpp:1724-2088, srtrjdat.h:955-1056
//   SR integral    cpp/src/core/srradint.cpp:1214-1604 (method 0), srradint.h:884-919
//   end terms      cpp/src/core/srradint.cpp:398-871
//
// Precision: trajectory, tables and the per-chunk phase reference are FP64;
// the per-point phase increment (|t| < ~2 revolutions) and the integrand sum
// are FP32: 64-point FMA block sums, added with Kahan compensation within a
// 256-point chunk, chunk sums in FP64.  A condition check of every point's
// FP32 sum sends ill-conditioned points to an FP64 recomputation.
#include "srw_sr_dev.h"

// ------------------------------------------------------------------------
// Field and RK4 (FP64)
// ------------------------------------------------------------------------
__device__ __forceinline__ double radical_one_plus_small(double x)
{   // gmmeth.h:390-394; the degree-8 polynomial is evaluated in Estrin form
    // to shorten the dependency chain of the serial RK4 loop
    if (fabs(x) > 0.01) return sqrt(1. + x);
    double x2 = x * x, x4 = x2 * x2;
    double p01 = fma(x, 0.5, 1.), p23 = fma(x, 0.0625, -0.125);
    double p45 = fma(x, 0.02734375, -0.0390625), p67 = fma(x, 0.01611328125, -0.0205078125);
    double lo = fma(x2, p23, p01), hi = fma(x2, p67, p45);
    return fma(x4 * x2, hi, lo);
}

// cos(pi a) and sin(pi a) for the field harmonics: exact reduction to
// r = a - rint(a), then Taylor series of cos/sin(pi r) (|pi r| <= pi/2,
// truncation below 1e-17) evaluated in Estrin form.
__device__ __forceinline__ void sincospi_fast(double a, double& s, double& c)
{
    double n = rint(a);
    double x = (a - n) * 3.141592653589793;
    double x2 = x * x, x4 = x2 * x2, x8 = x4 * x4;
    // cos: 1 - x2/2 + x4/24 - ... + x20/20!
    double c0 = fma(x2, -0.5, 1.);
    double c1 = fma(x2, -1.3888888888888889e-03, 4.1666666666666664e-02);
    double c2 = fma(x2, -2.7557319223985893e-07, 2.4801587301587302e-05);
    double c3 = fma(x2, -1.1470745597729725e-11, 2.0876756987868100e-09);
    double c4 = fma(x4, 4.1103176233121648e-19, fma(x2, -1.5619206968586225e-16, 4.7794773323873853e-14));
    double clo = fma(x4, c1, c0), chi = fma(x4, c3, c2);
    c = fma(x8, fma(x8, c4, chi), clo);
    // sin: x (1 - x2/6 + ... - x20/21!)
    double s0 = fma(x2, -1.6666666666666666e-01, 1.);
    double s1 = fma(x2, -1.9841269841269841e-04, 8.3333333333333333e-03);
    double s2 = fma(x2, -2.5052108385441720e-08, 2.7557319223985893e-06);
    double s3 = fma(x2, -7.6471637318198164e-13, 1.6059043836821613e-10);
    double s4 = fma(x4, 1.9572941063391263e-20, fma(x2, -8.2206352466243295e-18, 2.8114572543455206e-15));
    double slo = fma(x4, s1, s0), shi = fma(x4, s3, s2);
    s = x * fma(x8, fma(x8, s4, shi), slo);
    if (((long long)n) & 1) { s = -s; c = -c; }
}

__device__ __forceinline__ double cosh_small(double a)
{
    if (fabs(a) > 0.1) return cosh(a);
    double a2 = a * a;
    return 1. + 0.5 * a2 * (1. + a2 / 12. * (1. + a2 / 30. * (1. + a2 / 56. * (1. + a2 / 90.))));
}

// srmagfld.h:401-502 with the per-harmonic constants precomputed on the host
// (srwd_harm); cos/sin of k*z use cospi/sinpi of the argument in half turns.
__device__ __forceinline__ void harm_addB(const srwd_harm& h, double halfTotTerm, double yr, double zr, double* B)
{
    double zrr = zr + h.shift;
    if (zrr < -halfTotTerm || zrr > halfTotTerm) return;
    double Bm = h.Bm0;
    if (yr != 0.) Bm *= cosh_small(h.kw * yr);
    double dB;
    if (zrr < -h.halfThis || zrr > h.halfThis) {
        bool lft = zrr < -h.halfThis;
        double za = lft ? zrr + h.halfThis : zrr - h.halfThis;
        double za2 = za * za;
        dB = (h.kw * Bm * za) * (1. - h.c4 * za2) * exp(-h.p * za2);
        dB *= lft ? h.sL : h.sR;
    } else {
        double sn, cs;
        sincospi_fast(h.kwpi * zrr, sn, cs);
        dB = Bm * ((h.sym == 1) ? cs : sn);
    }
    if (h.isX) B[0] += dB; else B[1] += dB;
}

__device__ __forceinline__ void und_addB(const srwd_und& u, double yr, double zr, double* B)
{
    if (zr < -u.halfTotTerm || zr > u.halfTotTerm) return;
    for (int i = 0; i < u.nh; i++) harm_addB(u.h[i], u.halfTotTerm, yr, zr, B);
}

__device__ void field_B(const srwd_field& f, double x, double y, double z, double* B)
{
    B[0] = B[1] = B[2] = 0.;
    (void)x;
    for (int i = 0; i < f.n; i++) und_addB(f.u[i], y - f.u[i].cy, z - f.u[i].cz, B);
}

struct rk_ctx {
    const srwd_field* f;
    double gEm2, mult;
    bool one;              // single undulator with a single harmonic: constants in registers
    srwd_harm h;
    double halfTotTerm, cy, cz;
};

__device__ __forceinline__ void rk_deriv(const rk_ctx& c, const double* F, double* dF)
{   // srgtrjdt.h:279-293
    double xd = F[1], yd = F[3];
    double zd = radical_one_plus_small(-(c.gEm2 + xd * xd + yd * yd));
    double B[3];
    if (c.one) {
        B[0] = B[1] = B[2] = 0.;
        double zr = F[4] - c.cz;
        if (zr >= -c.halfTotTerm && zr <= c.halfTotTerm) harm_addB(c.h, c.halfTotTerm, F[2] - c.cy, zr, B);
    } else field_B(*c.f, F[0], F[2], F[4], B);
    dF[0] = xd;
    dF[1] = c.mult * (yd * B[2] - zd * B[1]);
    dF[2] = yd;
    dF[3] = c.mult * (zd * B[0] - xd * B[2]);
    dF[4] = zd;
}

__device__ void rk4_step(const rk_ctx& c, double* y, const double* d, double h)
{   // gmintrk.h:169-200
    double hh = 0.5 * h, h6 = h / 6., yt[5], dyt[5], dym[5];
    for (int i = 0; i < 5; i++) yt[i] = y[i] + hh * d[i];
    rk_deriv(c, yt, dyt);
    for (int i = 0; i < 5; i++) yt[i] = y[i] + hh * dyt[i];
    rk_deriv(c, yt, dym);
    for (int i = 0; i < 5; i++) { yt[i] = y[i] + h * dym[i]; dym[i] += dyt[i]; }
    rk_deriv(c, yt, dyt);
    for (int i = 0; i < 5; i++) y[i] += h6 * (d[i] + dyt[i] + 2. * dym[i]);
}

// Integrates np points from xmin to xmax starting at y0 (gmintrk.h:129-165);
// point j is stored at output index base + dir*j.
__device__ void rk_chain(const rk_ctx& c, const double* y0, double xmin, double xmax, long long np,
                         double* oX, double* oXp, double* oY, double* oYp, double* oZ,
                         long long base, int dir)
{
    double step = (xmax - xmin) / (double)(np - 1);
    double y[5], d[5];
    for (int k = 0; k < 5; k++) y[k] = y0[k];
    for (long long j = 0; j < np; j++) {
        long long o = base + dir * j;
        oX[o] = y[0]; oXp[o] = y[1]; oY[o] = y[2]; oYp[o] = y[3];
        if (oZ) oZ[o] = y[4];
        if (j != np - 1) { rk_deriv(c, y, d); rk4_step(c, y, d, step); }
    }
}

// "Arrive" at xmax from 0 with n points, returning the end state.
__device__ void rk_arrive(const rk_ctx& c, double* y, double xmax, long long n)
{
    double step = xmax / (double)(n - 1), d[5];
    for (long long j = 0; j < n - 1; j++) { rk_deriv(c, y, d); rk4_step(c, y, d, step); }
}

// One thread per (trajectory, side): side 0 fills the part below ct = 0,
// side 1 the part above it (srgtrjdt.cpp:90-340; the "rotated" branches are
// unreachable because ctStart < ctEnd is enforced by the caller).
__global__ void __launch_bounds__(64) k_traj_rk4(const srwd_traj_job* __restrict__ jobs,
                                                 const srwd_field* __restrict__ fields, int nTraj,
                                                 double* X, double* Xp, double* Y, double* Yp, double* Z,
                                                 long long ld)
{
    int t = blockIdx.x * blockDim.x + threadIdx.x;
    int tr = t >> 1, side = t & 1;
    if (tr >= nTraj) return;
    const srwd_traj_job J = jobs[tr];
    rk_ctx c;
    c.f = &fields[J.field];
    c.gEm2 = 1. / J.gammaInt / J.gammaInt;
    c.mult = J.mult;
    c.one = (c.f->n == 1 && c.f->u[0].nh == 1);
    if (c.one) {
        c.h = c.f->u[0].h[0];
        c.halfTotTerm = c.f->u[0].halfTotTerm;
        c.cy = c.f->u[0].cy;
        c.cz = c.f->u[0].cz;
    }
    double* oX = X + tr * ld; double* oXp = Xp + tr * ld; double* oY = Y + tr * ld;
    double* oYp = Yp + tr * ld; double* oZ = Z ? Z + tr * ld : nullptr;
    long long ns = J.np;
    double sMin = J.ctStart, sMax = J.ctEnd;
    double sStep = (ns <= 1) ? 0 : (sMax - sMin) / (ns - 1);
    double tol = (sMax - sMin) * 1.E-12;
    bool left = sMin < -tol, right = sMax > tol;
    double ic[5];
    for (int k = 0; k < 5; k++) ic[k] = J.init[k];
    if (left && sMax < -tol) {
        if (side != 0) return;
        long long auxNp = (long long)(int)fabs(sMax / sStep) + 1;
        if (auxNp <= 1) auxNp = 2;
        rk_arrive(c, ic, sMax, auxNp);
        if (ns <= 1) { oX[0] = ic[0]; oXp[0] = ic[1]; oY[0] = ic[2]; oYp[0] = ic[3]; if (oZ) oZ[0] = ic[4]; }
        else rk_chain(c, ic, sMax, sMin, ns, oX, oXp, oY, oYp, oZ, ns - 1, -1);
        return;
    }
    if (right && sMin > tol) {
        if (side != 1) return;
        long long auxNp = (long long)(int)fabs(sMin / sStep) + 1;
        if (auxNp <= 1) auxNp = 2;
        rk_arrive(c, ic, sMin, auxNp);
        if (ns <= 1) { oX[0] = ic[0]; oXp[0] = ic[1]; oY[0] = ic[2]; oYp[0] = ic[3]; if (oZ) oZ[0] = ic[4]; }
        else rk_chain(c, ic, sMin, sMax, ns, oX, oXp, oY, oYp, oZ, 0, 1);
        return;
    }
    long long is0 = 0;
    double s0Act = 0.;
    if (left) {
        is0 = (int)fabs((-sMin + tol) / sStep);
        if (is0 >= ns) is0 = ns - 1;
        s0Act = sMin + is0 * sStep;
        if (s0Act < -tol) rk_arrive(c, ic, s0Act, 2);
    }
    if (side == 0) {
        if (!left) return;
        long long nsL = is0 + 1;
        if (nsL <= 1) { oX[0] = ic[0]; oXp[0] = ic[1]; oY[0] = ic[2]; oYp[0] = ic[3]; if (oZ) oZ[0] = ic[4]; }
        else rk_chain(c, ic, s0Act, sMin, nsL, oX, oXp, oY, oYp, oZ, nsL - 1, -1);
    } else {
        if (!right) return;
        long long nsR = ns - is0;
        if (nsR < 1) nsR = 1;
        if (nsR <= 1) { oX[is0] = ic[0]; oXp[is0] = ic[1]; oY[is0] = ic[2]; oYp[is0] = ic[3]; if (oZ) oZ[is0] = ic[4]; }
        else rk_chain(c, ic, s0Act, sMax, nsR, oX, oXp, oY, oYp, oZ, is0, 1);
    }
}

// ------------------------------------------------------------------------
// Parallel-in-s RK4.  A chain of n points y_{j+1} = y_j + D(y_j) (D = one RK4
// increment, rk4_step) is the fixed point of y = y_0 + prefix-sum(D(y)).  One
// workgroup iterates that map: every thread forms the increments of its
// contiguous segment of steps from the current iterate, a block scan turns
// them into positions, and the loop stops when no component moved by more
// than 1e-15 of its magnitude (the planar-undulator chains here reach an exact
// fixed point in 7-8 sweeps).  The result equals the serial recurrence up to
// the association order of the sums.  A chain that has not converged after
// SRWD_PIC_MAXIT sweeps is recomputed serially by one thread, so every
// trajectory ends with the serial answer or its converged equivalent.
// ------------------------------------------------------------------------
#define SRWD_PIC_THREADS 512
#define SRWD_PIC_MAXIT 48

// rk_deriv with the field path fixed at compile time: ONE = 1 for a single
// undulator with a single harmonic (constants in registers), 0 otherwise.
template <int ONE>
__device__ __forceinline__ void rk_deriv_t(const rk_ctx& c, const double* F, double* dF)
{   // srgtrjdt.h:279-293
    double xd = F[1], yd = F[3];
    double zd = radical_one_plus_small(-(c.gEm2 + xd * xd + yd * yd));
    double B[3];
    if (ONE) {
        B[0] = B[1] = B[2] = 0.;
        double zr = F[4] - c.cz;
        if (zr >= -c.halfTotTerm && zr <= c.halfTotTerm) harm_addB(c.h, c.halfTotTerm, F[2] - c.cy, zr, B);
    } else field_B(*c.f, F[0], F[2], F[4], B);
    dF[0] = xd;
    dF[1] = c.mult * (yd * B[2] - zd * B[1]);
    dF[2] = yd;
    dF[3] = c.mult * (zd * B[0] - xd * B[2]);
    dF[4] = zd;
}

template <int ONE>
__device__ __forceinline__ void rk4_increment(const rk_ctx& c, const double* y, double h, double* D)
{   // the update of rk4_step, returned instead of added
    double d[5], hh = 0.5 * h, h6 = h / 6., yt[5], dyt[5], dym[5];
    rk_deriv_t<ONE>(c, y, d);
    for (int i = 0; i < 5; i++) yt[i] = y[i] + hh * d[i];
    rk_deriv_t<ONE>(c, yt, dyt);
    for (int i = 0; i < 5; i++) yt[i] = y[i] + hh * dyt[i];
    rk_deriv_t<ONE>(c, yt, dym);
    for (int i = 0; i < 5; i++) { yt[i] = y[i] + h * dym[i]; dym[i] += dyt[i]; }
    rk_deriv_t<ONE>(c, yt, dyt);
    for (int i = 0; i < 5; i++) D[i] = h6 * (d[i] + dyt[i] + 2. * dym[i]);
}

// Exclusive block scan of 5 doubles per thread (SRWD_PIC_THREADS / 64 waves).
__device__ __forceinline__ void block_excl_scan5(const double* v, double* ex, double* lds /* [waves][5] */)
{
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    double inc[5];
    for (int k = 0; k < 5; k++) {
        double x = v[k];
        for (int off = 1; off < 64; off <<= 1) {
            double t = __shfl_up(x, off, 64);
            if (lane >= off) x += t;
        }
        inc[k] = x;
    }
    if (lane == 63) for (int k = 0; k < 5; k++) lds[w * 5 + k] = inc[k];
    __syncthreads();
    for (int k = 0; k < 5; k++) {
        double e = __shfl_up(inc[k], 1, 64);
        if (lane == 0) e = 0.;
        for (int q = 0; q < w; q++) e += lds[q * 5 + k];
        ex[k] = e;
    }
    __syncthreads();
}

// Block max-reduction of 10 non-negative doubles (5 changes, 5 magnitudes);
// every thread receives the result.
__device__ __forceinline__ void block_max10(double* v, double* lds /* [waves][10] */)
{
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    for (int k = 0; k < 10; k++) {
        double x = v[k];
        for (int off = 32; off > 0; off >>= 1) x = fmax(x, __shfl_xor(x, off, 64));
        v[k] = x;
    }
    if (lane == 0) for (int k = 0; k < 10; k++) lds[w * 10 + k] = v[k];
    __syncthreads();
    for (int k = 0; k < 10; k++) {
        double x = lds[k];
        for (int q = 1; q < SRWD_PIC_THREADS / 64; q++) x = fmax(x, lds[q * 10 + k]);
        v[k] = x;
    }
    __syncthreads();
}

// Storage of a chain in an iterate buffer.  Step j of a chain with `per`
// steps per thread is owned by thread j/per; its end point j+1 is kept at
// (j % per)*T + j/per, so the threads of a wave touch consecutive words, and
// point 0 at per*T.  A buffer component needs per*T + 1 doubles.
__device__ __forceinline__ long long pic_at(long long j, long long per)
{
    if (j == 0) return per * SRWD_PIC_THREADS;
    long long s = j - 1;
    return (s % per) * SRWD_PIC_THREADS + s / per;
}

// Solves one chain of n points from y0 with step h (component k of point j at
// buf[k*cap + pic_at(j, per)]) using the two iterate buffers Ya and Yb, and
// returns the one that holds the result.  All threads of the block call it
// with the same arguments; n - 1 <= cap - 1 must hold.
template <int ONE>
__device__ __forceinline__ double* picard_chain(const rk_ctx& c, const double* y0, double h, long long n,
                                                double* Ya, double* Yb, long long cap, double* lds,
                                                long long& perOut)
{
    const int t = threadIdx.x;
    const long long nSteps = n - 1;
    const long long per = nSteps > 0 ? (nSteps + SRWD_PIC_THREADS - 1) / SRWD_PIC_THREADS : 1;
    perOut = per;
    {   // initial iterate: free flight along the initial velocity.  The
        // accelerations are left out: extrapolated over a whole chain they
        // carry the orbit far off axis, where the field's cosh growth makes
        // the next iterate overflow.
        double d0[5];
        rk_deriv_t<ONE>(c, y0, d0);
        d0[1] = 0.; d0[3] = 0.;
        for (long long j = t; j < n; j += SRWD_PIC_THREADS) {
            long long a = pic_at(j, per);
            for (int k = 0; k < 5; k++) Ya[k * cap + a] = (j == 0) ? y0[k] : y0[k] + (double)j * h * d0[k];
        }
    }
    __syncthreads();
    if (nSteps <= 0) return Ya;
    const long long j0 = t * per, j1 = (j0 + per < nSteps) ? j0 + per : nSteps;
    bool done = false;
    for (int it = 0; it < SRWD_PIC_MAXIT && !done; it++) {
        double tot[5] = {0., 0., 0., 0., 0.};
        for (long long j = j0; j < j1; j++) {
            double y[5], D[5];
            long long a = pic_at(j, per), b = pic_at(j + 1, per);
            for (int k = 0; k < 5; k++) y[k] = Ya[k * cap + a];
            rk4_increment<ONE>(c, y, h, D);
            for (int k = 0; k < 5; k++) { Yb[k * cap + b] = D[k]; tot[k] += D[k]; }
        }
        double ex[5];
        block_excl_scan5(tot, ex, lds);
        double m[10] = {0., 0., 0., 0., 0., 0., 0., 0., 0., 0.};
        double acc[5];
        for (int k = 0; k < 5; k++) acc[k] = y0[k] + ex[k];
        for (long long j = j0; j < j1; j++) {
            long long b = pic_at(j + 1, per);
            for (int k = 0; k < 5; k++) {
                acc[k] += Yb[k * cap + b];
                double dv = fabs(acc[k] - Ya[k * cap + b]);
                // fmax drops a NaN: a non-finite iterate is recorded as infinite instead
                m[k] = isfinite(dv) ? fmax(m[k], dv) : INFINITY;
                m[5 + k] = fmax(m[5 + k], fabs(acc[k]));
                Yb[k * cap + b] = acc[k];
            }
        }
        if (t == 0) for (int k = 0; k < 5; k++) Yb[k * cap + per * SRWD_PIC_THREADS] = y0[k];
        block_max10(m, lds + 5 * (SRWD_PIC_THREADS / 64));
        double* sw = Ya; Ya = Yb; Yb = sw;
        bool diverged = false;
        done = true;
        for (int k = 0; k < 5; k++) {
            done = done && (m[k] <= 1e-15 * m[5 + k]);
            diverged = diverged || !isfinite(m[k]);
        }
        if (diverged) break;   // every thread holds the same m: the serial recurrence takes over
    }
    if (!done) {
        // serial recurrence, one thread
        if (t == 0) {
            double y[5], D[5];
            for (int k = 0; k < 5; k++) { y[k] = y0[k]; Ya[k * cap + pic_at(0, per)] = y[k]; }
            for (long long j = 0; j < nSteps; j++) {
                rk4_increment<ONE>(c, y, h, D);
                long long b = pic_at(j + 1, per);
                for (int k = 0; k < 5; k++) { y[k] += D[k]; Ya[k * cap + b] = y[k]; }
            }
        }
        __syncthreads();
    }
    return Ya;
}

__device__ __forceinline__ void rk_ctx_setup(rk_ctx& c, const srwd_traj_job& J, const srwd_field* fields)
{
    c.f = &fields[J.field];
    c.gEm2 = 1. / J.gammaInt / J.gammaInt;
    c.mult = J.mult;
    c.one = (c.f->n == 1 && c.f->u[0].nh == 1);
    if (c.one) {
        c.h = c.f->u[0].h[0];
        c.halfTotTerm = c.f->u[0].halfTotTerm;
        c.cy = c.f->u[0].cy;
        c.cz = c.f->u[0].cz;
    }
}

// Stores chain point j of buffer R at output index base + dir*j.
__device__ __forceinline__ void picard_store(const double* R, long long cap, long long per, long long n,
                                             long long base, int dir, double* oX, double* oXp, double* oY,
                                             double* oYp, double* oZ)
{
    for (long long j = threadIdx.x; j < n; j += SRWD_PIC_THREADS) {
        long long o = base + dir * j, a = pic_at(j, per);
        oX[o] = R[a]; oXp[o] = R[cap + a]; oY[o] = R[2 * cap + a]; oYp[o] = R[3 * cap + a];
        if (oZ) oZ[o] = R[4 * cap + a];
    }
}

// Advances y by a chain of n points with step h (the reference's "arrive"
// integrations), in pieces that fit the scratch buffers.
template <int ONE>
__device__ void picard_arrive(const rk_ctx& c, double* y, double h, long long n, double* Ya, double* Yb,
                              long long cap, double* lds)
{
    long long left = n;
    if (cap < 2) return;
    while (left > 1) {
        long long m = left < cap ? left : cap, per;
        double* R = picard_chain<ONE>(c, y, h, m, Ya, Yb, cap, lds, per);
        __syncthreads();
        for (int k = 0; k < 5; k++) y[k] = R[k * cap + pic_at(m - 1, per)];
        __syncthreads();
        left -= m - 1;
    }
}

// Same trajectory as k_traj_rk4 (srgtrjdt.cpp:90-340), each chain solved by
// picard_chain.  Two workgroups per trajectory (blockIdx.y): when the range
// spans ct = 0, the upstream and downstream chains are independent and
// workgroup 0 integrates the upstream one while workgroup 1 integrates the
// downstream one; otherwise workgroup 0 does all the work.  scratch holds
// 20*cap doubles per trajectory (10*cap per workgroup), cap =
// ceil((ld - 1)/T)*T + 1.
template <int ONE>
__global__ void __launch_bounds__(SRWD_PIC_THREADS) k_traj_picard(const srwd_traj_job* __restrict__ jobs,
                                                                   const srwd_field* __restrict__ fields,
                                                                   double* X, double* Xp, double* Y,
                                                                   double* Yp, double* Z, long long ld,
                                                                   double* scratch, long long cap)
{
    __shared__ double lds[15 * (SRWD_PIC_THREADS / 64)];
    const int tr = blockIdx.x;
    const srwd_traj_job J = jobs[tr];
    rk_ctx c;
    rk_ctx_setup(c, J, fields);
    double* oX = X + tr * ld; double* oXp = Xp + tr * ld; double* oY = Y + tr * ld;
    double* oYp = Yp + tr * ld; double* oZ = Z ? Z + tr * ld : nullptr;
    const int side = blockIdx.y;
    double* Ya = scratch + ((size_t)tr * 2 + side) * 10 * cap;
    double* Yb = Ya + 5 * cap;
    const long long ns = J.np;
    const double sMin = J.ctStart, sMax = J.ctEnd;
    const double sStep = (ns <= 1) ? 0 : (sMax - sMin) / (ns - 1);
    const double tol = (sMax - sMin) * 1.E-12;
    const bool left = sMin < -tol, right = sMax > tol;
    double ic[5];
    for (int k = 0; k < 5; k++) ic[k] = J.init[k];
    if ((left && sMax < -tol) || (right && sMin > tol)) {
        // the whole range lies on one side of ct = 0
        if (side) return;
        const bool below = left && sMax < -tol;
        const double sNear = below ? sMax : sMin, sFar = below ? sMin : sMax;
        long long auxNp = (long long)(int)fabs(sNear / sStep) + 1;
        if (auxNp <= 1) auxNp = 2;
        picard_arrive<ONE>(c, ic, sNear / (double)(auxNp - 1), auxNp, Ya, Yb, cap, lds);
        if (ns <= 1) {
            if (threadIdx.x == 0) {
                oX[0] = ic[0]; oXp[0] = ic[1]; oY[0] = ic[2]; oYp[0] = ic[3]; if (oZ) oZ[0] = ic[4];
            }
            return;
        }
        long long per;
        double* R = picard_chain<ONE>(c, ic, (sFar - sNear) / (double)(ns - 1), ns, Ya, Yb, cap, lds, per);
        if (below) picard_store(R, cap, per, ns, ns - 1, -1, oX, oXp, oY, oYp, oZ);
        else picard_store(R, cap, per, ns, 0, 1, oX, oXp, oY, oYp, oZ);
        return;
    }
    long long is0 = 0;
    double s0Act = 0.;
    if (left) {
        is0 = (int)fabs((-sMin + tol) / sStep);
        if (is0 >= ns) is0 = ns - 1;
        s0Act = sMin + is0 * sStep;
        if (s0Act < -tol) {   // one RK4 step from 0 to s0Act, computed by every thread alike
            double D[5];
            rk4_increment<ONE>(c, ic, s0Act, D);
            for (int k = 0; k < 5; k++) ic[k] += D[k];
        }
    }
    if (left && side == 0) {
        long long nsL = is0 + 1;
        if (nsL <= 1) {
            if (threadIdx.x == 0) {
                oX[0] = ic[0]; oXp[0] = ic[1]; oY[0] = ic[2]; oYp[0] = ic[3]; if (oZ) oZ[0] = ic[4];
            }
        } else {
            long long per;
            double* R = picard_chain<ONE>(c, ic, (sMin - s0Act) / (double)(nsL - 1), nsL, Ya, Yb, cap, lds, per);
            picard_store(R, cap, per, nsL, nsL - 1, -1, oX, oXp, oY, oYp, oZ);
        }
    }
    // the downstream chain: workgroup 1, or workgroup 0 when there is no upstream one
    if (right && side == (left ? 1 : 0)) {
        long long nsR = ns - is0;
        if (nsR < 1) nsR = 1;
        if (nsR <= 1) {
            if (threadIdx.x == 0) {
                oX[is0] = ic[0]; oXp[is0] = ic[1]; oY[is0] = ic[2]; oYp[is0] = ic[3]; if (oZ) oZ[is0] = ic[4];
            }
        } else {
            long long per;
            double* R = picard_chain<ONE>(c, ic, (sMax - s0Act) / (double)(nsR - 1), nsR, Ya, Yb, cap, lds, per);
            picard_store(R, cap, per, nsR, is0, 1, oX, oXp, oY, oYp, oZ);
        }
    }
}

// ------------------------------------------------------------------------
// Integrals of the squared angles (FP64); the interpolation itself is in
// srw_sr_dev.h
// ------------------------------------------------------------------------
__device__ __forceinline__ double interp_cub_half(const double* f, int i)
{   // gminterp.h:608-613
    if (i < 0) return 0.0625 * (5. * f[0] + 15. * f[1] - 5. * f[2] + f[3]);
    else if (i == 0) return 0.0625 * (-f[0] + 9. * f[1] + 9. * f[2] - f[3]);
    else return 0.0625 * (f[0] - 5. * f[1] + 15. * f[2] + 5. * f[3]);
}

// Cumulative Simpson of the squared angle, zeroed at s0 (srtrjdat.cpp:1724-1805).
// One 256-thread block per (trajectory, plane): the interval terms are formed
// in parallel and summed with a block prefix scan.  Interval i (1-based) uses
// the half-step cubic interpolant centred on nodes (i-1, i); as in the
// reference, the first interval's left end term is 0.
#define SRWD_QT_THREADS 256
__global__ void __launch_bounds__(SRWD_QT_THREADS) k_quad_terms(const srwd_trj_info* __restrict__ info,
                                                                const double* Xp, const double* Yp, long long ld,
                                                                double* IBx, double* IBz)
{
    const int tr = blockIdx.x >> 1, pl = blockIdx.x & 1;
    const srwd_trj_info I = info[tr];
    const double* ang = (pl == 0 ? Xp : Yp) + tr * ld;
    double* out = (pl == 0 ? IBx : IBz) + tr * ld;
    const long long np = I.np, n1 = np - 1;
    const double stepMult = 0.333333333333 * (0.5 * I.sStep);
    const int T = SRWD_QT_THREADS;
    long long seg = (n1 + T - 1) / T;
    long long i0 = 1 + threadIdx.x * seg, i1 = i0 + seg;
    if (i1 > n1 + 1) i1 = n1 + 1;
    auto term = [&](long long i) {
        const double* p;
        int mid;
        if (i == 1) { p = ang; mid = -1; }
        else if (i < n1) { p = ang + (i - 2); mid = 0; }
        else { p = ang + (n1 - 3); mid = 1; }
        double a = interp_cub_half(p, mid);
        double f0 = (i == 1) ? 0. : ang[i - 1] * ang[i - 1];
        return stepMult * (f0 + 4. * a * a + ang[i] * ang[i]);
    };
    double loc = 0.;
    for (long long i = i0; i < i1; i++) loc += term(i);
    __shared__ double sc[SRWD_QT_THREADS];
    __shared__ double sv;
    sc[threadIdx.x] = loc;
    __syncthreads();
    for (int off = 1; off < T; off <<= 1) {
        double v = threadIdx.x >= off ? sc[threadIdx.x - off] : 0.;
        __syncthreads();
        sc[threadIdx.x] += v;
        __syncthreads();
    }
    double run = threadIdx.x > 0 ? sc[threadIdx.x - 1] : 0.;
    if (threadIdx.x == 0) out[0] = 0.;
    for (long long i = i0; i < i1; i++) { run += term(i); out[i] = run; }
    __syncthreads();
    if (threadIdx.x == 0) {
        double sStep = I.sStep;
        long long idx = (long long)((I.s0 - I.sStart) / sStep);
        if (idx >= np - 1) idx = np - 2; else if (idx < 0) idx = 0;
        double rel = I.s0 - (I.sStart + sStep * idx);
        long long off;
        if (idx < 2) { off = 0; rel -= (2 - idx) * sStep; }
        else if (idx < np - 3) off = idx - 2;
        else if (idx < np - 2) { off = idx - 3; rel += sStep; }
        else { off = idx - 4; rel += 2 * sStep; }
        double c[6];
        coefs5(sStep, out + off, c);
        sv = c[0] + rel * (c[1] + rel * (c[2] + rel * (c[3] + rel * (c[4] + rel * c[5]))));
    }
    __syncthreads();
    double v = sv;
    for (long long j = threadIdx.x; j <= n1; j += T) out[j] -= v;
}

__device__ __forceinline__ double int_point_s(const srwd_trj_info& I, long long i)
{
    return (i == I.nInt - 1) ? I.sIntegFin : I.sIntegStart + (double)i * I.h;
}

// Integration-point tables.  One thread per (trajectory, integration point).
// Writes the FP32 record, the FP64 phase base Q_i and, for the first point of
// each chunk, the FP64 chunk reference.
__global__ void __launch_bounds__(256) k_sr_tables(const srwd_trj_info* __restrict__ info,
                                                   const double* X, const double* Y, const double* IBx,
                                                   const double* IBz, long long ldT,
                                                   srwd_rec* rec, srwd_chunk* chunk, double* Qout,
                                                   long long ldI, long long ldC, long long maxInt, double* T64)
{
    long long gi = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    int tr = blockIdx.y;
    if (gi >= maxInt) return;
    const srwd_trj_info I = info[tr];
    if (gi >= I.nInt) return;
    bool byDiv = I.nInt < 10000;   // srradint.h:1046-1050: manual "faster" below 10000 points
    const double* cx = X + tr * ldT; const double* cz = Y + tr * ldT;
    const double* ix = IBx + tr * ldT; const double* iz = IBz + tr * ldT;
    double bmx = -1. / I.betaNorm, bmz = 1. / I.betaNorm;
    auto eval = [&](long long i, double& s, double& Xv, double& Zv, double& Btx, double& Btz, double& IB) {
        s = int_point_s(I, i);
        double IBxv, IBzv;
        if (I.verNZ) { trj_pt p = eval_plane(I, cx, ix, bmx, s, byDiv); Xv = p.crd; Btx = p.bt; IBxv = p.ib; }
        else { double d = s - I.s0; Btx = I.dxds0; Xv = I.x0 + I.dxds0 * d; IBxv = I.dxds0 * I.dxds0 * d; }
        if (I.horNZ) { trj_pt p = eval_plane(I, cz, iz, bmz, s, byDiv); Zv = p.crd; Btz = p.bt; IBzv = p.ib; }
        else { double d = s - I.s0; Btz = I.dzds0; Zv = I.z0 + I.dzds0 * d; IBzv = I.dzds0 * I.dzds0 * d; }
        IB = IBxv + IBzv;
    };
    double s, Xv, Zv, Btx, Btz, IB;
    eval(gi, s, Xv, Zv, Btx, Btz, IB);
    double inv = 1. / (I.yObs - s);
    double U = Xv * inv, V = Zv * inv;
    double Q = s * I.gammaEm2 + IB + (Xv * Xv + Zv * Zv) * inv;
    long long c0 = (gi / SRWD_CHUNK) * SRWD_CHUNK;
    double invC, UC, VC;
    if (c0 == gi) { invC = inv; UC = U; VC = V; }
    else {
        double s2, X2, Z2, bt1, bt2, ib2;
        eval(c0, s2, X2, Z2, bt1, bt2, ib2);
        invC = 1. / (I.yObs - s2); UC = X2 * invC; VC = Z2 * invC;
    }
    // composite-rule weight in units of h/15 (srradint.cpp:1576-1588)
    double w = (gi == 0 || gi == I.nInt - 1) ? 7. : ((gi & 1) ? 16. : 14.);
    // w*A_x = b*(Px - xo) with b = w*inv^2 and the projected position
    // Px = X + Btx*(yObs - s); the record keeps b and b*Px (products in FP64)
    srwd_rec R;
    R.dInv = (float)(inv - invC);
    R.dU = (float)(U - UC);
    R.dV = (float)(V - VC);
    double bd = w * inv * inv, Pxd = Xv + Btx * (I.yObs - s), Pzd = Zv + Btz * (I.yObs - s);
    R.pad0 = 0.f;
    R.pad1 = 0.f;
    R.b = (float)bd;
    R.bPx = (float)(bd * Pxd);
    R.bPz = (float)(bd * Pzd);
    rec[tr * ldI + gi] = R;
    Qout[tr * ldI + gi] = Q;
    // FP64 trajectory values at the point, for the FP64 recomputation of
    // ill-conditioned points (k_sr_fp64)
    if (T64) {
        double* t = T64 + (tr * ldI + gi) * SRWD_T64;
        t[0] = Xv; t[1] = Zv; t[2] = Btx; t[3] = Btz; t[4] = s * I.gammaEm2 + IB; t[5] = inv;
    }
    if (c0 == gi) {
        srwd_chunk C; C.inv = inv; C.U = U; C.V = V; C.pad = 0;
        chunk[tr * ldC + gi / SRWD_CHUNK] = C;
    }
}

// Edge data at the integration limits (srtrjdat.h:1006-1034).  One thread per (trajectory, end).
__global__ void k_sr_edges(const srwd_trj_info* __restrict__ info, int nTraj, const double* X, const double* Y,
                           const double* IBx, const double* IBz, long long ldT, srwd_edge* edges)
{
    int t = blockIdx.x * blockDim.x + threadIdx.x;
    int tr = t >> 1, end = t & 1;
    if (tr >= nTraj) return;
    const srwd_trj_info I = info[tr];
    double s = end ? I.sIntegFin : I.sIntegStart;
    trj_pt px = eval_plane(I, X + tr * ldT, IBx + tr * ldT, -1. / I.betaNorm, s, true);
    trj_pt pz = eval_plane(I, Y + tr * ldT, IBz + tr * ldT, 1. / I.betaNorm, s, true);
    double* v = edges[tr].v[end];
    v[0] = px.dbds; v[1] = px.b; v[2] = px.bt; v[3] = px.crd; v[4] = px.ib;
    v[5] = pz.dbds; v[6] = pz.b; v[7] = pz.bt; v[8] = pz.crd; v[9] = pz.ib;
}

// Phase table A[row][i] = frac(K'*Q_i) with K' the phase constant in
// revolutions per metre for the row's photon energy.
__global__ void __launch_bounds__(256) k_phase_table(const double* __restrict__ Q, long long ldI,
                                                     const int* __restrict__ rowTraj,
                                                     const double* __restrict__ rowE,
                                                     const long long* __restrict__ rowN, float* A,
                                                     long long maxInt)
{
    long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    int row = blockIdx.y;
    if (i >= maxInt || i >= rowN[row]) return;
    double k = SRWD_KPH(rowE[row]) * (0.5 / M_PI);
    double t = k * Q[rowTraj[row] * ldI + i];
    A[row * ldI + i] = (float)(t - rint(t));
}

// ------------------------------------------------------------------------
// SR integral main kernel
// ------------------------------------------------------------------------
// Block = 256 observation points of one item; blockIdx.y = split of the
// integration range.  Per-point data of the trajectory are uniform across the
// block and are read with scalar loads.
// The phase of point (x_o, z_o) at s is Tc + A[s] + alpha dInv + beta dU +
// gam dV (revolutions), with Tc the point's FP64 phase at the chunk reference
// reduced to [-1/2, 1/2]; the loop leaves Tc out and each chunk sum is turned
// by exp(2 pi i Tc) once.
#define SRWD_BLOCK 64
#define SRWD_ROWS 4

// Kahan summation step: sum += v, with c holding the negated low-order part
// lost so far (the compensated total is sum - c).
__device__ __forceinline__ void kahan_add(float& sum, float& c, float v)
{
    float y = v - c;
    float t = sum + y;
    c = (t - sum) - y;
    sum = t;
}
static_assert(SRWD_CHUNK == 256, "k_sr_integral stages one chunk per 256-thread block");

__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(8, 8))) k_sr_integral(const srwd_item* __restrict__ items,
                                                     const int2* __restrict__ blocks,
                                                     const srwd_trj_info* __restrict__ info,
                                                     const srwd_rec* __restrict__ rec,
                                                     const srwd_chunk* __restrict__ chunks,
                                                     const float* __restrict__ A,
                                                     long long ldI, long long ldC, int nSplit,
                                                     long long nPointsPad, double* __restrict__ part)
{
    const int2 bk = blocks[blockIdx.x];
    const srwd_item& IT = items[bk.x];
    const int local = bk.y * 256 + threadIdx.x;
    const long long flat = IT.pointBase + local;
    const bool valid = local < IT.npts;
    const srwd_trj_info& I = info[IT.traj];
    int ix = valid ? local % IT.nxc : 0, iz = valid ? IT.iz0 + local / IT.nxc : 0;
    double xo = IT.xStart + ix * IT.xStep, zo = IT.zStart + iz * IT.zStep;
    double kph = SRWD_KPH(IT.eph) * (0.5 / M_PI);
    double r2 = xo * xo + zo * zo;
    float alpha = (float)(kph * r2), beta = (float)(-2. * kph * xo), gam = (float)(-2. * kph * zo);
    const float fxo = (float)xo, fzo = (float)zo;
    long long nInt = I.nInt;
    long long nChunk = (nInt + SRWD_CHUNK - 1) / SRWD_CHUNK;
    long long cBeg = (nChunk * blockIdx.y) / nSplit, cEnd = (nChunk * (blockIdx.y + 1)) / nSplit;
    const srwd_rec* R = rec + (long long)IT.traj * ldI;
    const srwd_chunk* CH = chunks + (long long)IT.traj * ldC;
    const float* Ar = A + (long long)IT.arow * ldI;
    double ox = 0., oxi = 0., oz = 0., ozi = 0.;
    // Per chunk the block stages in LDS, for each of its points s, the data
    // its lanes combine with their own x_o, z_o (blockDim.x == SRWD_CHUNK):
    // * when the block's points lie on at most SRWD_ROWS mesh rows, one float4
    //   per (row, s) = {A[s] + gam_row dV[s], b, b Px, b Pz}: the row's part of
    //   the phase is formed once per block instead of once per lane;
    // * otherwise one float4 per s = {A[s], b, b Px, b Pz}.
    __shared__ float4 sA[SRWD_ROWS][SRWD_CHUNK];
    const int first = bk.y * 256;
    const int last = min(first + 255, IT.npts - 1);
    const int rLo = IT.iz0 + first / IT.nxc, nRows = IT.iz0 + last / IT.nxc - rLo + 1;
    const bool byRow = nRows <= SRWD_ROWS;
    const int rs = byRow && valid ? iz - rLo : 0;
    const float nfxo = -fxo, nfzo = -fzo;
    for (long long c = cBeg; c < cEnd; c++) {
        const srwd_chunk C = CH[c];
        double tc = kph * (r2 * C.inv - 2. * xo * C.U - 2. * zo * C.V);
        float Tc = (float)(tc - rint(tc));
        long long i0 = c * SRWD_CHUNK, i1 = i0 + SRWD_CHUNK;
        if (i1 > nInt) i1 = nInt;
        const int n = (int)(i1 - i0);
        const srwd_rec* Rc = R + i0;
        __syncthreads();
        if ((int)threadIdx.x < n) {
            const srwd_rec& q = Rc[threadIdx.x];
            const float a = Ar[i0 + threadIdx.x];
            if (byRow) {
                for (int rr = 0; rr < nRows; rr++) {
                    // the same float as the lane's gam for that row
                    float gr = (float)(-2. * kph * (IT.zStart + (rLo + rr) * IT.zStep));
                    sA[rr][threadIdx.x] = make_float4(fmaf(gr, q.dV, a), q.b, q.bPx, q.bPz);
                }
            } else {
                sA[0][threadIdx.x] = make_float4(a, q.b, q.bPx, q.bPz);
            }
        }
        __syncthreads();
        // chunk sums of b (P - r_o) exp(2 pi i (t - Tc)): 64-point FP32 block
        // sums added with Kahan compensation (c* carry the lost low parts),
        // then turned by exp(2 pi i Tc) into FP64
        float q0 = 0.f, q1 = 0.f, q2 = 0.f, q3 = 0.f;
        float c0 = 0.f, c1 = 0.f, c2 = 0.f, c3 = 0.f;
        const float4* S = sA[rs];
        for (int b0 = 0; b0 < n; b0 += SRWD_BLOCK) {
            float bx = 0.f, bxi = 0.f, bz = 0.f, bzi = 0.f;
            const int b1 = b0 + SRWD_BLOCK < n ? b0 + SRWD_BLOCK : n;
            if (byRow) {
#pragma unroll 4
                for (int k = b0; k < b1; k++) {
                    const srwd_rec r = Rc[k];        // scalar loads: dInv, dU
                    const float4 L4 = S[k];
                    float t = fmaf(alpha, r.dInv, L4.x);
                    t = fmaf(beta, r.dU, t);
                    float cs = __builtin_amdgcn_cosf(t), sn = __builtin_amdgcn_sinf(t);
                    float wax = fmaf(L4.y, nfxo, L4.z), waz = fmaf(L4.y, nfzo, L4.w);   // b (P - r_o)
                    bx = fmaf(wax, cs, bx); bxi = fmaf(wax, sn, bxi);
                    bz = fmaf(waz, cs, bz); bzi = fmaf(waz, sn, bzi);
                }
            } else {
#pragma unroll 4
                for (int k = b0; k < b1; k++) {
                    const srwd_rec r = Rc[k];        // scalar loads: dInv, dU, dV
                    const float4 L4 = S[k];
                    float t = fmaf(alpha, r.dInv, L4.x);
                    t = fmaf(beta, r.dU, t);
                    t = fmaf(gam, r.dV, t);
                    float cs = __builtin_amdgcn_cosf(t), sn = __builtin_amdgcn_sinf(t);
                    float wax = fmaf(L4.y, nfxo, L4.z), waz = fmaf(L4.y, nfzo, L4.w);
                    bx = fmaf(wax, cs, bx); bxi = fmaf(wax, sn, bxi);
                    bz = fmaf(waz, cs, bz); bzi = fmaf(waz, sn, bzi);
                }
            }
            kahan_add(q0, c0, bx); kahan_add(q1, c1, bxi); kahan_add(q2, c2, bz); kahan_add(q3, c3, bzi);
        }
        double rxr = (double)q0 - (double)c0, rxi = (double)q1 - (double)c1;
        double rzr = (double)q2 - (double)c2, rzi = (double)q3 - (double)c3;
        double cT = __builtin_amdgcn_cosf(Tc), sT = __builtin_amdgcn_sinf(Tc);
        ox += rxr * cT - rxi * sT; oxi += rxr * sT + rxi * cT;
        oz += rzr * cT - rzi * sT; ozi += rzr * sT + rzi * cT;
    }
    if (!valid) return;
    double* o = part + ((long long)blockIdx.y * nPointsPad + flat) * 4;
    o[0] = ox; o[1] = oxi; o[2] = oz; o[3] = ozi;
}

// Terminating terms, condition check and stores of one point (defined with
// the finalize pass below).
__device__ void sr_finish_point(const srwd_item& IT, const srwd_trj_info& I, const srwd_edge& E, int termin,
                                int ix, int iz, const double* S, float2* Ex, float2* Ez, int* warn,
                                double sumSqX, double sumSqZ, double condTol, int* fpList, int* fpCount,
                                int listEntry, const srwd_req& Rq, const double* coord, const long long* last,
                                srwd_patch* patch = nullptr, int* patchCount = nullptr, int patchCap = 0);

// The same integral for SRWD_TILE_E consecutive photon energies of one
// request at once.  The phase of energy k is linear in k (K_k = K_0 + k dK,
// and the phase table rows obey A_k = A_0 + k A_dK mod 1), so per s a lane
// forms two phases, t_0 and the step d, and gets the phasors z_k =
// exp(2 pi i (t_0 + k d)) in chains of four: z_h and z_{h+1} from the
// hardware sine and cosine, z_{h+2}, z_{h+3} from the three-term recurrence
// z_{k+1} = 2 cos(2 pi d) z_k - z_{k-1} (2 FMAs each).  A rounding error in
// the recurrence grows at most linearly along a chain (as sin(k d)/sin(d)),
// so short chains keep it near the FP32 rounding of the phase itself.
// 64-point FP32 block sums are turned by each energy's FP64 chunk phase and
// accumulated in FP64.
// FUSED: each block then finishes its points itself (terminating terms,
// condition check, stores; the finalize pass of the two-pass form) and, once
// its stores are visible to the system, sets its completion flag
// (F.blockDone[blockIdx.x] = F.epoch); nSplit is 1.
template <int E, bool FUSED>
__global__ void __launch_bounds__(256) k_sr_integral_tile(const srwd_item* __restrict__ items,
                                                          const int2* __restrict__ blocks,
                                                          const srwd_trj_info* __restrict__ info,
                                                          const srwd_rec* __restrict__ rec,
                                                          const srwd_chunk* __restrict__ chunks,
                                                          const float* __restrict__ A,
                                                          long long ldI, long long ldC, int nSplit,
                                                          long long nPointsPad, double* __restrict__ part,
                                                          const srwd_fuse F)
{
    static_assert(E % 4 == 0, "energy tiles are chains of 4");
    const int2 bk = blocks[blockIdx.x];
    const srwd_item& IT = items[bk.x];
    const int local = bk.y * 256 + threadIdx.x;
    const bool valid = local < IT.npts;
    const srwd_trj_info& I = info[IT.traj];
    int ix = valid ? local % IT.nxc : 0, iz = valid ? IT.iz0 + local / IT.nxc : 0;
    double xo = IT.xStart + ix * IT.xStep, zo = IT.zStart + iz * IT.zStep;
    const double k0 = SRWD_KPH(IT.eph) * (0.5 / M_PI), kd = SRWD_KPH(IT.de) * (0.5 / M_PI);
    double r2 = xo * xo + zo * zo;
    const float a0 = (float)(k0 * r2), b0 = (float)(-2. * k0 * xo), g0 = (float)(-2. * k0 * zo);
    const float ad = (float)(kd * r2), bd = (float)(-2. * kd * xo), gd = (float)(-2. * kd * zo);
    const float nfxo = -(float)xo, nfzo = -(float)zo;
    long long nInt = I.nInt;
    long long nChunk = (nInt + SRWD_CHUNK - 1) / SRWD_CHUNK;
    long long cBeg = (nChunk * blockIdx.y) / nSplit, cEnd = (nChunk * (blockIdx.y + 1)) / nSplit;
    const srwd_rec* R = rec + (long long)IT.traj * ldI;
    const srwd_chunk* CH = chunks + (long long)IT.traj * ldC;
    const float* A0 = A + (long long)IT.arow * ldI;
    const float* Ad = A + (long long)IT.drow * ldI;
    double out[E][4];
#pragma unroll
    for (int e = 0; e < E; e++) out[e][0] = out[e][1] = out[e][2] = out[e][3] = 0.;
    // per (row, s): {A_0 + g0_row dV, A_dK + gd_row dV, b Px, b Pz} (rows as in
    // k_sr_integral), and b
    __shared__ float4 sA[SRWD_ROWS][SRWD_CHUNK];
    __shared__ float sB[SRWD_CHUNK];
    __shared__ float2 sRot[E][SRWD_CHUNK];
    const int first = bk.y * 256;
    const int last = min(first + 255, IT.npts - 1);
    const int rLo = IT.iz0 + first / IT.nxc, nRows = IT.iz0 + last / IT.nxc - rLo + 1;
    const bool byRow = nRows <= SRWD_ROWS;
    const int rs = byRow && valid ? iz - rLo : 0;
    for (long long c = cBeg; c < cEnd; c++) {
        const srwd_chunk C = CH[c];
        const double qc = r2 * C.inv - 2. * xo * C.U - 2. * zo * C.V;
        long long i0 = c * SRWD_CHUNK, i1 = i0 + SRWD_CHUNK;
        if (i1 > nInt) i1 = nInt;
        const int n = (int)(i1 - i0);
        const srwd_rec* Rc = R + i0;
        __syncthreads();
        // each energy's chunk phase, exactly (FP64), as a rotation (kept in
        // LDS, per lane, to save registers)
#pragma unroll
        for (int e = 0; e < E; e++) {
            double tc = SRWD_KPH(items[bk.x + e].eph) * (0.5 / M_PI) * qc;
            float Tc = (float)(tc - rint(tc));
            sRot[e][threadIdx.x] = make_float2(__builtin_amdgcn_cosf(Tc), __builtin_amdgcn_sinf(Tc));
        }
        if ((int)threadIdx.x < n) {
            const srwd_rec& q = Rc[threadIdx.x];
            const float u0 = A0[i0 + threadIdx.x], ud = Ad[i0 + threadIdx.x];
            sB[threadIdx.x] = q.b;
            if (byRow) {
                for (int rr = 0; rr < nRows; rr++) {
                    // the same floats as the lanes' g0, gd for that row
                    const double zr = IT.zStart + (rLo + rr) * IT.zStep;
                    const float g0r = (float)(-2. * k0 * zr), gdr = (float)(-2. * kd * zr);
                    sA[rr][threadIdx.x] = make_float4(fmaf(g0r, q.dV, u0), fmaf(gdr, q.dV, ud), q.bPx, q.bPz);
                }
            } else {
                sA[0][threadIdx.x] = make_float4(u0, ud, q.bPx, q.bPz);
            }
        }
        __syncthreads();
        const float4* S = sA[rs];
        for (int bb = 0; bb < n; bb += SRWD_BLOCK) {
            float acc[E][4];
#pragma unroll
            for (int e = 0; e < E; e++) acc[e][0] = acc[e][1] = acc[e][2] = acc[e][3] = 0.f;
            const int b1 = bb + SRWD_BLOCK < n ? bb + SRWD_BLOCK : n;
            // one evaluation of E energies at s index k; the dV terms are
            // folded into the staged rows when byRow
            auto eval = [&](int k, bool rowTerms) {
                const srwd_rec r = Rc[k];        // scalar loads: dInv, dU, dV
                const float4 L4 = S[k];
                const float b = sB[k];
                float t = fmaf(a0, r.dInv, L4.x), d = fmaf(ad, r.dInv, L4.y);
                t = fmaf(b0, r.dU, t);
                d = fmaf(bd, r.dU, d);
                if (rowTerms) {
                    t = fmaf(g0, r.dV, t);
                    d = fmaf(gd, r.dV, d);
                }
                const float c2 = 2.f * __builtin_amdgcn_cosf(d);
                const float wax = fmaf(b, nfxo, L4.z), waz = fmaf(b, nfzo, L4.w);   // b (P - r_o)
                // chains of 4 energies: the first two phasors of a chain
                // directly, the others by z_{k+1} = 2 cos(2 pi d) z_k - z_{k-1}
#pragma unroll
                for (int h = 0; h < E; h += 4) {
                    const float th = h == 0 ? t : fmaf((float)h, d, t), th1 = th + d;
                    float pr = __builtin_amdgcn_cosf(th), pi = __builtin_amdgcn_sinf(th);
                    float qr = __builtin_amdgcn_cosf(th1), qi = __builtin_amdgcn_sinf(th1);
#pragma unroll
                    for (int k = 0; k < 4; k++) {
                        const int e = h + k;
                        if (k >= 2) {
                            const float nr = fmaf(c2, qr, -pr), ni = fmaf(c2, qi, -pi);
                            pr = qr; pi = qi;
                            qr = nr; qi = ni;
                        }
                        const float zr = k == 0 ? pr : qr, zi = k == 0 ? pi : qi;
                        acc[e][0] = fmaf(wax, zr, acc[e][0]);
                        acc[e][1] = fmaf(wax, zi, acc[e][1]);
                        acc[e][2] = fmaf(waz, zr, acc[e][2]);
                        acc[e][3] = fmaf(waz, zi, acc[e][3]);
                    }
                }
            };
            // four s points in flight per wave (more independent work
            // between the dependent phase, sine and recurrence steps at two
            // waves per SIMD)
            if (byRow) {
#pragma unroll 4
                for (int k = bb; k < b1; k++) eval(k, false);
            } else {
#pragma unroll 2
                for (int k = bb; k < b1; k++) eval(k, true);
            }
#pragma unroll
            for (int e = 0; e < E; e++) {
                const float2 rt = sRot[e][threadIdx.x];
                const double ct = rt.x, st = rt.y;
                const double xr = acc[e][0], xi = acc[e][1], zr = acc[e][2], zi = acc[e][3];
                out[e][0] += xr * ct - xi * st; out[e][1] += xr * st + xi * ct;
                out[e][2] += zr * ct - zi * st; out[e][3] += zr * st + zi * ct;
            }
        }
    }
    if (valid) {
#pragma unroll
        for (int e = 0; e < E; e++) {
            double* o = part + ((long long)blockIdx.y * nPointsPad + items[bk.x + e].pointBase + local) * 4;
            o[0] = out[e][0]; o[1] = out[e][1]; o[2] = out[e][2]; o[3] = out[e][3];
        }
    }
    if constexpr (FUSED) {
        if (valid) {
#pragma unroll 1
            for (int e = 0; e < E; e++) {
                const srwd_item& Ie = items[bk.x + e];
                const double* pp = part + (Ie.pointBase + local) * 4;   // this thread's own sums
                const double S[4] = {pp[0], pp[1], pp[2], pp[3]};
                const srwd_req Rq = F.reqs[Ie.traj];
                double sumSqX = 0., sumSqZ = 0.;
                if (F.cond) {
                    const double* c = F.cond + 5 * Ie.traj;
                    const double xr = F.coord[Rq.xOff + ix], zr = F.coord[Rq.zOff + iz];
                    sumSqX = c[0] * xr * xr - 2. * xr * c[1] + c[3];
                    sumSqZ = c[0] * zr * zr - 2. * zr * c[2] + c[4];
                }
                sr_finish_point(Ie, info[Ie.traj], F.edges[Ie.traj], F.termin, ix, iz, S, F.Ex, F.Ez, F.warn, sumSqX,
                                sumSqZ, F.condTol, F.cond ? F.fpList : nullptr, F.fpCount,
                                ((Ie.blockBase + bk.y) << 8) | (int)threadIdx.x, Rq, F.coord, F.last);
            }
        }
        __syncthreads();
        if (threadIdx.x == 0) {
            // the block's stores reach memory before its flag says done
            __threadfence_system();
            __hip_atomic_store(F.blockDone + blockIdx.x, F.epoch, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_SYSTEM);
        }
    }
}

// ------------------------------------------------------------------------
// Finalisation: partial sums, terminating terms (srw_sr_dev.h), symmetry
// ------------------------------------------------------------------------
// Error model of the FP32 sums.  Each term's phase is formed in FP32 within
// two revolutions of its chunk reference and its sine and cosine come from
// the hardware FP32 instructions, so a term carries an error of about
// SRWD_EPS_TERM times its magnitude; over many s points these add like a
// random walk, to SRWD_EPS_TERM * sqrt(sum of squared terms).  A point where
// that estimate exceeds condTol times the magnitude of its field (both
// components, terminating terms included) is recomputed in FP64 (k_sr_fp64):
// the points near the nodes of the pattern, where the sums cancel.
// Calibrated against FP64 recomputation of every point (201 x 201 meshes at
// the ends and the centre of the +/-5 % band of the first harmonic; 48
// requests of harmonics 1-5, K 0.5-2.5): an estimate built on 1e-6 still
// bounded every measured error, 2e-6 leaves a factor 2.4 of margin.
#define SRWD_EPS_TERM 2e-6
// Energy-tile items: the phasor of energy k carries k times the rounding of
// the step's phase and the recurrence's roundings on top.
#define SRWD_EPS_TERM_TILE 4e-6

// Terminating terms, the derivative correction, the float store and the
// mirrored parts of one point; S = sum over the integration points of
// w_i * (A_x, A_z) * exp(i phi) with w_i in units of h/15.
// Rq, coord: the point's request (observation coordinates as the reference's
// loop forms them); last: the edge-derivative state table (nullptr without
// terminating terms).
__device__ void sr_finish_point(const srwd_item& IT, const srwd_trj_info& I, const srwd_edge& E, int termin,
                                int ix, int iz, const double* S, float2* Ex, float2* Ez, int* warn,
                                double sumSqX, double sumSqZ, double condTol, int* fpList, int* fpCount,
                                int listEntry, const srwd_req& Rq, const double* coord, const long long* last,
                                srwd_patch* patch, int* patchCount, int patchCap)
{
    double xo = coord[Rq.xOff + ix], zo = coord[Rq.zOff + iz];
    double h = I.h;
    double w = IT.actNorm * h / 15.;
    cplx ex = {w * S[0], w * S[1]}, ez = {w * S[2], w * S[3]};
    if (termin > 0) {
        cplx RL[2] = {{0, 0}, {0, 0}}, RR[2] = {{0, 0}, {0, 0}}, DL[2] = {{0, 0}, {0, 0}}, DR[2] = {{0, 0}, {0, 0}};
        bool bad = false;
        const long long p = ((long long)iz * IT.nxc + ix) * IT.ne + IT.ie;
        if (termin == 1 || termin == 2) {
            int r = residual(I, E.v[0], I.sIntegStart, xo, I.yObs, zo, IT.eph, IT.actNorm, RL, DL, true, 1e-4);
            bad |= r == SRWD_RESID_BAD;
            if (r != SRWD_RESID_OK) carried_derivative(I, E.v[0], I.sIntegStart, I.yObs, Rq, coord, last, 0, p, DL);
        }
        if (termin == 1 || termin == 3) {
            int r = residual(I, E.v[1], I.sIntegFin, xo, I.yObs, zo, IT.eph, IT.actNorm, RR, DR);
            bad |= r == SRWD_RESID_BAD;
            if (r != SRWD_RESID_OK) carried_derivative(I, E.v[1], I.sIntegFin, I.yObs, Rq, coord, last, 1, p, DR);
        }
        if (bad) atomicOr(warn, 1);
        double wd = IT.actNorm * h * h / 15.;
        ex.re += RR[0].re - RL[0].re + wd * (DL[0].re - DR[0].re);
        ex.im += RR[0].im - RL[0].im + wd * (DL[0].im - DR[0].im);
        ez.re += RR[1].re - RL[1].re + wd * (DL[1].re - DR[1].re);
        ez.im += RR[1].im - RL[1].im + wd * (DL[1].im - DR[1].im);
    }
    if (fpList) {
        // condition check of the FP32 sums against the point's field
        // magnitude: hand the point to the FP64 pass
        const double eps = IT.tiled ? SRWD_EPS_TERM_TILE : SRWD_EPS_TERM;
        const double e2 = eps * eps * w * w, t2 = condTol * condTol;
        const double mag2 = ex.re * ex.re + ex.im * ex.im + ez.re * ez.re + ez.im * ez.im;
        if (e2 * (fmax(sumSqX, 0.) + fmax(sumSqZ, 0.)) > t2 * mag2) {
            int k = atomicAdd(fpCount, 1);
            fpList[k] = listEntry;
            return;
        }
    }
    float2 vx = make_float2((float)ex.re, (float)ex.im), vz = make_float2((float)ez.re, (float)ez.im);
    long long ne = IT.ne;
    auto put = [&](int jx, int jz, float sx, float sz) {
        long long o = IT.outOffset + ((long long)jz * IT.nx + jx) * ne;
        const float2 a = make_float2(sx * vx.x, sx * vx.y), c = make_float2(sz * vz.x, sz * vz.y);
        Ex[o] = a;
        Ez[o] = c;
        if (patch) {
            const int k = atomicAdd(patchCount, 1);
            if (k < patchCap) patch[k] = srwd_patch{o, a, c};
        }
    };
    put(ix, iz, 1.f, 1.f);
    // symmetric parts (srradint.cpp:3264-3350): mirror in z flips Ez, in x flips Ex
    int hz = IT.nz >> 1, hx = IT.nx >> 1;
    bool mz = IT.symZ && iz < hz, mx = IT.symX && ix < hx;
    if (mz) put(ix, IT.nz - 1 - iz, 1.f, -1.f);
    if (mx) put(IT.nx - 1 - ix, iz, -1.f, 1.f);
    if (mz && mx) put(IT.nx - 1 - ix, IT.nz - 1 - iz, -1.f, -1.f);
}

// Sums of squared terms per trajectory, from the integration records:
// cond[0..4] = sum b^2, sum b bPx, sum b bPz, sum bPx^2, sum bPz^2, so that
// the squared x terms of point (xo, zo) add up to
// cond[0] xo^2 - 2 xo cond[1] + cond[3] (and the z terms likewise).
__global__ void __launch_bounds__(256) k_cond_sums(const srwd_trj_info* __restrict__ info,
                                                   const srwd_rec* __restrict__ rec, long long ldI, double* cond)
{
    int tr = blockIdx.x;
    long long n = info[tr].nInt;
    const srwd_rec* R = rec + (long long)tr * ldI;
    double v[5] = {0, 0, 0, 0, 0};
    for (long long i = threadIdx.x; i < n; i += blockDim.x) {
        double bb = R[i].b, px = R[i].bPx, pz = R[i].bPz;
        v[0] += bb * bb; v[1] += bb * px; v[2] += bb * pz; v[3] += px * px; v[4] += pz * pz;
    }
    __shared__ double sh[5][256];
    for (int k = 0; k < 5; k++) sh[k][threadIdx.x] = v[k];
    __syncthreads();
    for (int off = 128; off > 0; off >>= 1) {
        if ((int)threadIdx.x < off)
            for (int k = 0; k < 5; k++) sh[k][threadIdx.x] += sh[k][threadIdx.x + off];
        __syncthreads();
    }
    if (threadIdx.x < 5) cond[5 * tr + threadIdx.x] = sh[threadIdx.x][0];
}

// One workgroup per (entry, part): an entry is (first item, 256-point block)
// of a group of g items (consecutive photon energies of one request: the
// energy tiles, g = SRWD_TILE_E; or single items, g = 1).  Lane t takes
// energy t % g of point t / g of the part, so that a wave stores runs of g
// consecutive energies (the field is stored photon energy fastest).
__global__ void __launch_bounds__(256) k_sr_finalize(const srwd_item* __restrict__ items,
                                                     const int2* __restrict__ entries, int g,
                                                     const srwd_trj_info* __restrict__ info,
                                                     const srwd_edge* __restrict__ edges, int termin,
                                                     const double* __restrict__ part, int nSplit,
                                                     long long nPointsPad, float2* Ex, float2* Ez, int* warn,
                                                     const double* __restrict__ cond, double condTol, int* fpList,
                                                     int* fpCount, const srwd_req* __restrict__ reqs,
                                                     const double* __restrict__ coord,
                                                     const long long* __restrict__ last)
{
    const int2 bk = entries[blockIdx.x / g];
    const int q = blockIdx.x % g, e = threadIdx.x % g;
    const int p = q * (256 / g) + threadIdx.x / g;   // point of the 256-point block
    const srwd_item& IT = items[bk.x + e];
    const int local = bk.y * 256 + p;
    if (local >= IT.npts) return;
    const long long flat = IT.pointBase + local;
    const srwd_trj_info& I = info[IT.traj];
    int ix = local % IT.nxc, iz = IT.iz0 + local / IT.nxc;
    double S[4] = {0, 0, 0, 0};
    for (int k = 0; k < nSplit; k++) {
        const double* pp = part + ((long long)k * nPointsPad + flat) * 4;
        S[0] += pp[0]; S[1] += pp[1]; S[2] += pp[2]; S[3] += pp[3];
    }
    const srwd_req Rq = reqs[IT.traj];
    double sumSqX = 0., sumSqZ = 0.;
    if (cond) {
        const double* c = cond + 5 * IT.traj;
        double xo = coord[Rq.xOff + ix], zo = coord[Rq.zOff + iz];
        sumSqX = c[0] * xo * xo - 2. * xo * c[1] + c[3];
        sumSqZ = c[0] * zo * zo - 2. * zo * c[2] + c[4];
    }
    // list entry as the FP64 pass decodes it: the item's band-relative block and the point in it
    sr_finish_point(IT, I, edges[IT.traj], termin, ix, iz, S, Ex, Ez, warn, sumSqX, sumSqZ, condTol,
                    cond ? fpList : nullptr, fpCount, ((IT.blockBase + bk.y) << 8) | p, Rq, coord, last);
}

// FP64 sums of one point over the integration points of its trajectory
// table T (n entries of SRWD_T64 doubles): the trajectory interpolated at
// each point as the reference's manual methods do (srradint.cpp:1214-1258,
// 1379-1604), the phase reduced and its sine and cosine taken with the
// reference's Taylor series, the weights 7, 16, 14, ..., 16, 7 in units of
// h/15.  With T and n uniform over the wave the table is read by scalar
// loads, once per wave.
// List length from which k_sr_fp64 gives each point one lane (below it, one
// wave): about where a serial walk of 10^4 integration points per lane costs
// as much as the per-point table reads of the wave-per-point form.
#define SRWD_FP64_LANE_MIN 32768

// Branch-free form of cos_sin_taylor (the same operations on the taken path,
// so the same results).
__device__ __forceinline__ void cos_sin_taylor_sel(double x, double& c, double& s)
{
    const double PI = 3.141592653590, TwoPI = 2. * PI, ThreePIdTwo = 1.5 * PI, HalfPI = 0.5 * PI;
    const double One_dTwoPI = 1. / TwoPI;
    const double a2c = -0.5, a4c = 1. / 24., a6c = -1. / 720., a8c = 1. / 40320., a10c = -1. / 3628800.;
    const double a3s = -1. / 6., a5s = 1. / 120., a7s = -1. / 5040., a9s = 1. / 362880., a11s = -1. / 39916800.;
    x -= TwoPI * (double)(int)(x * One_dTwoPI);
    x = x < 0. ? x + TwoPI : x;
    const bool hi = x > ThreePIdTwo, chs = !hi && x > HalfPI;
    x = hi ? x - TwoPI : (chs ? x - PI : x);
    double x2 = x * x;
    c = 1. + x2 * (a2c + x2 * (a4c + x2 * (a6c + x2 * (a8c + x2 * a10c))));
    s = x * (1. + x2 * (a3s + x2 * (a5s + x2 * (a7s + x2 * (a9s + x2 * a11s)))));
    c = chs ? -c : c;
    s = chs ? -s : s;
}

template <typename TP>
__device__ __forceinline__ void fp64_point_sums(TP T, int n, double xo, double zo, double kph, double* acc)
{
#pragma unroll 2
    for (int i = 0; i < n; i++) {
        const int o = i * SRWD_T64;
        double Xv = T[o], Zv = T[o + 1], Btx = T[o + 2], Btz = T[o + 3], sQ = T[o + 4], inv = T[o + 5];
        double dx = xo - Xv, dz = zo - Zv;
        double ph = kph * (sQ + (dx * dx + dz * dz) * inv);
        double c, sn;
        cos_sin_taylor_sel(ph, c, sn);
        double wt = (i == 0 || i == n - 1) ? 7. : ((i & 1) ? 16. : 14.);
        double ax = wt * ((Btx - dx * inv) * inv), az = wt * ((Btz - dz * inv) * inv);
        acc[0] += ax * c; acc[1] += ax * sn; acc[2] += az * c; acc[3] += az * sn;
    }
}

__device__ __forceinline__ const double* wave_first_ptr(const double* p)
{
    const unsigned long long v = (unsigned long long)p;
    const unsigned lo = __builtin_amdgcn_readfirstlane((unsigned)v);
    const unsigned hi = __builtin_amdgcn_readfirstlane((unsigned)(v >> 32));
    return (const double*)(((unsigned long long)hi << 32) | lo);
}

// FP64 recomputation of the points the condition check handed over.  A fixed
// grid of waves walks the list (its length is read on the device).
// * Long lists: one lane per point, 64 consecutive list entries per wave.
//   Entries of one finalize wave are contiguous in the list and neighbouring
//   blocks belong to the same request, so a wave's points nearly always
//   share a trajectory; the table is then read once per wave (scalar loads)
//   instead of once per point.  A wave whose points span trajectories reads
//   per lane.
// * Short lists (fewer points than SRWD_FP64_LANE_MIN), where one lane per
//   point would leave the waves too few to hide the latency of a serial walk
//   over the trajectory: one wave per point, the integration points split
//   over its lanes.
__global__ void __launch_bounds__(256) k_sr_fp64(const srwd_item* __restrict__ items,
                                                 const int* __restrict__ blockItem,
                                                 const srwd_trj_info* __restrict__ info,
                                                 const srwd_edge* __restrict__ edges, int termin,
                                                 const double* __restrict__ T64, long long ldI, float2* Ex,
                                                 float2* Ez, int* warn, const int* __restrict__ fpList,
                                                 const int* __restrict__ fpCount, const srwd_req* __restrict__ reqs,
                                                 const double* __restrict__ coord,
                                                 const long long* __restrict__ last, const double* __restrict__ sums,
                                                 srwd_patch* patch, int* patchCount, int patchCap)
{
    const int lane = threadIdx.x & 63;
    const int nWaves = gridDim.x * (blockDim.x >> 6);
    const int count = *fpCount;
    if (count < SRWD_FP64_LANE_MIN) {
        for (int q = blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6); q < count; q += nWaves) {
            const int entry = fpList[q];
            const int blk = entry >> 8, thr = entry & 255;
            const srwd_item& IT = items[blockItem[blk]];
            const int local = (blk - IT.blockBase) * 256 + thr;
            const srwd_trj_info I = info[IT.traj];
            int ix = local % IT.nxc, iz = IT.iz0 + local / IT.nxc;
            const srwd_req& Rq = reqs[IT.traj];
            double xo = coord[Rq.xOff + ix], zo = coord[Rq.zOff + iz];
            double kph = SRWD_KPH(IT.eph);
            const double* T = T64 + (long long)IT.traj * ldI * SRWD_T64;
            double acc[4] = {0, 0, 0, 0};
            for (long long i = lane; i < I.nInt; i += 64) {
                const double* t = T + i * SRWD_T64;
                double Xv = t[0], Zv = t[1], Btx = t[2], Btz = t[3], sQ = t[4], inv = t[5];
                double dx = xo - Xv, dz = zo - Zv;
                double ph = kph * (sQ + (dx * dx + dz * dz) * inv);
                double c, sn;
                cos_sin_taylor(ph, c, sn);
                double wt = (i == 0 || i == I.nInt - 1) ? 7. : ((i & 1) ? 16. : 14.);
                double ax = wt * ((Btx - dx * inv) * inv), az = wt * ((Btz - dz * inv) * inv);
                acc[0] += ax * c; acc[1] += ax * sn; acc[2] += az * c; acc[3] += az * sn;
            }
            for (int m = 0; m < 4; m++)
                for (int off = 32; off > 0; off >>= 1) acc[m] += __shfl_down(acc[m], off);
            if (lane == 0)
                sr_finish_point(IT, I, edges[IT.traj], termin, ix, iz, acc, Ex, Ez, warn, 0., 0., 0., nullptr,
                                nullptr, 0, Rq, coord, last, patch, patchCount, patchCap);
        }
        return;
    }
    if (sums == nullptr) {
        // long list without a sums buffer: one lane per point, sums here
        for (int seg = blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6); seg * 64 < count; seg += nWaves) {
            const int q = seg * 64 + lane;
            if (q < count) {
                const int entry = fpList[q];
                const int blk = entry >> 8, thr = entry & 255;
                const srwd_item& IT = items[blockItem[blk]];
                const int local = (blk - IT.blockBase) * 256 + thr;
                const srwd_trj_info I = info[IT.traj];
                int ix = local % IT.nxc, iz = IT.iz0 + local / IT.nxc;
                const srwd_req& Rq = reqs[IT.traj];
                double xo = coord[Rq.xOff + ix], zo = coord[Rq.zOff + iz];
                double kph = SRWD_KPH(IT.eph);
                const double* T = T64 + (long long)IT.traj * ldI * SRWD_T64;
                const int n = (int)I.nInt;
                double acc[4] = {0, 0, 0, 0};
                if (__all(IT.traj == __builtin_amdgcn_readfirstlane(IT.traj)))
                    fp64_point_sums((const __attribute__((address_space(4))) double*)wave_first_ptr(T),
                                    __builtin_amdgcn_readfirstlane(n), xo, zo, kph, acc);
                else
                    fp64_point_sums(T, n, xo, zo, kph, acc);
                sr_finish_point(IT, I, edges[IT.traj], termin, ix, iz, acc, Ex, Ez, warn, 0., 0., 0., nullptr,
                                nullptr, 0, Rq, coord, last, patch, patchCount, patchCap);
            }
        }
        return;
    }
    // long lists: the sums come from k_sr_fp64_sums (in sums, 4 per entry)
    for (int q = blockIdx.x * blockDim.x + threadIdx.x; q < count; q += gridDim.x * blockDim.x) {
        const int entry = fpList[q];
        const int blk = entry >> 8, thr = entry & 255;
        const srwd_item& IT = items[blockItem[blk]];
        const int local = (blk - IT.blockBase) * 256 + thr;
        const srwd_trj_info I = info[IT.traj];
        int ix = local % IT.nxc, iz = IT.iz0 + local / IT.nxc;
        const srwd_req& Rq = reqs[IT.traj];
        const double acc[4] = {sums[4 * (long long)q], sums[4 * (long long)q + 1], sums[4 * (long long)q + 2],
                               sums[4 * (long long)q + 3]};
        sr_finish_point(IT, I, edges[IT.traj], termin, ix, iz, acc, Ex, Ez, warn, 0., 0., 0., nullptr, nullptr, 0,
                        Rq, coord, last, patch, patchCount, patchCap);
    }
}

// The integration sums of a long FP64 list, one lane per point, 64
// consecutive list entries per wave, into sums (4 per entry).  Entries of one
// finalize wave are contiguous in the list and neighbouring blocks belong to
// the same request, so a wave's points nearly always share a trajectory; the
// table is then read once per wave (scalar loads) instead of once per point.
// A wave whose points span trajectories reads per lane.  Kept apart from the
// terminating terms of k_sr_fp64 so that it runs with few registers.
__global__ void __launch_bounds__(256) k_sr_fp64_sums(const srwd_item* __restrict__ items,
                                                      const int* __restrict__ blockItem,
                                                      const srwd_trj_info* __restrict__ info,
                                                      const double* __restrict__ T64, long long ldI,
                                                      const int* __restrict__ fpList,
                                                      const int* __restrict__ fpCount,
                                                      const srwd_req* __restrict__ reqs,
                                                      const double* __restrict__ coord, double* __restrict__ sums)
{
    const int lane = threadIdx.x & 63;
    const int nWaves = gridDim.x * (blockDim.x >> 6);
    const int count = *fpCount;
    if (count < SRWD_FP64_LANE_MIN || sums == nullptr) return;
    for (int seg = blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6); seg * 64 < count; seg += nWaves) {
        const int q = seg * 64 + lane;
        if (q < count) {
            const int entry = fpList[q];
            const int blk = entry >> 8, thr = entry & 255;
            const srwd_item& IT = items[blockItem[blk]];
            const int local = (blk - IT.blockBase) * 256 + thr;
            const int traj = IT.traj;
            int ix = local % IT.nxc, iz = IT.iz0 + local / IT.nxc;
            const srwd_req& Rq = reqs[traj];
            double xo = coord[Rq.xOff + ix], zo = coord[Rq.zOff + iz];
            double kph = SRWD_KPH(IT.eph);
            const double* T = T64 + (long long)traj * ldI * SRWD_T64;
            const int n = (int)info[traj].nInt;
            double acc[4] = {0, 0, 0, 0};
            if (__all(traj == __builtin_amdgcn_readfirstlane(traj)))
                fp64_point_sums((const __attribute__((address_space(4))) double*)wave_first_ptr(T),
                                __builtin_amdgcn_readfirstlane(n), xo, zo, kph, acc);
            else
                fp64_point_sums(T, n, xo, zo, kph, acc);
            double* o = sums + 4 * (long long)q;
            o[0] = acc[0]; o[1] = acc[1]; o[2] = acc[2]; o[3] = acc[3];
        }
    }
}

// ------------------------------------------------------------------------
// Edge-derivative state of the reference's point loop
// ------------------------------------------------------------------------
// The derivative of the integrand at an integration edge enters each point's
// integral through the end correction (srradint.cpp:1583-1586).  The
// reference's loop (rows, columns, photon energies of the computed region)
// sets it only at points where the far-field expansion at that edge is
// accepted (:841-846) and keeps the previous point's value elsewhere.
// k_edge_blocks: per block of SRWD_EDGE_BLOCK points and edge, the last
// accepted point of the block (-1: none).
__global__ void __launch_bounds__(SRWD_EDGE_BLOCK) k_edge_blocks(const srwd_req* __restrict__ reqs,
                                                                 const int* __restrict__ blkReq,
                                                                 const srwd_trj_info* __restrict__ info,
                                                                 const srwd_edge* __restrict__ edges,
                                                                 const double* __restrict__ coord, int termin,
                                                                 long long* last)
{
    const long long b = blockIdx.x;
    const int r = blkReq[b];
    const srwd_req Rq = reqs[r];
    const srwd_trj_info& I = info[r];
    const long long q = (b - Rq.blk0) * SRWD_EDGE_BLOCK + threadIdx.x;
    long long v[2] = {-1, -1};
    if (q < Rq.nPts) {
        double xo, zo, eph;
        req_point(Rq, coord, q, xo, zo, eph);
        edge_geo g;
        if (termin == 1 || termin == 2) {
            edge_geometry(I, edges[r].v[0], I.sIntegStart, xo, I.yObs, zo, eph, g);
            if (g.ok2) v[0] = q;
        }
        if (termin == 1 || termin == 3) {
            edge_geometry(I, edges[r].v[1], I.sIntegFin, xo, I.yObs, zo, eph, g);
            if (g.ok2) v[1] = q;
        }
    }
    __shared__ long long sh[2][SRWD_EDGE_BLOCK / 64];
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    for (int e = 0; e < 2; e++) {
        long long x = v[e];
        for (int off = 32; off > 0; off >>= 1) x = max(x, (long long)__shfl_xor(x, off, 64));
        if (lane == 0) sh[e][wv] = x;
    }
    __syncthreads();
    if (threadIdx.x < 2) {
        long long x = -1;
        for (int k = 0; k < SRWD_EDGE_BLOCK / 64; k++) x = max(x, sh[threadIdx.x][k]);
        last[2 * b + threadIdx.x] = x;
    }
}

// Running maximum over the blocks of each request (one workgroup per request):
// afterwards last[2b + e] is the last accepted point at or before the end of
// block b.
__global__ void __launch_bounds__(256) k_edge_scan(const srwd_req* __restrict__ reqs, long long* last)
{
    const srwd_req Rq = reqs[blockIdx.x];
    const long long nb = (Rq.nPts + SRWD_EDGE_BLOCK - 1) / SRWD_EDGE_BLOCK;
    __shared__ long long sh[2][256];
    long long carry[2] = {-1, -1};
    for (long long c0 = 0; c0 < nb; c0 += 256) {
        const long long b = c0 + threadIdx.x;
        for (int e = 0; e < 2; e++) sh[e][threadIdx.x] = b < nb ? last[2 * (Rq.blk0 + b) + e] : -1;
        __syncthreads();
        for (int off = 1; off < 256; off <<= 1) {
            long long t0 = threadIdx.x >= (unsigned)off ? sh[0][threadIdx.x - off] : -1;
            long long t1 = threadIdx.x >= (unsigned)off ? sh[1][threadIdx.x - off] : -1;
            __syncthreads();
            sh[0][threadIdx.x] = max(sh[0][threadIdx.x], t0);
            sh[1][threadIdx.x] = max(sh[1][threadIdx.x], t1);
            __syncthreads();
        }
        if (b < nb)
            for (int e = 0; e < 2; e++) last[2 * (Rq.blk0 + b) + e] = max(carry[e], sh[e][threadIdx.x]);
        for (int e = 0; e < 2; e++) carry[e] = max(carry[e], sh[e][255]);
        __syncthreads();
    }
}

// ------------------------------------------------------------------------
// Launch wrappers (internal C ABI)
// ------------------------------------------------------------------------
extern "C" {

int srwd_launch_traj(const srwd_traj_job* dJobs, const srwd_field* dFields, int nTraj, double* dX,
                     double* dXp, double* dY, double* dYp, double* dZ, long long ld, double* dScratch,
                     int singleHarmonic, hipStream_t st)
{
    if (dScratch && ld >= 16) {   // picard_arrive needs buffers of at least 2 points
        if (singleHarmonic)
            hipLaunchKernelGGL(k_traj_picard<1>, dim3(nTraj, 2), dim3(SRWD_PIC_THREADS), 0, st, dJobs, dFields, dX, dXp,
                               dY, dYp, dZ, ld, dScratch, srwd_pic_cap(ld));
        else
            hipLaunchKernelGGL(k_traj_picard<0>, dim3(nTraj, 2), dim3(SRWD_PIC_THREADS), 0, st, dJobs, dFields, dX, dXp,
                               dY, dYp, dZ, ld, dScratch, srwd_pic_cap(ld));
    } else {
        int threads = 2 * nTraj;
        hipLaunchKernelGGL(k_traj_rk4, dim3((threads + 63) / 64), dim3(64), 0, st, dJobs, dFields, nTraj,
                           dX, dXp, dY, dYp, dZ, ld);
    }
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

int srwd_launch_tables(const srwd_trj_info* dInfo, int nTraj, long long maxInt, const double* dX,
                       const double* dXp, const double* dY, const double* dYp, long long ldT,
                       double* dIBx, double* dIBz, srwd_rec* dRec, srwd_chunk* dChunk, double* dQ,
                       long long ldI, long long ldC, srwd_edge* dEdge, double* dT64, hipStream_t st)
{
    hipLaunchKernelGGL(k_quad_terms, dim3(2 * nTraj), dim3(SRWD_QT_THREADS), 0, st, dInfo, dXp, dYp, ldT, dIBx,
                       dIBz);
    hipLaunchKernelGGL(k_sr_tables, dim3((unsigned)((maxInt + 255) / 256), nTraj), dim3(256), 0, st, dInfo,
                       dX, dY, dIBx, dIBz, ldT, dRec, dChunk, dQ, ldI, ldC, maxInt, dT64);
    hipLaunchKernelGGL(k_sr_edges, dim3((2 * nTraj + 63) / 64), dim3(64), 0, st, dInfo, nTraj, dX, dY,
                       dIBx, dIBz, ldT, dEdge);
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

int srwd_launch_phase_table(const double* dQ, long long ldI, const int* dRowTraj, const double* dRowE,
                            const long long* dRowN, int nRows, long long maxInt, float* dA, hipStream_t st)
{
    hipLaunchKernelGGL(k_phase_table, dim3((unsigned)((maxInt + 255) / 256), nRows), dim3(256), 0, st, dQ,
                       ldI, dRowTraj, dRowE, dRowN, dA, maxInt);
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

int srwd_launch_sr(const srwd_item* dItems, const int2* dBlocks, int nBlocks, int tiled, const srwd_trj_info* dInfo,
                   const srwd_rec* dRec, const srwd_chunk* dChunk, const float* dA, long long ldI,
                   long long ldC, int nSplit, long long nPointsPad, double* dPart, hipStream_t st)
{
    if (nBlocks <= 0) return 0;
    const srwd_fuse none{};
    if (tiled == 8)
        hipLaunchKernelGGL((k_sr_integral_tile<8, false>), dim3(nBlocks, nSplit), dim3(256), 0, st, dItems, dBlocks,
                           dInfo, dRec, dChunk, dA, ldI, ldC, nSplit, nPointsPad, dPart, none);
    else if (tiled == 4)
        hipLaunchKernelGGL((k_sr_integral_tile<4, false>), dim3(nBlocks, nSplit), dim3(256), 0, st, dItems, dBlocks,
                           dInfo, dRec, dChunk, dA, ldI, ldC, nSplit, nPointsPad, dPart, none);
    else if (tiled)
        return -1;
    else
        hipLaunchKernelGGL(k_sr_integral, dim3(nBlocks, nSplit), dim3(256), 0, st, dItems, dBlocks, dInfo, dRec,
                           dChunk, dA, ldI, ldC, nSplit, nPointsPad, dPart);
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

int srwd_launch_sr_fused(const srwd_item* dItems, const int2* dBlocks, int nBlocks, int tiled,
                         const srwd_trj_info* dInfo, const srwd_rec* dRec, const srwd_chunk* dChunk, const float* dA,
                         long long ldI, long long ldC, long long nPointsPad, double* dPart, const srwd_fuse& F,
                         hipStream_t st)
{
    if (nBlocks <= 0) return 0;
    if (tiled == 8)
        hipLaunchKernelGGL((k_sr_integral_tile<8, true>), dim3(nBlocks, 1), dim3(256), 0, st, dItems, dBlocks, dInfo,
                           dRec, dChunk, dA, ldI, ldC, 1, nPointsPad, dPart, F);
    else if (tiled == 4)
        hipLaunchKernelGGL((k_sr_integral_tile<4, true>), dim3(nBlocks, 1), dim3(256), 0, st, dItems, dBlocks, dInfo,
                           dRec, dChunk, dA, ldI, ldC, 1, nPointsPad, dPart, F);
    else
        return -1;
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

int srwd_launch_finalize(const srwd_item* dItems, const int2* dEntries, int nEntries, int g,
                         const srwd_trj_info* dInfo, const srwd_edge* dEdge, int termin, const double* dPart,
                         int nSplit, long long nPointsPad, float* dEx, float* dEz, int* dWarn, const double* dCond,
                         double condTol, int* dFpList, int* dFpCount, const srwd_req* dReq, const double* dCoord,
                         const long long* dLast, hipStream_t st)
{
    if (nEntries <= 0) return 0;
    if (g < 1 || 256 % g) return -1;
    hipLaunchKernelGGL(k_sr_finalize, dim3((unsigned)nEntries * g), dim3(256), 0, st, dItems, dEntries, g, dInfo,
                       dEdge, termin, dPart, nSplit, nPointsPad, (float2*)dEx, (float2*)dEz, dWarn, dCond, condTol,
                       dFpList, dFpCount, dReq, dCoord, dLast);
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

int srwd_launch_cond_sums(const srwd_trj_info* dInfo, int nTraj, const srwd_rec* dRec, long long ldI, double* dCond,
                          hipStream_t st)
{
    hipLaunchKernelGGL(k_cond_sums, dim3(nTraj), dim3(256), 0, st, dInfo, dRec, ldI, dCond);
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

int srwd_launch_sr_fp64(const srwd_item* dItems, const int* dBlockItem, const srwd_trj_info* dInfo,
                        const srwd_edge* dEdge, int termin, const double* dT64, long long ldI, float* dEx,
                        float* dEz, int* dWarn, const int* dFpList, const int* dFpCount, const srwd_req* dReq,
                        const double* dCoord, const long long* dLast, double* dSums, srwd_patch* dPatch,
                        int* dPatchCount, int patchCap, hipStream_t st)
{
    if (dSums)
        hipLaunchKernelGGL(k_sr_fp64_sums, dim3(2048), dim3(256), 0, st, dItems, dBlockItem, dInfo, dT64, ldI, dFpList,
                           dFpCount, dReq, dCoord, dSums);
    hipLaunchKernelGGL(k_sr_fp64, dim3(2048), dim3(256), 0, st, dItems, dBlockItem, dInfo, dEdge, termin, dT64, ldI,
                       (float2*)dEx, (float2*)dEz, dWarn, dFpList, dFpCount, dReq, dCoord, dLast,
                       (const double*)dSums, dPatch, dPatchCount, patchCap);
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

int srwd_launch_edge_state(const srwd_req* dReq, int nReq, const int* dBlkReq, long long nBlk,
                           const srwd_trj_info* dInfo, const srwd_edge* dEdge, const double* dCoord, int termin,
                           long long* dLast, hipStream_t st)
{
    if (nBlk <= 0 || nReq <= 0) return 0;
    hipLaunchKernelGGL(k_edge_blocks, dim3((unsigned)nBlk), dim3(SRWD_EDGE_BLOCK), 0, st, dReq, dBlkReq, dInfo, dEdge,
                       dCoord, termin, dLast);
    hipLaunchKernelGGL(k_edge_scan, dim3(nReq), dim3(256), 0, st, dReq, dLast);
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

}  // extern "C"

// ------------------------------------------------------------------------
// Zero-field flags and wavefront radius
// ------------------------------------------------------------------------
// srtrjdat.cpp:2048-2062: a plane is treated as moving when any position or
// angle sample is non-zero.  One block per trajectory.
__global__ void __launch_bounds__(256) k_nz_flags(srwd_trj_info* info, const double* X, const double* Xp,
                                                  const double* Y, const double* Yp, long long ld)
{
    int tr = blockIdx.x;
    long long np = info[tr].np;
    __shared__ int fv, fh;
    if (threadIdx.x == 0) { fv = 0; fh = 0; }
    __syncthreads();
    int v = 0, h = 0;
    for (long long i = threadIdx.x; i < np; i += blockDim.x) {
        long long o = tr * ld + i;
        v |= (X[o] != 0. || Xp[o] != 0.);
        h |= (Y[o] != 0. || Yp[o] != 0.);
    }
    if (v) atomicOr(&fv, 1);
    if (h) atomicOr(&fh, 1);
    __syncthreads();
    if (threadIdx.x == 0) { info[tr].verNZ = fv; info[tr].horNZ = fh; }
}

// Trajectory samples for the radius search (srtrjdat.cpp:1416-1460: the
// polynomials are used regardless of the zero-field flags).
__global__ void __launch_bounds__(256) k_rx_samples(const srwd_trj_info* __restrict__ info, const double* X,
                                                    const double* Y, const double* IBx, const double* IBz,
                                                    long long ldT, double* samp, long long ldS)
{
    long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    int tr = blockIdx.y;
    const srwd_trj_info I = info[tr];
    if (i >= I.np) return;
    double s = I.sStart + (double)i * I.sStep;
    trj_pt px = eval_plane(I, X + tr * ldT, IBx + tr * ldT, -1. / I.betaNorm, s, true);
    trj_pt pz = eval_plane(I, Y + tr * ldT, IBz + tr * ldT, 1. / I.betaNorm, s, true);
    double* o = samp + tr * ldS * 4;
    o[i] = px.bt; o[ldS + i] = pz.bt; o[2 * ldS + i] = px.crd; o[3 * ldS + i] = pz.crd;
}

// srradstr.cpp:2310-2501 (FindAverageDistanceToSource) with the full
// trajectory range.  mesh[4*tr] = xStart, xEnd, zStart, zEnd; out[4*tr] =
// Robs, RobsAbsErr, xc, zc.  One block per trajectory: each of the four
// misfit criteria keeps the first sample (scanning from the downstream end,
// then from the upstream end) whose misfit changes sign relative to the end
// sample.  Each thread visits its samples in increasing order and keeps its
// first hit per criterion; the block minimum follows from a wave reduction
// and one LDS exchange.
__global__ void __launch_bounds__(256) k_find_rx(const srwd_trj_info* __restrict__ info, const double* samp,
                                                 long long ldS, const double* mesh, double* out)
{
    const int tr = blockIdx.x;
    const srwd_trj_info I = info[tr];
    const double* bx = samp + tr * ldS * 4; const double* bz = bx + ldS;
    const double* xa = bx + 2 * ldS; const double* za = bx + 3 * ldS;
    const double xS = mesh[4 * tr], xE = mesh[4 * tr + 1], zS = mesh[4 * tr + 2], zE = mesh[4 * tr + 3];
    const long long N = I.np;
    const double sStart = I.sStart, sRange = (N - 1) * I.sStep, sEnd = sStart + sRange;
    const long long L = N - 1;
    const unsigned long long NONE = ~0ull;
    __shared__ unsigned long long first[4][8];
    const double R0b = I.yObs - sEnd, R0f = I.yObs - sStart;
    double crd[4] = {xS, xE, zS, zE};
    double ref[8];
    for (int k = 0; k < 4; k++) {
        bool zp = k >= 2;
        ref[k] = (crd[k] - (zp ? za[L] : xa[L])) * (1. / R0b) - (zp ? bz[L] : bx[L]);
        ref[4 + k] = (crd[k] - (zp ? za[0] : xa[0])) * (1. / R0f) - (zp ? bz[0] : bx[0]);
    }
    unsigned long long mine[8];
    for (int k = 0; k < 8; k++) mine[k] = NONE;
    for (long long is = 1 + threadIdx.x; is < N; is += blockDim.x) {
        double Rb = R0b + (double)is * I.sStep, Rf = R0f - (double)is * I.sStep;
        long long jb = L - is, jf = is;
        for (int k = 0; k < 4; k++) {
            bool zp = k >= 2;
            double mb = (crd[k] - (zp ? za[jb] : xa[jb])) * (1. / Rb) - (zp ? bz[jb] : bx[jb]);
            if (mb * ref[k] < 0. && mine[k] == NONE) mine[k] = (unsigned long long)is;
            double mf = (crd[k] - (zp ? za[jf] : xa[jf])) * (1. / Rf) - (zp ? bz[jf] : bx[jf]);
            if (mf * ref[4 + k] < 0. && mine[4 + k] == NONE) mine[4 + k] = (unsigned long long)is;
        }
    }
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    for (int k = 0; k < 8; k++) {
        unsigned long long v = mine[k];
        for (int off = 32; off > 0; off >>= 1) {
            unsigned long long o = __shfl_xor(v, off, 64);
            v = o < v ? o : v;
        }
        if (lane == 0) first[w][k] = v;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int k = 0; k < 8; k++)
            for (int q = 1; q < 4; q++) first[0][k] = first[q][k] < first[0][k] ? first[q][k] : first[0][k];
    }
    if (threadIdx.x != 0) return;
    const double VL = 1.E+23;
    double Rv[8];
    for (int k = 0; k < 4; k++) {
        Rv[k] = first[0][k] == NONE ? VL : R0b + (double)first[0][k] * I.sStep;
        Rv[4 + k] = first[0][4 + k] == NONE ? VL : R0f - (double)first[0][4 + k] * I.sStep;
    }
    double minRx = Rv[0] < Rv[1] ? Rv[0] : Rv[1], minRz = Rv[2] < Rv[3] ? Rv[2] : Rv[3];
    double minR = minRx < minRz ? minRx : minRz;
    double maxRx = Rv[4] < Rv[5] ? Rv[4] : Rv[5], maxRz = Rv[6] < Rv[7] ? Rv[6] : Rv[7];
    double maxR = (minRx < minRz) ? maxRx : maxRz;
    double avg = 0.5 * (minR + maxR);
    if (minR == VL) avg = (maxR == VL) ? VL : maxR;
    else if (maxR == VL) avg = minR;
    double Robs, ysrc;
    if (avg != VL) { Robs = avg; ysrc = I.yObs - Robs; }
    else {
        ysrc = (sStart < 0. && sEnd > 0.) ? 0.35 * sRange : sStart + 0.75 * sRange;
        Robs = I.yObs - ysrc;
    }
    long long iy = (long long)((ysrc - sStart) / I.sStep + 0.00001);
    if (iy < 0) iy = 0;
    if (iy >= N) iy = N - 1;
    out[4 * tr] = Robs; out[4 * tr + 1] = 0.25 * sRange; out[4 * tr + 2] = xa[iy]; out[4 * tr + 3] = za[iy];
}

extern "C" {
int srwd_launch_nz(srwd_trj_info* dInfo, int nTraj, const double* dX, const double* dXp, const double* dY,
                   const double* dYp, long long ldT, hipStream_t st)
{
    hipLaunchKernelGGL(k_nz_flags, dim3(nTraj), dim3(256), 0, st, dInfo, dX, dXp, dY, dYp, ldT);
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

int srwd_launch_rx(const srwd_trj_info* dInfo, int nTraj, long long maxNp, const double* dX, const double* dY,
                   const double* dIBx, const double* dIBz, long long ldT, double* dSamp, const double* dMesh,
                   double* dRx, hipStream_t st)
{
    hipLaunchKernelGGL(k_rx_samples, dim3((unsigned)((maxNp + 255) / 256), nTraj), dim3(256), 0, st, dInfo, dX,
                       dY, dIBx, dIBz, ldT, dSamp, ldT);
    hipLaunchKernelGGL(k_find_rx, dim3(nTraj), dim3(256), 0, st, dInfo, dSamp, ldT, dMesh, dRx);
    return hipGetLastError() == hipSuccess ? 0 : -1;
}
}

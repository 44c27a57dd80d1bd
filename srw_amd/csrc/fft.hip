// fft.hip - SRW-convention 2D FFT and drift-space propagation for gfx950.
//
// Reference algorithms (paths relative to the reference source root):
//   centred 2D FFT   cpp/src/ext/genmath/gmfft.cpp:295-893, gmfft.h:232-472
//   drift, mode 0    cpp/src/core/sroptdrf.h:348-409, 615-644
//   drift, mode 3    cpp/src/core/sroptdrf.cpp:726-1116, sroptdrf.h:806-878
//
// SRW's centred transform out[k] = sum_n f[n] exp(-+ i 2pi (k-N/2)(n-N/2)/N)
// equals (-1)^(k+N/2) * DFT[(-1)^n f][k]; for a drift the two middle sign
// factors cancel, so the propagation is
//   sign*phase1 -> DFT2 -> kernel*scale -> IDFT2 -> sign*phase3.
// Forward line transforms are decimation-in-frequency (natural order in,
// bit-reversed out) and inverse ones decimation-in-time (bit-reversed in,
// natural out), so no reordering pass exists: between the passes the x axis
// is stored bit-reversed.  Each pass holds whole lines in LDS and runs three
// radix-2 stages per LDS round trip in registers.
//
// Passes over one N x N slice (HBM sweeps):
//   A  rows   : pre-multiply, forward DIF along x
//   B  columns: forward DIF along y, Fresnel kernel, inverse DIT along y
//   C  rows   : inverse DIT along x, post-multiply (or intensity output)
#include "srw_fft.h"
#include "srw_fft_ops.h"

#include <math.h>

namespace {

// Global loads issued back to back per thread before their values are used.
#define SRWF_BATCH 8

// LDS position of element i of a line: one pad slot after every 8 elements,
// so the strided accesses of the radix-8 rounds (strides 1, 2, 8 elements)
// fall on distinct banks.  A line of N elements spans line_pitch(N) slots.
__device__ __forceinline__ int pad_idx(int i) { return i + (i >> 3); }
__host__ __device__ __forceinline__ int line_pitch(int n) { return n + n / 8 + 1; }

__device__ __forceinline__ int bitrev(int v, int logn)
{
    return (int)(__builtin_bitreverse32((unsigned)v) >> (32 - logn));
}

// LDS twiddle table: tw[k] = exp(-i 2pi k / N) for k < N/2, then for each
// H = 2^lH <= N/8 the H twiddles w_{2H}^t (t < H) stored contiguously at
// N/2 + H - 1 + t.  The lanes of a stage read consecutive or repeated
// entries instead of entries N/2H apart, which would fall on one LDS bank.
__host__ __device__ __forceinline__ int tw_slots(int n) { return n / 2 + n / 4 - 1; }

__device__ __forceinline__ int tw_index(int logn, int lH, int t)
{
    return (lH >= logn - 2) ? (t << (logn - 1 - lH)) : (1 << (logn - 1)) + (1 << lH) - 1 + t;
}

// Twiddle w_{2H}^t = exp(-+i 2pi t / 2H) = tw[t * (N / 2H)] (tw[k] = exp(-i 2pi k / N)).
template <bool INV>
__device__ __forceinline__ float2 twid(const float2* tw, int k)
{
    float2 w = tw[k];
    if (INV) w.y = -w.y;
    return w;
}

// One register round of DIF stages with half sizes H = b*2^(K-1) ... b
// (b = 2^lb) on groups of 2^K elements spaced b apart.  Lines of N = 2^logn
// elements with pitch P.  All sizes are powers of two, so the index
// arithmetic is shifts and masks.
template <int K, bool INV>
__device__ __forceinline__ void dif_round(float2* s, const float2* tw, int logn, int lb, int nlines, int P, int tid,
                                          int nthr)
{
    const int R = 1 << K;
    const int b = 1 << lb;
    const int lg = logn - K;   // log2 of the groups per line
    for (int G = tid; G < (nlines << lg); G += nthr) {
        int line = G >> lg, g = G & ((1 << lg) - 1);
        int j = ((g >> lb) << (lb + K)) + (g & (b - 1));
        float2* L = s + line * P;
        float2 x[R];
#pragma unroll
        for (int m = 0; m < R; m++) x[m] = L[pad_idx(j + m * b)];
#pragma unroll
        for (int r = K - 1; r >= 0; r--) {
            const int lH = lb + r;   // H = 2^lH
#pragma unroll
            for (int m = 0; m < R; m++) {
                if (m & (1 << r)) continue;
                int p = j + m * b;
                float2 a = x[m], c = x[m + (1 << r)];
                float2 w = twid<INV>(tw, tw_index(logn, lH, p & ((1 << lH) - 1)));
                x[m] = make_float2(a.x + c.x, a.y + c.y);
                x[m + (1 << r)] = cmulf(make_float2(a.x - c.x, a.y - c.y), w);
            }
        }
#pragma unroll
        for (int m = 0; m < R; m++) L[pad_idx(j + m * b)] = x[m];
    }
}

// One register round of DIT (inverse) stages with half sizes b ... b*2^(K-1).
template <int K>
__device__ __forceinline__ void dit_round(float2* s, const float2* tw, int logn, int lb, int nlines, int P, int tid,
                                          int nthr)
{
    const int R = 1 << K;
    const int b = 1 << lb;
    const int lg = logn - K;
    for (int G = tid; G < (nlines << lg); G += nthr) {
        int line = G >> lg, g = G & ((1 << lg) - 1);
        int j = ((g >> lb) << (lb + K)) + (g & (b - 1));
        float2* L = s + line * P;
        float2 x[R];
#pragma unroll
        for (int m = 0; m < R; m++) x[m] = L[pad_idx(j + m * b)];
#pragma unroll
        for (int r = 0; r < K; r++) {
            const int lH = lb + r;
#pragma unroll
            for (int m = 0; m < R; m++) {
                if (m & (1 << r)) continue;
                int p = j + m * b;
                float2 w = twid<true>(tw, tw_index(logn, lH, p & ((1 << lH) - 1)));
                float2 a = x[m], c = cmulf(x[m + (1 << r)], w);
                x[m] = make_float2(a.x + c.x, a.y + c.y);
                x[m + (1 << r)] = make_float2(a.x - c.x, a.y - c.y);
            }
        }
#pragma unroll
        for (int m = 0; m < R; m++) L[pad_idx(j + m * b)] = x[m];
    }
}

// Full forward (or inverse-direction) DIF over nlines lines in LDS.
template <bool INV>
__device__ void lds_dif(float2* s, const float2* tw, int logn, int nlines, int P, int tid, int nthr)
{
    int done = 0;
    while (done < logn) {
        int rem = logn - done;
        int K = rem >= 3 ? 3 : rem;
        int lb = logn - done - K;   // log2 of the smallest half size of this round
        __syncthreads();
        if (K == 3) dif_round<3, INV>(s, tw, logn, lb, nlines, P, tid, nthr);
        else if (K == 2) dif_round<2, INV>(s, tw, logn, lb, nlines, P, tid, nthr);
        else dif_round<1, INV>(s, tw, logn, lb, nlines, P, tid, nthr);
        done += K;
    }
    __syncthreads();
}

// Full inverse DIT (bit-reversed input) over nlines lines in LDS.
__device__ void lds_dit(float2* s, const float2* tw, int logn, int nlines, int P, int tid, int nthr)
{
    int done = 0;
    while (done < logn) {
        int rem = logn - done;
        int K = rem >= 3 ? 3 : rem;
        int lb = done;
        __syncthreads();
        if (K == 3) dit_round<3>(s, tw, logn, lb, nlines, P, tid, nthr);
        else if (K == 2) dit_round<2>(s, tw, logn, lb, nlines, P, tid, nthr);
        else dit_round<1>(s, tw, logn, lb, nlines, P, tid, nthr);
        done += K;
    }
    __syncthreads();
}

__device__ __forceinline__ void load_twiddles(float2* stw, const float2* tw, int n)
{
    for (int i = threadIdx.x; i < n / 2; i += blockDim.x) stw[i] = tw[i];
    // the per-stage tables for H = 1, 2, ..., n/8
    const int logn = 31 - __clz(n);
    for (int i = threadIdx.x; i < n / 4 - 1; i += blockDim.x) {
        int lH = 31 - __clz(i + 1), t = i + 1 - (1 << lH);
        stw[n / 2 + i] = tw[t << (logn - 1 - lH)];
    }
}

// ---------------------------------------------------------------------------
// Pass A: rows, pre-multiply, forward DIF along x.
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(1024) k_rows_fwd(float2* data, long long sliceStride, const srwf_slice* __restrict__ sp,
                                                  const float2* __restrict__ tw, int logn, int ny, int rowsPerWG,
                                                  int preOp, int inv)
{
    extern __shared__ float2 smem[];
    const int N = 1 << logn;
    const int P = line_pitch(N);
    float2* stw = smem;
    float2* s = smem + tw_slots(N);
    load_twiddles(stw, tw, N);
    const int slice = blockIdx.y;
    const srwf_slice& S = sp[slice];
    float2* base = data + slice * sliceStride;
    const int row0 = blockIdx.x * rowsPerWG;
    const int nrows = min(rowsPerWG, ny - row0);
    const int total = nrows * N;
    for (int e0 = threadIdx.x; e0 < total; e0 += SRWF_BATCH * blockDim.x) {
        float2 vb[SRWF_BATCH];   // all loads of a batch in flight before any is used
#pragma unroll
        for (int q = 0; q < SRWF_BATCH; q++) {
            int e = e0 + q * blockDim.x;
            if (e < total) vb[q] = base[(long long)row0 * N + e];
        }
#pragma unroll
        for (int q = 0; q < SRWF_BATCH; q++) {
            int e = e0 + q * blockDim.x;
            if (e >= total) break;
            int r = e >> logn, ix = e & (N - 1), iy = row0 + r;
            s[r * P + pad_idx(ix)] = srwf_pre(S, preOp, ix, iy, N, ny, vb[q]);
        }
    }
    if (inv) lds_dif<true>(s, stw, logn, nrows, P, threadIdx.x, blockDim.x);
    else lds_dif<false>(s, stw, logn, nrows, P, threadIdx.x, blockDim.x);
    for (int e = threadIdx.x; e < nrows * N; e += blockDim.x) {
        int r = e >> logn, ix = e & (N - 1);
        base[(long long)(row0 + r) * N + ix] = s[r * P + pad_idx(ix)];
    }
}

// ---------------------------------------------------------------------------
// Pass B: columns.  op 0: angular-representation kernel (mode 0);
// op 1: analytic quadratic-phase kernel (mode 3); op 2: forward only with the
// centred-output post-multiply and natural-order output (standalone FFT).
// Column positions hold bit-reversed x frequencies.
// ---------------------------------------------------------------------------
// Transforms the cpw columns c0.. held in LDS (forward, kernel, inverse) and
// stores them, or for op 2 writes the standalone output.
__device__ __forceinline__ void cols_block(float2* s, const float2* stw, const srwf_slice& S, float2* base, float2* out,
                                           long long outOff, int c0, int cpw, int lc, int logny, int lognx, int P,
                                           int op, int inv)
{
    const int NY = 1 << logny, NX = 1 << lognx;
    if (op == 2 && inv) lds_dif<true>(s, stw, logny, cpw, P, threadIdx.x, blockDim.x);
    else lds_dif<false>(s, stw, logny, cpw, P, threadIdx.x, blockDim.x);
    if (op == 0 || op == 1) {
        for (int e = threadIdx.x; e < NY * cpw; e += blockDim.x) {
            int c = e >> logny, r = e & (NY - 1);
            int kx = bitrev(c0 + c, lognx), ky = bitrev(r, logny);
            s[c * P + pad_idx(r)] = srwf_mid(S, op, kx, ky, NX, NY, s[c * P + pad_idx(r)]);
        }
        lds_dit(s, stw, logny, cpw, P, threadIdx.x, blockDim.x);
        for (int e = threadIdx.x; e < NY * cpw; e += blockDim.x) {
            int r = e >> lc, c = e & (cpw - 1);
            base[(long long)r * NX + c0 + c] = s[c * P + pad_idx(r)];
        }
    } else {
        // standalone transform: (-1)^(kx+ky+NX/2+NY/2) * scale * post-shift, natural order
        for (int e = threadIdx.x; e < NY * cpw; e += blockDim.x) {
            int r = e >> lc, c = e & (cpw - 1);
            int ky = r, kx = bitrev(c0 + c, lognx);
            out[outOff + (long long)ky * NX + kx] = srwf_std_out(S, kx, ky, NX, NY, s[c * P + pad_idx(bitrev(ky, logny))]);
        }
    }
}

// pairs = 0: one block of colsPerWG columns per workgroup.  pairs = 1: two
// adjacent blocks, read together as full rows of 2 * colsPerWG columns (whole
// 128-byte lines for 8 columns); the first block goes to LDS, the second to
// registers (at most SRWF_PAIR_REGS values per thread) until the first is done.
#define SRWF_PAIR_REGS 16
__global__ void __launch_bounds__(1024) k_cols(float2* data, long long sliceStride, const srwf_slice* __restrict__ sp,
                                              const float2* __restrict__ twy, int logny, int lognx, int colsPerWG,
                                              int pairs, int op, int inv, float2* out)
{
    extern __shared__ float2 smem[];
    const int NY = 1 << logny, NX = 1 << lognx;
    const int lc = 31 - __clz(colsPerWG);   // colsPerWG is a power of two
    const int P = line_pitch(NY);
    float2* stw = smem;
    float2* s = smem + tw_slots(NY);
    load_twiddles(stw, twy, NY);
    const int slice = blockIdx.y;
    const srwf_slice& S = sp[slice];
    float2* base = data + slice * sliceStride;
    const long long outOff = slice * sliceStride;
    if (pairs) {
        const int w = 2 * colsPerWG;            // columns read per row
        const int c0 = blockIdx.x * w;
        const int nq = NY * w / blockDim.x;     // loads per thread (host: <= SRWF_PAIR_REGS, exact)
        const int cHold = threadIdx.x & (w - 1);   // the same column for all of a thread's loads
        float2 reg[SRWF_PAIR_REGS];
#pragma unroll
        for (int q = 0; q < SRWF_PAIR_REGS; q++) {
            if (q < nq) {
                int r = (threadIdx.x + q * blockDim.x) >> (lc + 1);
                float2 v = base[(long long)r * NX + c0 + cHold];
                if (cHold < colsPerWG) s[cHold * P + pad_idx(r)] = v;
                else reg[q] = v;
            }
        }
        cols_block(s, stw, S, base, out, outOff, c0, colsPerWG, lc, logny, lognx, P, op, inv);
        __syncthreads();
#pragma unroll
        for (int q = 0; q < SRWF_PAIR_REGS; q++) {
            if (q < nq && cHold >= colsPerWG) {
                int r = (threadIdx.x + q * blockDim.x) >> (lc + 1);
                s[(cHold - colsPerWG) * P + pad_idx(r)] = reg[q];
            }
        }
        cols_block(s, stw, S, base, out, outOff, c0 + colsPerWG, colsPerWG, lc, logny, lognx, P, op, inv);
        return;
    }
    // XCD-aware order.  Workgroups go round-robin to the 8 XCDs, each with
    // its own L2; the L = 128 / (8 colsPerWG) column blocks that share a
    // row's 128-byte lines are given to workgroups 8 ids apart, which land on
    // the same XCD at nearly the same time, so each line is fetched once.
    int g = blockIdx.x;
    const int L = colsPerWG >= 16 ? 1 : 16 / colsPerWG;
    if (L > 1 && gridDim.x % (8 * L) == 0) {
        const int xcd = blockIdx.x & 7, k = blockIdx.x >> 3;
        g = ((k / L) * 8 + xcd) * L + (k % L);
    }
    const int c0 = g * colsPerWG;
    const int total = NY * colsPerWG;
    for (int e0 = threadIdx.x; e0 < total; e0 += SRWF_BATCH * blockDim.x) {
        float2 vb[SRWF_BATCH];
#pragma unroll
        for (int q = 0; q < SRWF_BATCH; q++) {
            int e = e0 + q * blockDim.x;
            if (e < total) {
                int r = e >> lc, c = e & (colsPerWG - 1);
                vb[q] = base[(long long)r * NX + c0 + c];
            }
        }
#pragma unroll
        for (int q = 0; q < SRWF_BATCH; q++) {
            int e = e0 + q * blockDim.x;
            if (e >= total) break;
            int r = e >> lc, c = e & (colsPerWG - 1);
            s[c * P + pad_idx(r)] = vb[q];
        }
    }
    cols_block(s, stw, S, base, out, outOff, c0, colsPerWG, lc, logny, lognx, P, op, inv);
}

// ---------------------------------------------------------------------------
// Pass C: rows, inverse DIT along x, post-multiply.  postOp 0: sign only;
// 1: sign and new quadratic phase (mode 3, sroptdrf.h:846-860);
// 2: as 1 but writes only |Ex|^2+|Ez|^2 of slice pairs (2k, 2k+1).
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(1024) k_rows_inv(float2* data, long long sliceStride, const srwf_slice* __restrict__ sp,
                                                  const float2* __restrict__ tw, int logn, int ny, int rowsPerWG,
                                                  int postOp, float* inten, long long intenStride)
{
    extern __shared__ float2 smem[];
    const int N = 1 << logn;
    const int P = line_pitch(N);
    float2* stw = smem;
    float2* s = smem + tw_slots(N);
    load_twiddles(stw, tw, N);
    const int pair = (postOp == 2);
    const int slice0 = blockIdx.y << pair;
    const int nsl = 1 + pair;
    const int row0 = blockIdx.x * rowsPerWG;
    const int nrows = min(rowsPerWG, ny - row0);
    const int total = nsl * nrows * N;
    for (int e0 = threadIdx.x; e0 < total; e0 += SRWF_BATCH * blockDim.x) {
        float2 vb[SRWF_BATCH];
#pragma unroll
        for (int q = 0; q < SRWF_BATCH; q++) {
            int e = e0 + q * blockDim.x;
            if (e < total) {
                int l = e >> logn, ix = e & (N - 1);
                int k = l / nrows, r = l - k * nrows;
                vb[q] = data[(slice0 + k) * sliceStride + (long long)(row0 + r) * N + ix];
            }
        }
#pragma unroll
        for (int q = 0; q < SRWF_BATCH; q++) {
            int e = e0 + q * blockDim.x;
            if (e >= total) break;
            int l = e >> logn, ix = e & (N - 1);
            s[l * P + pad_idx(ix)] = vb[q];
        }
    }
    lds_dit(s, stw, logn, nsl * nrows, P, threadIdx.x, blockDim.x);
    const srwf_slice& S = sp[slice0];
    for (int e = threadIdx.x; e < nrows * N; e += blockDim.x) {
        int r = e >> logn, ix = e & (N - 1), iy = row0 + r;
        float2 m = srwf_post(S, postOp, ix, iy);
        if (postOp == 2) {
            float2 a = s[r * P + pad_idx(ix)], b = s[(nrows + r) * P + pad_idx(ix)];
            // the unit-modulus phase does not change the intensity
            inten[blockIdx.y * intenStride + (long long)iy * N + ix] = a.x * a.x + a.y * a.y + b.x * b.x + b.y * b.y;
        } else {
            for (int k = 0; k < nsl; k++) {
                float2 v = cmulf(s[(k * nrows + r) * P + pad_idx(ix)], m);
                data[(slice0 + k) * sliceStride + (long long)iy * N + ix] = v;
            }
        }
    }
}

// Gather / scatter between SRW's [iy][ix][ie] layout and slice-major
// [ie][iy][ix] for one component.
__global__ void k_gather(const float2* __restrict__ src, float2* dst, int ne, long long nxy)
{
    long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nxy * ne) return;
    long long ie = i / nxy, p = i - ie * nxy;
    dst[i] = src[p * ne + ie];
}

__global__ void k_scatter(const float2* __restrict__ src, float2* dst, int ne, long long nxy)
{
    long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nxy * ne) return;
    long long ie = i / nxy, p = i - ie * nxy;
    dst[p * ne + ie] = src[i];
}

// Both components between SRW's [iy][ix][ie] order (ex, ez) and the
// slice-major order d[(2*ie + c)*nxy + p], p = iy*nx + ix, through 64 x 64
// LDS tiles: the energy-fastest side is read (or written) along ie, the
// slice side along p, both coalesced.  Grid (ceil(nxy/64), ceil(ne/64), 2).
__global__ void __launch_bounds__(256) k_to_slices(const float2* __restrict__ ex, const float2* __restrict__ ez,
                                                   float2* d, int ne, long long nxy)
{
    __shared__ float2 tile[64][65];
    const int c = blockIdx.z, tx = threadIdx.x & 63, ty = threadIdx.x >> 6;
    const float2* src = c ? ez : ex;
    const long long p0 = (long long)blockIdx.x * 64;
    const int e0 = blockIdx.y * 64;
    for (int r = ty; r < 64; r += 4) {
        long long p = p0 + r;
        int e = e0 + tx;
        if (p < nxy && e < ne) tile[r][tx] = src[p * ne + e];
    }
    __syncthreads();
    for (int r = ty; r < 64; r += 4) {
        int e = e0 + r;
        long long p = p0 + tx;
        if (p < nxy && e < ne) d[(2LL * e + c) * nxy + p] = tile[tx][r];
    }
}

__global__ void __launch_bounds__(256) k_from_slices(const float2* __restrict__ d, float2* ex, float2* ez, int ne,
                                                     long long nxy)
{
    __shared__ float2 tile[64][65];
    const int c = blockIdx.z, tx = threadIdx.x & 63, ty = threadIdx.x >> 6;
    float2* dst = c ? ez : ex;
    const long long p0 = (long long)blockIdx.x * 64;
    const int e0 = blockIdx.y * 64;
    for (int r = ty; r < 64; r += 4) {
        int e = e0 + r;
        long long p = p0 + tx;
        if (p < nxy && e < ne) tile[tx][r] = d[(2LL * e + c) * nxy + p];
    }
    __syncthreads();
    for (int r = ty; r < 64; r += 4) {
        long long p = p0 + r;
        int e = e0 + tx;
        if (p < nxy && e < ne) dst[p * ne + e] = tile[r][tx];
    }
}

int lds_bytes(int n, int lines) { return (int)(sizeof(float2) * (tw_slots(n) + (size_t)lines * line_pitch(n))); }

}  // namespace

extern "C" {

int srwf_init_attributes()
{
    int lim = 160 * 1024;
    if (hipFuncSetAttribute((const void*)k_cols, hipFuncAttributeMaxDynamicSharedMemorySize, lim) != hipSuccess) return -1;
    if (hipFuncSetAttribute((const void*)k_rows_fwd, hipFuncAttributeMaxDynamicSharedMemorySize, lim) != hipSuccess) return -1;
    if (hipFuncSetAttribute((const void*)k_rows_inv, hipFuncAttributeMaxDynamicSharedMemorySize, lim) != hipSuccess) return -1;
    return 0;
}

// Tile shapes (measured on the 1024 x 1024 x 402-slice drift):
// * rows: 2048 elements (two rows of 1024) per 256-thread workgroup;
// * columns: 8 columns per 1024-thread workgroup from 1024 rows up (80 KB of
//   LDS, two workgroups per CU overlap loading and transforming), 16 columns
//   (128-byte row segments) below.
#define SRWF_ROW_ELEMS 2048
#define SRWF_ROWS_THREADS 256
#define SRWF_COLS_BIG_THREADS 1024
int srwf_rows_per_wg(int logn)
{
    int n = 1 << logn, r = SRWF_ROW_ELEMS / n;
    return r < 1 ? 1 : (r > 16 ? 16 : r);
}
int srwf_cols_per_wg(int logn) { int n = 1 << logn; return n >= 1024 ? 8 : 16; }

void srwf_default_tiles(int lognx, int logny, int postOp, srwf_tiles* t)
{
    t->rpwF = srwf_rows_per_wg(lognx);
    t->thrF = SRWF_ROWS_THREADS;
    int cpw = srwf_cols_per_wg(logny), nx = 1 << lognx;
    t->cpw = cpw > nx ? nx : cpw;
    t->thrC = ((1 << logny) * t->cpw >= 8192) ? SRWF_COLS_BIG_THREADS : 256;
    int rpw = srwf_rows_per_wg(lognx);
    if (postOp == 2) rpw = rpw > 1 ? rpw / 2 : 1;
    t->rpwI = rpw;
    t->thrI = SRWF_ROWS_THREADS;
    t->pairC = 0;
}

int srwf_tiles_fit(int lognx, int logny, int postOp, const srwf_tiles* t)
{
    const int lim = 160 * 1024, nx = 1 << lognx, ny = 1 << logny;
    auto okThr = [](int v) { return v == 256 || v == 512 || v == 1024; };
    if (!okThr(t->thrF) || !okThr(t->thrC) || !okThr(t->thrI)) return 0;
    if (t->rpwF < 1 || t->rpwF > ny || lds_bytes(nx, t->rpwF) > lim) return 0;
    if (t->rpwI < 1 || t->rpwI > ny || lds_bytes(nx, (postOp == 2 ? 2 : 1) * t->rpwI) > lim) return 0;
    if (t->cpw < 1 || t->cpw > nx || (t->cpw & (t->cpw - 1)) != 0 || lds_bytes(ny, t->cpw) > lim) return 0;
    if (t->pairC != 0 && t->pairC != 1) return 0;
    if (t->pairC) {
        // two blocks per workgroup; each thread keeps one column's loads in registers
        int w = 2 * t->cpw;
        if (w > nx || t->thrC % w != 0 || (long long)ny * w % t->thrC != 0 || ny * w / t->thrC > SRWF_PAIR_REGS)
            return 0;
    }
    return 1;
}

static srwf_tiles tiles_or_default(int lognx, int logny, int postOp, const srwf_tiles* t)
{
    srwf_tiles d;
    if (t && srwf_tiles_fit(lognx, logny, postOp, t)) return *t;
    srwf_default_tiles(lognx, logny, postOp, &d);
    return d;
}

int srwf_launch_rows_fwd(float2* data, long long sliceStride, int nSlices, const srwf_slice* dSp, const float2* tw,
                         int logn, int ny, int preOp, int inv, const srwf_tiles* tiles, hipStream_t st)
{
    int logny = 31 - __builtin_clz((unsigned)ny);
    srwf_tiles T = tiles_or_default(logn, logny, 0, tiles);
    int rpw = T.rpwF;
    dim3 grid((ny + rpw - 1) / rpw, nSlices);
    hipLaunchKernelGGL(k_rows_fwd, grid, dim3(T.thrF), lds_bytes(1 << logn, rpw), st, data, sliceStride, dSp, tw, logn,
                       ny, rpw, preOp, inv);
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

int srwf_launch_cols(float2* data, long long sliceStride, int nSlices, const srwf_slice* dSp, const float2* twy,
                     int logny, int lognx, int op, int inv, float2* out, const srwf_tiles* tiles, hipStream_t st)
{
    srwf_tiles T = tiles_or_default(lognx, logny, 0, tiles);
    int cpw = T.cpw;
    int nx = 1 << lognx;
    dim3 grid(nx / (cpw * (T.pairC ? 2 : 1)), nSlices);
    hipLaunchKernelGGL(k_cols, grid, dim3(T.thrC), lds_bytes(1 << logny, cpw), st, data, sliceStride, dSp, twy, logny,
                       lognx, cpw, T.pairC, op, inv, out);
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

int srwf_launch_rows_inv(float2* data, long long sliceStride, int nSlices, const srwf_slice* dSp, const float2* tw,
                         int logn, int ny, int postOp, float* inten, long long intenStride, const srwf_tiles* tiles,
                         hipStream_t st)
{
    int logny = 31 - __builtin_clz((unsigned)ny);
    srwf_tiles T = tiles_or_default(logn, logny, postOp, tiles);
    int rpw = T.rpwI;
    int nl = (postOp == 2 ? 2 : 1) * rpw;
    dim3 grid((ny + rpw - 1) / rpw, postOp == 2 ? nSlices / 2 : nSlices);
    hipLaunchKernelGGL(k_rows_inv, grid, dim3(T.thrI), lds_bytes(1 << logn, nl), st, data, sliceStride, dSp, tw, logn,
                       ny, rpw, postOp, inten, intenStride);
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

int srwf_launch_slices(float2* ex, float2* ez, float2* d, int ne, long long nxy, int toSlices, hipStream_t st)
{
    dim3 grid((unsigned)((nxy + 63) / 64), (unsigned)((ne + 63) / 64), 2);
    if (toSlices) hipLaunchKernelGGL(k_to_slices, grid, dim3(256), 0, st, ex, ez, d, ne, nxy);
    else hipLaunchKernelGGL(k_from_slices, grid, dim3(256), 0, st, d, ex, ez, ne, nxy);
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

int srwf_launch_gather(const float2* src, float2* dst, int ne, long long nxy, int scatter, hipStream_t st)
{
    long long n = nxy * ne;
    dim3 grid((unsigned)((n + 255) / 256));
    if (scatter) hipLaunchKernelGGL(k_scatter, grid, dim3(256), 0, st, src, dst, ne, nxy);
    else hipLaunchKernelGGL(k_gather, grid, dim3(256), 0, st, src, dst, ne, nxy);
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

}  // extern "C"

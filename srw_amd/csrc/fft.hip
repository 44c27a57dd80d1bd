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

#include <math.h>

namespace {

__device__ __forceinline__ float2 cmulf(float2 a, float2 b)
{
    return make_float2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}

__device__ __forceinline__ float2 cis_rev(double rev)
{   // exp(i 2pi rev) with the argument reduced in FP64
    float r = (float)(rev - rint(rev));
    return make_float2(__builtin_amdgcn_cosf(r), __builtin_amdgcn_sinf(r));
}

__device__ __forceinline__ int bitrev(int v, int logn)
{
    return (int)(__builtin_bitreverse32((unsigned)v) >> (32 - logn));
}

// Twiddle w_{2H}^t = exp(-+i 2pi t / 2H) = tw[t * (N / 2H)] (tw[k] = exp(-i 2pi k / N)).
template <bool INV>
__device__ __forceinline__ float2 twid(const float2* tw, int k)
{
    float2 w = tw[k];
    if (INV) w.y = -w.y;
    return w;
}

// One register round of DIF stages with half sizes H = b*2^(K-1) ... b on
// groups of 2^K elements spaced b apart.  Lines of N elements with pitch P.
template <int K, bool INV>
__device__ __forceinline__ void dif_round(float2* s, const float2* tw, int logn, int b, int nlines, int P, int tid,
                                          int nthr)
{
    const int N = 1 << logn;
    const int R = 1 << K;
    const int groupsPerLine = N >> K;
    const int hTop = b << (K - 1);
    for (int G = tid; G < nlines * groupsPerLine; G += nthr) {
        int line = G / groupsPerLine, g = G - line * groupsPerLine;
        int j = (g / b) * (2 * hTop) + (g % b);
        float2* L = s + line * P;
        float2 x[R];
#pragma unroll
        for (int m = 0; m < R; m++) x[m] = L[j + m * b];
#pragma unroll
        for (int r = K - 1; r >= 0; r--) {
            int H = b << r;
            int tstep = N / (2 * H);
#pragma unroll
            for (int m = 0; m < R; m++) {
                if (m & (1 << r)) continue;
                int p = j + m * b;
                float2 a = x[m], c = x[m + (1 << r)];
                float2 w = twid<INV>(tw, (p & (H - 1)) * tstep);
                x[m] = make_float2(a.x + c.x, a.y + c.y);
                x[m + (1 << r)] = cmulf(make_float2(a.x - c.x, a.y - c.y), w);
            }
        }
#pragma unroll
        for (int m = 0; m < R; m++) L[j + m * b] = x[m];
    }
}

// One register round of DIT (inverse) stages with half sizes b ... b*2^(K-1).
template <int K>
__device__ __forceinline__ void dit_round(float2* s, const float2* tw, int logn, int b, int nlines, int P, int tid,
                                          int nthr)
{
    const int N = 1 << logn;
    const int R = 1 << K;
    const int groupsPerLine = N >> K;
    const int hTop = b << (K - 1);
    for (int G = tid; G < nlines * groupsPerLine; G += nthr) {
        int line = G / groupsPerLine, g = G - line * groupsPerLine;
        int j = (g / b) * (2 * hTop) + (g % b);
        float2* L = s + line * P;
        float2 x[R];
#pragma unroll
        for (int m = 0; m < R; m++) x[m] = L[j + m * b];
#pragma unroll
        for (int r = 0; r < K; r++) {
            int H = b << r;
            int tstep = N / (2 * H);
#pragma unroll
            for (int m = 0; m < R; m++) {
                if (m & (1 << r)) continue;
                int p = j + m * b;
                float2 w = twid<true>(tw, (p & (H - 1)) * tstep);
                float2 a = x[m], c = cmulf(x[m + (1 << r)], w);
                x[m] = make_float2(a.x + c.x, a.y + c.y);
                x[m + (1 << r)] = make_float2(a.x - c.x, a.y - c.y);
            }
        }
#pragma unroll
        for (int m = 0; m < R; m++) L[j + m * b] = x[m];
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
        int b = 1 << (logn - done - K);   // smallest half size of this round
        __syncthreads();
        if (K == 3) dif_round<3, INV>(s, tw, logn, b, nlines, P, tid, nthr);
        else if (K == 2) dif_round<2, INV>(s, tw, logn, b, nlines, P, tid, nthr);
        else dif_round<1, INV>(s, tw, logn, b, nlines, P, tid, nthr);
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
        int b = 1 << done;
        __syncthreads();
        if (K == 3) dit_round<3>(s, tw, logn, b, nlines, P, tid, nthr);
        else if (K == 2) dit_round<2>(s, tw, logn, b, nlines, P, tid, nthr);
        else dit_round<1>(s, tw, logn, b, nlines, P, tid, nthr);
        done += K;
    }
    __syncthreads();
}

__device__ __forceinline__ void load_twiddles(float2* stw, const float2* tw, int n)
{
    for (int i = threadIdx.x; i < n / 2; i += blockDim.x) stw[i] = tw[i];
}

// ---------------------------------------------------------------------------
// Pass A: rows, pre-multiply, forward DIF along x.
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(256) k_rows_fwd(float2* data, long long sliceStride, const srwf_slice* __restrict__ sp,
                                                  const float2* __restrict__ tw, int logn, int ny, int rowsPerWG,
                                                  int preOp, int inv)
{
    extern __shared__ float2 smem[];
    const int N = 1 << logn;
    const int P = N + 1;
    float2* stw = smem;
    float2* s = smem + N / 2;
    load_twiddles(stw, tw, N);
    const int slice = blockIdx.y;
    const srwf_slice& S = sp[slice];
    float2* base = data + slice * sliceStride;
    const int row0 = blockIdx.x * rowsPerWG;
    const int nrows = min(rowsPerWG, ny - row0);
    for (int e = threadIdx.x; e < nrows * N; e += blockDim.x) {
        int r = e / N, ix = e - r * N, iy = row0 + r;
        float2 v = base[(long long)iy * N + ix];
        if (preOp == 1) {   // quadratic phase removal, sroptdrf.h:812-821
            double rx = S.xStart + ix * S.xStep - S.xc, rz = S.zStart + iy * S.zStep - S.zc;
            v = cmulf(v, cis_rev(S.pre_rev * (S.invRx * rx * rx + S.invRz * rz * rz)));
        } else if (preOp == 2) {   // linear-phase pre-shift of a non-centred input (gmfft.h:250-276)
            v = cmulf(v, cis_rev(S.sh0x * ((ix - (N >> 1)) * S.xStep) + S.sh0y * ((iy - (ny >> 1)) * S.zStep)));
        }
        if ((ix + iy) & 1) { v.x = -v.x; v.y = -v.y; }
        s[r * P + ix] = v;
    }
    if (inv) lds_dif<true>(s, stw, logn, nrows, P, threadIdx.x, blockDim.x);
    else lds_dif<false>(s, stw, logn, nrows, P, threadIdx.x, blockDim.x);
    for (int e = threadIdx.x; e < nrows * N; e += blockDim.x) {
        int r = e / N, ix = e - r * N;
        base[(long long)(row0 + r) * N + ix] = s[r * P + ix];
    }
}

// ---------------------------------------------------------------------------
// Pass B: columns.  op 0: angular-representation kernel (mode 0);
// op 1: analytic quadratic-phase kernel (mode 3); op 2: forward only with the
// centred-output post-multiply and natural-order output (standalone FFT).
// Column positions hold bit-reversed x frequencies.
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(256) k_cols(float2* data, long long sliceStride, const srwf_slice* __restrict__ sp,
                                              const float2* __restrict__ twy, int logny, int lognx, int colsPerWG,
                                              int op, int inv, float2* out)
{
    extern __shared__ float2 smem[];
    const int NY = 1 << logny, NX = 1 << lognx;
    const int P = NY + 1;
    float2* stw = smem;
    float2* s = smem + NY / 2;
    load_twiddles(stw, twy, NY);
    const int slice = blockIdx.y;
    const srwf_slice& S = sp[slice];
    float2* base = data + slice * sliceStride;
    const int c0 = blockIdx.x * colsPerWG;
    for (int e = threadIdx.x; e < NY * colsPerWG; e += blockDim.x) {
        int r = e / colsPerWG, c = e - r * colsPerWG;
        s[c * P + r] = base[(long long)r * NX + c0 + c];
    }
    if (op == 2 && inv) lds_dif<true>(s, stw, logny, colsPerWG, P, threadIdx.x, blockDim.x);
    else lds_dif<false>(s, stw, logny, colsPerWG, P, threadIdx.x, blockDim.x);
    if (op == 0 || op == 1) {
        for (int e = threadIdx.x; e < NY * colsPerWG; e += blockDim.x) {
            int c = e / NY, r = e - c * NY;
            int kx = bitrev(c0 + c, lognx), ky = bitrev(r, logny);
            double qx = (kx - (NX >> 1)) * S.qStepX, qz = (ky - (NY >> 1)) * S.qStepZ;
            double rev;
            if (op == 0) {   // sroptdrf.h:615-644
                double q2 = qx * qx + qz * qz, cq = S.c1 * q2;
                rev = S.c0_rev * q2 * (1. + cq + 2. * cq * cq) + S.path_rev;
            } else {         // sroptdrf.h:835-845
                rev = S.c0x_rev * qx * qx + S.c0z_rev * qz * qz + S.phT_rev;
            }
            float2 w = cis_rev(rev);
            w.x *= S.midScale; w.y *= S.midScale;
            s[c * P + r] = cmulf(s[c * P + r], w);
        }
        lds_dit(s, stw, logny, colsPerWG, P, threadIdx.x, blockDim.x);
        for (int e = threadIdx.x; e < NY * colsPerWG; e += blockDim.x) {
            int r = e / colsPerWG, c = e - r * colsPerWG;
            base[(long long)r * NX + c0 + c] = s[c * P + r];
        }
    } else {
        // standalone transform: (-1)^(kx+ky+NX/2+NY/2) * scale * post-shift, natural order
        for (int e = threadIdx.x; e < NY * colsPerWG; e += blockDim.x) {
            int r = e / colsPerWG, c = e - r * colsPerWG;
            int ky = r, kx = bitrev(c0 + c, lognx);
            float2 v = s[c * P + bitrev(ky, logny)];
            float sg = ((kx + ky + (NX >> 1) + (NY >> 1)) & 1) ? -S.midScale : S.midScale;
            v.x *= sg; v.y *= sg;
            if (S.postShift) {
                double rev = S.sh1x * ((kx - (NX >> 1)) * S.qStepX) + S.sh1y * ((ky - (NY >> 1)) * S.qStepZ);
                v = cmulf(v, cis_rev(rev));
            }
            out[slice * sliceStride + (long long)ky * NX + kx] = v;
        }
    }
}

// ---------------------------------------------------------------------------
// Pass C: rows, inverse DIT along x, post-multiply.  postOp 0: sign only;
// 1: sign and new quadratic phase (mode 3, sroptdrf.h:846-860);
// 2: as 1 but writes only |Ex|^2+|Ez|^2 of slice pairs (2k, 2k+1).
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(256) k_rows_inv(float2* data, long long sliceStride, const srwf_slice* __restrict__ sp,
                                                  const float2* __restrict__ tw, int logn, int ny, int rowsPerWG,
                                                  int postOp, float* inten, long long intenStride)
{
    extern __shared__ float2 smem[];
    const int N = 1 << logn;
    const int P = N + 1;
    float2* stw = smem;
    float2* s = smem + N / 2;
    load_twiddles(stw, tw, N);
    const int pair = (postOp == 2);
    const int slice0 = blockIdx.y << pair;
    const int nsl = 1 + pair;
    const int row0 = blockIdx.x * rowsPerWG;
    const int nrows = min(rowsPerWG, ny - row0);
    for (int e = threadIdx.x; e < nsl * nrows * N; e += blockDim.x) {
        int l = e / N, ix = e - l * N;
        int k = l / nrows, r = l - k * nrows;
        s[l * P + ix] = data[(slice0 + k) * sliceStride + (long long)(row0 + r) * N + ix];
    }
    lds_dit(s, stw, logn, nsl * nrows, P, threadIdx.x, blockDim.x);
    const srwf_slice& S = sp[slice0];
    for (int e = threadIdx.x; e < nrows * N; e += blockDim.x) {
        int r = e / N, ix = e - r * N, iy = row0 + r;
        float2 m = make_float2(((ix + iy) & 1) ? -1.f : 1.f, 0.f);
        if (postOp >= 1) {
            double rx = S.nxStart + ix * S.nxStep - S.xc, rz = S.nzStart + iy * S.nzStep - S.zc;
            float2 w = cis_rev(S.post_rev * (S.invRxL * rx * rx + S.invRzL * rz * rz) + S.path3_rev);
            m = make_float2(m.x * w.x, m.x * w.y);
        }
        if (postOp == 2) {
            float2 a = s[r * P + ix], b = s[(nrows + r) * P + ix];
            // the unit-modulus phase does not change the intensity
            inten[blockIdx.y * intenStride + (long long)iy * N + ix] = a.x * a.x + a.y * a.y + b.x * b.x + b.y * b.y;
        } else {
            for (int k = 0; k < nsl; k++) {
                float2 v = cmulf(s[(k * nrows + r) * P + ix], m);
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

int lds_bytes(int n, int lines) { return (int)(sizeof(float2) * ((n / 2) + (size_t)lines * (n + 1))); }

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

int srwf_rows_per_wg(int logn) { int n = 1 << logn; return n >= 2048 ? 1 : (2048 / n < 16 ? 2048 / n : 16); }
int srwf_cols_per_wg(int logn) { int n = 1 << logn; return n >= 2048 ? 4 : (n >= 1024 ? 8 : 16); }

int srwf_launch_rows_fwd(float2* data, long long sliceStride, int nSlices, const srwf_slice* dSp, const float2* tw,
                         int logn, int ny, int preOp, int inv, hipStream_t st)
{
    int rpw = srwf_rows_per_wg(logn);
    dim3 grid((ny + rpw - 1) / rpw, nSlices);
    hipLaunchKernelGGL(k_rows_fwd, grid, dim3(256), lds_bytes(1 << logn, rpw), st, data, sliceStride, dSp, tw, logn,
                       ny, rpw, preOp, inv);
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

int srwf_launch_cols(float2* data, long long sliceStride, int nSlices, const srwf_slice* dSp, const float2* twy,
                     int logny, int lognx, int op, int inv, float2* out, hipStream_t st)
{
    int cpw = srwf_cols_per_wg(logny);
    int nx = 1 << lognx;
    if (cpw > nx) cpw = nx;
    dim3 grid(nx / cpw, nSlices);
    hipLaunchKernelGGL(k_cols, grid, dim3(256), lds_bytes(1 << logny, cpw), st, data, sliceStride, dSp, twy, logny,
                       lognx, cpw, op, inv, out);
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

int srwf_launch_rows_inv(float2* data, long long sliceStride, int nSlices, const srwf_slice* dSp, const float2* tw,
                         int logn, int ny, int postOp, float* inten, long long intenStride, hipStream_t st)
{
    int rpw = srwf_rows_per_wg(logn);
    if (postOp == 2) rpw = rpw > 1 ? rpw / 2 : 1;
    int nl = (postOp == 2 ? 2 : 1) * rpw;
    dim3 grid((ny + rpw - 1) / rpw, postOp == 2 ? nSlices / 2 : nSlices);
    hipLaunchKernelGGL(k_rows_inv, grid, dim3(256), lds_bytes(1 << logn, nl), st, data, sliceStride, dSp, tw, logn,
                       ny, rpw, postOp, inten, intenStride);
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

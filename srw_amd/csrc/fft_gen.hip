// fft_gen.hip - SRW-convention 2D FFT and drift passes for even sizes that
// are not powers of two (N = 2^a 3^b 5^c 7^d 11^e 13^f, 8 <= N <= 2048: the
// sizes SRW's resampling produces), for gfx950.
//
// Reference algorithms (paths relative to the reference source root):
//   centred 2D FFT   cpp/src/ext/genmath/gmfft.cpp:295-893 (FFTW, any even N)
//   drift            cpp/src/core/sroptdrf.h:348-409, 615-644, 806-878
//
// Line transforms are mixed-radix Stockham autosort transforms in LDS
// (natural order in and out; Govindaraju et al., "High performance discrete
// Fourier transforms on graphics processors", SC'08): stage s with radix R
// reads element j + r*N/R, multiplies it by w_{Ns R}^{(j mod Ns) r}, takes the
// R-point DFT and writes it to (j div Ns) Ns R + (j mod Ns) + q Ns, ping-ponging
// between two LDS lines.  Twiddles come from a table of exp(-2 pi i m/N).
// The drift passes (HBM sweeps over one slice):
//   rows       : pre-multiply, forward transform along x
//   transpose  : [ny][nx] -> [nx][ny]
//   columns    : forward along y, kernel, inverse along y (one line per workgroup)
//   transpose  : back
//   rows       : inverse along x, post-multiply (or intensity output)
// The pointwise operations are those of the power-of-2 passes (srw_fft_ops.h).
#include "srw_fft.h"
#include "srw_fft_ops.h"

namespace {

template <int R>
__device__ __forceinline__ void gstage(const float2* in, float2* out, const float2* tw, int N, int Ns, bool inv,
                                       int nlines, int tid, int nthr)
{
    const int M = N / R;                 // butterflies per line
    const int tstep = N / (Ns * R);      // twiddle index step of w_{Ns R}
    for (int G = tid; G < nlines * M; G += nthr) {
        const int line = G / M, j = G - line * M;
        const float2* a = in + line * N;
        float2* b = out + line * N;
        const int jm = j % Ns;
        float2 v[R];
#pragma unroll
        for (int p = 0; p < R; p++) v[p] = a[j + p * M];
#pragma unroll
        for (int p = 1; p < R; p++) {
            float2 w = tw[jm * p * tstep];
            if (inv) w.y = -w.y;
            v[p] = cmulf(v[p], w);
        }
        const int base = (j / Ns) * Ns * R + jm;
#pragma unroll
        for (int q = 0; q < R; q++) {
            float2 acc = v[0];
#pragma unroll
            for (int p = 1; p < R; p++) {
                float2 w = tw[((p * q) % R) * M];   // w_R^{pq}
                if (inv) w.y = -w.y;
                float2 t = cmulf(v[p], w);
                acc.x += t.x; acc.y += t.y;
            }
            b[base + q * Ns] = acc;
        }
    }
}

// Full transform of nlines lines of length pl.n held in buf[0]; returns the
// buffer that holds the result (buf[0] or buf[1]).
__device__ float2* gtransform(float2* b0, float2* b1, const float2* tw, const srwf_gplan& pl, bool inv, int nlines,
                              int tid, int nthr)
{
    float2* in = b0;
    float2* out = b1;
    int Ns = 1;
    for (int s = 0; s < pl.nst; s++) {
        __syncthreads();
        const int R = pl.R[s];
        switch (R) {
        case 2: gstage<2>(in, out, tw, pl.n, Ns, inv, nlines, tid, nthr); break;
        case 3: gstage<3>(in, out, tw, pl.n, Ns, inv, nlines, tid, nthr); break;
        case 4: gstage<4>(in, out, tw, pl.n, Ns, inv, nlines, tid, nthr); break;
        case 5: gstage<5>(in, out, tw, pl.n, Ns, inv, nlines, tid, nthr); break;
        case 7: gstage<7>(in, out, tw, pl.n, Ns, inv, nlines, tid, nthr); break;
        case 11: gstage<11>(in, out, tw, pl.n, Ns, inv, nlines, tid, nthr); break;
        default: gstage<13>(in, out, tw, pl.n, Ns, inv, nlines, tid, nthr); break;
        }
        Ns *= R;
        float2* t = in; in = out; out = t;
    }
    __syncthreads();
    return in;
}

// Row pass, one row (or one row of a slice pair) per workgroup.
// dir 0: srwf_pre then the forward (inv = 0) or inverse-direction transform;
// dir 1: inverse transform then srwf_post (postOp 2 writes |Ex|^2+|Ez|^2 of
// slice pairs).  LDS: twiddles, two line buffers per line.
__global__ void __launch_bounds__(256) k_grows(float2* data, long long sliceStride, const srwf_slice* __restrict__ sp,
                                               const float2* __restrict__ twg, srwf_gplan pl, int ny, int dir,
                                               int preOp, int inv, int postOp, float* inten, long long intenStride)
{
    extern __shared__ float2 gsm[];
    const int N = pl.n;
    const int nl = (dir == 1 && postOp == 2) ? 2 : 1;
    float2* tw = gsm;
    float2* b0 = gsm + N;
    float2* b1 = b0 + nl * N;
    for (int i = threadIdx.x; i < N; i += blockDim.x) tw[i] = twg[i];
    const int slice0 = blockIdx.y * nl, row = blockIdx.x;
    for (int e = threadIdx.x; e < nl * N; e += blockDim.x) {
        int k = e / N, ix = e - k * N;
        float2 v = data[(slice0 + k) * sliceStride + (long long)row * N + ix];
        if (dir == 0) v = srwf_pre(sp[slice0 + k], preOp, ix, row, N, ny, v);
        b0[e] = v;
    }
    float2* r = gtransform(b0, b1, tw, pl, dir == 1 ? true : inv != 0, nl, threadIdx.x, blockDim.x);
    const srwf_slice& S = sp[slice0];
    for (int ix = threadIdx.x; ix < N; ix += blockDim.x) {
        if (dir == 0) {
            data[slice0 * sliceStride + (long long)row * N + ix] = r[ix];
            continue;
        }
        float2 m = srwf_post(S, postOp, ix, row);
        if (postOp == 2) {
            float2 a = r[ix], b = r[N + ix];
            inten[blockIdx.y * intenStride + (long long)row * N + ix] = a.x * a.x + a.y * a.y + b.x * b.x + b.y * b.y;
        } else {
            for (int k = 0; k < nl; k++)
                data[(slice0 + k) * sliceStride + (long long)row * N + ix] = cmulf(r[k * N + ix], m);
        }
    }
}

// Column pass on transposed slices [NX][NY]: line kx of slice blockIdx.y.
// op 0 / 1: forward, kernel srwf_mid, inverse, back in place; op 2: forward
// (or inverse direction) and the standalone output written to out in the
// natural [ky][kx] layout.
__global__ void __launch_bounds__(256) k_gcols(float2* dataT, long long sliceStride, const srwf_slice* __restrict__ sp,
                                               const float2* __restrict__ twg, srwf_gplan pl, int NX, int op, int inv,
                                               float2* out)
{
    extern __shared__ float2 gsm[];
    const int NY = pl.n;
    float2* tw = gsm;
    float2* b0 = gsm + NY;
    float2* b1 = b0 + NY;
    for (int i = threadIdx.x; i < NY; i += blockDim.x) tw[i] = twg[i];
    const int slice = blockIdx.y, kx = blockIdx.x;
    float2* line = dataT + slice * sliceStride + (long long)kx * NY;
    for (int i = threadIdx.x; i < NY; i += blockDim.x) b0[i] = line[i];
    const srwf_slice& S = sp[slice];
    float2* r = gtransform(b0, b1, tw, pl, op == 2 ? inv != 0 : false, 1, threadIdx.x, blockDim.x);
    if (op == 2) {
        for (int ky = threadIdx.x; ky < NY; ky += blockDim.x)
            out[slice * sliceStride + (long long)ky * NX + kx] = srwf_std_out(S, kx, ky, NX, NY, r[ky]);
        return;
    }
    for (int ky = threadIdx.x; ky < NY; ky += blockDim.x) r[ky] = srwf_mid(S, op, kx, ky, NX, NY, r[ky]);
    float2* o = (r == b0) ? b1 : b0;
    float2* r2 = gtransform(r, o, tw, pl, true, 1, threadIdx.x, blockDim.x);
    for (int i = threadIdx.x; i < NY; i += blockDim.x) line[i] = r2[i];
}

// Transpose of each [rows][cols] slice into [cols][rows], 32 x 32 tiles.
__global__ void __launch_bounds__(256) k_gtranspose(const float2* __restrict__ src, float2* dst, int rows, int cols,
                                                    long long sliceStride)
{
    __shared__ float2 t[32][33];
    const long long off = blockIdx.z * sliceStride;
    const int c0 = blockIdx.x * 32, r0 = blockIdx.y * 32;
    const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;
    for (int k = ty; k < 32; k += 8) {
        int r = r0 + k, c = c0 + tx;
        if (r < rows && c < cols) t[k][tx] = src[off + (long long)r * cols + c];
    }
    __syncthreads();
    for (int k = ty; k < 32; k += 8) {
        int c = c0 + k, r = r0 + tx;
        if (r < rows && c < cols) dst[off + (long long)c * rows + r] = t[tx][k];
    }
}

}  // namespace

extern "C" {

int srwf_gplan_make(int n, srwf_gplan* pl)
{
    if (n < 8 || n > SRWF_GEN_MAX || (n & 1)) return -1;
    int m = n, k = 0;
    const int rad[7] = {4, 2, 3, 5, 7, 11, 13};
    for (int i = 0; i < 7; i++)
        while (m % rad[i] == 0) {
            if (k >= 16) return -1;
            pl->R[k++] = rad[i];
            m /= rad[i];
        }
    if (m != 1) return -1;
    pl->n = n;
    pl->nst = k;
    return 0;
}

int srwf_gen_init_attributes()
{
    int lim = 160 * 1024;
    if (hipFuncSetAttribute((const void*)k_grows, hipFuncAttributeMaxDynamicSharedMemorySize, lim) != hipSuccess) return -1;
    if (hipFuncSetAttribute((const void*)k_gcols, hipFuncAttributeMaxDynamicSharedMemorySize, lim) != hipSuccess) return -1;
    return 0;
}

int srwf_launch_grows(float2* data, long long sliceStride, int nSlices, const srwf_slice* dSp, const float2* twg,
                      const srwf_gplan& pl, int ny, int dir, int preOp, int inv, int postOp, float* inten,
                      long long intenStride, hipStream_t st)
{
    int nl = (dir == 1 && postOp == 2) ? 2 : 1;
    if (nl == 2 && (nSlices & 1)) return -1;
    size_t lds = sizeof(float2) * (size_t)pl.n * (1 + 2 * nl);
    hipLaunchKernelGGL(k_grows, dim3(ny, nSlices / nl), dim3(256), lds, st, data, sliceStride, dSp, twg, pl, ny, dir,
                       preOp, inv, postOp, inten, intenStride);
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

int srwf_launch_gcols(float2* dataT, long long sliceStride, int nSlices, const srwf_slice* dSp, const float2* twg,
                      const srwf_gplan& pl, int NX, int op, int inv, float2* out, hipStream_t st)
{
    size_t lds = sizeof(float2) * (size_t)pl.n * 3;
    hipLaunchKernelGGL(k_gcols, dim3(NX, nSlices), dim3(256), lds, st, dataT, sliceStride, dSp, twg, pl, NX, op, inv,
                       out);
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

int srwf_launch_gtranspose(const float2* src, float2* dst, int rows, int cols, long long sliceStride, int nSlices,
                           hipStream_t st)
{
    dim3 grid((cols + 31) / 32, (rows + 31) / 32, nSlices);
    hipLaunchKernelGGL(k_gtranspose, grid, dim3(256), 0, st, src, dst, rows, cols, sliceStride);
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

}  // extern "C"

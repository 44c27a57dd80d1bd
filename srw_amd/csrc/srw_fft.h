// srw_fft.h - per-slice parameters and launch wrappers of the FFT / drift passes.
#pragma once
#include <hip/hip_runtime.h>

// Parameters of one N x N complex slice (one field component at one photon
// energy).  Phases are in revolutions; see fft.hip for the pass structure.
struct srwf_slice {
    double xStart, xStep, zStart, zStep;        // input coordinate mesh
    double nxStart, nxStep, nzStart, nzStep;    // output coordinate mesh (mode 3 rescaled)
    double xc, zc;                              // wavefront centre
    double invRx, invRz, invRxL, invRzL;        // 1/R before and 1/(R+L) after the drift
    double pre_rev, post_rev, path3_rev;        // -pi/lambda, +pi/lambda (per 2pi), absolute-path term
    double qStepX, qStepZ;                      // angular-representation steps 1/(N dx)
    double c0_rev, c1, path_rev;                // mode 0 kernel
    double c0x_rev, c0z_rev, phT_rev;           // mode 3 kernel
    double sh0x, sh0y, sh1x, sh1y;              // linear-phase shifts of the standalone FFT
    float midScale;
    int postShift;
    int ny;
    int pad;
};

// Tile shapes of the three passes: rows per workgroup and threads of the
// forward row pass, columns per workgroup and threads of the column pass,
// rows per workgroup and threads of the inverse row pass, and whether the
// column pass takes two adjacent column blocks per workgroup.  Any shape that
// fits (srwf_tiles_fit) gives the same values; only the work assignment
// changes.
struct srwf_tiles {
    int rpwF, thrF, cpw, thrC, rpwI, thrI;
    int pairC;   // column pass reads two adjacent blocks per workgroup (full 2*cpw-column rows)
};

// Mixed-radix plan of an even line length n = 2^a 3^b 5^c 7^d 11^e 13^f,
// 8 <= n <= SRWF_GEN_MAX: the radices of the Stockham stages in order.
#define SRWF_GEN_MAX 2048
struct srwf_gplan {
    int n, nst;
    int R[16];
};

extern "C" {
// Plan of a line length; -1 when n is odd, out of range or has another prime factor.
int srwf_gplan_make(int n, srwf_gplan* pl);
int srwf_gen_init_attributes();
// Mixed-radix passes (fft_gen.hip).  twg: exp(-2 pi i m / n), m < n, of the
// line length of the pass.  Row pass dir 0: srwf_pre (preOp) then the
// forward transform (or the inverse direction when inv); dir 1: inverse
// transform then srwf_post (postOp; 2 writes the intensity of slice pairs).
int srwf_launch_grows(float2* data, long long sliceStride, int nSlices, const srwf_slice* dSp, const float2* twg,
                      const srwf_gplan& pl, int ny, int dir, int preOp, int inv, int postOp, float* inten,
                      long long intenStride, hipStream_t st);
// Column pass on transposed [NX][NY] slices: op 0 / 1 forward, srwf_mid,
// inverse, in place; op 2 forward (or inverse direction) and srwf_std_out
// into out ([NY][NX] per slice).
int srwf_launch_gcols(float2* dataT, long long sliceStride, int nSlices, const srwf_slice* dSp, const float2* twg,
                      const srwf_gplan& pl, int NX, int op, int inv, float2* out, hipStream_t st);
int srwf_launch_gtranspose(const float2* src, float2* dst, int rows, int cols, long long sliceStride, int nSlices,
                           hipStream_t st);
int srwf_rows_per_wg(int logn);
int srwf_cols_per_wg(int logn);
void srwf_default_tiles(int lognx, int logny, int postOp, srwf_tiles* t);
// 1 when the shape is valid for nx = 2^lognx, ny = 2^logny (LDS, thread counts).
int srwf_tiles_fit(int lognx, int logny, int postOp, const srwf_tiles* t);
// tiles: nullptr (or a shape that does not fit) selects the default shape.
int srwf_launch_rows_fwd(float2* data, long long sliceStride, int nSlices, const srwf_slice* dSp, const float2* tw,
                         int logn, int ny, int preOp, int inv, const srwf_tiles* tiles, hipStream_t st);
int srwf_launch_cols(float2* data, long long sliceStride, int nSlices, const srwf_slice* dSp, const float2* twy,
                     int logny, int lognx, int op, int inv, float2* out, const srwf_tiles* tiles, hipStream_t st);
int srwf_launch_rows_inv(float2* data, long long sliceStride, int nSlices, const srwf_slice* dSp, const float2* tw,
                         int logn, int ny, int postOp, float* inten, long long intenStride, const srwf_tiles* tiles,
                         hipStream_t st);
int srwf_launch_gather(const float2* src, float2* dst, int ne, long long nxy, int scatter, hipStream_t st);
// Both field components between SRW's [iy][ix][ie] order (ex, ez) and the
// slice-major d[(2*ie + c)*nxy + p]: toSlices = 1 reads ex/ez and writes d,
// 0 reads d and writes ex/ez.
int srwf_launch_slices(float2* ex, float2* ez, float2* d, int ne, long long nxy, int toSlices, hipStream_t st);
int srwf_init_attributes();
}

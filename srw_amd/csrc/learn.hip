// learn.hip - learned importance-sampling proposal for the macro-electron sum.
//
// q(g) = alpha * p(g) + (1 - alpha) * sum_k pi_k N(g; mu_k, diag sigma_k^2)
// over the standardized 6-D electron phase space g (p = N(0, I)); the mixture
// parameters are the output of an MLP 6 -> 64 (tanh) -> 64 (tanh) -> 104
// whose input is a 6-D descriptor of the beam and beamline.  Training
// minimises -sum_j w^_j log q(g_j) with w^_j proportional to
// p(g_j) c_j / q_old(g_j), c_j the electron's L2 intensity norm (the target
// q* ~ p*c is the variance-optimal proposal for the L2 error of the image).
// All kernels are FP32 and single-workgroup except the per-sample gradient.
#include "srw_learn.h"

#include <math.h>

namespace {

__device__ __forceinline__ float tanh_f(float x) { return tanhf(x); }

// Forward pass for one input vector; stores activations for the backward pass.
__global__ void __launch_bounds__(128) k_mlp_forward(const float* __restrict__ P, const float* __restrict__ in,
                                                     float* act)
{
    __shared__ float c[SRWL_IN], h1[SRWL_H], h2[SRWL_H];
    const float *W1 = P + SRWL_OFF_W1, *b1 = P + SRWL_OFF_B1, *W2 = P + SRWL_OFF_W2, *b2 = P + SRWL_OFF_B2;
    const float *W3 = P + SRWL_OFF_W3, *b3 = P + SRWL_OFF_B3;
    int t = threadIdx.x;
    if (t < SRWL_IN) c[t] = in[t];
    __syncthreads();
    if (t < SRWL_H) {
        float s = b1[t];
        for (int i = 0; i < SRWL_IN; i++) s += W1[t * SRWL_IN + i] * c[i];
        h1[t] = tanh_f(s);
    }
    __syncthreads();
    if (t < SRWL_H) {
        float s = b2[t];
        for (int i = 0; i < SRWL_H; i++) s += W2[t * SRWL_H + i] * h1[i];
        h2[t] = tanh_f(s);
    }
    __syncthreads();
    if (t < SRWL_OUT) {
        float s = b3[t];
        for (int i = 0; i < SRWL_H; i++) s += W3[t * SRWL_H + i] * h2[i];
        act[SRWL_ACT_OUT + t] = s;
    }
    if (t < SRWL_H) { act[SRWL_ACT_H1 + t] = h1[t]; act[SRWL_ACT_H2 + t] = h2[t]; }
    if (t < SRWL_IN) act[SRWL_ACT_IN + t] = c[t];
}

// Mixture parameters from the network output (log sigma clamped to [-3, 1]).
__device__ __forceinline__ void gmm_params(const float* o, int k, float& logit, const float*& mu, float* ls,
                                           float* lsMask)
{
    logit = o[k];
    mu = o + SRWL_K + k * SRWL_D;
    for (int d = 0; d < SRWL_D; d++) {
        float v = o[SRWL_K + SRWL_K * SRWL_D + k * SRWL_D + d];
        lsMask[d] = (v > -3.f && v < 1.f) ? 1.f : 0.f;
        ls[d] = fminf(fmaxf(v, -3.f), 1.f);
    }
}

// Per-sample gradient of -w^_j log q(g_j) with respect to the network output,
// accumulated over the batch with atomics.  One thread per sample.
__global__ void __launch_bounds__(256) k_gmm_grad(const float* __restrict__ act, const float* __restrict__ g,
                                                  const float* __restrict__ wh, int nb, float alpha, float* gout,
                                                  float* loss)
{
    int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= nb) return;
    const float* o = act + SRWL_ACT_OUT;
    const float* x = g + j * SRWL_D;
    const float LOG2PI = 1.8378770664093453f;
    // log p and log N_k
    float lp = 0.f;
    for (int d = 0; d < SRWL_D; d++) lp += -0.5f * x[d] * x[d];
    lp -= 0.5f * SRWL_D * LOG2PI;
    float lmax = -1e30f, lse = 0.f;
    for (int k = 0; k < SRWL_K; k++) lmax = fmaxf(lmax, o[k]);
    for (int k = 0; k < SRWL_K; k++) lse += expf(o[k] - lmax);
    float logZ = lmax + logf(lse);
    float lN[SRWL_K];
    for (int k = 0; k < SRWL_K; k++) {
        float logit; const float* mu; float ls[SRWL_D], msk[SRWL_D];
        gmm_params(o, k, logit, mu, ls, msk);
        float s = logit - logZ;
        for (int d = 0; d < SRWL_D; d++) {
            float z = (x[d] - mu[d]) * expf(-ls[d]);
            s += -0.5f * z * z - ls[d] - 0.5f * LOG2PI;
        }
        lN[k] = s;   // log(pi_k N_k)
    }
    float mmax = -1e30f;
    for (int k = 0; k < SRWL_K; k++) mmax = fmaxf(mmax, lN[k]);
    float ms = 0.f;
    for (int k = 0; k < SRWL_K; k++) ms += expf(lN[k] - mmax);
    float lm = mmax + logf(ms);                                   // log m(g)
    float top = fmaxf(lm, lp);
    float lq = top + logf(alpha * expf(lp - top) + (1.f - alpha) * expf(lm - top));
    float w = wh[j];
    atomicAdd(loss, -w * lq);
    // d(-w log q)/d o: -w (1-alpha) m / q * d log m / d o
    float f = -w * (1.f - alpha) * expf(lm - lq);
    for (int k = 0; k < SRWL_K; k++) {
        float r = expf(lN[k] - lm);                               // responsibility pi_k N_k / m
        float logit; const float* mu; float ls[SRWL_D], msk[SRWL_D];
        gmm_params(o, k, logit, mu, ls, msk);
        float pik = expf(logit - logZ);
        atomicAdd(&gout[k], f * (r - pik));
        for (int d = 0; d < SRWL_D; d++) {
            float is2 = expf(-2.f * ls[d]);
            float dz = x[d] - mu[d];
            atomicAdd(&gout[SRWL_K + k * SRWL_D + d], f * r * dz * is2);
            atomicAdd(&gout[SRWL_K + SRWL_K * SRWL_D + k * SRWL_D + d], f * r * (dz * dz * is2 - 1.f) * msk[d]);
        }
    }
}

// Backward pass of the MLP for one input (gradient of the output in gout),
// writes the parameter gradient.
__global__ void __launch_bounds__(128) k_mlp_backward(const float* __restrict__ P, const float* __restrict__ act,
                                                      const float* __restrict__ gout, float* G)
{
    __shared__ float go[SRWL_OUT], d2[SRWL_H], d1[SRWL_H];
    const float *W2 = P + SRWL_OFF_W2, *W3 = P + SRWL_OFF_W3;
    const float *c = act + SRWL_ACT_IN, *h1 = act + SRWL_ACT_H1, *h2 = act + SRWL_ACT_H2;
    int t = threadIdx.x;
    if (t < SRWL_OUT) go[t] = gout[t];
    __syncthreads();
    if (t < SRWL_OUT) {
        for (int i = 0; i < SRWL_H; i++) G[SRWL_OFF_W3 + t * SRWL_H + i] = go[t] * h2[i];
        G[SRWL_OFF_B3 + t] = go[t];
    }
    if (t < SRWL_H) {
        float s = 0.f;
        for (int o = 0; o < SRWL_OUT; o++) s += W3[o * SRWL_H + t] * go[o];
        d2[t] = s * (1.f - h2[t] * h2[t]);
    }
    __syncthreads();
    if (t < SRWL_H) {
        for (int i = 0; i < SRWL_H; i++) G[SRWL_OFF_W2 + t * SRWL_H + i] = d2[t] * h1[i];
        G[SRWL_OFF_B2 + t] = d2[t];
        float s = 0.f;
        for (int o = 0; o < SRWL_H; o++) s += W2[o * SRWL_H + t] * d2[o];
        d1[t] = s * (1.f - h1[t] * h1[t]);
    }
    __syncthreads();
    if (t < SRWL_H) {
        for (int i = 0; i < SRWL_IN; i++) G[SRWL_OFF_W1 + t * SRWL_IN + i] = d1[t] * c[i];
        G[SRWL_OFF_B1 + t] = d1[t];
    }
}

// Adam step (Kingma & Ba, 2015) on all parameters.
__global__ void k_adam(float* P, const float* __restrict__ G, float* m, float* v, int n, float lr, float b1, float b2,
                       float eps, int step)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    float g = G[i];
    float mi = b1 * m[i] + (1.f - b1) * g, vi = b2 * v[i] + (1.f - b2) * g * g;
    m[i] = mi; v[i] = vi;
    float mh = mi / (1.f - powf(b1, (float)step)), vh = vi / (1.f - powf(b2, (float)step));
    P[i] -= lr * mh / (sqrtf(vh) + eps);
}

}  // namespace

extern "C" {
int srwl_launch_mlp_forward(const float* P, const float* in, float* act, hipStream_t st)
{
    hipLaunchKernelGGL(k_mlp_forward, dim3(1), dim3(128), 0, st, P, in, act);
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

int srwl_launch_gmm_grad(const float* act, const float* g, const float* wh, int nb, float alpha, float* gout,
                         float* loss, hipStream_t st)
{
    hipLaunchKernelGGL(k_gmm_grad, dim3((nb + 255) / 256), dim3(256), 0, st, act, g, wh, nb, alpha, gout, loss);
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

int srwl_launch_mlp_backward(const float* P, const float* act, const float* gout, float* G, hipStream_t st)
{
    hipLaunchKernelGGL(k_mlp_backward, dim3(1), dim3(128), 0, st, P, act, gout, G);
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

int srwl_launch_adam(float* P, const float* G, float* m, float* v, int n, float lr, int step, hipStream_t st)
{
    hipLaunchKernelGGL(k_adam, dim3((n + 255) / 256), dim3(256), 0, st, P, G, m, v, n, lr, 0.9f, 0.999f, 1e-8f,
                       step);
    return hipGetLastError() == hipSuccess ? 0 : -1;
}
}

// srw_learn.h - layout of the learned proposal network (FP32 parameters).
#pragma once
#include <hip/hip_runtime.h>

#define SRWL_IN 6            // beam / beamline descriptor
#define SRWL_H 64            // hidden width
#define SRWL_K 8             // mixture components
#define SRWL_D 6             // phase-space dimension
#define SRWL_OUT (SRWL_K + 2 * SRWL_K * SRWL_D)   // logits, means, log sigmas = 104

#define SRWL_OFF_W1 0
#define SRWL_OFF_B1 (SRWL_OFF_W1 + SRWL_H * SRWL_IN)
#define SRWL_OFF_W2 (SRWL_OFF_B1 + SRWL_H)
#define SRWL_OFF_B2 (SRWL_OFF_W2 + SRWL_H * SRWL_H)
#define SRWL_OFF_W3 (SRWL_OFF_B2 + SRWL_H)
#define SRWL_OFF_B3 (SRWL_OFF_W3 + SRWL_OUT * SRWL_H)
#define SRWL_NPARAM (SRWL_OFF_B3 + SRWL_OUT)       // 11368

#define SRWL_ACT_IN 0
#define SRWL_ACT_H1 (SRWL_ACT_IN + SRWL_IN)
#define SRWL_ACT_H2 (SRWL_ACT_H1 + SRWL_H)
#define SRWL_ACT_OUT (SRWL_ACT_H2 + SRWL_H)
#define SRWL_NACT (SRWL_ACT_OUT + SRWL_OUT)

extern "C" {
int srwl_launch_mlp_forward(const float* P, const float* in, float* act, hipStream_t st);
int srwl_launch_gmm_grad(const float* act, const float* g, const float* wh, int nb, float alpha, float* gout,
                         float* loss, hipStream_t st);
int srwl_launch_mlp_backward(const float* P, const float* act, const float* gout, float* G, hipStream_t st);
int srwl_launch_adam(float* P, const float* G, float* m, float* v, int n, float lr, int step, hipStream_t st);
}

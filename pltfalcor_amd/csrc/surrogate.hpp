// The learned BSDF surrogate's device state, shared by the MLP kernels (mlp.hip) and the exact
// sample generator / evaluator (surrogate_data.hip).
#pragma once
#include <hip/hip_runtime.h>
#include <vector>
#include "../../include/plt_abi.h"

struct plt_surrogate {
    plt_context* ctx = nullptr;
    uint32_t max_batch = 0;
    // FP32 master parameters (the public order), gradients, Adam moments
    float *params = nullptr, *grads = nullptr, *adam_m = nullptr, *adam_v = nullptr;
    uint32_t step = 0;
    // bf16 operands: W (row-major [out x in_pad]) and Wt ([in x out_pad]) per layer
    __bf16 *W1 = nullptr, *W1t = nullptr, *W2 = nullptr, *W2t = nullptr, *W3 = nullptr, *W3t = nullptr, *W4 = nullptr, *W4t = nullptr;
    // activations: sample-major and feature-major (ld = max_batch) copies
    __bf16 *Xb = nullptr, *Xt = nullptr, *H1 = nullptr, *H1t = nullptr, *H2 = nullptr, *H2t = nullptr, *H3 = nullptr, *H3t = nullptr;
    __bf16 *dY = nullptr, *dYt = nullptr, *G3 = nullptr, *G3t = nullptr, *G2 = nullptr, *G2t = nullptr, *G1t = nullptr;
    float *Y = nullptr, *part = nullptr, *loss_part = nullptr;
    float* bias_part = nullptr;       // bias-gradient partials per batch tile
    uint32_t* train_state = nullptr;  // device-side training step state (plt_surrogate_train)
    // target statistics and the inputs' training box (plt_surrogate_set_stats): 16 + 16 + 22 + 22
    float* stats = nullptr;
    bool has_stats = false;
    uint32_t* counter = nullptr;  // out-of-distribution fallbacks of one evaluation
    std::vector<void*> owned;
};

namespace plt {
// Fused inference of n samples X [n x 22] (device): the raw outputs into Y [n x 16], or, with
// Sout, the decoded Mueller matrix applied to Sin [n x 4] for samples inside the training box
// (the others are left untouched for the exact pass).
int surrogate_infer(plt_surrogate* s, int n, const float* X, float* Y, const float* Sin, float* Sout, hipStream_t st);
// Raw outputs Y of the first min(*n_dev, cap) rows of X, the count read on the device.
int surrogate_infer_counted(plt_surrogate* s, const uint32_t* n_dev, uint32_t cap, const float* X, float* Y, hipStream_t st);
}  // namespace plt

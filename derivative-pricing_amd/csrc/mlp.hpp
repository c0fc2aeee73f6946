// Device-side entry points of the continuation-value network (mlp.hip), for pipelines that
// keep every buffer resident in HBM.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace mcdx {

constexpr int kMlpParams = 4417; // 2 -> 64 -> 64 -> 1

// Deterministic initialisation from Philox(seed): N(0,1) first layer and biases, He-scaled
// second layer, 1/8-scaled output layer, output bias = y_mean.
int mlp_init_device(float* w, uint64_t seed, float y_mean, hipStream_t s);
// `steps` Adam steps (zero initial moments) of 128 samples drawn by Philox(seed) from
// (xs, ts, ys)[0..n); samples whose target is NaN are redrawn (up to 8 times, then masked).
// loss[step] receives the mean squared error of the step's first 64 samples when non-null.
// With grad non-null: one pass over samples 0..127, gradients to grad, no update.
// avg_from >= 0: w receives the mean of the iterates after steps avg_from .. steps - 1
// (Polyak tail averaging) instead of the last iterate.
int mlp_train_device(float* w, const float* xs, const float* ts, const float* ys, uint64_t n,
                     int steps, float lr, float lr_decay, uint64_t seed, float* loss, float* grad,
                     hipStream_t s, int avg_from = -1);
// Features of an exercise-date grid, formed in the evaluation kernel: sample
// i = (d - 1) grid + k (d = 1 .. m - 1, k < grid) has features
// (feat_scale (xlo + h k - 1), 2 d / m - 1).
struct MlpEvalGrid {
    uint32_t grid; // 0: features from xs, ts
    int32_t m;
    double xlo, h, feat_scale;
};
// out[i] = network(x_i, t_i), i < n, on the MFMA forward kernel.
int mlp_eval_device(const float* w, const float* xs, const float* ts, uint64_t n, float* out,
                    hipStream_t s, const MlpEvalGrid* grid = nullptr);

} // namespace mcdx

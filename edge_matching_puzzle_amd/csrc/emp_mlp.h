// emp_mlp.h -- the candidate scorer: a 126-256-256-1 perceptron with BF16
// weights in the two dense layers and FP32 accumulation, trained and run by
// hand-written MFMA kernels (emp_mlp.hip).  It has no counterpart in the
// reference; it only orders candidates and never removes one.
#ifndef EMP_MLP_H
#define EMP_MLP_H

#include <hip/hip_runtime.h>

#include <cstdint>
#include <vector>

#include "emp_enumerate.h"

namespace emp {

constexpr uint32_t kMlpIn = 126;     // features per (situation, candidate)
constexpr uint32_t kMlpInPad = 128;  // padded to the MFMA K step
constexpr uint32_t kMlpHidden = 256;

// FP32 parameter blob, in this order:
//   W1 [kMlpInPad][kMlpHidden]   (rows 126, 127 unused, kept 0)
//   b1 [kMlpHidden]
//   W2 [kMlpHidden][kMlpHidden]
//   b2 [kMlpHidden]
//   w3 [kMlpHidden]
//   b3 [1]
constexpr uint32_t kMlpOffW1 = 0;
constexpr uint32_t kMlpOffB1 = kMlpOffW1 + kMlpInPad * kMlpHidden;
constexpr uint32_t kMlpOffW2 = kMlpOffB1 + kMlpHidden;
constexpr uint32_t kMlpOffB2 = kMlpOffW2 + kMlpHidden * kMlpHidden;
constexpr uint32_t kMlpOffW3 = kMlpOffB2 + kMlpHidden;
constexpr uint32_t kMlpOffB3 = kMlpOffW3 + kMlpHidden;
constexpr uint32_t kMlpParams = kMlpOffB3 + 1;

struct MlpTrainConfig {
    uint32_t epochs = 4;
    uint32_t batch = 1024;       // rows per step (multiple of 64)
    float lr = 1e-3f;            // Adam
    uint64_t seed = 1;           // batch order
};

// Mean squared error of score against target over the rows; the gradients of
// every parameter (same layout as the blob).  n <= 65536.
void mlp_gradients_on_device(hipStream_t stream, Workspace &ws, const float *params, const float *features,
                             const float *targets, uint64_t n, float *grads, float &loss);

// Adam on the MSE, batches drawn in a seeded order; params updated in place;
// loss_per_epoch[e] = mean batch loss of epoch e.
void mlp_train_on_device(hipStream_t stream, Workspace &ws, const float *features, const float *targets, uint64_t n,
                         const MlpTrainConfig &cfg, float *params, std::vector<float> &loss_per_epoch);

// Scores of n rows.
void mlp_score_on_device(hipStream_t stream, Workspace &ws, const float *params, const float *features, uint64_t n,
                         float *scores);

}  // namespace emp

#endif

// The learned precision selector's kernels (selector.hip) for the host driver (oosg.cpp).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace oosg
{

constexpr int SELECTOR_FEATURES = 8;
constexpr int SELECTOR_HIDDEN = 32;
constexpr int SELECTOR_PARAMETERS =
    SELECTOR_HIDDEN * SELECTOR_FEATURES + SELECTOR_HIDDEN + SELECTOR_HIDDEN * SELECTOR_HIDDEN + SELECTOR_HIDDEN +
    SELECTOR_HIDDEN + 1;

// p(x) for count × 8 features (device arrays); a1, a2 (count × 32, the hidden activations) and logit may be null.
hipError_t selectorForward(const float* params, const float* x, int64_t count, float* a1, float* a2,
                           float* probability, float* logit, hipStream_t stream);
// The gradient of the mean binary cross-entropy by the parameters, and the loss; work arrays a1, a2, d2, d1
// (count × 32), probability, logit, d3 (count).
hipError_t selectorGradient(const float* params, const float* x, const float* labels, int64_t count, float* a1,
                            float* a2, float* probability, float* logit, float* d3, float* d2, float* d1,
                            float* gradient, float* loss, hipStream_t stream);
// One Adam step (step counts from 1).
hipError_t selectorAdam(float* params, const float* gradient, float* m, float* v, float lr, int step,
                        hipStream_t stream);

} // namespace oosg

// Host interface of the differentiable rendering kernels.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>

#include "diff.h"

namespace wpt {

struct DiffContext {
    DevScene scene;
    CameraParams camera;
    DiffLayout layout;
    DiffParams params;
    DiffGrads grads;           // one vector: texture | roughness | emission
    uint32_t width, height, spp, sampleStart, maxDepth, rngStrategy;
    float* pathRadiance;       // SoA: 3 x (spp * width * height), index sample * pixels + pixel
    float* image;              // width * height * 3, mean over samples
    const float* imageGrad;    // dLoss / dimage, width * height * 3
    uint32_t* cursor;          // work claims
    uint2* spill;
    uint32_t spillThreads;
};

struct AdamStep {
    float* params;
    const float* grad;
    float* m;
    float* v;
    size_t count;
    size_t textureFloats, roughnessCount;  // section boundaries for the clamps
    float lr, beta1, beta2, epsilon;
    float bias1, bias2;                    // 1 - beta^t
    float textureRange[2], roughnessRange[2], emissionRange[2];
};

// Forward: renders every path, keeps each path's radiance, writes the image.
hipError_t enqueueDiffForward(const DiffContext& c, uint32_t blocks, uint32_t blockSize, hipStream_t stream);
// Adjoint: zeroes the gradient vector (gradFloats), replays every path.
hipError_t enqueueDiffAdjoint(const DiffContext& c, uint32_t blocks, uint32_t blockSize, size_t gradFloats,
    hipStream_t stream);
hipError_t enqueueAdam(const AdamStep& a, hipStream_t stream);
hipError_t diffOccupancy(uint32_t blockSize, int* forwardPerCu, int* adjointPerCu);

}  // namespace wpt

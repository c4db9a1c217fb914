// Near-separatrix classifier (3-D CNN) shared declarations.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <vector>

#include "ctx.hpp"

namespace msceer {

// Flat FP32 parameter vector: W1[27][16] b1[16] W2[27][32][16] b2[32]
// W3[27][16][32] b3[16] wh[16] bh.
constexpr int kW1 = 27 * 16, kW2 = 27 * 32 * 16, kW3 = 27 * 16 * 32;
constexpr int kOffW1 = 0, kOffB1 = kW1, kOffW2 = kOffB1 + 16, kOffB2 = kOffW2 + kW2;
constexpr int kOffW3 = kOffB2 + 32, kOffB3 = kOffW3 + kW3, kOffWh = kOffB3 + 16;
constexpr int kNumParams = kOffWh + 17;
constexpr int kGradBlocks = 512;           // work-groups of the weight-gradient kernels
constexpr int kPartWidth = kW2 + 32;       // widest per-block partial row
constexpr size_t kPackedBytes = 4096 + 2 * (size_t)(2 * kW2 + 2 * kW3);

// Device operands packed from the parameters (BF16 convolution weights, their
// mirrored transposes for the backward convolutions, FP32 biases and head).
struct CnnPacked {
    float* w1;      // [27][16], rounded to BF16
    float* b1;      // [16]
    float* b2;      // [32]
    float* b3;      // [16]
    float* head;    // wh[16], bh
    uint16_t* w2;   // [27][32][16]
    uint16_t* w2t;  // [27][16][32]: w2t[26-t][ci][co] = w2[t][co][ci]
    uint16_t* w3;   // [27][16][32]
    uint16_t* w3t;  // [27][32][16]
};

struct CnnState {
    DevBuf params, grads, m, v, pk, parts, loss;
    DevBuf tmp1, tmp2, xin, a1, a2, a3, prob, m0, m1;
    DevBuf dz3, dz2, dz1;
    int step = 0;
    int64_t vol_cap = 0;
    hipError_t ensure_volume(int64_t verts);
    hipError_t ensure_train(int64_t verts);
    CnnPacked packed() const;
    void release_all();
};

std::vector<float> cnn_initial_params(uint64_t seed);
hipError_t cnn_pack(CnnState& s, hipStream_t st);
// Normalisation + convolutions; keep_a3: store conv3's activations (training)
// instead of the head's probabilities.
hipError_t cnn_forward(CnnState& s, const float* values, int X, int Y, int Z, bool keep_a3, hipStream_t st);
// Loss and gradient of the last forward(keep_a3) into s.grads (loss in s.loss[0]).
hipError_t cnn_backward(CnnState& s, int X, int Y, int Z, const uint8_t* target, float pos_w, hipStream_t st);
hipError_t cnn_adam(CnnState& s, float lr, hipStream_t st);
// Mask from the last forward's probabilities: prob > thr, dilated (Chebyshev radius).
hipError_t cnn_mask(CnnState& s, int X, int Y, int Z, float thr, int dilation, uint8_t* out,
                    unsigned long long* count, hipStream_t st);

}  // namespace msceer

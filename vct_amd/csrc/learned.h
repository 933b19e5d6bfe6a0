// learned.h -- shapes and device buffers of the learned diffuse gather.
#pragma once

#include "vct_internal.h"

namespace vctk {

// 129 -> 64 -> 64 -> 64 -> 3; flat FP32 parameter vector in the order
// W1 (64x129, row-major out x in), b1, W2 (64x64), b2, W3 (64x64), b3,
// W4 (3x64), b4.
constexpr int kMlpIn = 160;  // 129 inputs padded to a multiple of 32 (MFMA K)
constexpr int kMlpW1 = 64 * 129;
constexpr int kMlpW2 = 64 * 64;
constexpr int kMlpW3 = 64 * 64;
constexpr int kMlpW4 = 3 * 64;
constexpr int kMlpParams = kMlpW1 + 64 + kMlpW2 + 64 + kMlpW3 + 64 + kMlpW4 + 3;  // 16835
constexpr int kMlpPartialStride = 16896;  // >= kMlpParams + 1 (loss), 256-aligned

struct MlpBuffers {
    float *master = nullptr, *m1 = nullptr, *m2 = nullptr;  // FP32 weights and Adam moments
    uint16_t *w1 = nullptr, *w2 = nullptr, *w3 = nullptr, *w4 = nullptr;  // BF16 operands
    uint16_t *w2t = nullptr, *w3t = nullptr, *w4t = nullptr;              // transposed BF16 operands
    float *partials = nullptr;  // [max_blocks][kMlpPartialStride]
    int max_blocks = 0;
    float out_scale = 1.0f;
};

hipError_t launch_label64(const PyrDesc &pyr, const vct_gbuffer_texel *pts, int npts, Extent ext, float4 *out,
                          hipStream_t s);
hipError_t launch_features(const PyrDesc &pyr, const vct_gbuffer_texel *pts, int npts, Extent ext, uint16_t *X,
                           hipStream_t s);
hipError_t launch_mlp_train_step(MlpBuffers &B, const uint16_t *X, const float4 *labels, int npts, int batch,
                                 uint32_t seed, int step, float lr, float b1, float b2, float eps, float *loss_out,
                                 hipStream_t s);
size_t mlp_train_smem_bytes();
// Learned diffuse gather over the G-buffer of t into t.out_ind; tiles whose
// predicted-luminance variance exceeds threshold are traced with the 6-cone
// gather instead.  fallback_list holds 1 + (number of 8x8 tiles) words: the
// count of such tiles, then their indices.
hipError_t launch_learned(const MlpBuffers &B, const PyrDesc &pyr, const TraceLaunch &t, float threshold,
                          uint32_t *fallback_list, hipStream_t s);
hipError_t launch_emit_operands(MlpBuffers &B, hipStream_t s);

}  // namespace vctk

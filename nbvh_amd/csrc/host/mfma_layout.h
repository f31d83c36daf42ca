// mfma_layout.h -- host-side arrangement of the MLP weights into v_mfma_f32_32x32x16 A-operand
// fragments (see mfma_layout.cpp for the index maps).
#pragma once

#include <cstdint>
#include <vector>

#include "../nbvh_internal.h"

namespace nbvh {

struct MfmaWeights {
    std::vector<uint16_t> frag16;  // FP16 fragments followed by BF16 fragments (same order)
    std::vector<uint8_t> frag8;    // FP8 e4m3 fragments of the hidden 64x64 layers
    float fp8_wscale[8];           // per hidden layer: weights were multiplied by this before e4m3
};

// Number of fragments (each 64 lanes x 8 elements) of the whole MLP.
uint32_t mfma_fragment_count(const nbvh_network_config& cfg);

// Feature (layer-1 input) index held by lane group g, k-step s, element e.
inline uint32_t mfma_l1_k(uint32_t K0, uint32_t g, uint32_t s, uint32_t e) { return (K0 / 2) * g + 8 * s + e; }

// Neuron (hidden-layer input) index held by lane group g, k-step s, element e: the rows that the
// 32x32 accumulator of the previous layer leaves in registers 8*(s%2)+e of row block s/2.
inline uint32_t mfma_h_k(uint32_t g, uint32_t s, uint32_t e)
{
    return 32 * (s / 2) + 16 * (s % 2) + 8 * (e / 4) + 4 * g + (e % 4);
}

MfmaWeights mfma_weight_layout(const nbvh_network_config& cfg, const std::vector<float>& mlp);

}  // namespace nbvh

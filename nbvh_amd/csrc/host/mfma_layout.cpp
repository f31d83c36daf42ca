// mfma_layout.cpp -- MLP weights as MFMA A-operand fragments.
//
// The MFMA kernels evaluate each layer transposed, H_next^T = W * H^T, with
// v_mfma_f32_32x32x16_{f16,bf16}: A = 32 neurons x 16 inputs (weights), B = 16 inputs x 32
// segments (activations), D = 32 neurons x 32 segments. Lane l supplies row/column l%32 and
// eight consecutive K elements of group g = l/32; the accumulator leaves neuron
// 8*(r/4) + 4*g + r%4 of its row block in register r of lane (g, segment). Because the
// contraction pairs A's element (g, e) with B's element (g, e), the weights can be permuted so
// that the accumulator registers of one layer ARE the B operand of the next (after ReLU and the
// conversion to 16 bits): no shuffle, no LDS round trip between layers.
//
// Fragment order: layer, row block rb (32 neurons), k-step s (16 inputs); within a fragment
// lane-major, 8 elements per lane (16 bytes for FP16/BF16: one ds_read_b128 per lane).
#include "mfma_layout.h"

#include <cmath>

#include "host_common.h"

namespace nbvh {

uint32_t mfma_fragment_count(const nbvh_network_config& cfg)
{
    const uint32_t K0 = cfg.n_points * cfg.grid.n_levels * cfg.grid.n_features_per_level;
    const uint32_t H  = cfg.n_neurons;
    uint32_t n = (H / 32) * (K0 / 16);                          // layer 1
    n += (cfg.n_hidden_layers - 1) * (H / 32) * (H / 16);       // hidden layers
    n += 1 * (H / 16);                                          // output layer: one row block (8 of 32 rows used)
    return n;
}

MfmaWeights mfma_weight_layout(const nbvh_network_config& cfg, const std::vector<float>& mlp)
{
    MfmaWeights out;
    const uint32_t K0 = cfg.n_points * cfg.grid.n_levels * cfg.grid.n_features_per_level;
    const uint32_t H  = cfg.n_neurons;
    for (int i = 0; i < 8; i++) out.fp8_wscale[i] = 1.f;
    // Shapes the MFMA kernels implement: K0 multiple of 16 and the feature split in halves.
    if (H != 64 || K0 % 16 != 0 || (K0 / 2) % cfg.grid.n_features_per_level != 0) return out;
    const uint32_t nf = mfma_fragment_count(cfg);
    std::vector<float> frag((size_t)nf * 512, 0.f);
    size_t f = 0, off = 0;
    uint32_t in = K0;
    for (uint32_t layer = 0; layer <= cfg.n_hidden_layers; layer++) {
        const bool last = layer == cfg.n_hidden_layers;
        const uint32_t rows = last ? cfg.n_outputs : H;
        const uint32_t rbs = last ? 1 : H / 32;
        for (uint32_t rb = 0; rb < rbs; rb++)
            for (uint32_t s = 0; s < in / 16; s++, f++)
                for (uint32_t lane = 0; lane < 64; lane++)
                    for (uint32_t e = 0; e < 8; e++) {
                        const uint32_t n = 32 * rb + lane % 32, g = lane / 32;
                        const uint32_t k = layer == 0 ? mfma_l1_k(K0, g, s, e) : mfma_h_k(g, s, e);
                        frag[f * 512 + lane * 8 + e] = n < rows ? mlp[off + (size_t)n * in + k] : 0.f;
                    }
        off += (size_t)rows * in;
        in = H;
    }
    out.frag16.resize((size_t)nf * 512 * 2);
    for (size_t i = 0; i < frag.size(); i++) {
        out.frag16[i]               = float_to_half(frag[i]);
        out.frag16[frag.size() + i] = float_to_bf16(frag[i]);
    }
    // FP8 e4m3 for the hidden 64x64 layers (layers 2..H): per-layer power-of-two scale that maps
    // the largest weight magnitude to [224, 448).
    const uint32_t l1_frags = (H / 32) * (K0 / 16), h_frags = (H / 32) * (H / 16);
    out.frag8.resize((size_t)(cfg.n_hidden_layers - 1) * h_frags * 512);
    for (uint32_t hl = 0; hl + 1 < cfg.n_hidden_layers && hl < 8; hl++) {
        const size_t base = (size_t)(l1_frags + hl * h_frags) * 512;
        float amax = 0.f;
        for (size_t i = 0; i < (size_t)h_frags * 512; i++) amax = std::fmax(amax, std::fabs(frag[base + i]));
        float sc = 1.f;
        if (amax > 0.f) {
            int e;
            std::frexp(448.f / amax, &e);  // 448/amax = m * 2^e, m in [0.5, 1)
            sc = std::ldexp(1.f, e - 1);
        }
        out.fp8_wscale[hl] = sc;
        for (size_t i = 0; i < (size_t)h_frags * 512; i++)
            out.frag8[(size_t)hl * h_frags * 512 + i] = float_to_e4m3(frag[base + i] * sc);
    }
    return out;
}

}  // namespace nbvh

// network.cpp -- hash-grid level table, parameter counts, seeded initialisation, number formats.
//
// The network is the reference's tcnn NetworkWithInputEncoding built at
// src/cuda/neural/neural_bvh_module.cu:170-262: a RepeatedComposite of one HashGrid shared by the
// P segment points (features concatenated point by point), then a FullyFusedMLP (no biases,
// ReLU hidden layers, Sigmoid on every output). tiny-cuda-nn itself is not in the reference tree;
// its public GridEncoding conventions are restated: per-level scale exp2(l*log2 b)*N_min - 1,
// resolution ceil(scale)+1, level size min(next_multiple(res^3, 8), 2^log2T), dense indexing
// while the stride stays within the level size, else the coherent prime hash (1, 2654435761,
// 805459861), entries of F consecutive features.
#include <cmath>
#include <cstring>
#include <vector>

#include "host_common.h"

namespace nbvh {

void network_validate(const nbvh_network_config& c)
{
    const auto& g = c.grid;
    if (g.n_levels == 0 || g.n_levels > (uint32_t)kMaxLevels) fail(NBVH_ERR_UNSUPPORTED, "grid: n_levels must be 1..32");
    if (!(g.n_features_per_level == 1 || g.n_features_per_level == 2 || g.n_features_per_level == 4 ||
          g.n_features_per_level == 8))
        fail(NBVH_ERR_UNSUPPORTED, "grid: n_features_per_level must be 1, 2, 4 or 8");
    if (g.log2_hashmap_size < 4 || g.log2_hashmap_size > 26) fail(NBVH_ERR_UNSUPPORTED, "grid: log2_hashmap_size must be 4..26");
    if (g.base_resolution < 1 || g.base_resolution > 4096) fail(NBVH_ERR_UNSUPPORTED, "grid: base_resolution must be 1..4096");
    if (!(g.per_level_scale >= 1.f && g.per_level_scale <= 4.f)) fail(NBVH_ERR_UNSUPPORTED, "grid: per_level_scale must be in [1, 4]");
    if (c.n_points < 1 || c.n_points > 4) fail(NBVH_ERR_UNSUPPORTED, "network: n_points must be 1..4");
    if (c.segment_distribution > NBVH_SEGMENT_UNIFORM)
        fail(NBVH_ERR_UNSUPPORTED, "network: segment_distribution must be 0, 1 or 2");
    if (c.n_neurons != 64) fail(NBVH_ERR_UNSUPPORTED, "network: n_neurons must be 64");
    if (c.n_hidden_layers < 1 || c.n_hidden_layers > 8) fail(NBVH_ERR_UNSUPPORTED, "network: n_hidden_layers must be 1..8");
    if (c.n_outputs != 8) fail(NBVH_ERR_UNSUPPORTED, "network: n_outputs must be 8 (vis, depth, normal3, albedo3)");
    if (c.depth_mode > NBVH_DEPTH_WORLD_T) fail(NBVH_ERR_UNSUPPORTED, "network: depth_mode must be LocalT or WorldT");
    if (c.depth_linear_activation > 1 || c.interpolate_depth > 1)
        fail(NBVH_ERR_INVALID_ARGUMENT, "network: depth_linear_activation and interpolate_depth are 0 or 1");
    if (!(c.depth_interpolation_eps >= 0.f) || !std::isfinite(c.depth_interpolation_eps))
        fail(NBVH_ERR_INVALID_ARGUMENT, "network: depth_interpolation_eps must be finite and >= 0");
}

std::vector<GridLevel> grid_levels(const nbvh_grid_config& g)
{
    std::vector<GridLevel> lv(g.n_levels);
    const float log2s = std::log2(g.per_level_scale);
    const uint32_t T  = 1u << g.log2_hashmap_size;
    uint32_t offset   = 0;
    for (uint32_t l = 0; l < g.n_levels; l++) {
        const float scale  = std::exp2((float)l * log2s) * (float)g.base_resolution - 1.0f;
        const uint32_t res = (uint32_t)std::ceil(scale) + 1u;
        const uint32_t maxp = 0xFFFFFFFFu / 2u;
        uint32_t params     = std::pow((float)res, 3.f) > (float)maxp ? maxp : res * res * res;
        params              = (params + 7u) / 8u * 8u;
        params              = params < T ? params : T;
        uint64_t stride     = 1;
        for (int d = 0; d < 3 && stride <= params; d++) stride *= res;
        const uint32_t dense = (uint32_t)!(params < stride);
        if (!dense && (params & (params - 1)) != 0) fail(NBVH_ERR_UNSUPPORTED, "grid: hashed level size not a power of two");
        lv[l] = GridLevel{scale, res, res * res, params, offset, dense};
        offset += params;
    }
    return lv;
}

uint64_t table_param_count(const nbvh_network_config& cfg)
{
    uint64_t entries = 0;
    for (const auto& L : grid_levels(cfg.grid)) entries += L.size;
    return entries * cfg.grid.n_features_per_level;
}

uint64_t mlp_param_count(const nbvh_network_config& cfg)
{
    const uint64_t K0 = (uint64_t)cfg.n_points * cfg.grid.n_levels * cfg.grid.n_features_per_level;
    const uint64_t H  = cfg.n_neurons;
    return H * K0 + (uint64_t)(cfg.n_hidden_layers - 1) * H * H + (uint64_t)cfg.n_outputs * H;
}

namespace {
struct SplitMix {
    uint64_t s;
    uint64_t next()
    {
        uint64_t z = (s += 0x9E3779B97F4A7C15ull);
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        return z ^ (z >> 31);
    }
    // Uniform in [-1, 1).
    double sym() { return (double)(next() >> 11) * (2.0 / 9007199254740992.0) - 1.0; }
};
}  // namespace

// Tables ~ U(+-table_range) (tcnn's grid initialisation range is 1e-4); MLP matrices
// Kaiming-uniform U(+-sqrt(6/fan_in)).
void network_init(const nbvh_network_config& cfg, uint64_t seed, float table_range, float* tables, float* mlp)
{
    SplitMix rt{seed ^ 0x7AB1E5ull};
    const uint64_t nt = table_param_count(cfg);
    for (uint64_t i = 0; i < nt; i++) tables[i] = (float)(rt.sym() * (double)table_range);
    SplitMix rm{seed ^ 0x3317A11ull};
    const uint32_t K0 = cfg.n_points * cfg.grid.n_levels * cfg.grid.n_features_per_level;
    const uint32_t H  = cfg.n_neurons;
    uint32_t in = K0;
    float* w = mlp;
    for (uint32_t layer = 0; layer <= cfg.n_hidden_layers; layer++) {
        const uint32_t out = layer == cfg.n_hidden_layers ? cfg.n_outputs : H;
        const double bound = std::sqrt(6.0 / (double)in);
        for (uint64_t i = 0; i < (uint64_t)out * in; i++) w[i] = (float)(rm.sym() * bound);
        w += (size_t)out * in;
        in = H;
    }
}

// tcnn's flat parameter vector of the reference's model (NetworkWithInputEncoding: the
// FullyFusedMLP's matrices first, then the grid's tables; each matrix row-major [out x in]; the
// output matrix padded to a multiple of 16 rows). Public tiny-cuda-nn convention; see the header.
uint64_t tcnn_param_count(const nbvh_network_config& cfg)
{
    const uint64_t K0 = (uint64_t)cfg.n_points * cfg.grid.n_levels * cfg.grid.n_features_per_level;
    const uint64_t K0p = (K0 + 15) / 16 * 16, H = cfg.n_neurons, Op = ((uint64_t)cfg.n_outputs + 15) / 16 * 16;
    return H * K0p + (uint64_t)(cfg.n_hidden_layers - 1) * H * H + Op * H + table_param_count(cfg);
}

void tcnn_params_split(const nbvh_network_config& cfg, const float* params, float* tables, float* mlp)
{
    const uint32_t K0 = cfg.n_points * cfg.grid.n_levels * cfg.grid.n_features_per_level;
    const uint32_t K0p = (K0 + 15) / 16 * 16, H = cfg.n_neurons, Op = (cfg.n_outputs + 15) / 16 * 16;
    const float* src = params;
    float* dst = mlp;
    for (uint32_t layer = 0; layer <= cfg.n_hidden_layers; layer++) {
        const uint32_t in = layer == 0 ? K0 : H, in_p = layer == 0 ? K0p : H;
        const uint32_t out = layer == cfg.n_hidden_layers ? cfg.n_outputs : H;
        const uint32_t out_p = layer == cfg.n_hidden_layers ? Op : H;
        for (uint32_t r = 0; r < out; r++)
            for (uint32_t c = 0; c < in; c++) *dst++ = src[(size_t)r * in_p + c];
        src += (size_t)out_p * in_p;
    }
    const uint64_t nt = table_param_count(cfg);
    for (uint64_t i = 0; i < nt; i++) tables[i] = src[i];
}

void tcnn_params_join(const nbvh_network_config& cfg, const float* tables, const float* mlp, float* params)
{
    const uint32_t K0 = cfg.n_points * cfg.grid.n_levels * cfg.grid.n_features_per_level;
    const uint32_t K0p = (K0 + 15) / 16 * 16, H = cfg.n_neurons, Op = (cfg.n_outputs + 15) / 16 * 16;
    const float* src = mlp;
    float* dst = params;
    for (uint32_t layer = 0; layer <= cfg.n_hidden_layers; layer++) {
        const uint32_t in = layer == 0 ? K0 : H, in_p = layer == 0 ? K0p : H;
        const uint32_t out = layer == cfg.n_hidden_layers ? cfg.n_outputs : H;
        const uint32_t out_p = layer == cfg.n_hidden_layers ? Op : H;
        for (uint32_t r = 0; r < out_p; r++)
            for (uint32_t c = 0; c < in_p; c++) dst[(size_t)r * in_p + c] = (r < out && c < in) ? *src++ : 0.f;
        dst += (size_t)out_p * in_p;
    }
    const uint64_t nt = table_param_count(cfg);
    for (uint64_t i = 0; i < nt; i++) dst[i] = tables[i];
}

// ----- number formats -----
uint16_t float_to_half(float f)
{
    uint32_t x;
    std::memcpy(&x, &f, 4);
    const uint32_t sign = (x >> 16) & 0x8000u;
    const uint32_t absx = x & 0x7FFFFFFFu;
    if (absx >= 0x7F800000u) return (uint16_t)(sign | (absx > 0x7F800000u ? 0x7E00u : 0x7C00u));
    if (absx >= 0x477FF000u) return (uint16_t)(sign | 0x7C00u);  // rounds to >= 65520 -> inf
    if (absx < 0x38800000u) {  // below 2^-14: subnormal half (or zero)
        float a;
        std::memcpy(&a, &absx, 4);
        const float scaled = a * 16777216.0f;  // a / 2^-24: exact
        const uint32_t m   = (uint32_t)std::nearbyint(scaled);
        return (uint16_t)(sign | m);
    }
    uint32_t mant = absx & 0x7FFFFFu;
    uint32_t e    = (absx >> 23) - 127 + 15;
    uint32_t h    = (e << 10) | (mant >> 13);
    const uint32_t rest = mant & 0x1FFFu;
    if (rest > 0x1000u || (rest == 0x1000u && (h & 1u))) h++;
    return (uint16_t)(sign | h);
}

float half_to_float(uint16_t h)
{
    const uint32_t sign = (uint32_t)(h & 0x8000u) << 16;
    const uint32_t e    = (h >> 10) & 0x1Fu;
    const uint32_t m    = h & 0x3FFu;
    float out;
    if (e == 0) {
        out = std::ldexp((float)m, -24);
        if (sign) out = -out;
        return out;
    }
    uint32_t x;
    if (e == 31)
        x = sign | 0x7F800000u | (m << 13);
    else
        x = sign | ((e - 15 + 127) << 23) | (m << 13);
    std::memcpy(&out, &x, 4);
    return out;
}

uint16_t float_to_bf16(float f)
{
    uint32_t x;
    std::memcpy(&x, &f, 4);
    if ((x & 0x7FFFFFFFu) > 0x7F800000u) return (uint16_t)((x >> 16) | 0x40u);
    const uint32_t lsb = (x >> 16) & 1u;
    x += 0x7FFFu + lsb;
    return (uint16_t)(x >> 16);
}

// OCP FP8 E4M3 (bias 7, max 448, no infinities, S.1111.111 = NaN); saturating.
uint8_t float_to_e4m3(float f)
{
    if (f != f) return 0x7F;
    const uint8_t sign = std::signbit(f) ? 0x80 : 0x00;
    float a = std::fabs(f);
    if (a >= 464.f) return sign | 0x7E;  // saturate to 448
    // Enumerate: e4m3 values are k * 2^(e-10) for normal, m * 2^-9 subnormal.
    if (a < std::ldexp(1.f, -6)) {       // subnormal range: step 2^-9
        const uint32_t m = (uint32_t)std::nearbyint(a * 512.f);
        if (m >= 8) return sign | 0x08;  // rounds up to the smallest normal
        return sign | (uint8_t)m;
    }
    int e;
    const float fr = std::frexp(a, &e);   // a = fr * 2^e, fr in [0.5, 1)
    int exp_field  = e - 1 + 7;           // a = 1.m * 2^(e-1)
    float mant     = (fr * 2.f - 1.f) * 8.f;
    uint32_t m     = (uint32_t)std::nearbyint(mant);
    if (m == 8) {
        m = 0;
        exp_field++;
    }
    if (exp_field > 15 || (exp_field == 15 && m == 7)) return sign | 0x7E;
    return sign | (uint8_t)((exp_field << 3) | m);
}

float e4m3_to_float(uint8_t v)
{
    const float s = (v & 0x80) ? -1.f : 1.f;
    const int e   = (v >> 3) & 0xF;
    const int m   = v & 0x7;
    if (e == 15 && m == 7) return NAN;
    if (e == 0) return s * std::ldexp((float)m, -9);
    return s * std::ldexp(1.f + (float)m / 8.f, e - 7);
}

}  // namespace nbvh

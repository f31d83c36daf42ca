// host_common.h -- host-side helpers of libnbvh_hip (C++).
#pragma once

#include <cstdint>
#include <stdexcept>
#include <string>
#include <vector>

#include "../nbvh_internal.h"

namespace nbvh {

// Error carrying an nbvh_status across the C++ layers; converted to a status code at the ABI.
struct Error : std::runtime_error {
    int32_t status;
    Error(int32_t s, const std::string& msg) : std::runtime_error(msg), status(s) {}
};

[[noreturn]] inline void fail(int32_t status, const std::string& msg) { throw Error(status, msg); }

void geosphere(uint32_t frequency, uint64_t seed, float amplitude, const float center[3], float radius,
               std::vector<float>& verts, std::vector<uint32_t>& tris);
void icosphere(uint32_t subdivisions, uint64_t seed, float amplitude, const float center[3], float radius,
               std::vector<float>& verts, std::vector<uint32_t>& tris);
void camera_look_at(const float eye[3], const float target[3], const float up[3], float fovy_deg, uint32_t w,
                    uint32_t h, nbvh_camera& cam);

// Binned-SAH BVH2; nodes[0] is the root, children of an interior node are adjacent.
void bvh_build(const float* verts, uint32_t n_verts, const uint32_t* tris, uint32_t n_tris, uint32_t max_leaf,
               std::vector<nbvh_node>& nodes, std::vector<uint32_t>& prims);

// N-BVH cut to at most n_leaves leaves (largest surface area first).
void nbvh_cut(const nbvh_node* bvh, uint32_t n_bvh, uint32_t n_leaves, std::vector<nbvh_node>& out);

// .nbvh (bvh-v2 serialisation) and structural validation.
void nbvh_parse(const void* bytes, size_t size, std::vector<nbvh_node>& out);
std::vector<uint8_t> nbvh_serialize(const nbvh_node* nodes, uint32_t n);
void nbvh_validate(const nbvh_node* nodes, uint32_t n);

// Dilated copy used by the kernels (neural_bvh2.cuh:73-78 evaluated once per node).
std::vector<nbvh_node> nbvh_dilate(const std::vector<nbvh_node>& nodes, float dilation);

// Network configuration helpers.
void network_validate(const nbvh_network_config& cfg);
std::vector<GridLevel> grid_levels(const nbvh_grid_config& g);
uint64_t table_param_count(const nbvh_network_config& cfg);
uint64_t mlp_param_count(const nbvh_network_config& cfg);
void network_init(const nbvh_network_config& cfg, uint64_t seed, float table_range, float* tables, float* mlp);
// tcnn's flat parameter order (MLP matrices with padding, then tables) <-> tables + MLP.
uint64_t tcnn_param_count(const nbvh_network_config& cfg);
void tcnn_params_split(const nbvh_network_config& cfg, const float* params, float* tables, float* mlp);
void tcnn_params_join(const nbvh_network_config& cfg, const float* tables, const float* mlp, float* params);

// Half / bfloat16 / FP8 e4m3 conversions (round to nearest even).
uint16_t float_to_half(float f);
float half_to_float(uint16_t h);
uint16_t float_to_bf16(float f);
uint8_t float_to_e4m3(float f);
float e4m3_to_float(uint8_t v);

}  // namespace nbvh

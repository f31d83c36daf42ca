// nbvh_io.cpp -- N-BVH cut, .nbvh serialisation, validation and node dilation.
#include <algorithm>
#include <cfloat>
#include <cstring>
#include <queue>
#include <vector>

#include "host_common.h"

namespace nbvh {

// N-BVH from a BVH2 (reference: traverse_and_build_neural_bvh cuts at a fixed depth,
// src/cuda/neural/neural_bvh_renderer_accel.cpp:25-88, then splits leaves by loss). Here the cut
// is chosen by size: starting from the root, the frontier node of largest surface area that is
// interior in the BVH is replaced by its two children until the frontier holds n_leaves nodes
// (ties broken by the lower BVH index). Frontier nodes become N-BVH leaves with the reference's
// flags: BVH leaves LEAF_FINAL, cut interior nodes LEAF_LOD(0), left = NO_NEURAL_CHILD.
// Nodes are numbered breadth first with the two children of every interior node adjacent.
void nbvh_cut(const nbvh_node* bvh, uint32_t n_bvh, uint32_t n_leaves, std::vector<nbvh_node>& out)
{
    if (n_bvh == 0 || n_leaves == 0) fail(NBVH_ERR_INVALID_ARGUMENT, "nbvh_cut: empty BVH or zero leaves");
    auto area = [&](uint32_t i) {
        const nbvh_node& n = bvh[i];
        const double dx = (double)n.aabb_max[0] - n.aabb_min[0], dy = (double)n.aabb_max[1] - n.aabb_min[1],
                     dz = (double)n.aabb_max[2] - n.aabb_min[2];
        return 2.0 * (dx * dy + dy * dz + dz * dx);
    };
    struct Item {
        double a;
        uint32_t idx;
        bool operator<(const Item& o) const { return a < o.a || (a == o.a && idx > o.idx); }
    };
    std::vector<char> expanded(n_bvh, 0);
    std::priority_queue<Item> pq;
    uint32_t frontier = 1;
    if (bvh[0].prim_count == 0) pq.push({area(0), 0});
    while (frontier < n_leaves && !pq.empty()) {
        const Item it = pq.top();
        pq.pop();
        const nbvh_node& n = bvh[it.idx];
        const uint32_t l = n.left_or_prim_offset;
        if (l + 1 >= n_bvh) fail(NBVH_ERR_FORMAT, "nbvh_cut: child index out of range");
        expanded[it.idx] = 1;
        frontier += 1;
        for (uint32_t c = l; c <= l + 1; c++)
            if (bvh[c].prim_count == 0) pq.push({area(c), c});
    }
    out.clear();
    std::vector<uint32_t> src{0};  // BVH index of each N-BVH node
    out.push_back(nbvh_node{});
    for (size_t i = 0; i < src.size(); i++) {
        const nbvh_node& b = bvh[src[i]];
        nbvh_node nd{};
        std::memcpy(nd.aabb_min, b.aabb_min, sizeof(nd.aabb_min));
        std::memcpy(nd.aabb_max, b.aabb_max, sizeof(nd.aabb_max));
        if (expanded[src[i]]) {
            nd.prim_count          = 0;
            nd.left_or_prim_offset = (uint32_t)src.size();
            src.push_back(b.left_or_prim_offset);
            src.push_back(b.left_or_prim_offset + 1);
            out.push_back(nbvh_node{});
            out.push_back(nbvh_node{});
        } else {
            nd.prim_count          = b.prim_count > 0 ? NBVH_LEAF_FINAL : NBVH_LEAF_LOD(0);
            nd.left_or_prim_offset = NBVH_NO_NEURAL_CHILD;
        }
        out[i] = nd;
    }
}

// Structural checks a traversal relies on: every node that can be traversed as interior at some
// LoD (left != NO_NEURAL_CHILD and prim_count != LEAF_FINAL) has its pair of children inside
// the array and after itself (no cycles), prim_count <= LOWEST_BVH_LOD (8).
void nbvh_validate(const nbvh_node* nodes, uint32_t n)
{
    if (n == 0) fail(NBVH_ERR_FORMAT, "N-BVH has no nodes");
    for (uint32_t i = 0; i < n; i++) {
        const nbvh_node& nd = nodes[i];
        if (nd.prim_count > NBVH_LEAF_LOD(NBVH_NUM_LODS - 1))
            fail(NBVH_ERR_FORMAT, "N-BVH node " + std::to_string(i) + ": prim_count out of range");
        const bool may_be_interior = nd.left_or_prim_offset != NBVH_NO_NEURAL_CHILD && nd.prim_count != NBVH_LEAF_FINAL;
        if (nd.prim_count == 0 && nd.left_or_prim_offset == NBVH_NO_NEURAL_CHILD)
            fail(NBVH_ERR_FORMAT, "N-BVH node " + std::to_string(i) + ": interior node without children");
        if (may_be_interior) {
            const uint64_t l = nd.left_or_prim_offset;
            if (l <= i || l + 1 >= n)
                fail(NBVH_ERR_FORMAT, "N-BVH node " + std::to_string(i) + ": child index out of range");
        }
        for (int c = 0; c < 3; c++) {
            if (!(nd.aabb_min[c] <= nd.aabb_max[c]))
                fail(NBVH_ERR_FORMAT, "N-BVH node " + std::to_string(i) + ": empty or NaN box");
        }
    }
}

// bvh::v2::Bvh<Node<float,3,64,32>>::serialize layout, little endian:
//   u64 num_nodes, u64 num_prim_ids,
//   num_nodes x { f32 bounds[6] = (min_x, max_x, min_y, max_y, min_z, max_z), u64 first_id, u64 prim_count },
//   num_prim_ids x u64
// (written by save_neural_bvh, src/cuda/neural/neural_bvh_renderer_accel.cpp:170-206).
void nbvh_parse(const void* bytes, size_t size, std::vector<nbvh_node>& out)
{
    const uint8_t* p = (const uint8_t*)bytes;
    if (!p || size < 16) fail(NBVH_ERR_FORMAT, ".nbvh: truncated header");
    uint64_t n_nodes, n_prims;
    std::memcpy(&n_nodes, p, 8);
    std::memcpy(&n_prims, p + 8, 8);
    if (n_nodes == 0 || n_nodes > (1ull << 31) || n_prims > (1ull << 40))
        fail(NBVH_ERR_FORMAT, ".nbvh: implausible node or primitive count");
    const uint64_t need = 16 + n_nodes * 40 + n_prims * 8;
    if (size != need)
        fail(NBVH_ERR_FORMAT, ".nbvh: size " + std::to_string(size) + " does not match header (" +
                                  std::to_string(need) + " bytes expected)");
    out.resize(n_nodes);
    for (uint64_t i = 0; i < n_nodes; i++) {
        const uint8_t* q = p + 16 + i * 40;
        float b[6];
        uint64_t first, count;
        std::memcpy(b, q, 24);
        std::memcpy(&first, q + 24, 8);
        std::memcpy(&count, q + 32, 8);
        if (first > 0xFFFFFFFFull || count > 0xFFFFFFFFull) fail(NBVH_ERR_FORMAT, ".nbvh: index field overflow");
        nbvh_node& nd = out[i];
        nd.aabb_min[0] = b[0];
        nd.aabb_max[0] = b[1];
        nd.aabb_min[1] = b[2];
        nd.aabb_max[1] = b[3];
        nd.aabb_min[2] = b[4];
        nd.aabb_max[2] = b[5];
        nd.left_or_prim_offset = (uint32_t)first;
        nd.prim_count          = (uint32_t)count;
    }
    nbvh_validate(out.data(), (uint32_t)out.size());
}

std::vector<uint8_t> nbvh_serialize(const nbvh_node* nodes, uint32_t n)
{
    std::vector<uint8_t> buf(16 + (size_t)n * 40 + 8, 0);
    const uint64_t nn = n, np = 1;  // one dummy primitive id, as save_neural_bvh writes
    std::memcpy(buf.data(), &nn, 8);
    std::memcpy(buf.data() + 8, &np, 8);
    for (uint32_t i = 0; i < n; i++) {
        uint8_t* q = buf.data() + 16 + (size_t)i * 40;
        const float b[6] = {nodes[i].aabb_min[0], nodes[i].aabb_max[0], nodes[i].aabb_min[1],
                            nodes[i].aabb_max[1], nodes[i].aabb_min[2], nodes[i].aabb_max[2]};
        const uint64_t first = nodes[i].left_or_prim_offset, count = nodes[i].prim_count;
        std::memcpy(q, b, 24);
        std::memcpy(q + 24, &first, 8);
        std::memcpy(q + 32, &count, 8);
    }
    return buf;
}

// NeuralBVH2Node dilation about the centre (neural_bvh2.cuh:73-78), computed once per node in
// float arithmetic in the reference's operation order.
std::vector<nbvh_node> nbvh_dilate(const std::vector<nbvh_node>& nodes, float dilation)
{
    std::vector<nbvh_node> out(nodes);
    for (auto& nd : out) {
        for (int c = 0; c < 3; c++) {
            const float ext    = nd.aabb_max[c] - nd.aabb_min[c];
            const float center = nd.aabb_min[c] + ext * 0.5f;
            const float e2     = ext * (1.f + dilation);
            nd.aabb_min[c]     = center - e2 * 0.5f;
            nd.aabb_max[c]     = center + e2 * 0.5f;
        }
    }
    return out;
}

}  // namespace nbvh

// bvh_build.cpp -- binned-SAH BVH2 over triangles.
//
// Replaces the reference's use of ext/bvh (madmann91/bvh v2, DefaultBuilder Quality::High,
// max leaf size 2: src/cuda/cuda_backend.cpp:270-310, include/cuda/cuda_backend.h:232) and the
// conversion to 32-byte BVH2Node (src/base/accel/bvh.cpp:133). Own design: 32-bin SAH per axis,
// object-median split when the bins cannot separate the centroids, leaves of at most max_leaf
// triangles, children of every interior node stored as an adjacent pair, root at index 0.
// Below depth 32 the split degrades to an object median so that the depth stays below 64
// (the traversal stack of the triangle kernels).
#include <algorithm>
#include <cfloat>
#include <cmath>
#include <vector>

#include "host_common.h"

namespace nbvh {
namespace {

struct Box {
    float mn[3] = {FLT_MAX, FLT_MAX, FLT_MAX};
    float mx[3] = {-FLT_MAX, -FLT_MAX, -FLT_MAX};
    void grow(const Box& b)
    {
        for (int c = 0; c < 3; c++) {
            mn[c] = std::min(mn[c], b.mn[c]);
            mx[c] = std::max(mx[c], b.mx[c]);
        }
    }
    void grow(const float p[3])
    {
        for (int c = 0; c < 3; c++) {
            mn[c] = std::min(mn[c], p[c]);
            mx[c] = std::max(mx[c], p[c]);
        }
    }
    double area() const
    {
        if (mx[0] < mn[0]) return 0.0;
        const double dx = (double)mx[0] - mn[0], dy = (double)mx[1] - mn[1], dz = (double)mx[2] - mn[2];
        return 2.0 * (dx * dy + dy * dz + dz * dx);
    }
};

constexpr int kBins = 32;
constexpr int kMedianDepth = 32;

struct Task {
    uint32_t node, begin, end, depth;
};

}  // namespace

void bvh_build(const float* verts, uint32_t n_verts, const uint32_t* tris, uint32_t n_tris, uint32_t max_leaf,
               std::vector<nbvh_node>& nodes, std::vector<uint32_t>& prims)
{
    if (n_tris == 0) fail(NBVH_ERR_INVALID_ARGUMENT, "bvh_build: no triangles");
    if (max_leaf == 0) fail(NBVH_ERR_INVALID_ARGUMENT, "bvh_build: max_leaf_size must be >= 1");
    std::vector<Box> boxes(n_tris);
    std::vector<float> cen((size_t)n_tris * 3);
    for (uint32_t t = 0; t < n_tris; t++) {
        for (int k = 0; k < 3; k++) {
            const uint32_t vi = tris[3 * t + k];
            if (vi >= n_verts) fail(NBVH_ERR_INVALID_ARGUMENT, "bvh_build: vertex index out of range");
            boxes[t].grow(verts + 3 * (size_t)vi);
        }
        for (int c = 0; c < 3; c++) cen[3 * (size_t)t + c] = 0.5f * (boxes[t].mn[c] + boxes[t].mx[c]);
    }
    prims.resize(n_tris);
    for (uint32_t t = 0; t < n_tris; t++) prims[t] = t;
    nodes.clear();
    nodes.reserve(2 * (size_t)n_tris);
    nodes.push_back(nbvh_node{});
    std::vector<Task> work{{0u, 0u, n_tris, 0u}};
    while (!work.empty()) {
        const Task task = work.back();
        work.pop_back();
        Box bb, cb;
        for (uint32_t i = task.begin; i < task.end; i++) {
            bb.grow(boxes[prims[i]]);
            cb.grow(&cen[3 * (size_t)prims[i]]);
        }
        nbvh_node& nd = nodes[task.node];
        for (int c = 0; c < 3; c++) {
            nd.aabb_min[c] = bb.mn[c];
            nd.aabb_max[c] = bb.mx[c];
        }
        const uint32_t count = task.end - task.begin;
        if (count <= max_leaf) {
            nd.prim_count          = count;
            nd.left_or_prim_offset = task.begin;
            continue;
        }
        uint32_t mid = task.begin + count / 2;
        int best_axis = -1;
        int best_bin = 0;
        if (task.depth < kMedianDepth) {
            double best_cost = DBL_MAX;
            for (int a = 0; a < 3; a++) {
                const float lo = cb.mn[a], hi = cb.mx[a];
                if (!(hi > lo)) continue;
                const float k = (float)kBins / (hi - lo);
                Box bin[kBins];
                uint32_t cnt[kBins] = {};
                for (uint32_t i = task.begin; i < task.end; i++) {
                    const uint32_t p = prims[i];
                    int b = (int)((cen[3 * (size_t)p + a] - lo) * k);
                    b = std::min(std::max(b, 0), kBins - 1);
                    bin[b].grow(boxes[p]);
                    cnt[b]++;
                }
                double right_area[kBins];
                uint32_t right_cnt[kBins];
                Box acc;
                uint32_t n = 0;
                for (int b = kBins - 1; b > 0; b--) {
                    acc.grow(bin[b]);
                    n += cnt[b];
                    right_area[b] = acc.area();
                    right_cnt[b]  = n;
                }
                Box lacc;
                uint32_t ln = 0;
                for (int b = 0; b < kBins - 1; b++) {
                    lacc.grow(bin[b]);
                    ln += cnt[b];
                    if (ln == 0 || right_cnt[b + 1] == 0) continue;
                    const double cost = lacc.area() * ln + right_area[b + 1] * right_cnt[b + 1];
                    if (cost < best_cost) {
                        best_cost = cost;
                        best_axis = a;
                        best_bin  = b + 1;
                    }
                }
            }
        }
        if (best_axis >= 0) {
            const int a = best_axis;
            const float lo = cb.mn[a];
            const float k = (float)kBins / (cb.mx[a] - lo);
            auto it = std::partition(prims.begin() + task.begin, prims.begin() + task.end, [&](uint32_t p) {
                int b = (int)((cen[3 * (size_t)p + a] - lo) * k);
                b = std::min(std::max(b, 0), kBins - 1);
                return b < best_bin;
            });
            mid = (uint32_t)(it - prims.begin());
            if (mid == task.begin || mid == task.end) best_axis = -1;
        }
        if (best_axis < 0) {
            // Object median along the widest centroid axis (ties broken by triangle id).
            int a = 0;
            for (int c = 1; c < 3; c++)
                if (cb.mx[c] - cb.mn[c] > cb.mx[a] - cb.mn[a]) a = c;
            mid = task.begin + count / 2;
            std::nth_element(prims.begin() + task.begin, prims.begin() + mid, prims.begin() + task.end,
                             [&](uint32_t x, uint32_t y) {
                                 const float cx = cen[3 * (size_t)x + a], cy = cen[3 * (size_t)y + a];
                                 return cx < cy || (cx == cy && x < y);
                             });
        }
        const uint32_t left = (uint32_t)nodes.size();
        nodes.push_back(nbvh_node{});
        nodes.push_back(nbvh_node{});
        nodes[task.node].prim_count          = 0;
        nodes[task.node].left_or_prim_offset = left;
        work.push_back({left + 1, mid, task.end, task.depth + 1});
        work.push_back({left, task.begin, mid, task.depth + 1});
    }
}

}  // namespace nbvh

// common.h -- device functions shared by the query kernels (HIP, gfx950).
//
// Every function states the reference code it follows. The HIP sources are compiled with
// -ffp-contract=off: multiply-adds are fused only where fmaf() is written, so the FP32 path
// evaluates exactly the operation sequence of the CPU restatement in oracle/.
#pragma once

#include <hip/hip_runtime.h>
#include <float.h>
#include <stdint.h>

#include "../nbvh_internal.h"

namespace nbvh {
namespace dev {

// glm::min / glm::max (NaN behaviour included).
__device__ __forceinline__ float gmin(float x, float y) { return (y < x) ? y : x; }
__device__ __forceinline__ float gmax(float x, float y) { return (x < y) ? y : x; }

struct Ray {
    float ox, oy, oz, tmin, dx, dy, dz, tmax;
};

__device__ __forceinline__ Ray load_ray(const nbvh_ray* rays, uint32_t i)
{
    const float4* p = reinterpret_cast<const float4*>(rays + i);
    const float4 a = p[0], b = p[1];
    return Ray{a.x, a.y, a.z, a.w, b.x, b.y, b.z, b.w};
}

struct Box {
    float mnx, mny, mnz, mxx, mxy, mxz;
    uint32_t prim_count, left;
};

__device__ __forceinline__ Box load_node(const nbvh_node* nodes, uint32_t i)
{
    const float4* p = reinterpret_cast<const float4*>(nodes + i);
    const float4 a = p[0], b = p[1];
    return Box{a.x, a.y, a.z, b.x, b.y, b.z, __float_as_uint(a.w), __float_as_uint(b.w)};
}

// N-BVH node sources. GlobalNodes reads HBM/L2; CachedNodes serves the first n_cached nodes
// from an LDS copy (the top of a breadth-first numbered tree: every traversal starts there).
struct GlobalNodes {
    const nbvh_node* nodes;
    __device__ __forceinline__ Box operator()(uint32_t i) const { return load_node(nodes, i); }
};

struct CachedNodes {
    const nbvh_node* nodes;
    const uint4* lds;  // 2 x 16 B per node
    uint32_t n_cached;
    __device__ __forceinline__ Box operator()(uint32_t i) const
    {
        if (i < n_cached) {
            const uint4 a = lds[2 * i], b = lds[2 * i + 1];
            return Box{__uint_as_float(a.x), __uint_as_float(a.y), __uint_as_float(a.z), __uint_as_float(b.x),
                       __uint_as_float(b.y), __uint_as_float(b.z), a.w, b.w};
        }
        return load_node(nodes, i);
    }
};

// Slab test, Majercik et al. 2018 (include/cuda/base/ray.cuh:116-135); tn/tf unclamped.
__device__ __forceinline__ void slab(const Box& b, const Ray& r, float& tn, float& tf)
{
    const float t0x = (b.mnx - r.ox) / r.dx, t1x = (b.mxx - r.ox) / r.dx;
    const float t0y = (b.mny - r.oy) / r.dy, t1y = (b.mxy - r.oy) / r.dy;
    const float t0z = (b.mnz - r.oz) / r.dz, t1z = (b.mxz - r.oz) / r.dz;
    const float sx = gmin(t0x, t1x), sy = gmin(t0y, t1y), sz = gmin(t0z, t1z);
    const float bx = gmax(t0x, t1x), by = gmax(t0y, t1y), bz = gmax(t0z, t1z);
    tn = gmax(gmax(sx, sy), sz);
    tf = gmin(gmin(bx, by), bz);
}

// Same slab test with the reciprocal direction replaced by the division (kept as division:
// the reference divides, and the FP32 path is bit-exact with it).

// NeuralBVH2::intersect (include/cuda/neural/accel/neural_bvh2.cuh:170-292), by default with the
// inference epsilons (the sampler of the training data passes the learning ones), on the boxes
// the node source returns (the query's are dilated by 1 + fixed_node_dilation about their centre
// once, on the host). Stack accessor S provides push/pop on whatever storage the kernel uses.
// Returns 1 = leaf segment found, 0 = stack exhausted, -1 = stack overflow, 2 = suspended after
// `budget` node visits with the next node pushed back (the next call resumes exactly there, so
// suspension never changes which segments are found or their order).
template <typename Nodes, typename Stack>
__device__ __forceinline__ int nbvh_next_leaf(const Nodes& nodes, uint32_t lod, const Ray& ray, Stack& st,
                                              float current_maxt, float& seg_mint, float& seg_maxt,
                                              uint32_t& seg_node, uint32_t budget = 0xFFFFFFFFu,
                                              float mint_eps = kInferenceMintEps, float maxt_eps = kInferenceMaxtEps)
{
    uint32_t idx = st.pop();
    while (true) {
        if (budget-- == 0u) {
            st.push(idx);  // at most the size before the pop that produced idx: never overflows
            return 2;
        }
        const Box node = nodes(idx);
        const bool leaf = (node.prim_count > 0u && node.prim_count <= 2u + lod) || node.left == NBVH_NO_NEURAL_CHILD;
        if (leaf) {
            float mint, maxt;
            slab(node, ray, mint, maxt);
            if (mint <= maxt) {
                bool valid = mint < current_maxt;
                if (mint < 0.f)
                    valid = valid && (maxt > maxt_eps);
                else
                    valid = valid && (mint > mint_eps);
                if (valid) {
                    seg_mint = mint;
                    seg_maxt = maxt;
                    seg_node = idx;
                    return 1;
                }
            }
            if (st.empty()) return 0;
            idx = st.pop();
            continue;
        }
        uint32_t c1 = node.left, c2 = node.left + 1u;
        float m1, m2;
        {
            float tn, tf;
            slab(nodes(c1), ray, tn, tf);
            float mint = gmax(-FLT_MAX, tn), maxt = gmin(current_maxt, tf);
            m1 = (mint <= maxt) ? mint : FLT_MAX;
            slab(nodes(c2), ray, tn, tf);
            mint = gmax(-FLT_MAX, tn);
            maxt = gmin(current_maxt, tf);
            m2   = (mint <= maxt) ? mint : FLT_MAX;
        }
        if (m1 > m2) {
            const float tm = m1;
            m1 = m2;
            m2 = tm;
            const uint32_t tc = c1;
            c1 = c2;
            c2 = tc;
        }
        if (m1 == FLT_MAX) {
            if (st.empty()) return 0;
            idx = st.pop();
        } else {
            idx = c1;
            if (m2 != FLT_MAX) {
                st.push(c2);
                if (st.size() >= kNeuralStackSize) return -1;
            }
        }
    }
}

// exp in float: Cody-Waite reduction by ln2 (hi/lo) and a degree-7 Taylor polynomial in
// Horner form (identical sequence in oracle/nbvh_oracle.cpp expf_cw).
__device__ __forceinline__ float expf_cw(float x)
{
    if (x != x) return x;
    if (x > 88.72283935546875f) return __builtin_huge_valf();
    if (x < -103.972076416015625f) return 0.f;
    const float n = rintf(x * 1.44269502162933349609375f);
    float r = fmaf(-n, 0.693145751953125f, x);
    r       = fmaf(-n, 1.428606765330187045037746429443359375e-06f, r);
    float p = 1.98412701138295233249664306640625e-04f;
    p = fmaf(p, r, 1.388888922519981861114501953125e-03f);
    p = fmaf(p, r, 8.3333335816860198974609375e-03f);
    p = fmaf(p, r, 4.1666667163372039794921875e-02f);
    p = fmaf(p, r, 1.666666716337203979492187500e-01f);
    p = fmaf(p, r, 0.5f);
    p = fmaf(p, r, 1.0f);
    p = fmaf(p, r, 1.0f);
    return ldexpf(p, (int)n);
}
__device__ __forceinline__ float sigmoidf_cw(float z) { return 1.f / (1.f + expf_cw(-z)); }

// sin and cos in float, written out so that the device and oracle/nbvh_oracle.cpp (sincosf_cw)
// run the same operation sequence: Cody-Waite reduction by pi/2 with a three-part constant
// (1.5703125 and a 12-bit second part, so q * part is exact for the small quadrant counts of
// the sampler's angles in [0, 2 pi)), then the Taylor polynomials of sin (to r^9) and cos (to
// r^10) on [-pi/4, pi/4] in Horner form, and the quadrant swap.
__device__ __forceinline__ void sincosf_cw(float x, float& s, float& c)
{
    const float q = rintf(x * 0.636619772367581343f);
    float r = fmaf(-q, 1.5703125f, x);
    r       = fmaf(-q, 4.838705062866211e-4f, r);
    r       = fmaf(-q, -4.371138828673793e-8f, r);
    const float r2 = r * r;
    float sp = fmaf(r2, 2.7557319223985893e-6f, -1.9841269841269841e-4f);
    sp = fmaf(sp, r2, 8.3333333333333332e-3f);
    sp = fmaf(sp, r2, -1.6666666666666666e-1f);
    const float sr = fmaf(sp, r2 * r, r);
    float cp = fmaf(r2, -2.7557319223985890e-7f, 2.4801587301587302e-5f);
    cp = fmaf(cp, r2, -1.3888888888888889e-3f);
    cp = fmaf(cp, r2, 4.1666666666666664e-2f);
    cp = fmaf(cp, r2, -0.5f);
    const float cr = fmaf(cp, r2, 1.0f);
    const int quadrant = (int)q & 3;
    s = quadrant == 0 ? sr : (quadrant == 1 ? cr : (quadrant == 2 ? -sr : -cr));
    c = quadrant == 0 ? cr : (quadrant == 1 ? -sr : (quadrant == 2 ? -cr : sr));
}

// 30-bit Morton code of a point given in [0,1]^3 (10 bits per axis, x in the lowest bit;
// coordinates clamped, NaN -> 0).
__device__ __forceinline__ uint32_t morton30(float x, float y, float z)
{
    const float q[3] = {x, y, z};
    uint32_t m = 0;
#pragma unroll
    for (int c = 0; c < 3; c++) {
        uint32_t v = (uint32_t)(fminf(fmaxf(q[c], 0.f), 1.f) * 1023.f);
        v = (v | (v << 16)) & 0x030000FFu;
        v = (v | (v << 8)) & 0x0300F00Fu;
        v = (v | (v << 4)) & 0x030C30C3u;
        v = (v | (v << 2)) & 0x09249249u;
        m |= v << c;
    }
    return m;
}

// LCG of get_lcg_rng / lcg_randomf (include/cuda/utils/lcg_rng.cuh:12-72).
__device__ __forceinline__ uint32_t murmur_mix(uint32_t hash, uint32_t k)
{
    k *= 0xcc9e2d51u;
    k = (k << 15) | (k >> 17);
    k *= 0x1b873593u;
    hash ^= k;
    return ((hash << 13) | (hash >> 19)) * 5u + 0xe6546b64u;
}
__device__ __forceinline__ uint32_t murmur_fmix(uint32_t h)
{
    h ^= h >> 16;
    h *= 0x85ebca6bu;
    h ^= h >> 13;
    h *= 0xc2b2ae35u;
    h ^= h >> 16;
    return h;
}
// LCG state of get_lcg_rng(index = sample, frame, linear).
__device__ __forceinline__ uint32_t lcg_seed(uint32_t sample, uint32_t frame, uint32_t linear)
{
    return murmur_fmix(murmur_mix(murmur_mix(0u, linear), 512u * frame + sample));
}
__device__ __forceinline__ float lcg_float(uint32_t& state)
{
    state = state * 1664525u + 1013904223u;
    return ldexpf((float)state, -32);
}

// Segment -> P encoded points in the neural scene box
// (src/cuda/neural/neural_bvh_module_kernels.cu:10-97,266-303; EquidistantExterior/Interior, and
// Uniform with t_k = lcg_randomf of the ray's stream `rng`, advanced P times).
__device__ __forceinline__ void encode_segment(const NetArgs& a, const Ray& r, float mint, float maxt, float* pts,
                                               uint32_t* rng = nullptr)
{
    if (mint < 0.f) {
        const float t = mint;
        mint = maxt;
        maxt = t;
    }
    const float pox = r.ox + mint * r.dx, poy = r.oy + mint * r.dy, poz = r.oz + mint * r.dz;
    const float pix = r.ox + maxt * r.dx, piy = r.oy + maxt * r.dy, piz = r.oz + maxt * r.dz;
    const float eox = (pox - a.enc_center[0]) / a.enc_extent[0] + 0.5f;
    const float eoy = (poy - a.enc_center[1]) / a.enc_extent[1] + 0.5f;
    const float eoz = (poz - a.enc_center[2]) / a.enc_extent[2] + 0.5f;
    const float eix = (pix - a.enc_center[0]) / a.enc_extent[0] + 0.5f;
    const float eiy = (piy - a.enc_center[1]) / a.enc_extent[1] + 0.5f;
    const float eiz = (piz - a.enc_center[2]) / a.enc_extent[2] + 0.5f;
#pragma unroll
    for (int k = 0; k < 4; k++) {
        if (k < (int)a.n_points) {
            const float t = (rng != nullptr && a.seg_uniform) ? lcg_float(*rng) : a.seg_t[k];
            const float w = 1.f - t;
            pts[3 * k + 0] = w * eox + t * eix;
            pts[3 * k + 1] = w * eoy + t * eiy;
            pts[3 * k + 2] = w * eoz + t * eiz;
        }
    }
}

// tcnn grid_index (public): dense stride indexing or coherent prime hash, modulo level size,
// for the 8 corners of cell g (corner c adds bit d of c to coordinate d).
// Hashed levels have a power-of-two size (a mask). A dense level has size >= res^3 with
// res >= 2; for a point inside the unit cube every cell coordinate is at most res - 1, so the
// corner indices stay below 2 * size and one conditional subtraction is the modulo (EXACT =
// false, branch-free). EXACT = true takes the full modulo for points anywhere.
template <bool EXACT>
__device__ __forceinline__ void grid_corners(const GridLevel& L, const uint32_t g[3], uint32_t idx[8])
{
    const uint32_t hy0 = g[1] * 2654435761u, hy1 = (g[1] + 1u) * 2654435761u;
    const uint32_t hz0 = g[2] * 805459861u, hz1 = (g[2] + 1u) * 805459861u;
    const uint32_t base = g[0] + g[1] * L.res + g[2] * L.res2;
#pragma unroll
    for (uint32_t c = 0; c < 8; c++) {
        const uint32_t h = ((g[0] + (c & 1u)) ^ ((c & 2u) ? hy1 : hy0) ^ ((c & 4u) ? hz1 : hz0)) & (L.size - 1u);
        const uint32_t i = base + (c & 1u) + ((c >> 1) & 1u) * L.res + ((c >> 2) & 1u) * L.res2;
        const uint32_t d = EXACT ? i % L.size : (i >= L.size ? i - L.size : i);
        idx[c] = L.dense ? d : h;
    }
}

// True when every coordinate is in [0, 1] (false for NaN): the fast grid indexing is exact.
template <int N>
__device__ __forceinline__ bool in_unit_cube(const float* x)
{
    bool ok = true;
#pragma unroll
    for (int i = 0; i < N; i++) ok = ok && (x[i] >= 0.f) && (x[i] <= 1.f);
    return ok;
}

// Cell and fractional position of a point in a level: pos = fmaf(scale, x, 0.5).
__device__ __forceinline__ void grid_cell(const GridLevel& L, float x, float y, float z, uint32_t g[3], float f[3])
{
    const float px = fmaf(L.scale, x, 0.5f), py = fmaf(L.scale, y, 0.5f), pz = fmaf(L.scale, z, 0.5f);
    const float fx = floorf(px), fy = floorf(py), fz = floorf(pz);
    g[0] = (uint32_t)(int)fx;
    g[1] = (uint32_t)(int)fy;
    g[2] = (uint32_t)(int)fz;
    f[0] = px - fx;
    f[1] = py - fy;
    f[2] = pz - fz;
}

// Corner weight in tcnn's order: ((1 * w_x) * w_y) * w_z.
__device__ __forceinline__ float corner_weight(uint32_t c, const float f[3])
{
    float w = (c & 1u) ? f[0] : 1.f - f[0];
    w *= (c & 2u) ? f[1] : 1.f - f[1];
    w *= (c & 4u) ? f[2] : 1.f - f[2];
    return w;
}

__device__ __forceinline__ float half_bits_to_float(uint32_t bits16)
{
    return (float)__builtin_bit_cast(_Float16, (uint16_t)bits16);
}

// Trilinear interpolation of the F features of one level at point (x, y, z):
// acc[f] = sum over corners c = 0..7 of fmaf(w_c, v_c[f], .), tcnn's corner order.
// Tables: FP32 (T = float) or IEEE half bits (T = uint16_t), entries of F consecutive values.
template <int F, typename T, bool EXACT>
__device__ __forceinline__ void interp_level(const GridLevel& L, const T* __restrict__ table, float x, float y, float z,
                                             float acc[F])
{
    uint32_t g[3];
    float f[3];
    grid_cell(L, x, y, z, g, f);
    uint32_t idx[8];
    grid_corners<EXACT>(L, g, idx);
    const T* base = table + (size_t)L.offset * F;
#pragma unroll
    for (int k = 0; k < F; k++) acc[k] = 0.f;
#pragma unroll
    for (uint32_t c = 0; c < 8; c++) {
        const float w = corner_weight(c, f);
        const T* v = base + (size_t)idx[c] * F;
        float val[F];
        if constexpr (sizeof(T) == 4) {
            if constexpr (F == 2) {
                const float2 q = *reinterpret_cast<const float2*>(v);
                val[0] = q.x;
                val[1] = q.y;
            } else if constexpr (F == 4) {
                const float4 q = *reinterpret_cast<const float4*>(v);
                val[0] = q.x;
                val[1] = q.y;
                val[2] = q.z;
                val[3] = q.w;
            } else {
#pragma unroll
                for (int k = 0; k < F; k++) val[k] = (float)v[k];
            }
        } else {
            if constexpr (F == 2) {
                const uint32_t q = *reinterpret_cast<const uint32_t*>(v);
                val[0] = half_bits_to_float(q & 0xFFFFu);
                val[1] = half_bits_to_float(q >> 16);
            } else if constexpr (F == 4) {
                const uint2 q = *reinterpret_cast<const uint2*>(v);
                val[0] = half_bits_to_float(q.x & 0xFFFFu);
                val[1] = half_bits_to_float(q.x >> 16);
                val[2] = half_bits_to_float(q.y & 0xFFFFu);
                val[3] = half_bits_to_float(q.y >> 16);
            } else {
#pragma unroll
                for (int k = 0; k < F; k++) val[k] = half_bits_to_float(v[k]);
            }
        }
#pragma unroll
        for (int k = 0; k < F; k++) acc[k] = fmaf(w, val[k], acc[k]);
    }
}

// Hash-grid features of P points: x[p*L*F + l*F + f].
template <int P, int L, int F, typename T, bool EXACT>
__device__ __forceinline__ void grid_features_impl(const NetArgs& a, const T* __restrict__ table, const float* pts,
                                                   float* x)
{
#pragma unroll
    for (int p = 0; p < P; p++) {
#pragma unroll
        for (int l = 0; l < L; l++) {
            float acc[F];
            interp_level<F, T, EXACT>(a.level[l], table, pts[3 * p], pts[3 * p + 1], pts[3 * p + 2], acc);
#pragma unroll
            for (int k = 0; k < F; k++) x[p * L * F + l * F + k] = acc[k];
        }
    }
}

// One wave-uniform decision per call: the branch-free indexing when every lane's points are in
// the unit cube (the normal case: segments inside the neural scene box), else the exact modulo.
template <int P, int L, int F, typename T>
__device__ __forceinline__ void grid_features(const NetArgs& a, const T* __restrict__ table, const float* pts, float* x)
{
    if (__all(in_unit_cube<3 * P>(pts)))
        grid_features_impl<P, L, F, T, false>(a, table, pts, x);
    else
        grid_features_impl<P, L, F, T, true>(a, table, pts, x);
}

// LocalT depth decode (include/cuda/neural/neural_bvh_module_common.cuh:32-57); FLT_MAX = rejected.
// The reference evaluates (segment_length + 1e-3) in double (double literal); FP32 here.
__device__ __forceinline__ float decode_depth(float u, float mint, float maxt)
{
    const float len = maxt - mint;
    if (mint < 0.f) {
        const float inward = u * (len + 1e-3f);
        if (inward < maxt - len * kDepthEpsilon - kFixedDepthEpsilon) return maxt - inward;
        return FLT_MAX;
    }
    return u * (len + 1e-3f) + mint;
}

// WorldT decode (neural_bvh_module_common.cuh:58-84): the same on the ray's span [smin, smax]
// through the neural scene's box; `inside` is the segment's negative mint.
__device__ __forceinline__ float decode_span(float u, float smin, float smax, bool inside)
{
    const float len = smax - smin;
    if (inside) {
        const float inward = u * (len + 1e-3f);
        if (inward < smax - len * kDepthEpsilon - kFixedDepthEpsilon) return smax - inward;
        return FLT_MAX;
    }
    return u * (len + 1e-3f) + smin;
}

// ---------------------------------------------------------------------------------------------
// Watertight ray/triangle (Woop, Benthin, Wald, JCGT 2(1) 2013) with FP64 recomputation of the
// edge functions when one is exactly zero; accepted interval [tmin, tmax) as
// BaseCudaBLAS::intersect_triangle (include/cuda/accel/bvh.cuh:98-136).
// ---------------------------------------------------------------------------------------------
struct TriSetup {
    int kx, ky, kz;
    float Sx, Sy, Sz;
};

__device__ __forceinline__ float comp(const Ray& r, int k, bool dir)
{
    return dir ? (k == 0 ? r.dx : (k == 1 ? r.dy : r.dz)) : (k == 0 ? r.ox : (k == 1 ? r.oy : r.oz));
}

__device__ __forceinline__ TriSetup tri_setup(const Ray& r)
{
    int kz = 0;
    if (fabsf(r.dy) > fabsf(r.dx)) kz = 1;
    if (fabsf(r.dz) > fabsf(comp(r, kz, true))) kz = 2;
    int kx = kz + 1 == 3 ? 0 : kz + 1;
    int ky = kx + 1 == 3 ? 0 : kx + 1;
    const float dkz = comp(r, kz, true);
    if (dkz < 0.f) {
        const int t = kx;
        kx = ky;
        ky = t;
    }
    TriSetup s;
    s.kx = kx;
    s.ky = ky;
    s.kz = kz;
    s.Sx = comp(r, kx, true) / dkz;
    s.Sy = comp(r, ky, true) / dkz;
    s.Sz = 1.f / dkz;
    return s;
}

__device__ __forceinline__ float pick(const float v[3], int k) { return k == 0 ? v[0] : (k == 1 ? v[1] : v[2]); }

__device__ __forceinline__ bool tri_intersect(const Ray& r, const TriSetup& s, const float* __restrict__ v0,
                                              const float* __restrict__ v1, const float* __restrict__ v2, float tmax,
                                              float& t_out, float& u_out, float& v_out)
{
    const float A[3] = {v0[0] - r.ox, v0[1] - r.oy, v0[2] - r.oz};
    const float B[3] = {v1[0] - r.ox, v1[1] - r.oy, v1[2] - r.oz};
    const float C[3] = {v2[0] - r.ox, v2[1] - r.oy, v2[2] - r.oz};
    const float Akz = pick(A, s.kz), Bkz = pick(B, s.kz), Ckz = pick(C, s.kz);
    const float Ax = pick(A, s.kx) - s.Sx * Akz, Ay = pick(A, s.ky) - s.Sy * Akz;
    const float Bx = pick(B, s.kx) - s.Sx * Bkz, By = pick(B, s.ky) - s.Sy * Bkz;
    const float Cx = pick(C, s.kx) - s.Sx * Ckz, Cy = pick(C, s.ky) - s.Sy * Ckz;
    float U = Cx * By - Cy * Bx;
    float V = Ax * Cy - Ay * Cx;
    float W = Bx * Ay - By * Ax;
    if (U == 0.f || V == 0.f || W == 0.f) {
        U = (float)((double)Cx * (double)By - (double)Cy * (double)Bx);
        V = (float)((double)Ax * (double)Cy - (double)Ay * (double)Cx);
        W = (float)((double)Bx * (double)Ay - (double)By * (double)Ax);
    }
    if ((U < 0.f || V < 0.f || W < 0.f) && (U > 0.f || V > 0.f || W > 0.f)) return false;
    const float det = U + V + W;
    if (det == 0.f) return false;
    const float Az = s.Sz * Akz, Bz = s.Sz * Bkz, Cz = s.Sz * Ckz;
    const float T   = (U * Az + V * Bz) + W * Cz;
    const float rcp = 1.f / det;
    const float t   = T * rcp;
    if (!(t >= r.tmin && t < tmax)) return false;
    t_out = t;
    u_out = V * rcp;
    v_out = W * rcp;
    return true;
}

// BVH2Node::intersect (include/base/accel/bvh.h:120-134): entry distance or FLT_MAX.
__device__ __forceinline__ float box_enter(const Box& b, const Ray& r, float tmax)
{
    float tn, tf;
    slab(b, r, tn, tf);
    const float mint = gmax(r.tmin, tn), maxt = gmin(tmax, tf);
    return (mint <= maxt) ? mint : FLT_MAX;
}

__device__ __forceinline__ void store_hit(nbvh_hit* hits, uint32_t i, const nbvh_hit& h)
{
    float4* p = reinterpret_cast<float4*>(hits + i);
    p[0] = make_float4(h.t, __uint_as_float(h.prim_id), __uint_as_float(h.node_id), __uint_as_float(h.info));
    p[1] = make_float4(h.normal_or_uv[0], h.normal_or_uv[1], h.normal_or_uv[2], h.logit);
    p[2] = make_float4(h.albedo[0], h.albedo[1], h.albedo[2], h.reserved);
}

// Part of claim index i of a multi-part launch: the last part whose begin <= i (parts sorted).
__device__ __forceinline__ uint32_t batch_part(const RayBatch& b, uint32_t i)
{
    uint32_t lo = 0, hi = b.n_parts;
    while (hi - lo > 1u) {
        const uint32_t mid = (lo + hi) >> 1;
        if (b.parts[mid].begin <= i)
            lo = mid;
        else
            hi = mid;
    }
    return lo;
}

// Claims of the launch: n, or the device-side count when the batch carries one (at most n).
__device__ __forceinline__ uint32_t batch_count(const RayBatch& b)
{
    if (b.live == nullptr) return b.n;
    const uint32_t c = *b.live;
    return c < b.n ? c : b.n;
}

// Claim index -> ray index of the launch (the coherence order, when the batch carries one).
__device__ __forceinline__ uint32_t batch_ray(const RayBatch& b, uint32_t claim)
{
    return b.order ? b.order[claim] : claim;
}

// Claim index -> (part, index within the part). PARTS: the launch has a part table (a
// compile-time switch, so single-array launches carry no part bookkeeping).
template <bool PARTS>
__device__ __forceinline__ void batch_locate(const RayBatch& b, uint32_t i, uint32_t& part, uint32_t& local)
{
    if constexpr (PARTS) {
        part  = batch_part(b, i);
        local = i - b.parts[part].begin;
    } else {
        part  = 0;
        local = i;
    }
}

template <bool PARTS>
__device__ __forceinline__ const nbvh_ray* batch_rays(const RayBatch& b, uint32_t part)
{
    if constexpr (PARTS) return b.parts[part].rays;
    return b.rays;
}

template <bool PARTS>
__device__ __forceinline__ nbvh_hit* batch_hits(const RayBatch& b, uint32_t part)
{
    if constexpr (PARTS) return b.parts[part].hits;
    return b.hits;
}

// Any-hit flag of a ray: per part in a multi-part launch, else the launch's.
template <bool PARTS>
__device__ __forceinline__ bool batch_any_hit(const RayBatch& b, uint32_t part, uint32_t launch_any_hit)
{
    if constexpr (PARTS) return b.parts[part].any_hit != 0u;
    return launch_any_hit != 0u;
}

}  // namespace dev
}  // namespace nbvh

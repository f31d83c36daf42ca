// neural_mfma.hip -- FP16 / BF16 / FP8 neural query and network inference on CDNA4 matrix cores.
//
// One wave evaluates 64 segments per iteration (the slots of the persistent wave scheduler):
//   * FP16 hash tables (scaled by a power of two), trilinear interpolation in FP32;
//   * the MLP as H^T = W * X^T with v_mfma_f32_32x32x16_{f16,bf16}: two 32-neuron row blocks x
//     two 32-segment column blocks per layer, FP32 accumulation, weights staged once per block
//     in LDS as ready-made A fragments (one ds_read_b128 per lane and MFMA);
//   * the accumulators of a layer become the B operand of the next after ReLU + conversion,
//     in registers (weights are pre-permuted on the host, host/mfma_layout.cpp);
//   * layer-1 B operands: each lane gathers the features of its own segment, one
//     v_permlane32_swap per register pair moves the halves into the operand layout.
#include <hip/hip_runtime.h>

#include "../device/wave_query.h"
#include "launch.h"

namespace nbvh {
namespace dev {

typedef _Float16 half8_t __attribute__((ext_vector_type(8)));
typedef __bf16 bf16x8_t __attribute__((ext_vector_type(8)));
typedef float floatx16 __attribute__((ext_vector_type(16)));

template <bool BF16>
struct Frag;
template <>
struct Frag<false> {
    typedef half8_t vec;
    typedef _Float16 elem;
    static __device__ __forceinline__ floatx16 mfma(vec a, vec b, floatx16 c)
    {
        return __builtin_amdgcn_mfma_f32_32x32x16_f16(a, b, c, 0, 0, 0);
    }
};
template <>
struct Frag<true> {
    typedef bf16x8_t vec;
    typedef __bf16 elem;
    static __device__ __forceinline__ floatx16 mfma(vec a, vec b, floatx16 c)
    {
        return __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, b, c, 0, 0, 0);
    }
};

// Number of fragments for K0 inputs, 64 hidden, 4 hidden layers, 8 outputs.
template <int K0>
struct MlpShape {
    static constexpr int KS1    = K0 / 16;        // layer-1 k-steps
    static constexpr int F_L1   = 2 * KS1;        // layer-1 fragments
    static constexpr int F_HID  = 2 * 4;          // fragments per hidden layer
    static constexpr int N_FRAG = F_L1 + 3 * F_HID + 4;
};

// v_permlane32_swap(x, y): lanes 32..63 of x <-> lanes 0..31 of y, on 16-byte fragments.
template <typename V>
__device__ __forceinline__ void swap_halves(V& x, V& y)
{
    uint4 a = __builtin_bit_cast(uint4, x), b = __builtin_bit_cast(uint4, y);
    const auto r0 = __builtin_amdgcn_permlane32_swap(a.x, b.x, false, false);
    const auto r1 = __builtin_amdgcn_permlane32_swap(a.y, b.y, false, false);
    const auto r2 = __builtin_amdgcn_permlane32_swap(a.z, b.z, false, false);
    const auto r3 = __builtin_amdgcn_permlane32_swap(a.w, b.w, false, false);
    a = make_uint4(r0[0], r1[0], r2[0], r3[0]);
    b = make_uint4(r0[1], r1[1], r2[1], r3[1]);
    x = __builtin_bit_cast(V, a);
    y = __builtin_bit_cast(V, b);
}

// Cache policy (buffer-load aux bits: 1 sc0, 2 nt) of the corner gathers of levels
// >= NBVH_GATHER_POLICY_FROM; the coarser levels use the default policy.
#ifndef NBVH_GATHER_POLICY_FROM
#define NBVH_GATHER_POLICY_FROM 64
#endif
#ifndef NBVH_GATHER_POLICY
#define NBVH_GATHER_POLICY 0
#endif

template <int F, int AUX>
__device__ __forceinline__ void load_corner(__amdgpu_buffer_rsrc_t tab, uint32_t off, uint32_t q[F / 2])
{
    if constexpr (F == 2) {
        q[0] = __builtin_amdgcn_raw_buffer_load_b32(tab, off, 0, AUX);
    } else {
        const auto v = __builtin_amdgcn_raw_buffer_load_b64(tab, off, 0, AUX);
        q[0] = v[0];
        q[1] = v[1];
    }
}

// Levels whose gathers are issued together (G x 8 loads in flight per lane before the first use).
#ifndef NBVH_GATHER_GROUP
#define NBVH_GATHER_GROUP 4
#endif

// Layer-1 B operands. Every lane computes all K0 features of its own segment (wave-uniform
// gather code); lo[s] holds features 8s..8s+7, hi[s] features K0/2+8s..K0/2+8s+7. Column block 0
// (segments 0..31) needs lo from lanes 0..31 and hi of the same segments in lanes 32..63;
// column block 1 (segments 32..63) needs lo of those segments in lanes 0..31 and hi in lanes
// 32..63. Swapping lanes 32..63 of lo with lanes 0..31 of hi (layer1_operands) produces exactly
// these operands, matching the weight order K(s, g, e) = g*K0/2 + 8s + e of host/mfma_layout.cpp.
//
// Features of all P points of this lane's segment, converted to 16 bits as each level completes:
// feature k = p*L*F + l*F + f goes to lo[k/8][k%8] (k < K0/2) or hi[(k-K0/2)/8][..]. The FP16
// tables are read through a buffer resource: 32-bit byte offsets (one VGPR per gather), and the
// hardware range check returns 0 for any offset past the table instead of faulting. Levels are
// processed in groups of G: all corner offsets of the group, then all its loads, then the
// corner weights and the FP32 accumulation over corners 0..7 (tcnn's trilinear weights).
// Byte offsets of the 8 corners of one level at point (x, y, z), and the fractional position.
// HASHED: the level is hashed (a power-of-two size), so only the coherent prime hash is formed:
// two multiplies per level, then one xor-and and one add-shift per corner. The hash of any
// position is in range, so this form is exact for points outside the unit cube too. Otherwise
// the general form (dense or hashed at run time, grid_corners) with its EXACT modulo switch.
template <int F, bool HASHED, bool EXACT>
__device__ __forceinline__ void level_offsets(const GridLevel& lv, float x, float y, float z, uint32_t off[8],
                                              float fr[3])
{
    uint32_t g[3];
    grid_cell(lv, x, y, z, g, fr);
    constexpr uint32_t kShift = F == 2 ? 2u : 3u;  // log2 of the entry bytes (F halves)
    if constexpr (HASHED) {
        const uint32_t hy0 = g[1] * 2654435761u, hy1 = hy0 + 2654435761u;
        const uint32_t hz0 = g[2] * 805459861u, hz1 = hz0 + 805459861u;
        const uint32_t yz[4] = {hy0 ^ hz0, hy1 ^ hz0, hy0 ^ hz1, hy1 ^ hz1};
        const uint32_t mask = lv.size - 1u, x1 = g[0] + 1u;
#pragma unroll
        for (int c = 0; c < 8; c++) off[c] = (((((c & 1) ? x1 : g[0]) ^ yz[c >> 1]) & mask) + lv.offset) << kShift;
    } else {
        uint32_t idx[8];
        grid_corners<EXACT>(lv, g, idx);
#pragma unroll
        for (int c = 0; c < 8; c++) off[c] = (idx[c] + lv.offset) << kShift;
    }
}

// One group of G levels of point p: all 8G corner offsets, then all 8G loads, then the corner
// weights and the FP32 accumulation over corners 0..7 (tcnn's trilinear weights); the features
// go to the layer-1 operand halves lo / hi as 16-bit values.
template <int P, int L, int F, int G, bool BF16, bool HASHED, bool EXACT>
__device__ __forceinline__ void gather_group(const NetArgs& a, __amdgpu_buffer_rsrc_t tab, const float* pt, int p,
                                             int l0, typename Frag<BF16>::vec* lo, typename Frag<BF16>::vec* hi)
{
    constexpr int HALF = P * L * F / 2;
    typedef typename Frag<BF16>::elem E;
    uint32_t off[G][8];
    float fr[G][3];
#pragma unroll
    for (int i = 0; i < G; i++) level_offsets<F, HASHED, EXACT>(a.level[l0 + i], pt[0], pt[1], pt[2], off[i], fr[i]);
    // One load per corner. (Measured: serving x-neighbour corners with one 16-byte load plus a
    // range-masked second load halves the throughput, so corners stay separate.)
    uint32_t q[G][8][F / 2];
#pragma unroll
    for (int i = 0; i < G; i++)
#pragma unroll
        for (int c = 0; c < 8; c++) {
            if (l0 + i >= NBVH_GATHER_POLICY_FROM)
                load_corner<F, NBVH_GATHER_POLICY>(tab, off[i][c], q[i][c]);
            else
                load_corner<F, 0>(tab, off[i][c], q[i][c]);
        }
#pragma unroll
    for (int i = 0; i < G; i++) {
        // tcnn's corner weights ((1 * w_x) * w_y) * w_z, the x-y products shared by the z pair.
        const float wx[2] = {1.f - fr[i][0], fr[i][0]}, wy[2] = {1.f - fr[i][1], fr[i][1]};
        const float wz[2] = {1.f - fr[i][2], fr[i][2]};
        float acc[F];
#pragma unroll
        for (int k = 0; k < F; k++) acc[k] = 0.f;
#pragma unroll
        for (int c = 0; c < 8; c++) {
            const float w = (wx[c & 1] * wy[(c >> 1) & 1]) * wz[c >> 2];
#pragma unroll
            for (int k = 0; k < F; k++) {
                const uint32_t bits = (k & 1) ? (q[i][c][k / 2] >> 16) : (q[i][c][k / 2] & 0xFFFFu);
                acc[k] = fmaf(w, half_bits_to_float(bits), acc[k]);
            }
        }
#pragma unroll
        for (int f = 0; f < F; f++) {
            const int k = p * L * F + (l0 + i) * F + f;
            if (k < HALF)
                lo[k / 8][k % 8] = (E)acc[f];
            else
                hi[(k - HALF) / 8][(k - HALF) % 8] = (E)acc[f];
        }
    }
}

// ND >= 0: the kernel is specialised for grids whose levels [ND, L) are hashed (n_dense == ND,
// checked at launch); groups of those levels take the hash-only form, the others the general
// form. ND < 0: every level takes the general form (any grid). EXACT: the full modulo for
// points outside the unit cube (dense levels only; the hash is exact everywhere).
template <int P, int L, int F, bool BF16, bool EXACT, int ND>
__device__ __forceinline__ void gather_operands(const NetArgs& a, __amdgpu_buffer_rsrc_t tab, const float* pts,
                                                typename Frag<BF16>::vec* lo, typename Frag<BF16>::vec* hi)
{
    constexpr int G = (L % NBVH_GATHER_GROUP == 0) ? NBVH_GATHER_GROUP : 1;
    static_assert(F == 2 || F == 4, "FP16 gather implemented for F = 2 and F = 4");
#pragma unroll
    for (int p = 0; p < P; p++) {
#pragma unroll
        for (int l0 = 0; l0 < L; l0 += G) {
            if (ND >= 0 && l0 >= ND)  // folded at compile time after unrolling
                gather_group<P, L, F, G, BF16, true, false>(a, tab, pts + 3 * p, p, l0, lo, hi);
            else
                gather_group<P, L, F, G, BF16, false, EXACT>(a, tab, pts + 3 * p, p, l0, lo, hi);
        }
    }
}

template <int P, int L, int F, bool BF16, int ND>
__device__ __forceinline__ void layer1_operands(const NetArgs& a, const float* pts,
                                                typename Frag<BF16>::vec lo[P * L * F / 16],
                                                typename Frag<BF16>::vec hi[P * L * F / 16])
{
    constexpr int K0 = P * L * F;
    const __amdgpu_buffer_rsrc_t tab =
        __builtin_amdgcn_make_buffer_rsrc((void*)a.tables_f16, (short)0, (int)a.tables_f16_bytes, 0x00020000);
    if (__all(in_unit_cube<3 * P>(pts)))
        gather_operands<P, L, F, BF16, false, ND>(a, tab, pts, lo, hi);
    else
        gather_operands<P, L, F, BF16, true, ND>(a, tab, pts, lo, hi);
#pragma unroll
    for (int s = 0; s < K0 / 16; s++) swap_halves(lo[s], hi[s]);
}

// The 8 output rows of this lane's own segment (column lane%32 of column block g), times inv.
// Rows 0..3 of a column block sit in lane group 0, rows 4..7 in lane group 1 (registers 0..3).
__device__ __forceinline__ void gather_logits(const floatx16 out[2], uint32_t g, float inv, float z[8])
{
    float send[4], recv[4];
#pragma unroll
    for (int r = 0; r < 4; r++) {
        send[r] = g ? out[0][r] : out[1][r];
        recv[r] = __shfl_xor(send[r], 32);
    }
#pragma unroll
    for (int r = 0; r < 4; r++) {
        z[r]     = (g ? recv[r] : out[0][r]) * inv;
        z[4 + r] = (g ? out[1][r] : recv[r]) * inv;
    }
}

// ReLU + conversion of the accumulators acc[rb][cb] into the B operands of the next layer.
template <bool BF16>
__device__ __forceinline__ void acc_to_b(const floatx16 acc[2][2], typename Frag<BF16>::vec bh[2][4])
{
    typedef typename Frag<BF16>::elem E;
#pragma unroll
    for (int cb = 0; cb < 2; cb++)
#pragma unroll
        for (int s = 0; s < 4; s++)
#pragma unroll
            for (int e = 0; e < 8; e++) {
                const float v = acc[s / 2][cb][8 * (s % 2) + e];
                bh[cb][s][e] = (E)(v > 0.f ? v : 0.f);
            }
}

// The whole network for the 64 segments of a wave. own[3P]: this lane's segment points.
// Returns this lane's 8 logits (descaled). wfrag: LDS fragments (16 B per lane each).
template <int P, int L, int F, bool BF16, int ND>
__device__ __forceinline__ void mfma_network(const NetArgs& a, const uint4* __restrict__ wfrag, const float* own,
                                             float z[8])
{
    constexpr int K0 = P * L * F;
    typedef MlpShape<K0> S;
    typedef Frag<BF16> FR;
    typedef typename FR::vec V;
    const uint32_t lane = threadIdx.x & 63u;
    const uint32_t g    = lane >> 5;
    V b1[2][S::KS1];
    layer1_operands<P, L, F, BF16, ND>(a, own, b1[0], b1[1]);
    floatx16 acc[2][2];
#pragma unroll
    for (int rb = 0; rb < 2; rb++)
#pragma unroll
        for (int cb = 0; cb < 2; cb++)
#pragma unroll
            for (int r = 0; r < 16; r++) acc[rb][cb][r] = 0.f;
#pragma unroll
    for (int rb = 0; rb < 2; rb++)
#pragma unroll
        for (int s = 0; s < S::KS1; s++) {
            const V w = __builtin_bit_cast(V, wfrag[(rb * S::KS1 + s) * 64 + lane]);
            acc[rb][0] = FR::mfma(w, b1[0][s], acc[rb][0]);
            acc[rb][1] = FR::mfma(w, b1[1][s], acc[rb][1]);
        }
    int fbase = S::F_L1;
#pragma unroll
    for (int layer = 0; layer < 3; layer++) {
        V bh[2][4];
        acc_to_b<BF16>(acc, bh);
#pragma unroll
        for (int rb = 0; rb < 2; rb++)
#pragma unroll
            for (int cb = 0; cb < 2; cb++)
#pragma unroll
                for (int r = 0; r < 16; r++) acc[rb][cb][r] = 0.f;
#pragma unroll
        for (int rb = 0; rb < 2; rb++)
#pragma unroll
            for (int s = 0; s < 4; s++) {
                const V w = __builtin_bit_cast(V, wfrag[(fbase + rb * 4 + s) * 64 + lane]);
                acc[rb][0] = FR::mfma(w, bh[0][s], acc[rb][0]);
                acc[rb][1] = FR::mfma(w, bh[1][s], acc[rb][1]);
            }
        fbase += S::F_HID;
    }
    floatx16 out[2];
    {
        V bh[2][4];
        acc_to_b<BF16>(acc, bh);
#pragma unroll
        for (int cb = 0; cb < 2; cb++)
#pragma unroll
            for (int r = 0; r < 16; r++) out[cb][r] = 0.f;
#pragma unroll
        for (int s = 0; s < 4; s++) {
            const V w = __builtin_bit_cast(V, wfrag[(fbase + s) * 64 + lane]);
            out[0] = FR::mfma(w, bh[0][s], out[0]);
            out[1] = FR::mfma(w, bh[1][s], out[1]);
        }
    }
    gather_logits(out, g, a.inv_table_scale, z);
}

template <int P, int L, int F, bool BF16>
__device__ __forceinline__ void stage_weights(const NetArgs& a, uint4* wfrag)
{
    constexpr int NF = MlpShape<P * L * F>::N_FRAG;
    const uint4* src = reinterpret_cast<const uint4*>(a.mlp_frag) + (BF16 ? (size_t)NF * 64 : 0);
    for (uint32_t i = threadIdx.x; i < (uint32_t)NF * 64; i += blockDim.x) wfrag[i] = src[i];
    __syncthreads();
}

// ---------------------------------------------------------------------------------------------
// FP8 variant: FP16 first and output layers, e4m3 hidden layers (v_mfma_f32_32x32x16_fp8_fp8).
// Weights of hidden layer l carry a per-layer power-of-two scale (host, amax -> [224, 448)).
// Activations entering an e4m3 layer are scaled per segment (per MFMA column) by a power of two
// that brings the segment's largest activation into [224, 448): the scale depends on the segment
// alone, so results do not depend on what else the wave evaluates. The bias-free ReLU MLP is
// positively homogeneous, so every scale is undone exactly on the logits.
// ---------------------------------------------------------------------------------------------
typedef long fp8x8_t;  // one K = 16 operand per lane: 8 e4m3 values, element e in byte e

struct Fp8Weights {
    static constexpr int kHidden = 3;       // hidden 64x64 layers in e4m3
    static constexpr int kFragsPerLayer = 8;
};

// 2^k with amax * 2^k in [224, 448); 1 for zero, subnormal or non-finite amax.
__device__ __forceinline__ float fp8_column_scale(float amax)
{
    const uint32_t b = __float_as_uint(amax);
    const int e = (int)(b >> 23) - 127;  // amax >= 0 (post-ReLU maximum)
    if (!(amax > 0.f) || e == -127 || e == 128) return 1.f;
    int k = ((b & 0x7FFFFFu) < 0x600000u ? 8 : 7) - e;  // mantissa < 1.75: 2^8 side, else 2^7
    k = k < -126 ? -126 : (k > 127 ? 127 : k);
    return __uint_as_float((uint32_t)(k + 127) << 23);
}

// Per column block: the largest (post-ReLU) activation of this lane's column, over both lane halves.
__device__ __forceinline__ void column_scales(const floatx16 acc[2][2], float sx[2])
{
#pragma unroll
    for (int cb = 0; cb < 2; cb++) {
        float m = 0.f;
#pragma unroll
        for (int rb = 0; rb < 2; rb++)
#pragma unroll
            for (int r = 0; r < 16; r++) m = fmaxf(m, acc[rb][cb][r]);
        m = fmaxf(m, __shfl_xor(m, 32));
        sx[cb] = fp8_column_scale(m);
    }
}

__device__ __forceinline__ fp8x8_t pack_fp8(const float v[8])
{
    int lo = 0, hi = 0;
    lo = __builtin_amdgcn_cvt_pk_fp8_f32(v[0], v[1], lo, false);
    lo = __builtin_amdgcn_cvt_pk_fp8_f32(v[2], v[3], lo, true);
    hi = __builtin_amdgcn_cvt_pk_fp8_f32(v[4], v[5], hi, false);
    hi = __builtin_amdgcn_cvt_pk_fp8_f32(v[6], v[7], hi, true);
    return (fp8x8_t)(((uint64_t)(uint32_t)hi << 32) | (uint32_t)lo);
}

// LDS images of the FP8-variant weights: FP16 layer-1 and output fragments, e4m3 hidden ones.
template <int K0>
struct Fp8Lds {
    uint4 l1[MlpShape<K0>::F_L1 * 64];
    uint4 out[4 * 64];
    uint2 hid[Fp8Weights::kHidden * Fp8Weights::kFragsPerLayer * 64];
};

template <int K0>
__device__ __forceinline__ void stage_weights_fp8(const NetArgs& a, Fp8Lds<K0>& w)
{
    typedef MlpShape<K0> S;
    const uint4* f16 = reinterpret_cast<const uint4*>(a.mlp_frag);
    const uint2* f8  = reinterpret_cast<const uint2*>(a.mlp_frag_fp8);
    for (uint32_t i = threadIdx.x; i < (uint32_t)S::F_L1 * 64; i += blockDim.x) w.l1[i] = f16[i];
    for (uint32_t i = threadIdx.x; i < 4u * 64; i += blockDim.x) w.out[i] = f16[(size_t)(S::F_L1 + 3 * S::F_HID) * 64 + i];
    for (uint32_t i = threadIdx.x; i < (uint32_t)Fp8Weights::kHidden * Fp8Weights::kFragsPerLayer * 64; i += blockDim.x)
        w.hid[i] = f8[i];
    __syncthreads();
}

template <int P, int L, int F, int ND>
__device__ __forceinline__ void mfma_network_fp8(const NetArgs& a, const Fp8Lds<P * L * F>& w, const float* own,
                                                 float z[8])
{
    constexpr int K0 = P * L * F;
    typedef MlpShape<K0> S;
    typedef Frag<false> FR;
    typedef typename FR::vec V;
    typedef typename FR::elem E;
    const uint32_t lane = threadIdx.x & 63u;
    const uint32_t g    = lane >> 5;
    V b1[2][S::KS1];
    layer1_operands<P, L, F, false, ND>(a, own, b1[0], b1[1]);
    floatx16 acc[2][2];
#pragma unroll
    for (int rb = 0; rb < 2; rb++)
#pragma unroll
        for (int cb = 0; cb < 2; cb++)
#pragma unroll
            for (int r = 0; r < 16; r++) acc[rb][cb][r] = 0.f;
#pragma unroll
    for (int rb = 0; rb < 2; rb++)
#pragma unroll
        for (int s = 0; s < S::KS1; s++) {
            const V wv = __builtin_bit_cast(V, w.l1[(rb * S::KS1 + s) * 64 + lane]);
            acc[rb][0] = FR::mfma(wv, b1[0][s], acc[rb][0]);
            acc[rb][1] = FR::mfma(wv, b1[1][s], acc[rb][1]);
        }
    float cum[2] = {a.table_scale, a.table_scale};  // product of every scale applied, per column block
#pragma unroll
    for (int layer = 0; layer < Fp8Weights::kHidden; layer++) {
        float sx[2];
        column_scales(acc, sx);
        fp8x8_t bh[2][4];
#pragma unroll
        for (int cb = 0; cb < 2; cb++) {
#pragma unroll
            for (int s = 0; s < 4; s++) {
                float v[8];
#pragma unroll
                for (int e = 0; e < 8; e++) v[e] = fmaxf(acc[s / 2][cb][8 * (s % 2) + e], 0.f) * sx[cb];
                bh[cb][s] = pack_fp8(v);
            }
            cum[cb] *= sx[cb] * a.fp8_wscale[layer];
        }
#pragma unroll
        for (int rb = 0; rb < 2; rb++)
#pragma unroll
            for (int cb = 0; cb < 2; cb++)
#pragma unroll
                for (int r = 0; r < 16; r++) acc[rb][cb][r] = 0.f;
#pragma unroll
        for (int rb = 0; rb < 2; rb++)
#pragma unroll
            for (int s = 0; s < 4; s++) {
                const fp8x8_t wv = __builtin_bit_cast(
                    fp8x8_t, w.hid[(layer * Fp8Weights::kFragsPerLayer + rb * 4 + s) * 64 + lane]);
                acc[rb][0] = __builtin_amdgcn_mfma_f32_32x32x16_fp8_fp8(wv, bh[0][s], acc[rb][0], 0, 0, 0);
                acc[rb][1] = __builtin_amdgcn_mfma_f32_32x32x16_fp8_fp8(wv, bh[1][s], acc[rb][1], 0, 0, 0);
            }
    }
    floatx16 out[2];
    {
        float sx[2];
        column_scales(acc, sx);
        V bh[2][4];
#pragma unroll
        for (int cb = 0; cb < 2; cb++) {
#pragma unroll
            for (int s = 0; s < 4; s++)
#pragma unroll
                for (int e = 0; e < 8; e++) bh[cb][s][e] = (E)(fmaxf(acc[s / 2][cb][8 * (s % 2) + e], 0.f) * sx[cb]);
            cum[cb] *= sx[cb];
        }
#pragma unroll
        for (int cb = 0; cb < 2; cb++)
#pragma unroll
            for (int r = 0; r < 16; r++) out[cb][r] = 0.f;
#pragma unroll
        for (int s = 0; s < 4; s++) {
            const V wv = __builtin_bit_cast(V, w.out[s * 64 + lane]);
            out[0] = FR::mfma(wv, bh[0][s], out[0]);
            out[1] = FR::mfma(wv, bh[1][s], out[1]);
        }
    }
    // cum is a product of powers of two: its reciprocal is exact.
    gather_logits(out, g, 1.f / cum[g], z);
}

// N-BVH nodes copied into LDS per workgroup (the first ones: the top of the breadth-first tree,
// all of it for up to this many nodes); 16 KB for 512 nodes keeps two workgroups per CU.
#ifndef NBVH_NODE_CACHE
#define NBVH_NODE_CACHE 512
#endif
constexpr uint32_t kNodeCache = NBVH_NODE_CACHE;

// Copy the first min(n_nodes, kNodeCache) nodes into LDS (the caller synchronises).
__device__ __forceinline__ uint32_t stage_nodes(const NetArgs& a, uint4* lds)
{
    const uint32_t n = a.n_nodes < kNodeCache ? a.n_nodes : kNodeCache;
    const uint4* src = reinterpret_cast<const uint4*>(a.nodes);
    for (uint32_t i = threadIdx.x; i < 2 * n; i += blockDim.x) lds[i] = src[i];
    return n;
}

// Waves per workgroup of the query kernel; the workgroup shares one LDS copy of the weights.
#ifndef NBVH_MFMA_BLOCK_WAVES
#define NBVH_MFMA_BLOCK_WAVES 4
#endif
constexpr int kWavesPerBlock = NBVH_MFMA_BLOCK_WAVES;
// Minimum waves per SIMD the register allocation must allow (latency hiding of the gathers).
#ifndef NBVH_MFMA_WAVES_PER_EU
#define NBVH_MFMA_WAVES_PER_EU 2
#endif

template <int P, int L, int F, bool BF16, bool PARTS, int ND>
__global__ __launch_bounds__(64 * kWavesPerBlock, NBVH_MFMA_WAVES_PER_EU) void neural_query_mfma_kernel(NetArgs a,
                                                                                RayBatch b,
                                                                                uint32_t* __restrict__ queue)
{
    constexpr int NF = MlpShape<P * L * F>::N_FRAG;
    __shared__ uint4 wfrag[NF * 64];
    __shared__ uint32_t stack_lds[kWavesPerBlock * kNeuralStackSize * 64];
    __shared__ uint4 node_lds[2 * (kNodeCache > 0 ? kNodeCache : 1)];
    const uint32_t n_cached = stage_nodes(a, node_lds);
    stage_weights<P, L, F, BF16>(a, wfrag);  // ends with the barrier
    const CachedNodes nodes{a.nodes, node_lds, n_cached};
    const uint32_t lane = threadIdx.x & 63u, wave = threadIdx.x >> 6;
    LdsStack st{stack_lds + wave * kNeuralStackSize * 64 + lane, 0};
    Slot s;
    s.state = SLOT_EMPTY;
    RayQueue q;
    q.init(queue, batch_count(b), a.image_width, a.queue_bands);
    while (fill_slots<PARTS>(a, nodes, s, st, b, q)) {
        float own[3 * P];
        if (s.state == SLOT_SEG) {
            encode_segment(a, s.ray, s.seg_mint, s.seg_maxt, own, &s.rng);
        } else {
#pragma unroll
            for (int c = 0; c < 3 * P; c++) own[c] = 0.5f;
        }
        float z[8];
        mfma_network<P, L, F, BF16, ND>(a, wfrag, own, z);
        if (s.state == SLOT_SEG) consume_segment<PARTS>(a, s, z, b);
    }
}

template <int P, int L, int F, bool BF16, int ND>
__global__ __launch_bounds__(64 * kWavesPerBlock) void network_mfma_kernel(NetArgs a, const float* __restrict__ in,
                                                                           uint32_t n, float* __restrict__ out,
                                                                           float* __restrict__ logits)
{
    constexpr int NF = MlpShape<P * L * F>::N_FRAG;
    __shared__ uint4 wfrag[NF * 64];
    stage_weights<P, L, F, BF16>(a, wfrag);
    // Persistent over chunks of blockDim.x inputs (the chunk start is block-uniform, so every wave
    // takes part in each MFMA); the weights are staged once per resident workgroup.
    for (uint32_t base = blockIdx.x * blockDim.x; base < n; base += gridDim.x * blockDim.x) {
        const uint32_t i = base + threadIdx.x;
        float own[3 * P];
#pragma unroll
        for (int c = 0; c < 3 * P; c++) own[c] = i < n ? in[(size_t)i * (3 * P) + c] : 0.5f;
        float z[8];
        mfma_network<P, L, F, BF16, ND>(a, wfrag, own, z);
        if (i < n) {
#pragma unroll
            for (int j = 0; j < 8; j++) {
                if (out) out[(size_t)i * 8 + j] = sigmoidf_cw(z[j]);
                if (logits) logits[(size_t)i * 8 + j] = z[j];
            }
        }
    }
}

template <int P, int L, int F, bool PARTS, int ND>
__global__ __launch_bounds__(64 * kWavesPerBlock, NBVH_MFMA_WAVES_PER_EU) void neural_query_fp8_kernel(NetArgs a,
                                                                               RayBatch b,
                                                                               uint32_t* __restrict__ queue)
{
    __shared__ Fp8Lds<P * L * F> w;
    __shared__ uint32_t stack_lds[kWavesPerBlock * kNeuralStackSize * 64];
    __shared__ uint4 node_lds[2 * (kNodeCache > 0 ? kNodeCache : 1)];
    const uint32_t n_cached = stage_nodes(a, node_lds);
    stage_weights_fp8<P * L * F>(a, w);  // ends with the barrier
    const CachedNodes nodes{a.nodes, node_lds, n_cached};
    const uint32_t lane = threadIdx.x & 63u, wave = threadIdx.x >> 6;
    LdsStack st{stack_lds + wave * kNeuralStackSize * 64 + lane, 0};
    Slot s;
    s.state = SLOT_EMPTY;
    RayQueue q;
    q.init(queue, batch_count(b), a.image_width, a.queue_bands);
    while (fill_slots<PARTS>(a, nodes, s, st, b, q)) {
        float own[3 * P];
        if (s.state == SLOT_SEG) {
            encode_segment(a, s.ray, s.seg_mint, s.seg_maxt, own, &s.rng);
        } else {
#pragma unroll
            for (int c = 0; c < 3 * P; c++) own[c] = 0.5f;
        }
        float z[8];
        mfma_network_fp8<P, L, F, ND>(a, w, own, z);
        if (s.state == SLOT_SEG) consume_segment<PARTS>(a, s, z, b);
    }
}

template <int P, int L, int F, int ND>
__global__ __launch_bounds__(64 * kWavesPerBlock) void network_fp8_kernel(NetArgs a, const float* __restrict__ in,
                                                                          uint32_t n, float* __restrict__ out,
                                                                          float* __restrict__ logits)
{
    __shared__ Fp8Lds<P * L * F> w;
    stage_weights_fp8<P * L * F>(a, w);
    for (uint32_t base = blockIdx.x * blockDim.x; base < n; base += gridDim.x * blockDim.x) {
        const uint32_t i = base + threadIdx.x;
        float own[3 * P];
#pragma unroll
        for (int c = 0; c < 3 * P; c++) own[c] = i < n ? in[(size_t)i * (3 * P) + c] : 0.5f;
        float z[8];
        mfma_network_fp8<P, L, F, ND>(a, w, own, z);
        if (i < n) {
#pragma unroll
            for (int j = 0; j < 8; j++) {
                if (out) out[(size_t)i * 8 + j] = sigmoidf_cw(z[j]);
                if (logits) logits[(size_t)i * 8 + j] = z[j];
            }
        }
    }
}

}  // namespace dev

namespace {

template <typename K>
uint32_t blocks_per_cu(K kernel, uint32_t threads)
{
    int b = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&b, kernel, (int)threads, 0) != hipSuccess || b <= 0) {
        (void)hipGetLastError();
        b = 1;
    }
    return (uint32_t)b;
}

// Dense-prefix length the kernels are specialised for (the hash-only gather of levels >= it):
// 4 covers the 1024x1024 workload's grid (16 levels, base 16, scale 1.5, 2^19) and the chess
// configuration (8 levels, base 8, scale 2, 2^18). Other grids take the general gather (ND = -1).
constexpr int kDensePrefix = 4;

template <int P, int L, int F, bool BF16, bool PARTS, int ND>
hipError_t launch_query_k(const NetArgs& a, const RayBatch& rb, uint32_t* queue, uint32_t n_cus, hipStream_t stream)
{
    auto k = dev::neural_query_mfma_kernel<P, L, F, BF16, PARTS, ND>;
    static const uint32_t bpc = blocks_per_cu(k, 64 * dev::kWavesPerBlock);
    const uint32_t per_block = 64 * dev::kWavesPerBlock;
    uint32_t blocks = (rb.n + per_block - 1) / per_block;
    const uint32_t cap = a.blocks_per_cu != 0u && a.blocks_per_cu < bpc ? a.blocks_per_cu : bpc;
    blocks = blocks < n_cus * cap ? blocks : n_cus * cap;
    hipLaunchKernelGGL(k, dim3(blocks), dim3(per_block), 0, stream, a, rb, queue);
    return hipGetLastError();
}

template <int P, int L, int F, bool BF16, int ND>
hipError_t launch_query_nd(const NetArgs& a, const RayBatch& rb, uint32_t* queue, uint32_t n_cus, hipStream_t stream)
{
    return rb.parts ? launch_query_k<P, L, F, BF16, true, ND>(a, rb, queue, n_cus, stream)
                    : launch_query_k<P, L, F, BF16, false, ND>(a, rb, queue, n_cus, stream);
}

template <int P, int L, int F, bool BF16>
hipError_t launch_query(const NetArgs& a, const RayBatch& rb, uint32_t* queue, uint32_t n_cus, hipStream_t stream)
{
    return a.n_dense == (uint32_t)kDensePrefix ? launch_query_nd<P, L, F, BF16, kDensePrefix>(a, rb, queue, n_cus, stream)
                                               : launch_query_nd<P, L, F, BF16, -1>(a, rb, queue, n_cus, stream);
}

template <int P, int L, int F, bool BF16, int ND>
hipError_t launch_network_nd(const NetArgs& a, const float* in, uint32_t n, float* out, float* logits, uint32_t n_cus,
                             hipStream_t stream)
{
    auto k = dev::network_mfma_kernel<P, L, F, BF16, ND>;
    static const uint32_t bpc = blocks_per_cu(k, 64 * dev::kWavesPerBlock);
    const uint32_t per_block = 64 * dev::kWavesPerBlock;
    uint32_t blocks = (n + per_block - 1) / per_block;
    blocks = blocks < n_cus * bpc ? blocks : n_cus * bpc;
    hipLaunchKernelGGL(k, dim3(blocks), dim3(per_block), 0, stream, a, in, n, out, logits);
    return hipGetLastError();
}

template <int P, int L, int F, bool BF16>
hipError_t launch_network(const NetArgs& a, const float* in, uint32_t n, float* out, float* logits, uint32_t n_cus,
                          hipStream_t stream)
{
    return a.n_dense == (uint32_t)kDensePrefix
               ? launch_network_nd<P, L, F, BF16, kDensePrefix>(a, in, n, out, logits, n_cus, stream)
               : launch_network_nd<P, L, F, BF16, -1>(a, in, n, out, logits, n_cus, stream);
}

template <int P, int L, int F, bool PARTS, int ND>
hipError_t launch_query_fp8_k(const NetArgs& a, const RayBatch& rb, uint32_t* queue, uint32_t n_cus,
                              hipStream_t stream)
{
    auto k = dev::neural_query_fp8_kernel<P, L, F, PARTS, ND>;
    static const uint32_t bpc = blocks_per_cu(k, 64 * dev::kWavesPerBlock);
    const uint32_t per_block = 64 * dev::kWavesPerBlock;
    uint32_t blocks = (rb.n + per_block - 1) / per_block;
    const uint32_t cap = a.blocks_per_cu != 0u && a.blocks_per_cu < bpc ? a.blocks_per_cu : bpc;
    blocks = blocks < n_cus * cap ? blocks : n_cus * cap;
    hipLaunchKernelGGL(k, dim3(blocks), dim3(per_block), 0, stream, a, rb, queue);
    return hipGetLastError();
}

template <int P, int L, int F, int ND>
hipError_t launch_query_fp8_nd(const NetArgs& a, const RayBatch& rb, uint32_t* queue, uint32_t n_cus,
                               hipStream_t stream)
{
    return rb.parts ? launch_query_fp8_k<P, L, F, true, ND>(a, rb, queue, n_cus, stream)
                    : launch_query_fp8_k<P, L, F, false, ND>(a, rb, queue, n_cus, stream);
}

template <int P, int L, int F>
hipError_t launch_query_fp8(const NetArgs& a, const RayBatch& rb, uint32_t* queue, uint32_t n_cus, hipStream_t stream)
{
    return a.n_dense == (uint32_t)kDensePrefix ? launch_query_fp8_nd<P, L, F, kDensePrefix>(a, rb, queue, n_cus, stream)
                                               : launch_query_fp8_nd<P, L, F, -1>(a, rb, queue, n_cus, stream);
}

template <int P, int L, int F, int ND>
hipError_t launch_network_fp8_nd(const NetArgs& a, const float* in, uint32_t n, float* out, float* logits,
                                 uint32_t n_cus, hipStream_t stream)
{
    auto k = dev::network_fp8_kernel<P, L, F, ND>;
    static const uint32_t bpc = blocks_per_cu(k, 64 * dev::kWavesPerBlock);
    const uint32_t per_block = 64 * dev::kWavesPerBlock;
    uint32_t blocks = (n + per_block - 1) / per_block;
    blocks = blocks < n_cus * bpc ? blocks : n_cus * bpc;
    hipLaunchKernelGGL(k, dim3(blocks), dim3(per_block), 0, stream, a, in, n, out, logits);
    return hipGetLastError();
}

template <int P, int L, int F>
hipError_t launch_network_fp8(const NetArgs& a, const float* in, uint32_t n, float* out, float* logits, uint32_t n_cus,
                              hipStream_t stream)
{
    return a.n_dense == (uint32_t)kDensePrefix
               ? launch_network_fp8_nd<P, L, F, kDensePrefix>(a, in, n, out, logits, n_cus, stream)
               : launch_network_fp8_nd<P, L, F, -1>(a, in, n, out, logits, n_cus, stream);
}

}  // namespace

#define NBVH_MFMA_SHAPES(X) X(3, 16, 2) X(3, 8, 4)

hipError_t query_mfma(const NetArgs& a, uint32_t precision, const RayBatch& rb, uint32_t* queue, uint32_t n_cus,
                      hipStream_t stream, bool* supported)
{
    *supported = true;
    if (a.n_hidden_layers == 4 && a.mlp_frag != nullptr && a.tables_f16_bytes != 0) {
#define X(P_, L_, F_)                                                                 \
    if (a.n_points == P_ && a.n_levels == L_ && a.n_features == F_) {                \
        if (precision == NBVH_PREC_FP16)                                             \
            return launch_query<P_, L_, F_, false>(a, rb, queue, n_cus, stream);          \
        if (precision == NBVH_PREC_BF16)                                             \
            return launch_query<P_, L_, F_, true>(a, rb, queue, n_cus, stream);           \
        if (precision == NBVH_PREC_FP8 && a.mlp_frag_fp8 != nullptr)                 \
            return launch_query_fp8<P_, L_, F_>(a, rb, queue, n_cus, stream);             \
    }
        NBVH_MFMA_SHAPES(X)
#undef X
    }
    *supported = false;
    return hipSuccess;
}

hipError_t network_mfma(const NetArgs& a, uint32_t precision, const float* in, uint32_t n, float* out, float* logits,
                        uint32_t n_cus, hipStream_t stream, bool* supported)
{
    *supported = true;
    if (a.n_hidden_layers == 4 && a.mlp_frag != nullptr && a.tables_f16_bytes != 0) {
#define X(P_, L_, F_)                                                                 \
    if (a.n_points == P_ && a.n_levels == L_ && a.n_features == F_) {                \
        if (precision == NBVH_PREC_FP16)                                             \
            return launch_network<P_, L_, F_, false>(a, in, n, out, logits, n_cus, stream);  \
        if (precision == NBVH_PREC_BF16)                                             \
            return launch_network<P_, L_, F_, true>(a, in, n, out, logits, n_cus, stream);   \
        if (precision == NBVH_PREC_FP8 && a.mlp_frag_fp8 != nullptr)                 \
            return launch_network_fp8<P_, L_, F_>(a, in, n, out, logits, n_cus, stream);     \
    }
        NBVH_MFMA_SHAPES(X)
#undef X
    }
    *supported = false;
    return hipSuccess;
}

}  // namespace nbvh

// nb_mlp.h -- learned far-field correction: network layout and device helpers
// shared by the training kernels (nb_mlp.hip) and the corrected traversal
// (nb_forces.hip).
//
// Network: 13 inputs -> 64 -> 64 -> 64 (SiLU) -> 3 outputs, BF16 weights and
// activations on the matrix cores (v_mfma_f32_16x16x{16,32}_bf16), FP32
// accumulation and bias.  A wavefront evaluates a tile of 16 samples; every
// layer is computed transposed, out^T = W^T in^T, so a layer's result (one
// 16x16 block per 16 features, lane (g, p) holding features 4g..4g+3 of
// sample p) is directly the next layer's B operand once the K index of the
// weights is permuted to match (slot_feature below): no transposition through
// LDS on the forward or backward path.
#pragma once

#include "nb_internal.h"

namespace nb {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef __bf16 bf16x4 __attribute__((ext_vector_type(4)));
typedef short s16x4 __attribute__((ext_vector_type(4)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

constexpr int kMlpIn = 13, kMlpHidden = 64, kMlpOut = 3, kMlpParams = 9411;
constexpr int kMlpTile = 16;  // samples per wavefront tile

// FP32 master weights on the device, padded to whole 16x16 blocks:
// W1[16][64] b1[64] W2[64][64] b2[64] W3[64][64] b3[64] W4[64][16] b4[16]
// (row = input index).  Padding entries are zero and stay zero in training.
constexpr int kPW1 = 0, kPB1 = 1024, kPW2 = 1088, kPB2 = 5184, kPW3 = 5248, kPB3 = 9344,
              kPW4 = 9408, kPB4 = 10432, kMlpPadded = 10448;

// Operand pack: BF16 weights pre-arranged in matrix-core lane order, one
// 16-byte slot per lane per operand (16x16x16 operands use the first 8 bytes).
enum : int {
    kOpA1 = 0,    // forward layer 1, [ob]          (4)   16x16x16
    kOpA2 = 4,    // forward layer 2, [ob][kb]      (8)   16x16x32
    kOpA3 = 12,   // forward layer 3, [ob][kb]      (8)
    kOpA4 = 20,   // forward layer 4, [kb]          (2)
    kOpB4T = 22,  // backward through W4, [ib]      (4)   16x16x16
    kOpA3T = 26,  // backward through W3, [ib][kb]  (8)
    kOpA2T = 34,  // backward through W2, [ib][kb]  (8)
    kOpCount = 42
};

// Feature index held in K slot i (0..7) of lane group g for the 32-wide K
// block kb: the two 16-feature output blocks 2kb and 2kb+1 of the previous
// layer, 4 values each, in the order the previous layer's result holds them.
__host__ __device__ inline int slot_feature(int kb, int g, int i) {
    return 32 * kb + (i < 4 ? 4 * g + i : 16 + 4 * g + (i - 4));
}

// Per internal cell, for the correction's features: normalised traceless
// quadrupole Q / (M e^2) {xx, yy, zz, xy, xz, yz}, geometric-centre offset
// o = (centre - com) / e, ln(M / M_root) / 16, and the edge e.
struct MlpCell {
    float4 a;  // qxx qyy qzz qxy
    float4 b;  // qxz qyz ox  oy
    float4 c;  // oz  lm  e   0
};

__device__ inline f32x4 mfma32(bf16x8 a, bf16x8 b, f32x4 c) {
    return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0);
}
__device__ inline f32x4 mfma16(bf16x4 a, bf16x4 b, f32x4 c) {
    return __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(__builtin_bit_cast(s16x4, a),
                                                      __builtin_bit_cast(s16x4, b), c, 0, 0, 0);
}

__device__ inline float mlp_sigmoid(float z) { return __builtin_amdgcn_rcpf(1.0f + __expf(-z)); }
__device__ inline float mlp_silu(float z) { return z * mlp_sigmoid(z); }
__device__ inline float mlp_dsilu(float z) {
    const float s = mlp_sigmoid(z);
    return s * (1.0f + z * (1.0f - s));
}

__device__ inline bf16x8 pack8(f32x4 lo, f32x4 hi) {
    bf16x8 r;
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        r[i] = (__bf16)lo[i];
        r[4 + i] = (__bf16)hi[i];
    }
    return r;
}
__device__ inline bf16x8 silu8(f32x4 lo, f32x4 hi) {
    bf16x8 r;
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        r[i] = (__bf16)mlp_silu(lo[i]);
        r[4 + i] = (__bf16)mlp_silu(hi[i]);
    }
    return r;
}

__device__ inline bf16x8 load_op8(const uint4 *pack, int op, int lane) {
    return __builtin_bit_cast(bf16x8, pack[op * 64 + lane]);
}
__device__ inline bf16x4 load_op4(const uint4 *pack, int op, int lane) {
    const uint4 u = pack[op * 64 + lane];
    return __builtin_bit_cast(bf16x4, make_uint2(u.x, u.y));
}
// bias values of lane group g for a 16-feature block starting at `base`
__device__ inline f32x4 load_bias(const float *master, int base, int g) {
    const float4 b = *reinterpret_cast<const float4 *>(master + base + 4 * g);
    return f32x4{b.x, b.y, b.z, b.w};
}

// Forward weights of one lane, held in registers.
struct MlpForward {
    bf16x4 a1[4];
    bf16x8 a2[4][2], a3[4][2], a4[2];
    f32x4 b1[4], b2[4], b3[4], b4;

    __device__ void load(const uint4 *pack, const float *master, int lane) {
        const int g = lane >> 4;
#pragma unroll
        for (int ob = 0; ob < 4; ++ob) {
            a1[ob] = load_op4(pack, kOpA1 + ob, lane);
#pragma unroll
            for (int kb = 0; kb < 2; ++kb) {
                a2[ob][kb] = load_op8(pack, kOpA2 + 2 * ob + kb, lane);
                a3[ob][kb] = load_op8(pack, kOpA3 + 2 * ob + kb, lane);
            }
            b1[ob] = load_bias(master, kPB1 + 16 * ob, g);
            b2[ob] = load_bias(master, kPB2 + 16 * ob, g);
            b3[ob] = load_bias(master, kPB3 + 16 * ob, g);
        }
        a4[0] = load_op8(pack, kOpA4, lane);
        a4[1] = load_op8(pack, kOpA4 + 1, lane);
        b4 = load_bias(master, kPB4, g);
    }
};

// Pre-activations and BF16 activations of one tile (kept for the backward pass).
struct MlpActs {
    f32x4 z1[4], z2[4], z3[4];
    bf16x8 h1[2], h2[2], h3[2];
};

// One tile forward.  xin: lane (g, p) holds features 4g..4g+3 of sample p.
// Returns the output block: lane (0, p) holds the 3 outputs of sample p
// (rows 3..15 are zero).
__device__ inline f32x4 mlp_forward_tile(const MlpForward &w, bf16x4 xin, MlpActs &a) {
#pragma unroll
    for (int ob = 0; ob < 4; ++ob) a.z1[ob] = mfma16(w.a1[ob], xin, w.b1[ob]);
    a.h1[0] = silu8(a.z1[0], a.z1[1]);
    a.h1[1] = silu8(a.z1[2], a.z1[3]);
#pragma unroll
    for (int ob = 0; ob < 4; ++ob)
        a.z2[ob] = mfma32(w.a2[ob][1], a.h1[1], mfma32(w.a2[ob][0], a.h1[0], w.b2[ob]));
    a.h2[0] = silu8(a.z2[0], a.z2[1]);
    a.h2[1] = silu8(a.z2[2], a.z2[3]);
#pragma unroll
    for (int ob = 0; ob < 4; ++ob)
        a.z3[ob] = mfma32(w.a3[ob][1], a.h2[1], mfma32(w.a3[ob][0], a.h2[0], w.b3[ob]));
    a.h3[0] = silu8(a.z3[0], a.z3[1]);
    a.h3[1] = silu8(a.z3[2], a.z3[3]);
    return mfma32(w.a4[1], a.h3[1], mfma32(w.a4[0], a.h3[0], w.b4));
}

// The 13 features of a (cell, target) pair from FP32 quantities: r = com -
// target, the cell's record m.  Lane group g receives features 4g..4g+3.
__device__ inline bf16x4 mlp_features(float rx, float ry, float rz, const MlpCell &m, int g,
                                      float &inv_d) {
    const float d2 = fmaf(rx, rx, fmaf(ry, ry, rz * rz));
    inv_d = __builtin_amdgcn_rsqf(d2);
    const float ux = rx * inv_d, uy = ry * inv_d, uz = rz * inv_d;
    const float par = fmaf(m.b.z, ux, fmaf(m.b.w, uy, m.c.x * uz));
    const float o2 = fmaf(m.b.z, m.b.z, fmaf(m.b.w, m.b.w, m.c.x * m.c.x));
    const float perp = __builtin_sqrtf(fmaxf(o2 - par * par, 0.f));
    float f0, f1, f2, f3;
    if (g == 0) {
        f0 = ux; f1 = uy; f2 = uz; f3 = m.c.z * inv_d;
    } else if (g == 1) {
        f0 = m.c.y; f1 = m.a.x; f2 = m.a.y; f3 = m.a.z;
    } else if (g == 2) {
        f0 = m.a.w; f1 = m.b.x; f2 = m.b.y; f3 = par;
    } else {
        f0 = perp; f1 = 0.f; f2 = 0.f; f3 = 0.f;
    }
    bf16x4 r;
    r[0] = (__bf16)f0;
    r[1] = (__bf16)f1;
    r[2] = (__bf16)f2;
    r[3] = (__bf16)f3;
    return r;
}

// Device weights of a network, for the corrected traversal.
struct MlpWeights {
    const uint4 *pack;
    const float *master;
};
MlpWeights mlp_weights(const nb_mlp *n);

// Correction arguments of the traversal.
struct CorrArgs {
    const uint4 *pack;
    const float *master;
    const MlpCell *mcell;  // per internal cell
    float tau2;            // correct cells with e^2 > tau2 * d_group^2
};

// quad: 6 FP64 per internal cell (needed); mcell, gquad: optional outputs
void launch_quadrupoles(const TreeBufs &t, const double *x, const double *y, const double *z,
                        const double *m, double *quad, MlpCell *mcell, float4 *gquad, double G,
                        hipStream_t s);
void launch_pair_walk(const TreeBufs &t, uint32_t nt, const uint32_t *targets, const double *x,
                      const double *y, const double *z, double theta, uint32_t *stack,
                      uint32_t stack_cap, uint64_t cap, uint32_t *pair_cell, uint32_t *pair_body,
                      unsigned long long *counter, hipStream_t s);
void launch_pair_targets(const TreeBufs &t, uint64_t count, const uint32_t *pair_cell,
                         const uint32_t *pair_body, const double *x, const double *y,
                         const double *z, const double *m, const double *quad,
                         const MlpCell *mcell, double G, double eps2, float *features,
                         float *targets, hipStream_t s);
void launch_bh_group_corr(const TreeBufs &t, const double *x, const double *y, const double *z,
                          const double *m, double eps2, double theta, const CorrArgs &corr,
                          const float4 *gquad, double *ax, double *ay, double *az,
                          uint32_t *interactions, hipStream_t s);

}  // namespace nb

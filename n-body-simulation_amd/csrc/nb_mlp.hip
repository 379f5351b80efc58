// nb_mlp.hip -- learned far-field correction for the Barnes-Hut traversal.
//
// A small network predicts, for an accepted octree cell and a target body,
// the residual of the monopole approximation:
//     correction = net(features) * G M e^2 / d^4
// (features and normalisation: nbody_capi.h, nb_mlp_pairs).  This file holds
//   - the octree's FP64 quadrupole moments (parallel-axis sums, deepest level
//     first) and the per-cell FP32 feature records;
//   - training pairs: (cell, body) pairs accepted by a per-body opening test,
//     with targets from the exact sum over the cell's bodies (the direct-sum
//     arithmetic restricted to the cell);
//   - the network on the matrix cores: inference, a fused forward+backward
//     pass per 16-sample tile, the weight-gradient products over the batch,
//     and Adam -- every reduction in a fixed order, so training is
//     deterministic;
//   - the nb_mlp_* entry points.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <numeric>
#include <string>
#include <vector>

#include "nb_mlp.h"

namespace nb {
namespace {

constexpr int kThreads = 256;

// ---- weights -----------------------------------------------------------------
// Operand op, lane l: the 8 BF16 K-slot values of that lane (see nb_mlp.h).
__global__ void __launch_bounds__(64) k_mlp_pack(const float *__restrict__ w, uint4 *pack) {
    const int op = blockIdx.x, lane = threadIdx.x, g = lane >> 4, r = lane & 15;
    float v[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    if (op < kOpA2) {
        const int ob = op - kOpA1;
        for (int i = 0; i < 4; ++i) v[i] = w[kPW1 + (4 * g + i) * 64 + 16 * ob + r];
    } else if (op < kOpA4) {
        const int base = op < kOpA3 ? kPW2 : kPW3;
        const int t = op - (op < kOpA3 ? kOpA2 : kOpA3), ob = t >> 1, kb = t & 1;
        for (int i = 0; i < 8; ++i) v[i] = w[base + slot_feature(kb, g, i) * 64 + 16 * ob + r];
    } else if (op < kOpB4T) {
        const int kb = op - kOpA4;
        for (int i = 0; i < 8; ++i) v[i] = w[kPW4 + slot_feature(kb, g, i) * 16 + r];
    } else if (op < kOpA3T) {
        const int ib = op - kOpB4T;
        for (int i = 0; i < 4; ++i) v[i] = w[kPW4 + (16 * ib + r) * 16 + 4 * g + i];
    } else {
        const int base = op < kOpA2T ? kPW3 : kPW2;
        const int t = op - (op < kOpA2T ? kOpA3T : kOpA2T), ib = t >> 1, kb = t & 1;
        for (int i = 0; i < 8; ++i) v[i] = w[base + (16 * ib + r) * 64 + slot_feature(kb, g, i)];
    }
    bf16x8 b;
    for (int i = 0; i < 8; ++i) b[i] = (__bf16)v[i];
    pack[op * 64 + lane] = __builtin_bit_cast(uint4, b);
}

// ---- inference -------------------------------------------------------------------
__global__ void __launch_bounds__(64) k_mlp_infer(uint64_t count, const float *__restrict__ F,
                                                  float *__restrict__ Y,
                                                  const uint4 *__restrict__ pack,
                                                  const float *__restrict__ master) {
    const int lane = threadIdx.x, g = lane >> 4, p = lane & 15;
    MlpForward w;
    w.load(pack, master, lane);
    for (uint64_t tile = blockIdx.x; tile * kMlpTile < count; tile += gridDim.x) {
        const uint64_t j = tile * kMlpTile + p;
        const bool valid = j < count;
        bf16x4 xin;
        for (int i = 0; i < 4; ++i) {
            const int f = 4 * g + i;
            xin[i] = (__bf16)((valid && f < kMlpIn) ? F[j * kMlpIn + f] : 0.f);
        }
        MlpActs a;
        const f32x4 y = mlp_forward_tile(w, xin, a);
        if (g == 0 && valid)
            for (int c = 0; c < kMlpOut; ++c) Y[j * kMlpOut + c] = y[c];
    }
}

// ---- training ---------------------------------------------------------------------
// Per-batch activation store, transposed ([feature][sample], BF16), so that the
// weight-gradient products read 8 consecutive samples per lane.
struct BatchStore {
    __bf16 *xt, *h1t, *h2t, *h3t;   // 16, 64, 64, 64 rows
    __bf16 *d1t, *d2t, *d3t, *dyt;  // 64, 64, 64, 16 rows
};

// sample index of batch position t: an affine permutation of the data set
__device__ inline uint64_t batch_sample(uint64_t count, uint64_t pa, uint64_t pc, uint64_t base,
                                        uint64_t t) {
    const uint64_t pos = (base + t) % count;
    return ((pa * pos) % count + pc) % count;  // pa, pos < 2^32: no overflow
}

// Uniformly random rotation (Shoemake's quaternion method) drawn from a
// splitmix64 stream keyed by `key`.
__device__ inline void random_rotation(uint64_t key, float R[9]) {
    float u[3];
    for (int k = 0; k < 3; ++k) {
        uint64_t z = (key += 0x9E3779B97F4A7C15ull);
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        z ^= z >> 31;
        u[k] = (float)(z >> 40) * 0x1.0p-24f;
    }
    const float a = sqrtf(1.f - u[0]), b = sqrtf(u[0]);
    const float t2 = 6.2831853f * u[1], t3 = 6.2831853f * u[2];
    const float w = a * __sinf(t2), x = a * __cosf(t2), y = b * __sinf(t3), z = b * __cosf(t3);
    R[0] = 1 - 2 * (y * y + z * z); R[1] = 2 * (x * y - w * z);     R[2] = 2 * (x * z + w * y);
    R[3] = 2 * (x * y + w * z);     R[4] = 1 - 2 * (x * x + z * z); R[5] = 2 * (y * z - w * x);
    R[6] = 2 * (x * z - w * y);     R[7] = 2 * (y * z + w * x);     R[8] = 1 - 2 * (x * x + y * y);
}

// Rotate one sample: the direction (f0..2), the quadrupole (f5..10, as the
// symmetric matrix R Q R^T) and the target; the other features are invariant.
__device__ inline void rotate_sample(const float R[9], float f[kMlpIn], float t[kMlpOut]) {
    float u[3], v[3];
    for (int r = 0; r < 3; ++r) u[r] = R[3 * r] * f[0] + R[3 * r + 1] * f[1] + R[3 * r + 2] * f[2];
    for (int r = 0; r < 3; ++r) v[r] = R[3 * r] * t[0] + R[3 * r + 1] * t[1] + R[3 * r + 2] * t[2];
    const float Q[9] = {f[5], f[8], f[9], f[8], f[6], f[10], f[9], f[10], f[7]};
    float RQ[9], Q2[9];
    for (int r = 0; r < 3; ++r)
        for (int c = 0; c < 3; ++c)
            RQ[3 * r + c] = R[3 * r] * Q[c] + R[3 * r + 1] * Q[3 + c] + R[3 * r + 2] * Q[6 + c];
    for (int r = 0; r < 3; ++r)
        for (int c = 0; c < 3; ++c)
            Q2[3 * r + c] = RQ[3 * r] * R[3 * c] + RQ[3 * r + 1] * R[3 * c + 1] + RQ[3 * r + 2] * R[3 * c + 2];
    for (int r = 0; r < 3; ++r) {
        f[r] = u[r];
        t[r] = v[r];
    }
    f[5] = Q2[0];
    f[6] = Q2[4];
    f[7] = Q2[8];
    f[8] = Q2[1];
    f[9] = Q2[2];
    f[10] = Q2[5];
}

// One tile of 16 samples: forward, loss, backward to every pre-activation
// gradient.  lossp[tile] = sum of squared errors of the tile.  aug_key != 0
// rotates every sample by a random rotation keyed by (aug_key, sample).
__global__ void __launch_bounds__(64) k_mlp_fwdbwd(
    uint32_t B, uint32_t nvalid, const float *__restrict__ F, const float *__restrict__ T,
    uint64_t count, uint64_t pa, uint64_t pc, uint64_t base, float coef, uint64_t aug_key,
    const uint4 *__restrict__ pack, const float *__restrict__ master, BatchStore st,
    float *lossp) {
    const int lane = threadIdx.x, g = lane >> 4, p = lane & 15;
    const uint32_t tile = blockIdx.x;
    const uint32_t t = tile * kMlpTile + p;
    const bool valid = t < nvalid;
    const uint64_t j = valid ? batch_sample(count, pa, pc, base, t) : 0;
    MlpForward w;
    w.load(pack, master, lane);
    float fe[kMlpIn], tv[kMlpOut];
    for (int f = 0; f < kMlpIn; ++f) fe[f] = valid ? F[j * kMlpIn + f] : 0.f;
    for (int c = 0; c < kMlpOut; ++c) tv[c] = valid ? T[j * kMlpOut + c] : 0.f;
    if (aug_key && valid) {
        float R[9];
        random_rotation(aug_key ^ (j * 0xD1B54A32D192ED03ull), R);
        rotate_sample(R, fe, tv);
    }
    // lane group g takes features 4g..4g+3 (constant indices: no scratch)
    const float s0 = g == 0 ? fe[0] : g == 1 ? fe[4] : g == 2 ? fe[8] : fe[12];
    const float s1 = g == 0 ? fe[1] : g == 1 ? fe[5] : g == 2 ? fe[9] : 0.f;
    const float s2 = g == 0 ? fe[2] : g == 1 ? fe[6] : g == 2 ? fe[10] : 0.f;
    const float s3 = g == 0 ? fe[3] : g == 1 ? fe[7] : g == 2 ? fe[11] : 0.f;
    bf16x4 xin;
    xin[0] = (__bf16)s0;
    xin[1] = (__bf16)s1;
    xin[2] = (__bf16)s2;
    xin[3] = (__bf16)s3;
    MlpActs a;
    const f32x4 y = mlp_forward_tile(w, xin, a);
    float sq = 0.f;
    bf16x4 dyb;
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        float dy = 0.f;
        if (g == 0 && i < kMlpOut && valid) {
            const float r = y[i] - tv[i];
            sq += r * r;
            dy = coef * r;
        }
        dyb[i] = (__bf16)dy;
    }
    for (int off = 32; off >= 1; off >>= 1) sq += __shfl_xor(sq, off, 64);
    if (lane == 0) lossp[tile] = sq;

    // backward: dh3 = W4 dy, dz3 = dh3 * silu'(z3), ...
    const f32x4 zero = {0.f, 0.f, 0.f, 0.f};
    f32x4 dz[4];
    bf16x8 d3[2], d2[2], d1[2];
#pragma unroll
    for (int ib = 0; ib < 4; ++ib) {
        const f32x4 dh = mfma16(load_op4(pack, kOpB4T + ib, lane), dyb, zero);
        for (int i = 0; i < 4; ++i) dz[ib][i] = dh[i] * mlp_dsilu(a.z3[ib][i]);
    }
    d3[0] = pack8(dz[0], dz[1]);
    d3[1] = pack8(dz[2], dz[3]);
#pragma unroll
    for (int ib = 0; ib < 4; ++ib) {
        const f32x4 dh = mfma32(load_op8(pack, kOpA3T + 2 * ib + 1, lane), d3[1],
                                mfma32(load_op8(pack, kOpA3T + 2 * ib, lane), d3[0], zero));
        for (int i = 0; i < 4; ++i) dz[ib][i] = dh[i] * mlp_dsilu(a.z2[ib][i]);
    }
    d2[0] = pack8(dz[0], dz[1]);
    d2[1] = pack8(dz[2], dz[3]);
#pragma unroll
    for (int ib = 0; ib < 4; ++ib) {
        const f32x4 dh = mfma32(load_op8(pack, kOpA2T + 2 * ib + 1, lane), d2[1],
                                mfma32(load_op8(pack, kOpA2T + 2 * ib, lane), d2[0], zero));
        for (int i = 0; i < 4; ++i) dz[ib][i] = dh[i] * mlp_dsilu(a.z1[ib][i]);
    }
    d1[0] = pack8(dz[0], dz[1]);
    d1[1] = pack8(dz[2], dz[3]);

    // store transposed: row = feature, column = batch position
    for (int i = 0; i < 4; ++i) {
        st.xt[(uint64_t)(4 * g + i) * B + t] = xin[i];
        st.dyt[(uint64_t)(4 * g + i) * B + t] = dyb[i];
    }
#pragma unroll
    for (int kb = 0; kb < 2; ++kb)
        for (int i = 0; i < 8; ++i) {
            const uint64_t row = (uint64_t)slot_feature(kb, g, i) * B + t;
            st.h1t[row] = a.h1[kb][i];
            st.h2t[row] = a.h2[kb][i];
            st.h3t[row] = a.h3[kb][i];
            st.d1t[row] = d1[kb][i];
            st.d2t[row] = d2[kb][i];
            st.d3t[row] = d3[kb][i];
        }
}

// Gradient tiles (16x16 blocks of the padded parameter layout):
//   0-3 W1 (ob)  4-7 b1  8-23 W2 (ib,ob)  24-27 b2  28-43 W3  44-47 b3
//   48-51 W4 (ib)  52 b4
constexpr int kGradTiles = 53;

// dW[in][out] = sum over the batch of act[in] * dz[out]; a bias block uses a
// row of ones for act.  grid = (kGradTiles, ksplit): each workgroup sums its
// share of the batch in order and writes its own partial.
__global__ void __launch_bounds__(64) k_mlp_wgrad(uint32_t B, BatchStore st, float *partial) {
    const int lane = threadIdx.x, g = lane >> 4, r = lane & 15;
    const int id = blockIdx.x;
    const __bf16 *at = nullptr, *bt = nullptr;
    int ib = 0, ob = 0, pbase = 0, ncols = 64;
    bool bias = false;
    if (id < 4) { at = st.xt; bt = st.d1t; ob = id; pbase = kPW1; }
    else if (id < 8) { bias = true; bt = st.d1t; ob = id - 4; pbase = kPB1; }
    else if (id < 24) { at = st.h1t; bt = st.d2t; ib = (id - 8) >> 2; ob = (id - 8) & 3; pbase = kPW2; }
    else if (id < 28) { bias = true; bt = st.d2t; ob = id - 24; pbase = kPB2; }
    else if (id < 44) { at = st.h2t; bt = st.d3t; ib = (id - 28) >> 2; ob = (id - 28) & 3; pbase = kPW3; }
    else if (id < 48) { bias = true; bt = st.d3t; ob = id - 44; pbase = kPB3; }
    else if (id < 52) { at = st.h3t; bt = st.dyt; ib = id - 48; pbase = kPW4; ncols = 16; }
    else { bias = true; bt = st.dyt; pbase = kPB4; }
    const uint32_t per = B / gridDim.y;  // a multiple of 32
    const uint32_t k0 = blockIdx.y * per, k1 = k0 + per;
    bf16x8 ones;
    for (int i = 0; i < 8; ++i) ones[i] = (__bf16)1.0f;
    f32x4 acc = {0.f, 0.f, 0.f, 0.f};
    for (uint32_t k = k0; k < k1; k += 32) {
        const bf16x8 a = bias ? ones
                              : __builtin_bit_cast(bf16x8, *reinterpret_cast<const uint4 *>(
                                                               at + (uint64_t)(16 * ib + r) * B + k + 8 * g));
        const bf16x8 b = __builtin_bit_cast(
            bf16x8, *reinterpret_cast<const uint4 *>(bt + (uint64_t)(16 * ob + r) * B + k + 8 * g));
        acc = mfma32(a, b, acc);
    }
    float *out = partial + (uint64_t)blockIdx.y * kMlpPadded;
    if (bias) {
        if (g == 0) out[pbase + 16 * ob + r] = acc[0];
    } else {
        for (int i = 0; i < 4; ++i) out[pbase + (16 * ib + 4 * g + i) * ncols + 16 * ob + r] = acc[i];
    }
}

// Sum the partials (fixed order); optionally apply Adam; loss of the batch.
__global__ void __launch_bounds__(kThreads) k_mlp_reduce(uint32_t ksplit,
                                                         const float *__restrict__ partial,
                                                         float *grad, const float *__restrict__ lossp,
                                                         uint32_t ntiles, float inv_n3,
                                                         float *loss_out, int adam, float *w,
                                                         float *m, float *v, float lr, float b1,
                                                         float b2, float eps, float c1, float c2) {
    const uint32_t j = blockIdx.x * kThreads + threadIdx.x;
    if (j == 0) {
        float s = 0.f;
        for (uint32_t t = 0; t < ntiles; ++t) s += lossp[t];
        *loss_out = s * inv_n3;
    }
    if (j >= (uint32_t)kMlpPadded) return;
    float gsum = 0.f;
    for (uint32_t s = 0; s < ksplit; ++s) gsum += partial[(uint64_t)s * kMlpPadded + j];
    if (grad) grad[j] = gsum;
    if (adam) {
        const float mj = b1 * m[j] + (1.f - b1) * gsum;
        const float vj = b2 * v[j] + (1.f - b2) * gsum * gsum;
        m[j] = mj;
        v[j] = vj;
        w[j] -= lr * (mj / c1) / (sqrtf(vj / c2) + eps);
    }
}

// ---- quadrupoles and feature records ------------------------------------------------
// Cells of depth `level`: Q = sum over children of Q_child + m_child (3 s s^T
// - |s|^2 I), s = child com - cell com (parallel axis); a body child has
// Q_child = 0 and s = x - com.  Launched deepest level first.  Optional
// outputs: the network's feature record (mcell) and the traversal's FP32
// record G*Q = {xx, yy, zz, xy}, {xz, yz, 0, 0} (gquad, two per cell).
__global__ void __launch_bounds__(kThreads) k_quad_level(
    int level, uint32_t icap, const TreeHdr *__restrict__ h, const double *__restrict__ x,
    const double *__restrict__ y, const double *__restrict__ z, const double *__restrict__ m,
    const int32_t *__restrict__ ref, const InternalSums *__restrict__ sums,
    const uint8_t *__restrict__ int_depth, const uint32_t *__restrict__ int_first,
    const uint32_t *__restrict__ sidx, double *quad, MlpCell *mcell, float4 *gquad, double G) {
    if ((uint32_t)level >= h->max_depth) return;  // no internal cell this deep
    const uint32_t I = min(h->internal_count, icap);
    const uint64_t stride = (uint64_t)gridDim.x * kThreads;
    for (uint64_t k = (uint64_t)blockIdx.x * kThreads + threadIdx.x; k < I; k += stride) {
        if (int_depth[k] != level) continue;
        const InternalSums s = sums[k];
        const double M = s.m;
        const double cx = s.x / M, cy = s.y / M, cz = s.z / M;
        double q[6] = {0, 0, 0, 0, 0, 0};
        for (int d = 0; d < 8; ++d) {
            const int32_t r = ref[1 + 8 * k + d];
            if (r == kRefEmpty) continue;
            double mc, sx, sy, sz;
            if (r >= 0) {
                const InternalSums c = sums[r];
                mc = c.m;
                sx = c.x / c.m - cx;
                sy = c.y / c.m - cy;
                sz = c.z / c.m - cz;
                for (int u = 0; u < 6; ++u) q[u] += quad[6 * (uint64_t)r + u];
            } else {
                const uint32_t b = body_of_ref(r);
                mc = m[b];
                sx = x[b] - cx;
                sy = y[b] - cy;
                sz = z[b] - cz;
            }
            const double s2 = sx * sx + sy * sy + sz * sz;
            q[0] += mc * (3 * sx * sx - s2);
            q[1] += mc * (3 * sy * sy - s2);
            q[2] += mc * (3 * sz * sz - s2);
            q[3] += mc * 3 * sx * sy;
            q[4] += mc * 3 * sx * sz;
            q[5] += mc * 3 * sy * sz;
        }
        for (int u = 0; u < 6; ++u) quad[6 * k + u] = q[u];
        if (gquad) {
            gquad[2 * k] = make_float4((float)(G * q[0]), (float)(G * q[1]), (float)(G * q[2]),
                                       (float)(G * q[3]));
            gquad[2 * k + 1] = make_float4((float)(G * q[4]), (float)(G * q[5]), 0.f, 0.f);
        }
        if (!mcell) continue;
        // geometric cell: replay the classification chain of the cell's first body
        const uint32_t b0 = sidx[int_first[k]];
        Cell c{h->minx, h->miny, h->minz, h->edge};
        for (int l = 0; l < level; ++l) descend(c, x[b0], y[b0], z[b0]);
        const double e = c.e, inv = 1.0 / (M * e * e);
        const double ox = (c.mx + e / 2 - cx) / e, oy = (c.my + e / 2 - cy) / e,
                     oz = (c.mz + e / 2 - cz) / e;
        const double lm = log(M / sums[0].m) / 16.0;
        MlpCell rec;
        rec.a = make_float4((float)(q[0] * inv), (float)(q[1] * inv), (float)(q[2] * inv),
                            (float)(q[3] * inv));
        rec.b = make_float4((float)(q[4] * inv), (float)(q[5] * inv), (float)ox, (float)oy);
        rec.c = make_float4((float)oz, (float)lm, (float)e, 0.f);
        mcell[k] = rec;
    }
}

// ---- training pairs -----------------------------------------------------------------
// Per target body: depth-first walk, accepting an internal cell when
// e / d < theta (d = distance to its centre of mass) and opening it
// otherwise; body leaves are skipped (they are exact).  Accepted cells are
// appended to the pair list (stack in global memory, stack_cap per target).
__global__ void __launch_bounds__(kThreads) k_pair_walk(
    uint32_t nt, const uint32_t *__restrict__ targets, uint32_t icap, const TreeHdr *__restrict__ h,
    const double *__restrict__ x, const double *__restrict__ y, const double *__restrict__ z,
    const int32_t *__restrict__ ref, const InternalSums *__restrict__ sums,
    const uint8_t *__restrict__ int_depth, double theta, uint32_t *stack, uint32_t stack_cap,
    uint64_t cap, uint32_t *pair_cell, uint32_t *pair_body, unsigned long long *counter) {
    const uint32_t t = blockIdx.x * kThreads + threadIdx.x;
    if (t >= nt) return;
    const uint32_t b = targets[t];
    const double px = x[b], py = y[b], pz = z[b];
    const double root_edge = h->edge;
    const uint32_t I = min(h->internal_count, icap);
    if (I == 0) return;
    uint32_t *stk = stack + (uint64_t)t * stack_cap;
    uint32_t sp = 0;
    stk[sp++] = 0;  // internal cell indices
    uint64_t budget = (uint64_t)I + 16;
    while (sp > 0 && budget-- > 0) {
        const uint32_t k = stk[--sp];
        const InternalSums s = sums[k];
        const double dx = s.x / s.m - px, dy = s.y / s.m - py, dz = s.z / s.m - pz;
        const double d2 = dx * dx + dy * dy + dz * dz;
        const double e = ldexp(root_edge, -(int)int_depth[k]);
        if (k != 0 && e * e < theta * theta * d2) {
            const unsigned long long slot = atomicAdd(counter, 1ull);
            if (slot < cap) {
                pair_cell[slot] = k;
                pair_body[slot] = b;
            }
            continue;
        }
        for (int d = 7; d >= 0; --d) {
            const int32_t r = ref[1 + 8ull * k + d];
            if (r >= 0 && (uint32_t)r < I) {
                if (sp >= stack_cap) return;  // cannot happen: stack_cap >= 7 * depth + 1
                stk[sp++] = (uint32_t)r;
            }
        }
    }
}

// Features and target of pair pi (cell k, body at p): the monopole of the
// cell is subtracted from the exact force a of the cell's bodies.
__device__ inline void pair_record(uint64_t pi, uint32_t k, double px, double py, double pz,
                                   double ax, double ay, double az,
                                   const InternalSums *__restrict__ sums,
                                   const double *__restrict__ quad,
                                   const MlpCell *__restrict__ mcell, double G, double eps2,
                                   float *features, float *targets) {
    const InternalSums s = sums[k];
    const double M = s.m;
    const double rx = s.x / M - px, ry = s.y / M - py, rz = s.z / M - pz;
    const double d2 = rx * rx + ry * ry + rz * rz;
    const double d = sqrt(d2);
    double den = d2 + eps2;
    den = den * den * den;
    den = 1.0 / sqrt(den);
    const MlpCell c = mcell[k];
    const double e = (double)c.c.z;
    const double ux = rx / d, uy = ry / d, uz = rz / d;
    const double qn = 1.0 / (M * e * e);
    float *f = features + pi * kMlpIn;
    f[0] = (float)ux;
    f[1] = (float)uy;
    f[2] = (float)uz;
    f[3] = (float)(e / d);
    f[4] = c.c.y;
    for (int u = 0; u < 6; ++u) f[5 + u] = (float)(quad[6ull * k + u] * qn);
    const double ox = c.b.z, oy = c.b.w, oz = c.c.x;
    const double par = ox * ux + oy * uy + oz * uz;
    const double perp2 = ox * ox + oy * oy + oz * oz - par * par;
    f[11] = (float)par;
    f[12] = (float)sqrt(perp2 > 0 ? perp2 : 0.0);
    const double scale = d2 * d2 / (G * M * e * e);
    float *tg = targets + pi * kMlpOut;
    tg[0] = (float)((G * ax - G * M * (rx * den)) * scale);
    tg[1] = (float)((G * ay - G * M * (ry * den)) * scale);
    tg[2] = (float)((G * az - G * M * (rz * den)) * scale);
}

// One wavefront per pair: exact force of the cell's bodies on the target
// (FP64, the direct-sum arithmetic), monopole, features and target.
__global__ void __launch_bounds__(64) k_pair_targets(
    uint64_t count, const uint32_t *__restrict__ pair_cell, const uint32_t *__restrict__ pair_body,
    const double *__restrict__ x, const double *__restrict__ y, const double *__restrict__ z,
    const double *__restrict__ m, const uint32_t *__restrict__ sidx,
    const InternalSums *__restrict__ sums, const uint32_t *__restrict__ int_first,
    const uint32_t *__restrict__ int_count, const double *__restrict__ quad,
    const MlpCell *__restrict__ mcell, double G, double eps2, float *features, float *targets) {
    const int lane = threadIdx.x;
    for (uint64_t pi = blockIdx.x; pi < count; pi += gridDim.x) {
        const uint32_t k = pair_cell[pi], b = pair_body[pi];
        const double px = x[b], py = y[b], pz = z[b];
        const uint32_t f0 = int_first[k], nc = int_count[k];
        double ax = 0, ay = 0, az = 0;
        for (uint32_t i = lane; i < nc; i += 64) {
            const uint32_t j = sidx[f0 + i];
            const double rx = x[j] - px, ry = y[j] - py, rz = z[j] - pz;
            double den = rx * rx + ry * ry + rz * rz + eps2;
            den = den * den * den;
            den = 1.0 / sqrt(den);
            ax += m[j] * (rx * den);
            ay += m[j] * (ry * den);
            az += m[j] * (rz * den);
        }
        for (int off = 32; off >= 1; off >>= 1) {
            ax += __shfl_xor(ax, off, 64);
            ay += __shfl_xor(ay, off, 64);
            az += __shfl_xor(az, off, 64);
        }
        // lane 0 writes the pair; no early `continue` of the other lanes, so
        // the wave stays together for the next pair's shuffles
        if (lane == 0)
            pair_record(pi, k, px, py, pz, ax, ay, az, sums, quad, mcell, G, eps2, features,
                        targets);
    }
}

}  // namespace

void launch_quadrupoles(const TreeBufs &t, const double *x, const double *y, const double *z,
                        const double *m, double *quad, MlpCell *mcell, float4 *gquad, double G,
                        hipStream_t s) {
    const uint32_t blocks = std::max<uint32_t>(1, std::min<uint32_t>(4096, ceil_div(t.icap, kThreads)));
    launch_max_depth(t, s);  // k_quad_level skips the levels below the deepest leaf
    for (int level = kMaxDepth - 1; level >= 0; --level)
        hipLaunchKernelGGL(k_quad_level, dim3(blocks), dim3(kThreads), 0, s, level, t.icap,
                           (const TreeHdr *)t.hdr, x, y, z, m, (const int32_t *)t.ref,
                           (const InternalSums *)t.int_sums, (const uint8_t *)t.int_depth,
                           (const uint32_t *)t.int_first, (const uint32_t *)t.sidx, quad, mcell,
                           gquad, G);
}

void launch_pair_walk(const TreeBufs &t, uint32_t nt, const uint32_t *targets, const double *x,
                      const double *y, const double *z, double theta, uint32_t *stack,
                      uint32_t stack_cap, uint64_t cap, uint32_t *pair_cell, uint32_t *pair_body,
                      unsigned long long *counter, hipStream_t s) {
    hipLaunchKernelGGL(k_pair_walk, dim3(ceil_div(nt, kThreads)), dim3(kThreads), 0, s, nt, targets,
                       t.icap, (const TreeHdr *)t.hdr, x, y, z, (const int32_t *)t.ref,
                       (const InternalSums *)t.int_sums, (const uint8_t *)t.int_depth, theta,
                       stack, stack_cap, cap, pair_cell, pair_body, counter);
}

void launch_pair_targets(const TreeBufs &t, uint64_t count, const uint32_t *pair_cell,
                         const uint32_t *pair_body, const double *x, const double *y,
                         const double *z, const double *m, const double *quad,
                         const MlpCell *mcell, double G, double eps2, float *features,
                         float *targets, hipStream_t s) {
    if (count == 0) return;
    const uint32_t blocks = (uint32_t)std::min<uint64_t>(count, 65536);
    hipLaunchKernelGGL(k_pair_targets, dim3(blocks), dim3(64), 0, s, count, pair_cell, pair_body,
                       x, y, z, m, (const uint32_t *)t.sidx, (const InternalSums *)t.int_sums,
                       (const uint32_t *)t.int_first, (const uint32_t *)t.int_count, quad, mcell,
                       G, eps2, features, targets);
}

}  // namespace nb

// ---- nb_mlp C ABI ---------------------------------------------------------------------
using namespace nb;

struct nb_mlp {
    int device = 0;
    hipStream_t stream = nullptr;
    float *master = nullptr, *adam_m = nullptr, *adam_v = nullptr;
    uint4 *pack = nullptr;
    std::string err;
};

namespace {

int mlp_fail(nb_mlp *n, int code, const std::string &msg) {
    if (n) n->err = msg;
    return code;
}

#define MLP_HIP(n, call)                                                                  \
    do {                                                                                  \
        hipError_t e_ = (call);                                                           \
        if (e_ != hipSuccess)                                                             \
            return mlp_fail(n, NB_E_HIP, std::string(#call) + ": " + hipGetErrorString(e_)); \
    } while (0)

// flat 9411-parameter layout <-> padded device layout
int flat_to_padded(int i) {
    // flat offsets: W1 at 0, then B1, W2, B2, W3, B3, W4, B4
    constexpr int fB1 = 832, fW2 = 896, fB2 = 4992, fW3 = 5056, fB3 = 9152, fW4 = 9216,
                  fB4 = 9408;
    if (i < fB1) return kPW1 + i;  // W1[13][64] rows keep their stride
    if (i < fW2) return kPB1 + (i - fB1);
    if (i < fB2) return kPW2 + (i - fW2);
    if (i < fW3) return kPB2 + (i - fB2);
    if (i < fB3) return kPW3 + (i - fW3);
    if (i < fW4) return kPB3 + (i - fB3);
    if (i < fB4) {
        const int q = i - fW4;  // W4[64][3]
        return kPW4 + (q / 3) * 16 + q % 3;
    }
    return kPB4 + (i - fB4);
}

// splitmix64 -> uniform in [0, 1)
uint64_t splitmix(uint64_t &s) {
    uint64_t z = (s += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}
double uniform01(uint64_t &s) { return (double)(splitmix(s) >> 11) * 0x1.0p-53; }

int refresh_pack(nb_mlp *n) {
    hipLaunchKernelGGL(k_mlp_pack, dim3(kOpCount), dim3(64), 0, n->stream, (const float *)n->master,
                       n->pack);
    MLP_HIP(n, hipGetLastError());
    return NB_OK;
}

// Device scratch of one training / gradient call.
struct Batch {
    uint32_t B = 0, ksplit = 1;
    BatchStore st{};
    float *partial = nullptr, *lossp = nullptr, *grad = nullptr, *loss = nullptr;
    std::vector<void *> owned;
    ~Batch() {
        for (void *p : owned) (void)hipFree(p);
    }
    template <typename T>
    bool alloc(T **p, uint64_t count) {
        void *q = nullptr;
        if (hipMalloc(&q, std::max<uint64_t>(count, 1) * sizeof(T)) != hipSuccess) return false;
        owned.push_back(q);
        *p = static_cast<T *>(q);
        return true;
    }
    bool init(uint32_t batch, uint32_t steps_of_loss) {
        B = (batch + 31) / 32 * 32;
        ksplit = std::max<uint32_t>(1, std::min<uint32_t>(B / 256, 16));
        while ((B / 32) % ksplit) --ksplit;  // each share a whole number of 32-sample steps
        const uint64_t rows16 = 16ull * B, rows64 = 64ull * B;
        return alloc(&st.xt, rows16) && alloc(&st.h1t, rows64) && alloc(&st.h2t, rows64) &&
               alloc(&st.h3t, rows64) && alloc(&st.d1t, rows64) && alloc(&st.d2t, rows64) &&
               alloc(&st.d3t, rows64) && alloc(&st.dyt, rows16) &&
               alloc(&partial, (uint64_t)ksplit * kMlpPadded) && alloc(&lossp, B / kMlpTile) &&
               alloc(&grad, kMlpPadded) && alloc(&loss, std::max<uint32_t>(steps_of_loss, 1));
    }
};

// enqueue forward+backward+gradient (+Adam) of one batch
void enqueue_step(nb_mlp *n, Batch &bt, uint32_t nvalid, const float *F, const float *T,
                  uint64_t count, uint64_t pa, uint64_t pc, uint64_t base, uint64_t aug_key,
                  float *loss_slot, bool adam, float lr, float b1, float b2, float eps, float c1,
                  float c2) {
    const float coef = 2.0f / (3.0f * (float)nvalid);
    hipLaunchKernelGGL(k_mlp_fwdbwd, dim3(bt.B / kMlpTile), dim3(64), 0, n->stream, bt.B, nvalid,
                       F, T, count, pa, pc, base, coef, aug_key, (const uint4 *)n->pack,
                       (const float *)n->master, bt.st, bt.lossp);
    hipLaunchKernelGGL(k_mlp_wgrad, dim3(kGradTiles, bt.ksplit), dim3(64), 0, n->stream, bt.B,
                       bt.st, bt.partial);
    hipLaunchKernelGGL(k_mlp_reduce, dim3(ceil_div(kMlpPadded, kThreads)), dim3(kThreads), 0,
                       n->stream, bt.ksplit, (const float *)bt.partial, adam ? nullptr : bt.grad,
                       (const float *)bt.lossp, bt.B / kMlpTile, 1.0f / (3.0f * (float)nvalid),
                       loss_slot, (int)adam, n->master, n->adam_m, n->adam_v, lr, b1, b2, eps, c1,
                       c2);
    if (adam)
        hipLaunchKernelGGL(k_mlp_pack, dim3(kOpCount), dim3(64), 0, n->stream,
                           (const float *)n->master, n->pack);
}

}  // namespace

namespace nb {
MlpWeights mlp_weights(const nb_mlp *n) { return MlpWeights{n->pack, n->master}; }
}  // namespace nb

extern "C" {

void nb_mlp_default_train_opts(nb_mlp_train_opts *o) {
    if (!o) return;
    o->epochs = 20;
    o->batch = 4096;
    o->lr = 2e-3f;
    o->lr_final = 5e-5f;
    o->beta1 = 0.9f;
    o->beta2 = 0.999f;
    o->eps = 1e-8f;
    o->augment = 1;
    o->seed = 1;
}

int nb_mlp_create(nb_mlp **out, uint64_t seed) {
    if (!out) return NB_E_INVALID_ARG;
    *out = nullptr;
    int count = 0;
    if (hipGetDeviceCount(&count) != hipSuccess || count == 0) return NB_E_NO_DEVICE;
    nb_mlp *n = new nb_mlp();
    auto bail = [&](int code) {
        nb_mlp_destroy(n);
        return code;
    };
    if (hipGetDevice(&n->device) != hipSuccess) return bail(NB_E_HIP);
    if (hipStreamCreateWithFlags(&n->stream, hipStreamNonBlocking) != hipSuccess) return bail(NB_E_HIP);
    if (hipMalloc((void **)&n->master, kMlpPadded * sizeof(float)) != hipSuccess ||
        hipMalloc((void **)&n->adam_m, kMlpPadded * sizeof(float)) != hipSuccess ||
        hipMalloc((void **)&n->adam_v, kMlpPadded * sizeof(float)) != hipSuccess ||
        hipMalloc((void **)&n->pack, kOpCount * 64 * sizeof(uint4)) != hipSuccess)
        return bail(NB_E_HIP);
    // Uniform(-sqrt(6/fan_in), sqrt(6/fan_in)) weights, zero biases
    std::vector<float> flat(kMlpParams, 0.f);
    uint64_t s = seed;
    auto fill = [&](int off, int fan_in, int fan_out) {
        const double a = std::sqrt(6.0 / fan_in);
        for (int i = 0; i < fan_in * fan_out; ++i) flat[off + i] = (float)((2 * uniform01(s) - 1) * a);
    };
    fill(0, 13, 64);
    fill(896, 64, 64);
    fill(5056, 64, 64);
    fill(9216, 64, 3);
    int rc = nb_mlp_set_params(n, flat.data());
    if (rc) return bail(rc);
    *out = n;
    return NB_OK;
}

int nb_mlp_destroy(nb_mlp *n) {
    if (!n) return NB_OK;
    if (n->stream) (void)hipStreamSynchronize(n->stream);
    (void)hipFree(n->master);
    (void)hipFree(n->adam_m);
    (void)hipFree(n->adam_v);
    (void)hipFree(n->pack);
    if (n->stream) (void)hipStreamDestroy(n->stream);
    delete n;
    return NB_OK;
}

const char *nb_mlp_last_error(const nb_mlp *n) { return n ? n->err.c_str() : "null network"; }

int nb_mlp_set_params(nb_mlp *n, const float *params) {
    if (!n || !params) return NB_E_INVALID_ARG;
    std::vector<float> pad(kMlpPadded, 0.f);
    for (int i = 0; i < kMlpParams; ++i) {
        if (!std::isfinite(params[i])) return mlp_fail(n, NB_E_INVALID_ARG, "non-finite parameter");
        pad[flat_to_padded(i)] = params[i];
    }
    MLP_HIP(n, hipMemcpyAsync(n->master, pad.data(), kMlpPadded * sizeof(float),
                              hipMemcpyHostToDevice, n->stream));
    MLP_HIP(n, hipMemsetAsync(n->adam_m, 0, kMlpPadded * sizeof(float), n->stream));
    MLP_HIP(n, hipMemsetAsync(n->adam_v, 0, kMlpPadded * sizeof(float), n->stream));
    int rc = refresh_pack(n);
    if (rc) return rc;
    MLP_HIP(n, hipStreamSynchronize(n->stream));
    return NB_OK;
}

int nb_mlp_get_params(nb_mlp *n, float *params) {
    if (!n || !params) return NB_E_INVALID_ARG;
    std::vector<float> pad(kMlpPadded);
    MLP_HIP(n, hipMemcpyAsync(pad.data(), n->master, kMlpPadded * sizeof(float),
                              hipMemcpyDeviceToHost, n->stream));
    MLP_HIP(n, hipStreamSynchronize(n->stream));
    for (int i = 0; i < kMlpParams; ++i) params[i] = pad[flat_to_padded(i)];
    return NB_OK;
}

int nb_mlp_forward(nb_mlp *n, uint64_t count, const float *features, float *out) {
    if (!n || (count && (!features || !out))) return NB_E_INVALID_ARG;
    if (count == 0) return NB_OK;
    float *F = nullptr, *Y = nullptr;
    MLP_HIP(n, hipMalloc((void **)&F, count * kMlpIn * sizeof(float)));
    if (hipMalloc((void **)&Y, count * kMlpOut * sizeof(float)) != hipSuccess) {
        (void)hipFree(F);
        return mlp_fail(n, NB_E_HIP, "hipMalloc");
    }
    int rc = NB_OK;
    if (hipMemcpyAsync(F, features, count * kMlpIn * sizeof(float), hipMemcpyHostToDevice,
                       n->stream) != hipSuccess)
        rc = NB_E_HIP;
    if (!rc) {
        const uint32_t tiles = (uint32_t)std::min<uint64_t>((count + kMlpTile - 1) / kMlpTile, 16384);
        hipLaunchKernelGGL(k_mlp_infer, dim3(tiles), dim3(64), 0, n->stream, count,
                           (const float *)F, Y, (const uint4 *)n->pack, (const float *)n->master);
        if (hipMemcpyAsync(out, Y, count * kMlpOut * sizeof(float), hipMemcpyDeviceToHost,
                           n->stream) != hipSuccess ||
            hipStreamSynchronize(n->stream) != hipSuccess)
            rc = NB_E_HIP;
    }
    (void)hipFree(F);
    (void)hipFree(Y);
    return rc ? mlp_fail(n, rc, "nb_mlp_forward: HIP error") : NB_OK;
}

int nb_mlp_grad(nb_mlp *n, uint32_t count, const float *features, const float *targets,
                float *grad, float *loss) {
    if (!n || count == 0 || !features || !targets || !grad) return NB_E_INVALID_ARG;
    Batch bt;
    float *F = nullptr, *T = nullptr;
    if (!bt.init(count, 1) || !bt.alloc(&F, (uint64_t)count * kMlpIn) ||
        !bt.alloc(&T, (uint64_t)count * kMlpOut))
        return mlp_fail(n, NB_E_HIP, "hipMalloc (gradient scratch)");
    MLP_HIP(n, hipMemcpyAsync(F, features, (uint64_t)count * kMlpIn * sizeof(float),
                              hipMemcpyHostToDevice, n->stream));
    MLP_HIP(n, hipMemcpyAsync(T, targets, (uint64_t)count * kMlpOut * sizeof(float),
                              hipMemcpyHostToDevice, n->stream));
    // identity order: pa = 1, pc = 0
    enqueue_step(n, bt, count, F, T, count, 1, 0, 0, 0, bt.loss, false, 0, 0, 0, 0, 1, 1);
    MLP_HIP(n, hipGetLastError());
    std::vector<float> pad(kMlpPadded);
    float L = 0.f;
    MLP_HIP(n, hipMemcpyAsync(pad.data(), bt.grad, kMlpPadded * sizeof(float),
                              hipMemcpyDeviceToHost, n->stream));
    MLP_HIP(n, hipMemcpyAsync(&L, bt.loss, sizeof(float), hipMemcpyDeviceToHost, n->stream));
    MLP_HIP(n, hipStreamSynchronize(n->stream));
    for (int i = 0; i < kMlpParams; ++i) grad[i] = pad[flat_to_padded(i)];
    if (loss) *loss = L;
    return NB_OK;
}

int nb_mlp_train(nb_mlp *n, uint64_t count, const float *features, const float *targets,
                 const nb_mlp_train_opts *opts, float *epoch_loss) {
    if (!n || count == 0 || count >= (1ull << 32) || !features || !targets || !opts ||
        opts->epochs == 0 || opts->batch == 0)
        return NB_E_INVALID_ARG;
    const uint32_t batch = (uint32_t)std::min<uint64_t>(opts->batch, count);
    const uint64_t per_epoch = count / batch;
    const uint64_t steps = per_epoch * opts->epochs;
    if (steps > (1u << 24)) return mlp_fail(n, NB_E_INVALID_ARG, "too many training steps");
    Batch bt;
    float *F = nullptr, *T = nullptr;
    if (!bt.init(batch, (uint32_t)steps) || !bt.alloc(&F, count * kMlpIn) ||
        !bt.alloc(&T, count * kMlpOut))
        return mlp_fail(n, NB_E_HIP, "hipMalloc (training scratch)");
    MLP_HIP(n, hipMemcpyAsync(F, features, count * kMlpIn * sizeof(float), hipMemcpyHostToDevice,
                              n->stream));
    MLP_HIP(n, hipMemcpyAsync(T, targets, count * kMlpOut * sizeof(float), hipMemcpyHostToDevice,
                              n->stream));
    MLP_HIP(n, hipMemsetAsync(n->adam_m, 0, kMlpPadded * sizeof(float), n->stream));
    MLP_HIP(n, hipMemsetAsync(n->adam_v, 0, kMlpPadded * sizeof(float), n->stream));
    uint64_t rs = opts->seed;
    uint64_t step = 0;
    for (uint32_t ep = 0; ep < opts->epochs; ++ep) {
        // a fresh affine permutation of the data set per epoch
        uint64_t pa = 1, pc = 0;
        if (count > 1) {
            do {
                pa = 1 + splitmix(rs) % (count - 1);
            } while (std::gcd(pa, count) != 1);
            pc = splitmix(rs) % count;
        }
        const uint64_t aug_key = opts->augment ? (splitmix(rs) | 1ull) : 0ull;
        for (uint64_t b = 0; b < per_epoch; ++b, ++step) {
            // cosine decay from lr to lr_final over all steps
            const double frac = steps > 1 ? (double)step / (double)(steps - 1) : 0.0;
            const float lr = (float)(opts->lr_final + 0.5 * (opts->lr - opts->lr_final) *
                                                          (1 + std::cos(M_PI * frac)));
            const float c1 = (float)(1 - std::pow((double)opts->beta1, (double)(step + 1)));
            const float c2 = (float)(1 - std::pow((double)opts->beta2, (double)(step + 1)));
            enqueue_step(n, bt, batch, F, T, count, pa, pc, b * batch, aug_key, bt.loss + step,
                         true, lr, opts->beta1, opts->beta2, opts->eps, c1, c2);
        }
        MLP_HIP(n, hipGetLastError());
    }
    std::vector<float> losses(steps);
    MLP_HIP(n, hipMemcpyAsync(losses.data(), bt.loss, steps * sizeof(float), hipMemcpyDeviceToHost,
                              n->stream));
    MLP_HIP(n, hipStreamSynchronize(n->stream));
    if (epoch_loss) {
        for (uint32_t ep = 0; ep < opts->epochs; ++ep) {
            double s = 0;
            for (uint64_t b = 0; b < per_epoch; ++b) s += losses[ep * per_epoch + b];
            epoch_loss[ep] = (float)(s / (double)per_epoch);
        }
    }
    for (float L : losses)
        if (!std::isfinite(L)) return mlp_fail(n, NB_E_INVALID_ARG, "training diverged (non-finite loss)");
    return NB_OK;
}

int nb_mlp_adam_step(nb_mlp *n, const float *grad, float lr, uint32_t t) {
    // One Adam update from a given gradient (beta 0.9 / 0.999, eps 1e-8); used
    // to check the optimiser against the CPU restatement.
    if (!n || !grad || t == 0) return NB_E_INVALID_ARG;
    Batch bt;
    if (!bt.alloc(&bt.partial, kMlpPadded) || !bt.alloc(&bt.lossp, 1) || !bt.alloc(&bt.loss, 1))
        return mlp_fail(n, NB_E_HIP, "hipMalloc");
    std::vector<float> pad(kMlpPadded, 0.f);
    for (int i = 0; i < kMlpParams; ++i) pad[flat_to_padded(i)] = grad[i];
    MLP_HIP(n, hipMemcpyAsync(bt.partial, pad.data(), kMlpPadded * sizeof(float),
                              hipMemcpyHostToDevice, n->stream));
    MLP_HIP(n, hipMemsetAsync(bt.lossp, 0, sizeof(float), n->stream));
    const float c1 = (float)(1 - std::pow(0.9, (double)t)), c2 = (float)(1 - std::pow(0.999, (double)t));
    hipLaunchKernelGGL(k_mlp_reduce, dim3(ceil_div(kMlpPadded, kThreads)), dim3(kThreads), 0,
                       n->stream, 1u, (const float *)bt.partial, nullptr, (const float *)bt.lossp,
                       1u, 1.0f, bt.loss, 1, n->master, n->adam_m, n->adam_v, lr, 0.9f, 0.999f,
                       1e-8f, c1, c2);
    int rc = refresh_pack(n);
    if (rc) return rc;
    MLP_HIP(n, hipStreamSynchronize(n->stream));
    return NB_OK;
}

}  // extern "C"

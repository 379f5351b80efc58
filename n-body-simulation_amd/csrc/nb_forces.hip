// nb_forces.hip -- direct-sum and Barnes-Hut force kernels for gfx950.
//
// FP64 kernels reproduce the reference's arithmetic operation for operation
// (this file is compiled with -ffp-contract=off, and sycl::rsqrt is evaluated
// as 1.0/sqrt(x), two correctly rounded IEEE operations):
//   k_direct_f64   NaiveAlgorithm.cpp:299-352 (sum over j in index order)
//   k_bh_exact     BarnesHutAlgorithm.cpp:319-393 (per-body stack traversal,
//                  same push order, same opening test)
// FP32 kernels do the pair arithmetic in FP32 (fmaf, v_rsq_f32) and fold the
// FP32 partial of every tile into FP64 accumulators:
//   k_direct_f32c  LDS-staged source tiles, two targets per lane, split j
//   k_bh_group     one wavefront per group (an octree cell of <= 64 bodies);
//                  a cell is opened if any body of the group needs it; lanes
//                  test 8 sibling sets at a time and accepted cells/bodies are
//                  batched in LDS and evaluated 64 at a time.
#include <algorithm>
#include <cmath>

#include "nb_mlp.h"

namespace nb {
namespace {

__device__ inline void raise_flag(TreeHdr *h, uint32_t f) { atomicOr(&h->status, f); }

constexpr uint32_t kXcds = 8;  // XCDs of an MI355X; workgroups are dealt round-robin

// ---------------------------------------------------------------------------
// FP64 direct sum, reference order.  The sum over j = 0..n-1 of a target must
// run in that order to be bit-exact, but only its three additions per pair
// are sequential; the pair terms m_j (r_ij den_ij), with their IEEE division
// and square root, are not.  A workgroup owns 64 targets (one per lane).
// Sources are staged in LDS kF64Stage at a time; for each chunk of kF64Chunk
// sources all sixteen waves compute the terms into one of two LDS tiles while
// waves 0, 1 and 2 add the x, y and z components of the previous tile in
// source order, eight terms' reads in flight ahead of their additions.
// (16,384 bodies: 1.26 ms with 8 waves and one read per addition by wave 0,
// 1.09 ms with 16 waves, 0.75 ms with the reads ahead, 0.72 ms with
// 48-source chunks instead of 32, 0.66 ms with inv_sqrt_ieee, 0.58 ms with
// the three components added by three waves instead of one.)
constexpr int kDirectF64Tile = 64;                  // targets per workgroup
constexpr int kF64Threads = 1024;                   // 16 waves computing pair terms
constexpr int kF64Chunk = 48;                       // sources per term tile
constexpr int kF64PerWave = kF64Chunk / (kF64Threads / 64);
constexpr int kF64Stage = 240;                      // sources per LDS stage
static_assert(kF64Stage % kF64Chunk == 0, "whole chunks per stage");
static_assert(kF64Threads >= 3 * 64 && kF64Chunk % (kF64Threads / 64) == 0,
              "three adder waves; whole terms per wave");

// Returns, in waves 0, 1 and 2, component 0, 1 or 2 (x, y, z) of the sum
// over j in order of m_j (r_ij den_ij) for the lane's target at (px, py, pz):
// the three sums are independent, so three waves on three SIMDs add them.
__device__ inline double direct_f64_body(uint32_t n, const double *__restrict__ x,
                                         const double *__restrict__ y,
                                         const double *__restrict__ z,
                                         const double *__restrict__ m, double eps2, double px,
                                         double py, double pz) {
    __shared__ double4 src[kF64Stage];
    __shared__ double tx[2][kF64Chunk][kDirectF64Tile], ty[2][kF64Chunk][kDirectF64Tile],
        tz[2][kF64Chunk][kDirectF64Tile];
    const uint32_t lane = threadIdx.x & 63u, w = threadIdx.x >> 6;
    double (*const own)[kF64Chunk][kDirectF64Tile] = w == 0 ? tx : (w == 1 ? ty : tz);
    double acc = 0;
    uint32_t pending = 0;  // sources in the tile waves 0-2 still have to add
    int buf = 0;
    // waves 0-2: add their component of tile `b` (cnt sources) in source order
    auto add_tile = [&](int b, uint32_t cnt) {
        uint32_t jj = 0;
        // eight terms' LDS reads in flight before their (ordered) additions
        for (; jj + 8 <= cnt; jj += 8) {
            double v[8];
#pragma unroll
            for (int u = 0; u < 8; ++u) v[u] = own[b][jj + u][lane];
#pragma unroll
            for (int u = 0; u < 8; ++u) acc += v[u];
        }
        for (; jj < cnt; ++jj) acc += own[b][jj][lane];
    };
    for (uint32_t k0 = 0; k0 < n; k0 += kF64Stage) {
        const uint32_t staged = min((uint32_t)kF64Stage, n - k0);
        __syncthreads();  // the previous stage's terms are computed
        for (uint32_t j = threadIdx.x; j < staged; j += kF64Threads)
            src[j] = make_double4(x[k0 + j], y[k0 + j], z[k0 + j], m[k0 + j]);
        __syncthreads();
        for (uint32_t c = 0; c < staged; c += kF64Chunk) {
            const uint32_t cnt = min((uint32_t)kF64Chunk, staged - c);
#pragma unroll
            for (int u = 0; u < kF64PerWave; ++u) {
                const uint32_t jj = w * kF64PerWave + u;  // the same for the whole wave
                if (jj < cnt) {
                    const double4 sj = src[c + jj];
                    const double r_x = sj.x - px;
                    const double r_y = sj.y - py;
                    const double r_z = sj.z - pz;
                    double denominator = (r_x * r_x) + (r_y * r_y) + (r_z * r_z) + eps2;
                    denominator = denominator * denominator * denominator;
                    denominator = inv_sqrt_ieee(denominator);
                    tx[buf][jj][lane] = sj.w * (r_x * denominator);
                    ty[buf][jj][lane] = sj.w * (r_y * denominator);
                    tz[buf][jj][lane] = sj.w * (r_z * denominator);
                }
            }
            // waves 0-2 add the previous tile (written before the last barrier)
            if (w < 3 && pending) add_tile(buf ^ 1, pending);
            __syncthreads();
            pending = cnt;
            buf ^= 1;
        }
    }
    if (w < 3 && pending) add_tile(buf ^ 1, pending);
    return acc;
}

__global__ void __launch_bounds__(kF64Threads) k_direct_f64(
    uint32_t n, const double *__restrict__ x, const double *__restrict__ y,
    const double *__restrict__ z, const double *__restrict__ m, double G, double eps2,
    double *ax, double *ay, double *az) {
    const uint32_t i = blockIdx.x * kDirectF64Tile + (threadIdx.x & 63u);
    const bool valid = i < n;
    const uint32_t ii = valid ? i : 0;
    const double acc = direct_f64_body(n, x, y, z, m, eps2, x[ii], y[ii], z[ii]);
    const uint32_t w = threadIdx.x >> 6;
    if (valid && w < 3) (w == 0 ? ax : (w == 1 ? ay : az))[i] = acc * G;
}

__global__ void __launch_bounds__(kF64Threads) k_direct_targets_f64(
    uint32_t n, const double *__restrict__ x, const double *__restrict__ y,
    const double *__restrict__ z, const double *__restrict__ m, double G, double eps2,
    uint32_t nt, const uint32_t *__restrict__ targets, double *ax, double *ay, double *az) {
    const uint32_t t = blockIdx.x * kDirectF64Tile + (threadIdx.x & 63u);
    const bool valid = t < nt;
    const uint32_t i = valid ? targets[t] : 0;
    const double acc = direct_f64_body(n, x, y, z, m, eps2, x[i], y[i], z[i]);
    const uint32_t w = threadIdx.x >> 6;
    if (valid && w < 3) (w == 0 ? ax : (w == 1 ? ay : az))[t] = acc * G;
}

// ---------------------------------------------------------------------------
// Split direct sums: FP32 pairs with FP64 tile accumulation (k_direct_f32c),
// and FP64 pairs with a corrected v_rsq_f64 (k_direct_f64f).
//
// grid.x: tiles of kSplitTargets targets (kSplitPerLane per lane); grid.y:
// source splits, each a contiguous source range.  Sources are staged in LDS
// kSplitTile at a time, converted from the FP64 state arrays on the way (no
// separate packing pass).  With one split a block writes its targets' sums;
// with more, each block stores its partial sums and k_split_reduce adds them
// in split order 0, 1, ..., so results do not depend on block timing.  (The
// partials can also be added by the last block of each tile to arrive, but
// the device-scope fence each block then needs writes back the L2: the
// 16,384-body FP32 sum took 232 us that way against 98 us.)
constexpr int kSplitThreads = 256;
constexpr int kSplitPerLane = 2;  // targets per lane (each LDS read feeds 2 pairs)
constexpr uint32_t kSplitTargets = kSplitThreads * kSplitPerLane;
constexpr int kSplitTile = 256;   // sources per LDS stage
constexpr uint32_t kSplitBlocks = 2048;  // blocks a launch aims for (2 per SIMD)
// sources per interleaved FP64 step (x kSplitPerLane targets): 124 VGPRs, four
// waves per SIMD; 2 (72 VGPRs, seven waves) and 8 (240) measured within 2 %,
// five waves forced (spilling) 15 % slower
constexpr int kF64fSources = 4;
static_assert(kSplitTile % kF64fSources == 0, "whole steps per stage");

__device__ inline void pair_f32(const float4 s, float px, float py, float pz, float eps2,
                                float &fx, float &fy, float &fz) {
    const float dx = s.x - px, dy = s.y - py, dz = s.z - pz;
    const float r2 = fmaf(dx, dx, fmaf(dy, dy, fmaf(dz, dz, eps2)));
    const float inv = __builtin_amdgcn_rsqf(r2);
    const float f = (s.w * inv) * (inv * inv);
    fx = fmaf(f, dx, fx);
    fy = fmaf(f, dy, fy);
    fz = fmaf(f, dz, fz);
}

// v^(-1/2) in FP64 from the hardware estimate y = v_rsq_f64(v), whose
// relative error is near 1e-8 (single-precision accuracy): with
// e = 1 - v y^2 (one fused multiply-add, so e is exact to the last bits),
// v^(-1/2) = y (1 - e)^(-1/2) = y (1 + e/2 + 3e^2/8 + 5e^3/16 + ...).  The
// correction to second order in e leaves a truncation error near
// 5e^3/16 ~ 1e-22, below FP64 rounding; five operations (a Newton step,
// first order in e, would leave ~1e-14).  Checked against long-double
// closed forms in the GPU tests.
__device__ inline double rsqrt_f64(double v) {
    const double y = __builtin_amdgcn_rsq(v);
    const double e = fma(-(v * y), y, 1.0);
    return fma(y * e, fma(e, 0.375, 0.5), y);
}

__device__ inline void pair_f64f(const double4 s, double px, double py, double pz, double eps2,
                                 double &fx, double &fy, double &fz) {
    const double dx = s.x - px, dy = s.y - py, dz = s.z - pz;
    const double r2 = fma(dx, dx, fma(dy, dy, fma(dz, dz, eps2)));
    const double inv = rsqrt_f64(r2);
    const double f = (s.w * inv) * (inv * inv);
    fx = fma(f, dx, fx);
    fy = fma(f, dy, fy);
    fz = fma(f, dz, fz);
}

__device__ inline uint32_t split_target(uint32_t q) {
    return blockIdx.x * kSplitTargets + q * kSplitThreads + threadIdx.x;
}

// End of a split direct sum for this lane's targets: with one split the
// block writes scale * sums; with more it stores its partial sums at
// partial[split][3][n] for k_split_reduce.
__device__ inline void split_finish(uint32_t n, double scale, const double (&sx)[kSplitPerLane],
                                    const double (&sy)[kSplitPerLane],
                                    const double (&sz)[kSplitPerLane], double *partial,
                                    double *out_x, double *out_y, double *out_z) {
    const bool direct = gridDim.y == 1;
    const uint64_t base = (uint64_t)blockIdx.y * 3ull * n;
#pragma unroll
    for (int q = 0; q < kSplitPerLane; ++q) {
        const uint32_t i = split_target(q);
        if (i >= n) continue;
        if (direct) {
            out_x[i] = sx[q] * scale;
            out_y[i] = sy[q] * scale;
            out_z[i] = sz[q] * scale;
        } else {
            partial[base + i] = sx[q];
            partial[base + n + i] = sy[q];
            partial[base + 2ull * n + i] = sz[q];
        }
    }
}

// out = scale * (sum over splits 0, 1, ... in that order) of the partial sums.
__global__ void __launch_bounds__(kSplitThreads) k_split_reduce(uint32_t n, uint32_t splits,
                                                                double scale,
                                                                const double *__restrict__ partial,
                                                                double *out_x, double *out_y,
                                                                double *out_z) {
    const uint32_t i = blockIdx.x * kSplitThreads + threadIdx.x;
    if (i >= n) return;
    double ox = 0, oy = 0, oz = 0;
    // eight splits' loads in flight before their (ordered) additions: one
    // wave per CU at 16,384 bodies, so the loads' latency is not hidden by
    // other waves (18.8 us with one split at a time)
    constexpr uint32_t kAhead = 8;
    uint32_t sp = 0;
    for (; sp + kAhead <= splits; sp += kAhead) {
        double vx[kAhead], vy[kAhead], vz[kAhead];
#pragma unroll
        for (uint32_t u = 0; u < kAhead; ++u) {
            const uint64_t b = (uint64_t)(sp + u) * 3ull * n;
            vx[u] = partial[b + i];
            vy[u] = partial[b + n + i];
            vz[u] = partial[b + 2ull * n + i];
        }
#pragma unroll
        for (uint32_t u = 0; u < kAhead; ++u) {
            ox += vx[u];
            oy += vy[u];
            oz += vz[u];
        }
    }
    for (; sp < splits; ++sp) {
        const uint64_t b = (uint64_t)sp * 3ull * n;
        ox += partial[b + i];
        oy += partial[b + n + i];
        oz += partial[b + 2ull * n + i];
    }
    out_x[i] = ox * scale;
    out_y[i] = oy * scale;
    out_z[i] = oz * scale;
}

// FP32 pairs on absolute FP32 positions with G*m as the source weight; each
// stage's FP32 sums are folded into FP64.
__global__ void __launch_bounds__(kSplitThreads) k_direct_f32c(
    uint32_t n, const double *__restrict__ x, const double *__restrict__ y,
    const double *__restrict__ z, const double *__restrict__ m, double G, float eps2,
    uint32_t chunk, double *partial, double *ax_out, double *ay_out, double *az_out) {
    __shared__ float4 tile[kSplitTile];
    float px[kSplitPerLane], py[kSplitPerLane], pz[kSplitPerLane];
    double ax[kSplitPerLane], ay[kSplitPerLane], az[kSplitPerLane];
#pragma unroll
    for (int q = 0; q < kSplitPerLane; ++q) {
        const uint32_t i = split_target(q);
        const uint32_t ii = i < n ? i : n - 1;
        px[q] = (float)x[ii];
        py[q] = (float)y[ii];
        pz[q] = (float)z[ii];
        ax[q] = ay[q] = az[q] = 0.0;
    }
    const uint32_t j0 = blockIdx.y * chunk;
    const uint32_t j1 = min(n, j0 + chunk);
    for (uint32_t k = j0; k < j1; k += kSplitTile) {
        const uint32_t j = k + threadIdx.x;
        if (j < j1)
            tile[threadIdx.x] =
                make_float4((float)x[j], (float)y[j], (float)z[j], (float)(G * m[j]));
        __syncthreads();
        float fx[kSplitPerLane], fy[kSplitPerLane], fz[kSplitPerLane];
#pragma unroll
        for (int q = 0; q < kSplitPerLane; ++q) fx[q] = fy[q] = fz[q] = 0.f;
        auto pairs = [&](int jj) {
            const float4 sj = tile[jj];
#pragma unroll
            for (int q = 0; q < kSplitPerLane; ++q)
                pair_f32(sj, px[q], py[q], pz[q], eps2, fx[q], fy[q], fz[q]);
        };
        const uint32_t cnt = min((uint32_t)kSplitTile, j1 - k);  // uniform
        if (cnt == (uint32_t)kSplitTile) {
#pragma unroll 8
            for (int jj = 0; jj < kSplitTile; ++jj) pairs(jj);
        } else {
            for (uint32_t jj = 0; jj < cnt; ++jj) pairs((int)jj);
        }
#pragma unroll
        for (int q = 0; q < kSplitPerLane; ++q) {
            ax[q] += (double)fx[q];
            ay[q] += (double)fy[q];
            az[q] += (double)fz[q];
        }
        __syncthreads();
    }
    split_finish(n, 1.0, ax, ay, az, partial, ax_out, ay_out, az_out);
}

// FP64 pairs (the reference's formula, NaiveAlgorithm.cpp:318-351, with the
// inverse square root from rsqrt_f64 and fused multiply-adds); G is applied
// to the sum, as in the reference (:348-351).
__global__ void __launch_bounds__(kSplitThreads) k_direct_f64f(
    uint32_t n, const double *__restrict__ x, const double *__restrict__ y,
    const double *__restrict__ z, const double *__restrict__ m, double G, double eps2,
    uint32_t chunk, double *partial, double *ax_out, double *ay_out, double *az_out) {
    __shared__ double4 tile[kSplitTile];
    double px[kSplitPerLane], py[kSplitPerLane], pz[kSplitPerLane];
    double ax[kSplitPerLane], ay[kSplitPerLane], az[kSplitPerLane];
#pragma unroll
    for (int q = 0; q < kSplitPerLane; ++q) {
        const uint32_t i = split_target(q);
        const uint32_t ii = i < n ? i : n - 1;
        px[q] = x[ii];
        py[q] = y[ii];
        pz[q] = z[ii];
        ax[q] = ay[q] = az[q] = 0.0;
    }
    const uint32_t j0 = blockIdx.y * chunk;
    const uint32_t j1 = min(n, j0 + chunk);
    for (uint32_t k = j0; k < j1; k += kSplitTile) {
        const uint32_t j = k + threadIdx.x;
        if (j < j1) tile[threadIdx.x] = make_double4(x[j], y[j], z[j], m[j]);
        __syncthreads();
        auto pairs = [&](int jj) {
            const double4 sj = tile[jj];
#pragma unroll
            for (int q = 0; q < kSplitPerLane; ++q)
                pair_f64f(sj, px[q], py[q], pz[q], eps2, ax[q], ay[q], az[q]);
        };
        // kF64fSources sources x kSplitPerLane targets, stage by stage across the
        // pairs: the compiler otherwise runs each pair's chain (the estimate,
        // its five-step correction, the force) to the end before the next
        // pair starts.  The same operations per pair as pair_f64f, and each
        // target's sum still runs over the sources in order: bit-identical.
        auto pairs_s = [&](int jj) {
            constexpr int P = kF64fSources * kSplitPerLane;
            double dx[P], dy[P], dz[P], r2[P], yv[P], w[P];
#pragma unroll
            for (int u = 0; u < kF64fSources; ++u) {
                const double4 sj = tile[jj + u];
#pragma unroll
                for (int q = 0; q < kSplitPerLane; ++q) {
                    const int p = u * kSplitPerLane + q;
                    dx[p] = sj.x - px[q];
                    dy[p] = sj.y - py[q];
                    dz[p] = sj.z - pz[q];
                    w[p] = sj.w;
                    r2[p] = fma(dx[p], dx[p], fma(dy[p], dy[p], fma(dz[p], dz[p], eps2)));
                }
            }
#pragma unroll
            for (int p = 0; p < P; ++p) yv[p] = rsqrt_f64(r2[p]);
#pragma unroll
            for (int u = 0; u < kF64fSources; ++u)
#pragma unroll
                for (int q = 0; q < kSplitPerLane; ++q) {
                    const int p = u * kSplitPerLane + q;
                    const double f = (w[p] * yv[p]) * (yv[p] * yv[p]);
                    ax[q] = fma(f, dx[p], ax[q]);
                    ay[q] = fma(f, dy[p], ay[q]);
                    az[q] = fma(f, dz[p], az[q]);
                }
        };
        // the split's last stage may be partial: only its staged sources
        const uint32_t cnt = min((uint32_t)kSplitTile, j1 - k);  // uniform
        if (cnt == (uint32_t)kSplitTile) {
            for (int jj = 0; jj < kSplitTile; jj += kF64fSources) pairs_s(jj);
        } else {
            for (uint32_t jj = 0; jj < cnt; ++jj) pairs((int)jj);
        }
        __syncthreads();
    }
    split_finish(n, G, ax, ay, az, partial, ax_out, ay_out, az_out);
}

// ---------------------------------------------------------------------------
// Reference-semantics Barnes-Hut traversal (FP64).  Per body a depth-first
// walk with its own stack (the older part in global memory, interleaved
// across bodies so that lanes at the same stack depth touch adjacent words);
// children are pushed in digit order 5,7,4,6,1,3,0,2 as in
// BarnesHutAlgorithm.cpp:370-385, and the last pushed is visited at once.
// The stack is kept as batches, one per opened cell (below).  The visit
// order, and so every sum, is the reference's.  1M-body disk at theta 0.5:
// 27.8 ms with one stack entry per child in memory; 16.9 ms with the top six
// batches as seven registers each; 15.2 ms with four such batches, the
// opening test decided from an estimate and a node's three loads issued
// together; with a batch as one word, 14.4, 13.3, 12.5, 12.5 and 12.7 ms
// with 2, 4, 8, 12 and 16 batches in registers.  Loading the record of the
// stack top one visit ahead was slower: 13.8 ms (85 VGPRs, five waves).
// Workgroups dealt to the XCDs in contiguous runs of bodies: 10.7 ms.
constexpr int kExactThreads = 256;

__global__ void __launch_bounds__(kExactThreads) k_bh_exact(
    uint32_t n, const uint32_t *__restrict__ order, const double *__restrict__ x,
    const double *__restrict__ y, const double *__restrict__ z, const double4 *__restrict__ cm,
    const double *__restrict__ edge, const int32_t *__restrict__ ref, double G, double eps2,
    double theta, uint32_t *stack, uint32_t cap, uint32_t icap, double *ax, double *ay,
    double *az, uint32_t *inter, TreeHdr *h) {
    // Workgroup b runs on XCD b mod 8: give each XCD a contiguous run of the
    // octree order, so that its L2 holds the records of one region of the tree.
    const uint32_t nblk = gridDim.x, xcd = blockIdx.x % kXcds;
    const uint32_t per = nblk / kXcds, extra = nblk % kXcds;
    const uint32_t lb = xcd * per + min(xcd, extra) + blockIdx.x / kXcds;
    const uint32_t id = lb * kExactThreads + threadIdx.x;
    if (id >= n) return;
    const uint32_t i = order ? order[id] : id;
    const int32_t self_ref = ref_of_body(i);
    const double pos_x_i = x[i], pos_y_i = y[i], pos_z_i = z[i];
    uint32_t *stk = stack + id;
    const uint64_t stride = n;
    double acc_x = 0, acc_y = 0, acc_z = 0;
    uint32_t cnt = 0;
    // The opening test edge * (1.0 / sqrt(d2)) < theta is decided from
    // edge * rsqrt_f64(d2), within 1e-14 of it, wherever that estimate is
    // farther than 2^-30 (relative) from theta; only the rest pay the IEEE
    // square root and division.  The decisions, and so the visit order and
    // every sum, are the reference's.
    const double theta_lo = theta * (1.0 - 0x1p-30), theta_hi = theta * (1.0 + 0x1p-30);
    // The stack holds batches: the unvisited children of one opened cell r,
    // pushed in digit order 5,7,4,6,1,3,0 and popped from the end, as one
    // word (r << 3) | count (cell indices stay below 2^26).  The top kBatches
    // batches live in registers, bw[0] on top, with no empty batch above a
    // full one; older batches go to the memory stack one word each.  Popping
    // a child is a shift of a constant; a memory pop returns a whole batch.
    constexpr int kBatches = 8;  // 66 VGPRs
    constexpr uint32_t kPopDigit = 0x0316475u;  // nibble c-1: the child popped at count c
    uint32_t bw[kBatches];
#pragma unroll
    for (int k = 0; k < kBatches; ++k) bw[k] = 0;
    uint32_t mtop = 0;      // batches in memory
    uint32_t mentries = 0;  // children they hold
    uint32_t node = 0;  // the root first
    uint64_t budget = 1ull + 8ull * (uint64_t)icap;
    for (;;) {
        if (budget-- == 0) {
            raise_flag(h, kFlagStack);
            break;
        }
        const double4 c = cm[node];
        const int32_t r = ref[node];
        const double e = edge[node];  // every slot below mcap has one, read or not
        // all three loads in flight together: left to itself the compiler
        // issues the centre of mass and the edge only after the tests on the
        // mass and the reference, a second memory round trip per visit
        asm volatile("" ::"v"(c.x), "v"(c.y), "v"(c.z), "v"(c.w), "v"(r), "v"(e));
        bool opened = false;
        if (c.w != 0 && r != self_ref) {
            const double d_x = c.x - pos_x_i;
            const double d_y = c.y - pos_y_i;
            const double d_z = c.z - pos_z_i;
            bool accept = r < 0;
            if (!accept) {
                const double d2 = d_x * d_x + d_y * d_y + d_z * d_z;
                const double est = e * rsqrt_f64(d2);
                if (est < theta_lo) {
                    accept = true;
                } else if (!(est > theta_hi)) {  // near theta, or not finite
                    const double d = 1.0 / sqrt(d2);
                    const double currentTheta = e * d;
                    accept = currentTheta < theta;
                }
            }
            if (accept) {
                double denominator = (d_x * d_x) + (d_y * d_y) + (d_z * d_z) + eps2;
                denominator = denominator * denominator * denominator;
                denominator = inv_sqrt_ieee(denominator);
                acc_x += c.w * (d_x * denominator);
                acc_y += c.w * (d_y * denominator);
                acc_z += c.w * (d_z * denominator);
                ++cnt;
            } else {
                uint32_t held = 0;
#pragma unroll
                for (int k = 0; k < kBatches; ++k) held += bw[k] & 7u;
                if (mentries + held + 8 > cap) {
                    raise_flag(h, kFlagStack);
                    break;
                }
                if (bw[0]) {
                    // the bottom batch goes to memory, the others down one
                    if (bw[kBatches - 1]) {
                        stk[(uint64_t)(mtop++) * stride] = bw[kBatches - 1];
                        mentries += bw[kBatches - 1] & 7u;
                    }
#pragma unroll
                    for (int k = kBatches - 1; k > 0; --k) bw[k] = bw[k - 1];
                }
                bw[0] = ((uint32_t)r << 3) | 7u;
                node = 8u * (uint32_t)r + 3u;  // child 2, pushed last
                opened = true;
            }
        }
        if (!opened) {
            if (bw[0] == 0) {  // every register batch is empty
                if (mtop == 0) break;
                bw[0] = stk[(uint64_t)(--mtop) * stride];
                mentries -= bw[0] & 7u;
            }
            const uint32_t c = bw[0] & 7u;
            node = (bw[0] & ~7u) + 1u + ((kPopDigit >> (4u * (c - 1u))) & 7u);
            if (c == 1) {  // batch 0 is done: the others up one
#pragma unroll
                for (int k = 0; k < kBatches - 1; ++k) bw[k] = bw[k + 1];
                bw[kBatches - 1] = 0;
            } else {
                bw[0] -= 1u;
            }
        }
    }
    ax[i] = acc_x * G;
    ay[i] = acc_y * G;
    az[i] = acc_z * G;
    if (inter) inter[i] = cnt;
}

// ---------------------------------------------------------------------------
// Group traversal (FP32 pair maths, FP64 accumulation).
//
// One wavefront walks the octree for one group: an octree cell with at most
// 64 bodies (a contiguous run of the octree order).  A cell is accepted when
// its centre of mass is farther than edge/theta from the group's bounding box
// -- so from every body of the group: the reference's opening test
// (BarnesHutAlgorithm.cpp:355-359) taken over the group; bodies in leaves are
// always taken exactly.  Each iteration pops up to 8 cells whose children
// must be examined, and lane l examines child (l & 7) of cell l >> 3: one
// 32-byte record load per lane.  Accepted cells and bodies are appended to an
// LDS list; every 128 entries the list is evaluated against the group's bodies.
// A group of s bodies uses lpb = floor(64/s) lanes per body, each taking
// every lpb-th list entry, so s * lpb of the 64 lanes do useful pairs (51 for
// s = 17, where rounding s up to a power of two would leave 34).
// Positions are relative to the group centre; the records' hi/lo split keeps
// the FP32 differences exact to the last FP32 bit.  Waves take groups from an
// atomic work queue.
constexpr int kGroup = 64;          // lanes per wavefront (a group has up to 2 x kGroup bodies)
static_assert(kGroupMax <= 2 * kGroup, "a lane holds at most two bodies of its group");
static_assert(kGroupLimit <= kGroup, "the tunable groups fit one body per lane");
constexpr int kStackCap = 1024;     // cells waiting to be opened (entry = k << 6 | depth)
// Child records per lane and walk step (8 kPop cells).  More loads in flight
// per wave cost registers: on a 1M-body disk kPop 2 and 4 made the traversal
// 7 % and 17 % slower than 1 (occupancy 6 -> 4 waves per SIMD).
constexpr int kPop = 1;
// List entries evaluated per flush, at most.  Each flush folds its FP32 sums
// into the FP64 accumulators and moves the list's tail to the front: 128
// entries halve those costs against 64.
constexpr int kFlushSpan = 2 * kGroup;
constexpr int kListCap = kFlushSpan + kGroup * kPop;
constexpr int kGroupBlocks = 8192;  // persistent waves
constexpr uint32_t kQueueStride = 32;  // one 128-byte line per group-queue counter
static_assert(kXcds * kQueueStride <= TreeBufs::kGroupQueueWords, "queue counters fit TreeBufs::gq");

// Every traversal starts with fresh queue cursors (a kernel, not a memset,
// so the sequence can be captured in a graph).

// The device-scope maximum with a read first: the read-modify-write (which
// serialises at one memory channel when every wave does it) only when it
// raises the value.
__device__ inline void raise_max(uint32_t *p, uint32_t v) {
    if (v > __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) atomicMax(p, v);
}

// Set bits of a wave-uniform mask in the lanes below this one (v_mbcnt: two
// VALU instructions, where a per-lane mask and popcount take four).
__device__ inline uint32_t count_below(uint64_t mask) {
    return __builtin_amdgcn_mbcnt_hi((uint32_t)(mask >> 32),
                                     __builtin_amdgcn_mbcnt_lo((uint32_t)mask, 0u));
}

// A value every lane of the wave holds alike, made visibly uniform (SGPR).
// The walks below are wave-collective (ballots, shuffles, one LDS list per
// wave): every branch that steers them must be uniform to the compiler, or
// control-flow structurisation may let parts of the wave run different
// iterations with partial exec masks.
__device__ inline uint32_t wave_uniform(uint32_t v) { return __builtin_amdgcn_readfirstlane(v); }

// lane 0's value of v, uniform
__device__ inline uint32_t lane0_uniform(uint32_t v) { return wave_uniform(__shfl(v, 0, 64)); }

// With the learned correction (kCorr), accepted cells that are close enough
// (e^2 > tau2 * d_group^2) are also queued in a second LDS list; every 16 such
// cells, the wave evaluates the network for the 16 cells x each body of the
// group (one 16-sample matrix-core tile per body) and adds the corrections of
// a body in a fixed order.
constexpr int kCorrTile = 16;
constexpr int kCellCap = kCorrTile + kGroup * kPop;

// One octree as the group walk sees it (body indices of sidx and of the node
// references index x, y, z, m, ax, ay, az, inter).
struct GroupTree {
    uint32_t n;
    const uint32_t *sidx;
    const double *x, *y, *z, *m;
    const NodeF32 *rec;
    uint32_t icap;
    float ro2_root;  // (root edge / theta)^2
    float e2_root;   // root edge^2
    double *ax, *ay, *az;
    uint32_t *inter;
    const float4 *gquad;  // G*Q per internal cell (quadrupole walks)
};

// The same for a stride known only at run time (lanes per body not a power
// of two).
__device__ __forceinline__ void flush_share_rt(const float4 *p, uint32_t stride, uint32_t t,
                                               float rx, float ry, float rz, float eps2, float &fx,
                                               float &fy, float &fz) {
    uint32_t k = 0;
    for (; k + 4 <= t; k += 4) {
        const float4 e[4] = {p[(k + 0) * stride], p[(k + 1) * stride], p[(k + 2) * stride],
                             p[(k + 3) * stride]};
#pragma unroll
        for (int u = 0; u < 4; ++u) pair_f32(e[u], rx, ry, rz, eps2, fx, fy, fz);
    }
    for (; k < t; ++k) pair_f32(p[k * stride], rx, ry, rz, eps2, fx, fy, fz);
}

// Pair forces of list entries p[0], p[LPB], p[2 LPB], ... (t entries, in that
// order) on the NB points (rx, ry, rz)[b].  LPB and the uniform trip count make
// the LDS offsets immediates and let four reads be in flight ahead of the
// maths; with NB = 2 every entry read feeds two pairs.
template <int LPB, int NB>
__device__ __forceinline__ void flush_share(const float4 *p, uint32_t t, const float (&rx)[NB],
                                            const float (&ry)[NB], const float (&rz)[NB],
                                            float eps2, float (&fx)[NB], float (&fy)[NB],
                                            float (&fz)[NB]) {
    uint32_t k = 0;
    for (; k + 4 <= t; k += 4) {
        const float4 e[4] = {p[(k + 0) * LPB], p[(k + 1) * LPB], p[(k + 2) * LPB],
                             p[(k + 3) * LPB]};
#pragma unroll
        for (int u = 0; u < 4; ++u)
#pragma unroll
            for (int b = 0; b < NB; ++b) pair_f32(e[u], rx[b], ry[b], rz[b], eps2, fx[b], fy[b], fz[b]);
    }
    for (; k < t; ++k) {
        const float4 e = p[k * LPB];
#pragma unroll
        for (int b = 0; b < NB; ++b) pair_f32(e, rx[b], ry[b], rz[b], eps2, fx[b], fy[b], fz[b]);
    }
}

// A list entry that adds exactly zero for any eps^2: its distance squared
// overflows to +inf, so rsqrt gives 0 and the force 0 * offset = 0.
__device__ inline float4 far_entry() { return make_float4(1e30f, 0.f, 0.f, 0.f); }

// A cell's monopole and traceless quadrupole, G folded into e.w = G M and
// Q = G sum m (3 s s^T - |s|^2 I) about the cell's centre of mass, passed as
// qa = {xx, yy, zz, xy}, qb = {xz, yz, -, -}.  With R = com - body (the pair
// offset), the potential -G M / |R| - R.Q.R / (2 |R|^5) gives
//   a = R (G M + 5/2 R.Q.R / R^2) / R^3 - Q.R / R^5,
// with R^2 softened by eps^2 as in the monopole.  32 VALU instructions (the
// monopole pair takes 13); 49 flop by the direct sum's counting.
__device__ inline void pair_quad(const float4 e, const float4 qa, const float4 qb, float px,
                                 float py, float pz, float eps2, float &fx, float &fy, float &fz) {
    const float dx = e.x - px, dy = e.y - py, dz = e.z - pz;
    const float r2 = fmaf(dx, dx, fmaf(dy, dy, fmaf(dz, dz, eps2)));
    const float inv = __builtin_amdgcn_rsqf(r2);
    const float inv2 = inv * inv;
    const float inv3 = inv * inv2;
    const float qx = fmaf(qa.x, dx, fmaf(qa.w, dy, qb.x * dz));
    const float qy = fmaf(qa.w, dx, fmaf(qa.y, dy, qb.y * dz));
    const float qz = fmaf(qb.x, dx, fmaf(qb.y, dy, qa.z * dz));
    const float rqr = fmaf(dx, qx, fmaf(dy, qy, dz * qz));
    const float b = inv3 * inv2;                                // R^-5
    const float a = fmaf((2.5f * rqr) * inv2, b, e.w * inv3);  // GM R^-3 + 5/2 RQR R^-7
    fx = fmaf(a, dx, fx);
    fy = fmaf(a, dy, fy);
    fz = fmaf(a, dz, fz);
    fx = fmaf(-b, qx, fx);
    fy = fmaf(-b, qy, fy);
    fz = fmaf(-b, qz, fz);
}

// Quadrupole pairs of cell entries p[0], p[LPB], ... (t of them, cell indices
// k[0], k[LPB], ...; the cells' G*Q records in gq[2k], gq[2k+1]).  Four
// entries' records are loaded ahead of their arithmetic.
template <int LPB>
__device__ __forceinline__ void qflush_share(const float4 *p, const uint32_t *k, uint32_t stride,
                                             uint32_t t, const float4 *__restrict__ gq, float rx,
                                             float ry, float rz, float eps2, float &fx, float &fy,
                                             float &fz) {
    const uint32_t st = LPB > 0 ? (uint32_t)LPB : stride;
    uint32_t u = 0;
    for (; u + 4 <= t; u += 4) {
        float4 e[4], qa[4], qb[4];
#pragma unroll
        for (int v = 0; v < 4; ++v) {
            e[v] = p[(u + v) * st];
            const uint32_t c = k[(u + v) * st];
            qa[v] = gq[2ull * c];
            qb[v] = gq[2ull * c + 1];
        }
#pragma unroll
        for (int v = 0; v < 4; ++v) pair_quad(e[v], qa[v], qb[v], rx, ry, rz, eps2, fx, fy, fz);
    }
    for (; u < t; ++u) {
        const uint32_t c = k[u * st];
        pair_quad(p[u * st], gq[2ull * c], gq[2ull * c + 1], rx, ry, rz, eps2, fx, fy, fz);
    }
}

// LDS of one wave's walk.
struct GroupLds {
    uint32_t *stk;
    float4 *lst;
    float4 *qlst;    // quadrupole: accepted cells' entries
    uint32_t *qk;    // quadrupole: their internal-cell indices
    float4 *cl_e;    // correction: accepted cell offset from the group centre, G*M
    uint32_t *cl_k;  // correction: accepted cell's internal index
    float4 *boff;    // correction: body offsets from the group centre
    float4 *bcorr;   // correction: learned correction per body
};

// Bodies of the first of two passes over each flushed list, for a group of
// s <= 64 bodies (s: one pass), and the list entries per flush.  A pass over
// sa bodies gives each 64 / sa lanes, so an entry costs 1 / (64 / sa) trips;
// the split minimising the trips of both passes is taken if it saves at least
// 12 % of the one-pass trips (a second pass has its own set-up, reduction and
// registers; 31 of the 64 sizes split).  26 bodies: 5 with 12 lanes, then 21
// with 3, 0.42 trips per entry instead of 0.5 with 2 lanes each (52 of 64
// lanes busy).  A flush takes a multiple
// of both passes' lanes per body (60 for 3 and 12), so that only a group's
// last flush has remainder trips.  Computed at compile time.
struct TwoPass {
    uint8_t first[kGroup + 1];
    uint8_t chunk[kGroup + 1];
};
static_assert(kFlushSpan <= 255, "flush spans fit TwoPass::chunk");
constexpr uint32_t gcd_u32(uint32_t a, uint32_t b) { return b ? gcd_u32(b, a % b) : a; }
constexpr TwoPass make_two_pass() {
    TwoPass t{};
    for (uint32_t s = 1; s <= (uint32_t)kGroup; ++s) {
        // compare sums of 1 / lanes-per-body exactly: a / b < c / d <=> a d < c b
        const uint32_t one = (uint32_t)kGroup / s;  // one pass: 1 / one trips per entry
        uint32_t best = s, bn = 1, bd = one;          // best cost bn / bd
        for (uint32_t a = 1; a < s; ++a) {
            const uint32_t l1 = (uint32_t)kGroup / a, l2 = (uint32_t)kGroup / (s - a);
            const uint32_t cn = l1 + l2, cd = l1 * l2;  // 1 / l1 + 1 / l2
            if (cn * bd < bn * cd && 100 * cn * one < 88 * cd) {
                best = a;
                bn = cn;
                bd = cd;
            }
        }
        t.first[s] = (uint8_t)best;
        const uint32_t l1 = (uint32_t)kGroup / best;
        const uint32_t l2 = best < s ? (uint32_t)kGroup / (s - best) : l1;
        const uint32_t lcm = l1 / gcd_u32(l1, l2) * l2;
        t.chunk[s] = (uint8_t)(lcm <= (uint32_t)kFlushSpan ? lcm * ((uint32_t)kFlushSpan / lcm)
                                                             : (uint32_t)kFlushSpan);
    }
    return t;
}
__constant__ constexpr TwoPass kTwoPassTable = make_two_pass();

// Forces on the s bodies of one group starting at sorted position g0 (both
// wave-uniform, checked by the caller).  NB = 1 (s <= 64): lane l holds body
// l % s and list share l / s; the lpb = 64 / s lanes of a body share its list,
// and the 64 - s * lpb lanes beyond repeat share 0 and their sums are dropped.
// NB = 2 (64 < s <= 128): a lane holds bodies lane and lane + 64 and takes
// the whole list for both.  Monopole walks of one body per lane split the
// group into two passes over every flushed list (kTwoPassTable): bodies
// [0, sa) as above with sa in place of s, then bodies [sa, s), lane l
// holding body sa + l % (s - sa) and share l / (s - sa).  *ncorr_out
// accumulates network evaluations.
template <bool kCorr, int NB, bool kQuad>
__device__ __forceinline__ void group_walk_n(const GroupTree &T, uint32_t g0, uint32_t s,
                                             float eps2, const CorrArgs &corr,
                                             const MlpForward &w, const GroupLds &L,
                                             uint32_t &peak, bool &overflow,
                                             uint32_t &ncorr_out, uint32_t &nquad_out) {
    static_assert(!kQuad || NB == 1, "quadrupole walks hold one body per lane");
    const uint32_t lane = threadIdx.x & 63u;
    const uint32_t *__restrict__ sidx = T.sidx;
    const double *__restrict__ x = T.x, *__restrict__ y = T.y, *__restrict__ z = T.z,
                 *__restrict__ m = T.m;
    const NodeF32 *__restrict__ rec = T.rec;
    const uint32_t icap = T.icap;
    const float ro2_root = T.ro2_root, e2_root = T.e2_root;
    double *ax = T.ax, *ay = T.ay, *az = T.az;
    uint32_t *inter = T.inter;
    uint32_t *stk = L.stk;
    float4 *lst = L.lst, *cl_e = L.cl_e, *boff = L.boff, *bcorr = L.bcorr;
    uint32_t *cl_k = L.cl_k;
    float4 *qlst = L.qlst;
    uint32_t *qk = L.qk;
    const float4 *__restrict__ gquad = T.gquad;
    {
        // (quadrupole walks keep one pass: their second list is shared out
        // by the first pass's mapping only)
        constexpr bool kTwoPass = !kQuad && NB == 1;
        // bodies of the first pass (all of them without a second one)
        const uint32_t sa = kTwoPass && NB == 1 ? (uint32_t)kTwoPassTable.first[s] : s;
        const uint32_t sb = s - sa;                             // second pass: 0 or more
        const uint32_t ss = NB == 2 ? (uint32_t)kGroup : sa;  // body slots of the first pass
        const uint32_t sub = NB == 2 ? lane : lane % ss;      // which body of the group
        const uint32_t q = NB == 2 ? 0u : lane / ss;          // which share of the list
        const uint32_t lpb = (uint32_t)kGroup / ss;           // lanes per body
        const uint32_t qs = q < lpb ? q : 0u;                 // the share this lane computes
        // the second pass: body sa + sub2 in share q2 of 64 / sb
        const uint32_t sub2 = sb ? lane % sb : 0u;
        const uint32_t q2 = sb ? lane / sb : 0u;
        const uint32_t lpb2 = sb ? (uint32_t)kGroup / sb : 1u;
        const uint32_t qs2 = q2 < lpb2 ? q2 : 0u;
        // list entries per flush: a multiple of lpb (124 to 128), so that only
        // the group's last flush has a remainder trip (of lpb2 too with two passes)
        const uint32_t chunk =
            sb ? (uint32_t)kTwoPassTable.chunk[s] : lpb * ((uint32_t)kFlushSpan / lpb);
        const bool pow2 = (ss & (ss - 1)) == 0;
        bool valid[NB];
        uint32_t bi[NB];
        double px[NB], py[NB], pz[NB];
#pragma unroll
        for (int b = 0; b < NB; ++b) {
            const uint32_t k = sub + (uint32_t)b * kGroup;
            valid[b] = k < (NB == 2 ? s : sa);
            // an unused slot repeats a body of the group (its box stays the group's)
            bi[b] = sidx[g0 + (valid[b] ? k : (b == 0 ? 0u : sub))];
            px[b] = x[bi[b]];
            py[b] = y[bi[b]];
            pz[b] = z[bi[b]];
        }
        // the second pass's body (every lane holds one of the sb bodies)
        uint32_t bi2 = bi[0];
        double px2 = px[0], py2 = py[0], pz2 = pz[0];
        if (sb) {
            bi2 = sidx[g0 + sa + sub2];
            px2 = x[bi2];
            py2 = y[bi2];
            pz2 = z[bi2];
        }

        // bounding box (FP64, exact min/max), centre and half extent
        double mnx = fmin(px[0], px2), mny = fmin(py[0], py2), mnz = fmin(pz[0], pz2);
        double mxx = fmax(px[0], px2), mxy = fmax(py[0], py2), mxz = fmax(pz[0], pz2);
#pragma unroll
        for (int b = 1; b < NB; ++b) {
            mnx = fmin(mnx, px[b]);
            mny = fmin(mny, py[b]);
            mnz = fmin(mnz, pz[b]);
            mxx = fmax(mxx, px[b]);
            mxy = fmax(mxy, py[b]);
            mxz = fmax(mxz, pz[b]);
        }
#pragma unroll
        for (int off = 32; off >= 1; off >>= 1) {
            mnx = fmin(mnx, __shfl_xor(mnx, off, 64));
            mny = fmin(mny, __shfl_xor(mny, off, 64));
            mnz = fmin(mnz, __shfl_xor(mnz, off, 64));
            mxx = fmax(mxx, __shfl_xor(mxx, off, 64));
            mxy = fmax(mxy, __shfl_xor(mxy, off, 64));
            mxz = fmax(mxz, __shfl_xor(mxz, off, 64));
        }
        const double cx = 0.5 * (mnx + mxx), cy = 0.5 * (mny + mxy), cz = 0.5 * (mnz + mxz);
        // half extents rounded up so that the FP32 box always contains the bodies
        const float hx = (float)(0.5 * (mxx - mnx)) * 1.000001f;
        const float hy = (float)(0.5 * (mxy - mny)) * 1.000001f;
        const float hz = (float)(0.5 * (mxz - mnz)) * 1.000001f;
        const float gxh = (float)cx, gyh = (float)cy, gzh = (float)cz;
        const float gxl = (float)(cx - (double)gxh), gyl = (float)(cy - (double)gyh),
                    gzl = (float)(cz - (double)gzh);
        // The body's own offset is computed with exactly the operations that
        // produce its leaf's list entry (record hi/lo of (x*m)/m, k_records),
        // so the self pair has a distance of exactly 0 and adds nothing; with
        // eps^2 = 1e-22 any rounding residual would be a huge spurious force.
        float rx[NB], ry[NB], rz[NB];
#pragma unroll
        for (int b = 0; b < NB; ++b) {
            const double mb = m[bi[b]];
            const double c3[3] = {(px[b] * mb) / mb, (py[b] * mb) / mb, (pz[b] * mb) / mb};
            const float gh[3] = {gxh, gyh, gzh}, gl[3] = {gxl, gyl, gzl};
            float r3[3];
#pragma unroll
            for (int c = 0; c < 3; ++c) {
                const float hi = (float)c3[c];
                const float lo = (float)(c3[c] - (double)hi);
                r3[c] = (hi - gh[c]) + (lo - gl[c]);
            }
            rx[b] = r3[0];
            ry[b] = r3[1];
            rz[b] = r3[2];
        }
        float rx2 = rx[0], ry2 = ry[0], rz2 = rz[0];
        if (sb) {
            const double mb = m[bi2];
            const double c3[3] = {(px2 * mb) / mb, (py2 * mb) / mb, (pz2 * mb) / mb};
            const float gh[3] = {gxh, gyh, gzh}, gl[3] = {gxl, gyl, gzl};
            float r3[3];
#pragma unroll
            for (int c = 0; c < 3; ++c) {
                const float hi = (float)c3[c];
                const float lo = (float)(c3[c] - (double)hi);
                r3[c] = (hi - gh[c]) + (lo - gl[c]);
            }
            rx2 = r3[0];
            ry2 = r3[1];
            rz2 = r3[2];
        }

        double acc_x[NB], acc_y[NB], acc_z[NB];
#pragma unroll
        for (int b = 0; b < NB; ++b) acc_x[b] = acc_y[b] = acc_z[b] = 0.0;
        double acc2_x = 0.0, acc2_y = 0.0, acc2_z = 0.0;  // the second pass's body
        uint32_t n_inter = 0;
        uint32_t sp = 0, ln = 0;  // wave-uniform stack and list sizes
        uint32_t cn = 0;          // wave-uniform size of the correction's cell list
        uint32_t ncorr = 0;       // network evaluations of this group
        uint32_t qn = 0;          // wave-uniform size of the quadrupole cell list
        uint32_t nquad = 0;       // quadrupole pairs of this group
        if constexpr (kCorr) {
#pragma unroll
            for (int b = 0; b < NB; ++b) {
                const uint32_t k = sub + (uint32_t)b * kGroup;
                if (valid[b] && q == 0) boff[k] = make_float4(rx[b], ry[b], rz[b], 0.f);
                bcorr[lane + (uint32_t)b * kGroup] = make_float4(0.f, 0.f, 0.f, 0.f);
            }
            if (kTwoPass && sb && q2 == 0) boff[sa + sub2] = make_float4(rx2, ry2, rz2, 0.f);
        }

        // Entries lst[base .. base + cnt) (cnt <= kFlushSpan): lane share qs takes
        // qs, qs + lpb, ...  Every share runs cnt / lpb trips; then shares
        // qs < cnt % lpb take one more entry and the others a zero-force
        // entry (a select, so the trip count stays uniform and nothing is
        // written to the list).  The reads stay below base + cnt + lpb, inside
        // the list.
        auto flush = [&](uint32_t base, uint32_t cnt) {
            const uint32_t full = cnt / lpb;
            const uint32_t rem = cnt - full * lpb;
            const float4 *p = lst + base + qs;
            float fx[NB], fy[NB], fz[NB];
#pragma unroll
            for (int b = 0; b < NB; ++b) fx[b] = fy[b] = fz[b] = 0.f;
            if constexpr (NB == 2) {
                flush_share<1, 2>(p, cnt, rx, ry, rz, eps2, fx, fy, fz);  // lpb = 1
            } else {
                switch (lpb) {
                case 1: flush_share<1, 1>(p, full, rx, ry, rz, eps2, fx, fy, fz); break;
                case 2: flush_share<2, 1>(p, full, rx, ry, rz, eps2, fx, fy, fz); break;
                case 3: flush_share<3, 1>(p, full, rx, ry, rz, eps2, fx, fy, fz); break;
                case 4: flush_share<4, 1>(p, full, rx, ry, rz, eps2, fx, fy, fz); break;
                case 8: flush_share<8, 1>(p, full, rx, ry, rz, eps2, fx, fy, fz); break;
                case 16: flush_share<16, 1>(p, full, rx, ry, rz, eps2, fx, fy, fz); break;
                case 32: flush_share<32, 1>(p, full, rx, ry, rz, eps2, fx, fy, fz); break;
                case 64: flush_share<64, 1>(p, full, rx, ry, rz, eps2, fx, fy, fz); break;
                default:
                    flush_share_rt(p, lpb, full, rx[0], ry[0], rz[0], eps2, fx[0], fy[0], fz[0]);
                    break;
                }
                if (rem) {
                    const float4 e = p[full * lpb];
                    pair_f32(qs < rem ? e : far_entry(), rx[0], ry[0], rz[0], eps2, fx[0], fy[0],
                             fz[0]);
                }
            }
#pragma unroll
            for (int b = 0; b < NB; ++b) {
                acc_x[b] += (double)fx[b];
                acc_y[b] += (double)fy[b];
                acc_z[b] += (double)fz[b];
            }
            if (kTwoPass && sb) {  // the same entries for the second pass's bodies
                const uint32_t full2 = cnt / lpb2;
                const uint32_t rem2 = cnt - full2 * lpb2;
                const float4 *p2 = lst + base + qs2;
                float gx2[1] = {0.f}, gy2[1] = {0.f}, gz2[1] = {0.f};
                const float qx[1] = {rx2}, qy[1] = {ry2}, qz[1] = {rz2};
                switch (lpb2) {  // the second pass's common strides (32, 21 or 16 bodies)
                case 2: flush_share<2, 1>(p2, full2, qx, qy, qz, eps2, gx2, gy2, gz2); break;
                case 3: flush_share<3, 1>(p2, full2, qx, qy, qz, eps2, gx2, gy2, gz2); break;
                case 4: flush_share<4, 1>(p2, full2, qx, qy, qz, eps2, gx2, gy2, gz2); break;
                default:
                    flush_share_rt(p2, lpb2, full2, rx2, ry2, rz2, eps2, gx2[0], gy2[0], gz2[0]);
                    break;
                }
                if (rem2) {
                    const float4 e = p2[full2 * lpb2];
                    pair_f32(qs2 < rem2 ? e : far_entry(), rx2, ry2, rz2, eps2, gx2[0], gy2[0],
                             gz2[0]);
                }
                acc2_x += (double)gx2[0];
                acc2_y += (double)gy2[0];
                acc2_z += (double)gz2[0];
            }
            n_inter += cnt;
        };

        // Quadrupole cells qlst[base .. base + cnt), shared out over the lanes
        // of a body as in flush.
        auto qflush = [&](uint32_t base, uint32_t cnt) {
            const uint32_t full = cnt / lpb;
            const uint32_t rem = cnt - full * lpb;
            const float4 *p = qlst + base + qs;
            const uint32_t *pk = qk + base + qs;
            float fx = 0.f, fy = 0.f, fz = 0.f;
            switch (lpb) {
            case 1: qflush_share<1>(p, pk, 1, full, gquad, rx[0], ry[0], rz[0], eps2, fx, fy, fz); break;
            case 2: qflush_share<2>(p, pk, 2, full, gquad, rx[0], ry[0], rz[0], eps2, fx, fy, fz); break;
            case 4: qflush_share<4>(p, pk, 4, full, gquad, rx[0], ry[0], rz[0], eps2, fx, fy, fz); break;
            case 8: qflush_share<8>(p, pk, 8, full, gquad, rx[0], ry[0], rz[0], eps2, fx, fy, fz); break;
            case 16: qflush_share<16>(p, pk, 16, full, gquad, rx[0], ry[0], rz[0], eps2, fx, fy, fz); break;
            case 32: qflush_share<32>(p, pk, 32, full, gquad, rx[0], ry[0], rz[0], eps2, fx, fy, fz); break;
            case 64: qflush_share<64>(p, pk, 64, full, gquad, rx[0], ry[0], rz[0], eps2, fx, fy, fz); break;
            default: qflush_share<0>(p, pk, lpb, full, gquad, rx[0], ry[0], rz[0], eps2, fx, fy, fz); break;
            }
            if (rem) {
                // shares past the remainder take a zero-force entry with a zero
                // quadrupole (a far entry with a real Q would give inf * 0)
                const bool real = qs < rem;
                const uint32_t c = real ? pk[full * lpb] : 0u;
                const float4 zero = make_float4(0.f, 0.f, 0.f, 0.f);
                const float4 qa = gquad[2ull * c], qb = gquad[2ull * c + 1];
                pair_quad(real ? p[full * lpb] : far_entry(), real ? qa : zero, real ? qb : zero,
                          rx[0], ry[0], rz[0], eps2, fx, fy, fz);
            }
            acc_x[0] += (double)fx;
            acc_y[0] += (double)fy;
            acc_z[0] += (double)fz;
            n_inter += cnt;
            nquad += cnt * s;
        };

        // learned correction of cells cl[base .. base+cnt) (cnt <= 16) for every body
        auto corr_flush = [&](uint32_t base, uint32_t cnt) {
            const int g = (int)(lane >> 4), p = (int)(lane & 15);
            const bool pv = (uint32_t)p < cnt;
            const float4 ce = cl_e[base + p];
            MlpCell mc{};
            if (pv) mc = corr.mcell[cl_k[base + p]];
            const float e2c = mc.c.z * mc.c.z;
            ncorr += cnt * s;
            bf16x4 zero4;
            for (int i = 0; i < 4; ++i) zero4[i] = (__bf16)0.f;
            for (uint32_t b = 0; b < s; ++b) {
                const float4 bo = boff[b];
                float inv_d;
                bf16x4 xin = mlp_features(ce.x - bo.x, ce.y - bo.y, ce.z - bo.z, mc, g, inv_d);
                if (!pv) xin = zero4;
                MlpActs act;
                const f32x4 yv = mlp_forward_tile(w, xin, act);
                const float inv2 = inv_d * inv_d;
                const float sc = (pv && g == 0) ? ce.w * e2c * (inv2 * inv2) : 0.f;
                float cx = yv[0] * sc, cy = yv[1] * sc, cz = yv[2] * sc;
                for (int off = 8; off >= 1; off >>= 1) {
                    cx += __shfl_xor(cx, off, 64);
                    cy += __shfl_xor(cy, off, 64);
                    cz += __shfl_xor(cz, off, 64);
                }
                if (lane == 0) {
                    float4 t = bcorr[b];
                    t.x += cx;
                    t.y += cy;
                    t.z += cz;
                    bcorr[b] = t;
                }
            }
        };

        // A node goes to the list (accepted cell, or body leaf: exact pair, the
        // self pair adds 0) or is opened (cell failing the MAC); an empty slot
        // does neither.  d2 = squared distance to the group box.  Computed for
        // every lane without branches: an empty slot's values are unused.
        auto classify = [&](const NodeF32 &r, int depth, float4 &entry, float &d2,
                            bool &to_list, bool &to_open) {
            const float dx = (r.x - gxh) + (r.lx - gxl);
            const float dy = (r.y - gyh) + (r.ly - gyl);
            const float dz = (r.z - gzh) + (r.lz - gzl);
            entry = make_float4(dx, dy, dz, r.gm);
            const float ex = fmaxf(fabsf(dx) - hx, 0.f);
            const float ey = fmaxf(fabsf(dy) - hy, 0.f);
            const float ez = fmaxf(fabsf(dz) - hz, 0.f);
            d2 = fmaf(ex, ex, fmaf(ey, ey, ez * ez));
            // (edge / theta)^2 = ro2_root 4^-depth, exact (a power-of-two
            // scaling); computed rather than read from an LDS table, which
            // put an LDS round trip on every walk step's chain of waits
            const bool far_enough = d2 > ldexpf(ro2_root, -2 * depth);
            const bool cell = r.ref >= 0;
            to_list = r.ref != kRefEmpty && (!cell || far_enough);
            to_open = cell && !far_enough;
        };

        {
            float4 e;
            float d2r;
            const NodeF32 root = rec[0];
            // every lane evaluates the root alike
            bool rl, ro;
            classify(root, 0, e, d2r, rl, ro);
            const int cls = (int)wave_uniform(rl ? 1u : (ro ? 2u : 0u));
            if (cls == 1) {
                if (lane == 0) lst[0] = e;
                ln = 1;
                if constexpr (kCorr) {
                    if (wave_uniform(root.ref >= 0 && e2_root > corr.tau2 * d2r)) {
                        if (lane == 0) {
                            cl_e[0] = e;
                            cl_k[0] = (uint32_t)root.ref;
                        }
                        cn = 1;
                    }
                }
            } else if (cls == 2) {
                if (lane == 0) stk[0] = ((uint32_t)root.ref << 6);
                sp = 1;
            }
            __syncthreads();
        }

        // Pop up to 8 kPop cells and load this lane's kPop child records
        // (child lane & 7 of the cells lane >> 3, + 8, + 16, ...): the kPop
        // loads of a lane are in flight together, so a walk waits on memory
        // once per kPop batches of cells instead of once per batch.  Every
        // lane loads (a lane without a cell reads a child of cell 0 and is
        // marked idle): a load under a branch would make the compiler wait
        // for it at the branch's end, before the list evaluation that is
        // meant to hide it.
        NodeF32 r[kPop];
        int depth[kPop];
        bool live[kPop];
        auto pop_and_load = [&]() -> uint32_t {
            const uint32_t take = sp < 8u * kPop ? sp : 8u * kPop;
            sp -= take;
#pragma unroll
            for (int k = 0; k < kPop; ++k) {
                const uint32_t c = (lane >> 3) + 8u * (uint32_t)k;
                live[k] = c < take;
                const uint32_t entry = live[k] ? stk[sp + c] : 0u;
                depth[k] = (int)(entry & 63u) + 1;
                r[k] = rec[1ull + 8ull * (entry >> 6) + (lane & 7u)];
            }
            __syncthreads();
            return take;
        };

        // A correct octree opens every internal cell at most once per group,
        // so icap iterations bound the walk; the cap turns a corrupt tree into
        // an error instead of an endless loop.
        uint32_t budget = icap + 16;
        uint32_t take = sp > 0 ? pop_and_load() : 0u;
        while (take > 0) {
            if (budget-- == 0) {
                overflow = true;
                break;
            }
#pragma unroll
            for (int k = 0; k < kPop; ++k) {
                float4 e;
                float d2;
                bool to_list, to_open;
                classify(r[k], depth[k], e, d2, to_list, to_open);
                to_list = to_list && live[k];
                to_open = to_open && live[k];
                if constexpr (kQuad) {
                    // accepted cells to the quadrupole list, body leaves stay
                    const bool tq = to_list && r[k].ref >= 0;
                    const uint64_t bq = __ballot(tq);
                    if (tq) {
                        const uint32_t p = qn + count_below(bq);
                        qlst[p] = e;
                        qk[p] = (uint32_t)r[k].ref;
                    }
                    qn += __popcll(bq);  // < chunk before: at most kListCap - 1 after
                    to_list = to_list && !tq;
                }
                const uint64_t bl = __ballot(to_list);
                const uint64_t bo = __ballot(to_open);
                if (to_list) lst[ln + count_below(bl)] = e;
                if (to_open) {
                    const uint32_t p = sp + count_below(bo);
                    if (p < kStackCap) stk[p] = ((uint32_t)r[k].ref << 6) | (uint32_t)depth[k];
                }
                if constexpr (kCorr) {
                    const bool cc = to_list && r[k].ref >= 0 &&
                                    ldexpf(e2_root, -2 * depth[k]) > corr.tau2 * d2;
                    const uint64_t bc = __ballot(cc);
                    if (cc) {
                        const uint32_t p = cn + count_below(bc);
                        cl_e[p] = e;
                        cl_k[p] = (uint32_t)r[k].ref;
                    }
                    cn += __popcll(bc);  // < kCorrTile before: at most kCellCap - 1 after
                }
                ln += __popcll(bl);  // < chunk <= kFlushSpan before: at most kListCap - 1 after
                sp += __popcll(bo);
            }
            peak = sp > peak ? sp : peak;
            if (sp > kStackCap) {
                overflow = true;
                break;
            }
            __syncthreads();
            // issue the next loads; the list evaluation below hides their latency
            take = sp > 0 ? pop_and_load() : 0u;
            if (ln >= chunk) {
                uint32_t done = 0;
                for (; ln - done >= chunk; done += chunk) flush(done, chunk);
                __syncthreads();
                const uint32_t rest = ln - done;  // < chunk <= done: no overlap
                for (uint32_t i = lane; i < rest; i += kGroup) lst[i] = lst[done + i];
                ln = rest;
                __syncthreads();
            }
            if constexpr (kQuad) {
                if (qn >= chunk) {
                    uint32_t done = 0;
                    for (; qn - done >= chunk; done += chunk) qflush(done, chunk);
                    __syncthreads();
                    const uint32_t rest = qn - done;  // < chunk <= done: no overlap
                    for (uint32_t i = lane; i < rest; i += kGroup) {
                        qlst[i] = qlst[done + i];
                        qk[i] = qk[done + i];
                    }
                    qn = rest;
                    __syncthreads();
                }
            }
            if constexpr (kCorr) {
                if (cn >= (uint32_t)kCorrTile) {
                    uint32_t base = 0;
                    for (; cn - base >= (uint32_t)kCorrTile; base += kCorrTile)
                        corr_flush(base, kCorrTile);
                    __syncthreads();
                    const uint32_t rest = cn - base;
                    if (lane < rest) {
                        cl_e[lane] = cl_e[base + lane];
                        cl_k[lane] = cl_k[base + lane];
                    }
                    cn = rest;
                    __syncthreads();
                }
            }
        }
        if (ln > 0) flush(0, ln);
        if constexpr (kQuad) {
            if (qn > 0) qflush(0, qn);
        }
        if constexpr (kCorr) {
            if (cn > 0) corr_flush(0, cn);
        }
        __syncthreads();
        // add up the list shares of each body (fixed order: deterministic)
        if (pow2) {
            for (uint32_t off = ss; off < (uint32_t)kGroup; off <<= 1) {
                acc_x[0] += __shfl_xor(acc_x[0], off, 64);
                acc_y[0] += __shfl_xor(acc_y[0], off, 64);
                acc_z[0] += __shfl_xor(acc_z[0], off, 64);
            }
        } else {
            // body sub's shares sit in lanes sub + j ss, j < lpb; lane sub adds them
            double tx = acc_x[0], ty = acc_y[0], tz = acc_z[0];
            for (uint32_t j = 1; j < lpb; ++j) {
                const int src = (int)(sub + j * ss);
                tx += __shfl(acc_x[0], src, 64);
                ty += __shfl(acc_y[0], src, 64);
                tz += __shfl(acc_z[0], src, 64);
            }
            acc_x[0] = tx;
            acc_y[0] = ty;
            acc_z[0] = tz;
        }
        if (kTwoPass && sb) {
            // body sa + sub2's shares sit in lanes sub2 + j sb, j < lpb2; lane sub2 adds them
            double tx = acc2_x, ty = acc2_y, tz = acc2_z;
            for (uint32_t j = 1; j < lpb2; ++j) {
                const int src = (int)(sub2 + j * sb);
                tx += __shfl(acc2_x, src, 64);
                ty += __shfl(acc2_y, src, 64);
                tz += __shfl(acc2_z, src, 64);
            }
            acc2_x = tx;
            acc2_y = ty;
            acc2_z = tz;
        }
        ncorr_out += ncorr;
        nquad_out += nquad;
        // Stored also after an overflow (the caller reports it and the values
        // are then undefined): a predicate holding `overflow` would let the
        // compiler route the storing lanes past the caller's exit test, a
        // per-lane split of the wave's loop.
#pragma unroll
        for (int b = 0; b < NB; ++b) {
            if (valid[b] && q == 0) {
                if constexpr (kCorr) {
                    const float4 c = bcorr[sub + (uint32_t)b * kGroup];
                    acc_x[b] += (double)c.x;
                    acc_y[b] += (double)c.y;
                    acc_z[b] += (double)c.z;
                }
                ax[bi[b]] = acc_x[b];
                ay[bi[b]] = acc_y[b];
                az[bi[b]] = acc_z[b];
                if (inter) inter[bi[b]] = n_inter;
            }
        }
        if (kTwoPass && sb && q2 == 0) {  // lanes 0 .. sb - 1
            if constexpr (kCorr) {
                const float4 c = bcorr[sa + sub2];
                acc2_x += (double)c.x;
                acc2_y += (double)c.y;
                acc2_z += (double)c.z;
            }
            ax[bi2] = acc2_x;
            ay[bi2] = acc2_y;
            az[bi2] = acc2_z;
            if (inter) inter[bi2] = n_inter;
        }
    }
}

// Forces on the bodies of one group grp = {first sorted position, count}
// (checked by the caller): one body per lane up to 64 bodies, two beyond.
template <bool kCorr, bool kQuad = false>
__device__ __forceinline__ void group_walk(const GroupTree &T, uint2 grp, float eps2,
                                           const CorrArgs &corr, const MlpForward &w,
                                           const GroupLds &L, uint32_t &peak, bool &overflow,
                                           uint32_t &ncorr_out, uint32_t &nquad_out) {
    const uint32_t g0 = wave_uniform(grp.x);
    const uint32_t s = wave_uniform(grp.y);
    if constexpr (kGroupMax > (uint32_t)kGroup && !kQuad) {
        if (s > (uint32_t)kGroup) {
            group_walk_n<kCorr, 2, false>(T, g0, s, eps2, corr, w, L, peak, overflow, ncorr_out,
                                          nquad_out);
            return;
        }
    }
    group_walk_n<kCorr, 1, kQuad>(T, g0, s, eps2, corr, w, L, peak, overflow, ncorr_out,
                                  nquad_out);
}

// a group {first, count} that a walk may trust
__device__ inline bool group_ok(uint2 grp, uint32_t n) {
    return grp.y != 0 && grp.y <= kGroupLimit && grp.x < n && grp.y <= n - grp.x;
}

// Before a traversal: reset the group queue, and order the last groups of
// each XCD's chunk (kTailGroups, at most two thirds of the chunk: sorting all
// of it scatters the walks that share L2 lines) by decreasing estimated cost,
// a group's size times the list length the previous traversal recorded for
// its first body.  The walks that end the launch are then the shortest
// instead of whichever groups come last in octree order (on a 1M-body disk
// the slowest waves ended 200 us after the queue ran dry, a fifth of the
// launch).  A counting sort by cost class (8 classes per factor of two, 256
// in all) in LDS; within a class the order is that of the LDS atomics.  Only
// the order in which groups start changes: each group's forces are the same.
// Without recorded lengths (all zero) the order stays.  One workgroup per
// chunk.
constexpr uint32_t kTailGroups = 4096;
constexpr int kTailThreads = 1024;
constexpr uint32_t kCostClasses = 256;
__device__ inline uint32_t cost_class(uint32_t v) {  // decreasing in v; 0 -> last class
    if (v == 0) return kCostClasses - 1;
    const uint32_t e = 31u - (uint32_t)__clz(v);  // floor(log2 v)
    const uint32_t f = e >= 3 ? (v >> (e - 3)) & 7u : (v << (3 - e)) & 7u;
    return kCostClasses - 2 - min(8 * e + f, kCostClasses - 2);
}
__global__ void __launch_bounds__(kTailThreads) k_group_queue_prepare(
    TreeHdr *h, uint32_t *gq, uint32_t n, uint2 *groups, const uint32_t *__restrict__ sidx,
    const uint32_t *__restrict__ inter) {
    __shared__ uint2 grp[kTailGroups];
    __shared__ uint8_t cls[kTailGroups];
    __shared__ uint32_t hist[kCostClasses], fill[kCostClasses];
    const uint32_t t = threadIdx.x, c = blockIdx.x;  // c: chunk (and XCD) index
    if (c == 0 && t == 0) {
        h->group_next = 0;
        h->corr_pairs = 0;
        h->quad_pairs = 0;
    }
    if (t == 0) gq[c * kQueueStride] = 0;
    if (!inter) return;
    const uint32_t ngroups = min(h->group_count, n);  // the traversal's chunk bounds
    const uint32_t lo = (uint32_t)((uint64_t)ngroups * c / kXcds);
    const uint32_t hi = (uint32_t)((uint64_t)ngroups * (c + 1) / kXcds);
    const uint32_t m = min(2 * (hi - lo) / 3, kTailGroups);
    if (m < 2) return;  // the same for the whole workgroup
    const uint32_t base = hi - m;
    for (uint32_t b = t; b < kCostClasses; b += kTailThreads) hist[b] = fill[b] = 0;
    __syncthreads();
    for (uint32_t i = t; i < m; i += kTailThreads) {
        const uint2 g = groups[base + i];
        grp[i] = g;
        uint32_t cost = 0;  // a group the walk would reject keeps cost 0
        if (group_ok(g, n)) {
            const uint64_t v = (uint64_t)g.y * inter[sidx[g.x]];
            cost = v < 0xffffffffull ? (uint32_t)v : 0xffffffffu;
        }
        const uint32_t k = cost_class(cost);
        cls[i] = (uint8_t)k;
        atomicAdd(&hist[k], 1u);
    }
    __syncthreads();
    if (hist[kCostClasses - 1] == m) return;  // no recorded lengths: keep the order
    if (t == 0) {  // exclusive prefix of the classes
        uint32_t run = 0;
        for (uint32_t b = 0; b < kCostClasses; ++b) {
            const uint32_t v = hist[b];
            hist[b] = run;
            run += v;
        }
    }
    __syncthreads();
    for (uint32_t i = t; i < m; i += kTailThreads) {
        const uint32_t k = cls[i];
        groups[base + hist[k] + atomicAdd(&fill[k], 1u)] = grp[i];
    }
}

// At most 96 VGPRs, five waves per SIMD, for the monopole and quadrupole walks
// (their two-pass flushes would otherwise take 103 and leave four; measured:
// 3 % faster on the 4M Plummer sphere, a forced six waves spill and lose
// 10 %).  The learned correction's walks hold the network's weights in
// registers and stay at one wave: the compiler's note that it cannot meet the
// target there is expected.
#pragma clang diagnostic push
#pragma clang diagnostic ignored "-Wpass-failed"
template <bool kCorr, bool kQuad>
__global__ void __launch_bounds__(kGroup) __attribute__((amdgpu_waves_per_eu(5))) k_bh_group(
    uint32_t n, const uint32_t *__restrict__ sidx, const double *__restrict__ x,
    const double *__restrict__ y, const double *__restrict__ z, const double *__restrict__ m,
    const NodeF32 *__restrict__ rec, const uint2 *__restrict__ groups, double theta,
    float eps2, uint32_t icap, double *ax, double *ay, double *az, uint32_t *inter,
    TreeHdr *h, uint32_t *gq, CorrArgs corr, const float4 *__restrict__ gquad) {
    __shared__ uint32_t stk[kStackCap];
    __shared__ float4 lst[kListCap];
    __shared__ float4 qlst[kQuad ? kListCap : 1];
    __shared__ uint32_t qk[kQuad ? kListCap : 1];
    __shared__ float4 cl_e[kCorr ? kCellCap : 1];
    __shared__ uint32_t cl_k[kCorr ? kCellCap : 1];
    // bcorr is cleared lane by lane (lane + 64 b): at least one wave of entries
    constexpr uint32_t kCorrSlots = kGroupMax > (uint32_t)kGroup ? kGroupMax : (uint32_t)kGroup;
    __shared__ float4 boff[kCorr ? kGroupLimit : 1];
    __shared__ float4 bcorr[kCorr ? kCorrSlots : 1];
    const GroupLds L{stk, lst, qlst, qk, cl_e, cl_k, boff, bcorr};
    const uint32_t lane = threadIdx.x;
    const uint32_t ngroups = min(h->group_count, n);
    // (edge/theta)^2 of the root; a child at depth d uses ro2_root * 4^-d
    const double ro_root = h->edge / theta;
    const GroupTree T{n, sidx, x, y, z, m, rec, icap, (float)(ro_root * ro_root),
                      (float)(h->edge * h->edge), ax, ay, az, inter, gquad};
    MlpForward w;
    if constexpr (kCorr) w.load(corr.pack, corr.master, (int)lane);
    uint32_t peak = 0, ncorr = 0, nquad = 0;
    bool overflow = false;

    // Groups come in octree order: waves running at the same time walk
    // neighbouring groups, whose interaction lists overlap, so the records
    // they read are shared in L2.
    //
    // The list is cut into kXcds contiguous chunks, one per XCD: workgroup b
    // runs on XCD b mod kXcds, so the waves of one XCD walk one region of the
    // tree and its records stay in that XCD's L2.  A wave whose chunk is empty
    // goes on with the next chunks, so the XCDs finish together.  The chunk
    // counters sit in separate cache lines.
    uint32_t c = blockIdx.x % kXcds, visited = 0;
    for (;;) {
        const uint32_t lo = (uint32_t)((uint64_t)ngroups * c / kXcds);
        const uint32_t hi = (uint32_t)((uint64_t)ngroups * (c + 1) / kXcds);
        uint32_t slot = 0;
        if (lane == 0) {
            uint32_t *q = gq + c * kQueueStride;
            slot = __hip_atomic_load(q, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            if (slot < hi - lo)  // no read-modify-write once the chunk is done
                slot = __hip_atomic_fetch_add(q, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        }
        const uint32_t k = lane0_uniform(slot);
        if (k >= hi - lo) {
            if (++visited == kXcds) break;
            c = c + 1 == kXcds ? 0u : c + 1;
            continue;
        }
        const uint2 grp = groups[lo + k];
        if (!group_ok(grp, n)) {
            overflow = true;  // corrupt group table: report instead of reading out of bounds
            break;
        }
        group_walk<kCorr, kQuad>(T, grp, eps2, corr, w, L, peak, overflow, ncorr, nquad);
        if (overflow) break;
    }
    if constexpr (kCorr) {
        if (lane == 0 && ncorr) atomicAdd(&h->corr_pairs, (unsigned long long)ncorr);
    }
    if constexpr (kQuad) {
        if (lane == 0 && nquad) atomicAdd(&h->quad_pairs, (unsigned long long)nquad);
    }
    if (overflow) raise_flag(h, kFlagStack);
    if (lane == 0) raise_max(&h->stack_peak, peak);
}
#pragma clang diagnostic pop

// ---- ensemble of independent systems -------------------------------------------
__global__ void __launch_bounds__(64) k_ens_items(uint32_t nq, const uint32_t *__restrict__ qorder,
                                                  const uint32_t *__restrict__ items,
                                                  uint32_t *item_off, uint32_t *cursor,
                                                  uint32_t *status) {
    if (threadIdx.x != 0) return;
    uint32_t run = 0;
    for (uint32_t q = 0; q < nq; ++q) {
        item_off[q] = run;
        run += items[qorder[q]];
    }
    item_off[nq] = run;
    *cursor = 0;
    (void)status;
}

// Direct sum for targets [tile*64*TPL, (tile+1)*64*TPL) of one system, TPL
// per lane: FP32 pairs on offsets from the tile's first body (FP64
// subtraction, so nearby pairs keep their precision; a body's own offset
// equals its source offset, so the self pair adds exactly 0), FP64 sum of each
// 64-source chunk.  Two targets per lane halve the LDS reads per pair.
template <int TPL>
__device__ inline void direct_tile(const double *__restrict__ x, const double *__restrict__ y,
                                   const double *__restrict__ z, const double *__restrict__ m,
                                   uint32_t n, uint32_t tile, double G, float eps2, double *ax,
                                   double *ay, double *az, uint32_t *inter, float4 *stage) {
    const uint32_t lane = threadIdx.x & 63u;
    const uint32_t t0 = tile * kEnsTile * TPL;
    const double rx = x[t0], ry = y[t0], rz = z[t0];
    float px[TPL], py[TPL], pz[TPL];
    double acc_x[TPL], acc_y[TPL], acc_z[TPL];
#pragma unroll
    for (int q = 0; q < TPL; ++q) {
        const uint32_t i = t0 + lane + (uint32_t)q * kEnsTile;
        const uint32_t ii = i < n ? i : t0;
        px[q] = (float)(x[ii] - rx);
        py[q] = (float)(y[ii] - ry);
        pz[q] = (float)(z[ii] - rz);
        acc_x[q] = acc_y[q] = acc_z[q] = 0.0;
    }
    for (uint32_t c = 0; c < n; c += kEnsTile) {
        const uint32_t j = c + lane;
        float4 sv = make_float4(0.f, 0.f, 0.f, 0.f);
        if (j < n)
            sv = make_float4((float)(x[j] - rx), (float)(y[j] - ry), (float)(z[j] - rz),
                             (float)(G * m[j]));
        __syncthreads();
        stage[lane] = sv;
        __syncthreads();
        const uint32_t cnt = min(kEnsTile, n - c);
        float fx[TPL], fy[TPL], fz[TPL];
#pragma unroll
        for (int q = 0; q < TPL; ++q) fx[q] = fy[q] = fz[q] = 0.f;
        for (uint32_t k = 0; k < cnt; ++k) {
            const float4 e = stage[k];
#pragma unroll
            for (int q = 0; q < TPL; ++q) pair_f32(e, px[q], py[q], pz[q], eps2, fx[q], fy[q], fz[q]);
        }
#pragma unroll
        for (int q = 0; q < TPL; ++q) {
            acc_x[q] += (double)fx[q];
            acc_y[q] += (double)fy[q];
            acc_z[q] += (double)fz[q];
        }
    }
#pragma unroll
    for (int q = 0; q < TPL; ++q) {
        const uint32_t i = t0 + lane + (uint32_t)q * kEnsTile;
        if (i < n) {
            ax[i] = acc_x[q];
            ay[i] = acc_y[q];
            az[i] = acc_z[q];
            inter[i] = n;
        }
    }
}

// Persistent: every wave takes items from the queue until it is empty.
__global__ void __launch_bounds__(kGroup) k_ens_forces(EnsForceArgs a) {
    __shared__ uint32_t stk[kStackCap];
    __shared__ float4 lst[kListCap];
    const GroupLds L{stk, lst, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr};
    const uint32_t lane = threadIdx.x;
    const uint32_t total = a.item_off[a.nq];
    const MlpForward w{};
    const CorrArgs corr{};
    uint32_t peak = 0, ncorr = 0, nquad = 0;
    bool overflow = false;
    for (;;) {
        uint32_t slot = 0;
        if (lane == 0)
            slot = __hip_atomic_fetch_add(a.cursor, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        const uint32_t it = lane0_uniform(slot);
        if (it >= total) break;
        // queue position q with item_off[q] <= it < item_off[q + 1]
        uint32_t lo = 0, hi = a.nq;
        while (hi - lo > 1) {
            const uint32_t mid = (lo + hi) >> 1;
            if (a.item_off[mid] <= it) lo = mid;
            else hi = mid;
        }
        const uint32_t s = a.qorder[lo];
        const EnsSystem S = a.sys[s];
        const uint32_t local = it - a.item_off[lo];
        const uint32_t bo = S.body_off;
        if (S.bh) {
            const double edge = a.fb.hdr[s].edge;
            const double ro = edge / a.theta;
            const GroupTree T{S.n, a.fb.sidx + S.fb_off, a.x + bo, a.y + bo, a.z + bo, a.m + bo,
                              a.fb.rec + S.node_off, S.icap, (float)(ro * ro),
                              (float)(edge * edge), a.ax + bo, a.ay + bo, a.az + bo,
                              a.inter + bo, nullptr};
            const uint2 grp = a.fb.groups[S.fb_off + local];
            if (!group_ok(grp, S.n)) {
                overflow = true;
                break;
            }
            group_walk<false>(T, grp, a.eps2, corr, w, L, peak, overflow, ncorr, nquad);
            if (overflow) break;
        } else if (a.direct_tpl == 2) {
            direct_tile<2>(a.x + bo, a.y + bo, a.z + bo, a.m + bo, S.n, local, a.G, a.eps2,
                           a.ax + bo, a.ay + bo, a.az + bo, a.inter + bo, lst);
        } else {
            direct_tile<1>(a.x + bo, a.y + bo, a.z + bo, a.m + bo, S.n, local, a.G, a.eps2,
                           a.ax + bo, a.ay + bo, a.az + bo, a.inter + bo, lst);
        }
    }
    if (overflow && lane == 0) atomicOr(a.status, kFlagStack);
}

}  // namespace

void launch_direct_f64(uint32_t n, const double *x, const double *y, const double *z,
                       const double *m, double G, double eps2, double *ax, double *ay,
                       double *az, hipStream_t s) {
    hipLaunchKernelGGL(k_direct_f64, dim3(ceil_div(n, kDirectF64Tile)), dim3(kF64Threads), 0, s,
                       n, x, y, z, m, G, eps2, ax, ay, az);
}

void launch_direct_targets_f64(uint32_t n, const double *x, const double *y, const double *z,
                               const double *m, double G, double eps2, uint32_t nt,
                               const uint32_t *targets, double *ax, double *ay, double *az,
                               hipStream_t s) {
    if (nt == 0) return;
    hipLaunchKernelGGL(k_direct_targets_f64, dim3(ceil_div(nt, kDirectF64Tile)),
                       dim3(kF64Threads), 0, s, n, x, y, z, m, G, eps2, nt, targets, ax, ay, az);
}

// About kSplitBlocks blocks per launch; rounding the split count down keeps
// the partial sums at most 3 * 8 * kSplitBlocks * kSplitTargets bytes (24 MiB)
// for every n.
uint32_t direct_split_count(uint32_t n) {
    const uint32_t tiles = ceil_div(n, kSplitTargets);
    const uint32_t splits = kSplitBlocks / tiles;
    const uint32_t max_splits = std::max<uint32_t>(1, n / kSplitTile);
    return std::max<uint32_t>(1, std::min(splits, max_splits));
}

// scratch: the partial sums [split][3][n] doubles (none with one split)
uint64_t direct_split_scratch_bytes(uint32_t n) {
    const uint32_t splits = direct_split_count(n);
    return splits > 1 ? 3ull * n * splits * sizeof(double) : 0;
}

void launch_direct_split(bool fp64, uint32_t n, const double *x, const double *y,
                         const double *z, const double *m, double G, double eps2, double *ax,
                         double *ay, double *az, double *partial, hipStream_t s) {
    if (n == 0) return;
    const uint32_t splits = direct_split_count(n);
    const uint32_t tiles = ceil_div(n, kSplitTargets);
    const uint32_t chunk = ceil_div(ceil_div(n, splits), kSplitTile) * kSplitTile;
    const dim3 grid(tiles, splits);
    if (fp64)
        hipLaunchKernelGGL(k_direct_f64f, grid, dim3(kSplitThreads), 0, s, n, x, y, z, m, G, eps2,
                           chunk, partial, ax, ay, az);
    else
        hipLaunchKernelGGL(k_direct_f32c, grid, dim3(kSplitThreads), 0, s, n, x, y, z, m, G,
                           (float)eps2, chunk, partial, ax, ay, az);
    if (splits > 1)
        hipLaunchKernelGGL(k_split_reduce, dim3(ceil_div(n, kSplitThreads)), dim3(kSplitThreads),
                           0, s, n, splits, fp64 ? G : 1.0, (const double *)partial, ax, ay, az);
}

void launch_bh_exact(const TreeBufs &t, const double *x, const double *y, const double *z,
                     double G, double eps2, double theta, int sort_bodies, uint32_t *stack,
                     uint32_t stack_cap, double *ax, double *ay, double *az,
                     uint32_t *interactions, hipStream_t s) {
    hipLaunchKernelGGL(k_bh_exact, dim3(ceil_div(t.n, kExactThreads)), dim3(kExactThreads), 0, s,
                       t.n, sort_bodies ? (const uint32_t *)t.sidx : nullptr, x, y, z,
                       (const double4 *)t.cm, (const double *)t.edge, (const int32_t *)t.ref, G,
                       eps2, theta, stack, stack_cap, t.icap, ax, ay, az, interactions, t.hdr);
}

void launch_bh_group(const TreeBufs &t, const double *x, const double *y, const double *z,
                     const double *m, double eps2, double theta, const float4 *gquad, double *ax,
                     double *ay, double *az, uint32_t *interactions, hipStream_t s) {
    const uint32_t blocks = std::min<uint32_t>(kGroupBlocks, t.n);
    hipLaunchKernelGGL(k_group_queue_prepare, dim3(kXcds), dim3(kTailThreads), 0, s, t.hdr, t.gq,
                       t.n, t.groups, (const uint32_t *)t.sidx, (const uint32_t *)interactions);
    if (gquad)
        hipLaunchKernelGGL((k_bh_group<false, true>), dim3(blocks), dim3(kGroup), 0, s, t.n,
                           (const uint32_t *)t.sidx, x, y, z, m, (const NodeF32 *)t.rec,
                           (const uint2 *)t.groups, theta, (float)eps2, t.icap, ax, ay, az,
                           interactions, t.hdr, t.gq, CorrArgs{}, gquad);
    else
        hipLaunchKernelGGL((k_bh_group<false, false>), dim3(blocks), dim3(kGroup), 0, s, t.n,
                           (const uint32_t *)t.sidx, x, y, z, m, (const NodeF32 *)t.rec,
                           (const uint2 *)t.groups, theta, (float)eps2, t.icap, ax, ay, az,
                           interactions, t.hdr, t.gq, CorrArgs{}, (const float4 *)nullptr);
}

void launch_bh_group_corr(const TreeBufs &t, const double *x, const double *y, const double *z,
                          const double *m, double eps2, double theta, const CorrArgs &corr,
                          const float4 *gquad, double *ax, double *ay, double *az,
                          uint32_t *interactions, hipStream_t s) {
    const uint32_t blocks = std::min<uint32_t>(kGroupBlocks, t.n);
    hipLaunchKernelGGL(k_group_queue_prepare, dim3(kXcds), dim3(kTailThreads), 0, s, t.hdr, t.gq,
                       t.n, t.groups, (const uint32_t *)t.sidx, (const uint32_t *)interactions);
    if (gquad)
        hipLaunchKernelGGL((k_bh_group<true, true>), dim3(blocks), dim3(kGroup), 0, s, t.n,
                           (const uint32_t *)t.sidx, x, y, z, m, (const NodeF32 *)t.rec,
                           (const uint2 *)t.groups, theta, (float)eps2, t.icap, ax, ay, az,
                           interactions, t.hdr, t.gq, corr, gquad);
    else
        hipLaunchKernelGGL((k_bh_group<true, false>), dim3(blocks), dim3(kGroup), 0, s, t.n,
                           (const uint32_t *)t.sidx, x, y, z, m, (const NodeF32 *)t.rec,
                           (const uint2 *)t.groups, theta, (float)eps2, t.icap, ax, ay, az,
                           interactions, t.hdr, t.gq, corr, (const float4 *)nullptr);
}

void launch_ens_items(uint32_t nq, const uint32_t *qorder, const uint32_t *items,
                      uint32_t *item_off, uint32_t *cursor, uint32_t *status, hipStream_t s) {
    hipLaunchKernelGGL(k_ens_items, dim3(1), dim3(64), 0, s, nq, qorder, items, item_off, cursor,
                       status);
}

void launch_ens_forces(const EnsForceArgs &a, uint32_t waves, hipStream_t s) {
    hipLaunchKernelGGL(k_ens_forces, dim3(std::max<uint32_t>(1, waves)), dim3(kGroup), 0, s, a);
}

}  // namespace nb

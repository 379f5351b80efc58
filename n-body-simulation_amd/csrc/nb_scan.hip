// nb_scan.hip -- device-wide exclusive scan and stable LSD radix sort.
//
// Both are written for 64-lane wavefronts: prefix sums use __shfl_up inside a
// wave and LDS across the 4 waves of a 256-thread block; the radix sort ranks
// keys stably with an 8-ballot "match any" on the 8-bit digit, so no
// per-thread histogram is needed.  The octree build uses the scan for node
// numbering and the sort for the octant keys (rocPRIM/hipCUB are not used).
#include "nb_internal.h"

namespace nb {
namespace {

constexpr int kThreads = 256;
constexpr int kWaves = kThreads / 64;
constexpr int kScanItems = 8;
constexpr int kScanTile = kThreads * kScanItems;
constexpr int kRadixItems = 16;
constexpr int kRadixTile = kThreads * kRadixItems;

__device__ inline uint32_t lane_id() { return threadIdx.x & 63u; }

__device__ inline uint32_t wave_inclusive_scan(uint32_t v) {
    const uint32_t lane = lane_id();
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        const uint32_t t = __shfl_up(v, d, 64);
        if (lane >= (uint32_t)d) v += t;
    }
    return v;
}

// Exclusive prefix of v over the block's 256 threads; *total = block sum.
__device__ inline uint32_t block_exclusive_scan(uint32_t v, uint32_t *sh, uint32_t *total) {
    const uint32_t inc = wave_inclusive_scan(v);
    const uint32_t w = threadIdx.x >> 6;
    if (lane_id() == 63) sh[w] = inc;
    __syncthreads();
    uint32_t off = 0, tot = 0;
#pragma unroll
    for (int i = 0; i < kWaves; ++i) {
        const uint32_t s = sh[i];
        off += (uint32_t)i < w ? s : 0u;
        tot += s;
    }
    __syncthreads();
    *total = tot;
    return off + inc - v;
}

__global__ void __launch_bounds__(kThreads) k_scan_reduce(const uint32_t *__restrict__ in,
                                                          uint32_t n, uint32_t *part) {
    __shared__ uint32_t sh[kWaves];
    const uint64_t base = (uint64_t)blockIdx.x * kScanTile;
    uint32_t s = 0;
#pragma unroll
    for (int k = 0; k < kScanItems; ++k) {
        const uint64_t i = base + (uint64_t)k * kThreads + threadIdx.x;
        if (i < n) s += in[i];
    }
    uint32_t tot;
    block_exclusive_scan(s, sh, &tot);
    if (threadIdx.x == 0) part[blockIdx.x] = tot;
}

// One block: exclusive scan of the nb tile sums in place; *total = grand total.
__global__ void __launch_bounds__(kThreads) k_scan_parts(uint32_t *part, uint32_t nb,
                                                         uint32_t *total) {
    __shared__ uint32_t sh[kWaves];
    const uint32_t per = (nb + kThreads - 1) / kThreads;
    const uint32_t b0 = threadIdx.x * per;
    uint32_t s = 0;
    for (uint32_t i = 0; i < per; ++i)
        if (b0 + i < nb) s += part[b0 + i];
    uint32_t tot;
    uint32_t off = block_exclusive_scan(s, sh, &tot);
    for (uint32_t i = 0; i < per; ++i) {
        if (b0 + i < nb) {
            const uint32_t v = part[b0 + i];
            part[b0 + i] = off;
            off += v;
        }
    }
    if (threadIdx.x == 0 && total) *total = tot;
}

// Each thread scans kScanItems consecutive values; out may alias in.
__global__ void __launch_bounds__(kThreads) k_scan_down(const uint32_t *in, uint32_t *out,
                                                        uint32_t n, const uint32_t *part) {
    __shared__ uint32_t sh[kWaves];
    const uint64_t base = (uint64_t)blockIdx.x * kScanTile + (uint64_t)threadIdx.x * kScanItems;
    uint32_t v[kScanItems];
    uint32_t s = 0;
#pragma unroll
    for (int k = 0; k < kScanItems; ++k) {
        const uint64_t i = base + k;
        v[k] = i < n ? in[i] : 0u;
        s += v[k];
    }
    uint32_t tot;
    uint32_t off = block_exclusive_scan(s, sh, &tot) + part[blockIdx.x];
#pragma unroll
    for (int k = 0; k < kScanItems; ++k) {
        const uint64_t i = base + k;
        if (i < n) out[i] = off;
        off += v[k];
    }
}

// ---- radix sort -------------------------------------------------------------

// Per-block 256-bin histogram of digit (key >> shift) & 255, stored bin-major
// (hist[d * nb + block]) so that one exclusive scan gives every block its
// output offset per digit in stable order.
__global__ void __launch_bounds__(kThreads) k_radix_hist(const uint64_t *__restrict__ keys,
                                                         uint32_t n, int shift, uint32_t *hist,
                                                         uint32_t nb) {
    __shared__ uint32_t h[256];
    h[threadIdx.x] = 0;
    __syncthreads();
    const uint64_t base = (uint64_t)blockIdx.x * kRadixTile;
#pragma unroll 4
    for (int r = 0; r < kRadixItems; ++r) {
        const uint64_t i = base + (uint64_t)r * kThreads + threadIdx.x;
        if (i < n) atomicAdd(&h[(uint32_t)(keys[i] >> shift) & 255u], 1u);
    }
    __syncthreads();
    hist[(uint64_t)threadIdx.x * nb + blockIdx.x] = h[threadIdx.x];
}

// Stable scatter of one tile of kRadixTile keys.  The tile is loaded at once
// (kRadixItems rounds of 256 keys in index order) and ranked round by round:
// a key's place among its digit's keys in the tile is (same digit in earlier
// rounds) + (in earlier waves) + (in lower lanes of its own wave).  The tile
// is then reordered in LDS by digit and written out digit run by digit run,
// so consecutive threads write consecutive addresses.
__global__ void __launch_bounds__(kThreads) k_radix_scatter(
    const uint64_t *__restrict__ keys, const uint32_t *__restrict__ vals,
    uint64_t *__restrict__ okeys, uint32_t *__restrict__ ovals, uint32_t n, int shift,
    const uint32_t *__restrict__ offs, uint32_t nb) {
    __shared__ uint32_t running[256];  // tile histogram, then the per-digit cursor
    __shared__ uint32_t dbase[256];    // first tile position of digit d
    __shared__ uint32_t gbase[256];    // output position of the tile's first digit-d key
    __shared__ uint32_t wsum[kWaves];
    __shared__ uint32_t wcnt[2][kWaves][256];
    __shared__ uint64_t lk[kRadixTile];
    __shared__ uint32_t lv[kRadixTile];
    const uint32_t tid = threadIdx.x, lane = lane_id(), w = tid >> 6;
    const uint64_t tile = (uint64_t)blockIdx.x * kRadixTile;
    const uint64_t rest = (uint64_t)n - tile;  // > 0: one block per started tile
    const uint32_t tn = rest < (uint64_t)kRadixTile ? (uint32_t)rest : (uint32_t)kRadixTile;
    uint64_t kk[kRadixItems];
    uint32_t vv[kRadixItems], dg[kRadixItems];
#pragma unroll
    for (int r = 0; r < kRadixItems; ++r) {
        const uint32_t j = (uint32_t)r * kThreads + tid;
        kk[r] = j < tn ? keys[tile + j] : ~0ull;
        vv[r] = j < tn ? vals[tile + j] : 0u;
    }
    running[tid] = 0;
    gbase[tid] = offs[(uint64_t)tid * nb + blockIdx.x];
#pragma unroll
    for (int b = 0; b < 2; ++b)
#pragma unroll
        for (int k = 0; k < kWaves; ++k) wcnt[b][k][tid] = 0;
    __syncthreads();
#pragma unroll
    for (int r = 0; r < kRadixItems; ++r) {
        dg[r] = (uint32_t)(kk[r] >> shift) & 255u;
        if ((uint32_t)r * kThreads + tid < tn) atomicAdd(&running[dg[r]], 1u);
    }
    __syncthreads();
    {
        // exclusive scan of the tile histogram: thread tid owns digit tid
        const uint32_t c = running[tid];
        uint32_t inc = c;
#pragma unroll
        for (int off = 1; off < 64; off <<= 1) {
            const uint32_t o = __shfl_up(inc, off, 64);
            if (lane >= (uint32_t)off) inc += o;
        }
        if (lane == 63) wsum[w] = inc;
        __syncthreads();
        uint32_t ex = inc - c;
        for (uint32_t k = 0; k < w; ++k) ex += wsum[k];
        dbase[tid] = ex;
        running[tid] = ex;
    }
    __syncthreads();
    const uint64_t lt_mask = lane == 0 ? 0ull : (~0ull >> (64u - lane));
#pragma unroll
    for (int r = 0; r < kRadixItems; ++r) {
        const int buf = r & 1;
        const bool valid = (uint32_t)r * kThreads + tid < tn;
        const uint32_t d = valid ? dg[r] : 255u;
        uint64_t peers = __ballot(valid);
#pragma unroll
        for (int b = 0; b < 8; ++b) {
            const bool bit = (d >> b) & 1u;
            const uint64_t bal = __ballot(bit);
            peers &= bit ? bal : ~bal;
        }
        const uint32_t rank = (uint32_t)__popcll(peers & lt_mask);
        if (valid && rank == 0) wcnt[buf][w][d] = (uint32_t)__popcll(peers);
        __syncthreads();
        {
            // thread tid owns digit tid: turn the per-wave counts into positions
            uint32_t run = running[tid];
#pragma unroll
            for (int k = 0; k < kWaves; ++k) {
                const uint32_t c = wcnt[buf][k][tid];
                wcnt[buf][k][tid] = run;
                run += c;
                wcnt[buf ^ 1][k][tid] = 0;  // next round's counters
            }
            running[tid] = run;
        }
        __syncthreads();
        if (valid) {
            const uint32_t p = wcnt[buf][w][d] + rank;
            lk[p] = kk[r];
            lv[p] = vv[r];
        }
    }
    __syncthreads();
    for (uint32_t j = tid; j < tn; j += kThreads) {
        const uint64_t key = lk[j];
        const uint32_t d = (uint32_t)(key >> shift) & 255u;
        const uint32_t pos = gbase[d] + (j - dbase[d]);
        okeys[pos] = key;
        ovals[pos] = lv[j];
    }
}

}  // namespace

uint32_t scan_scratch_words(uint32_t n) { return ceil_div(n ? n : 1, kScanTile) + 1; }

void scan_exclusive_u32(const uint32_t *in, uint32_t *out, uint32_t n, uint32_t *total_dev,
                        uint32_t *scratch, uint32_t scratch_cap, hipStream_t s) {
    if (n == 0) return;
    const uint32_t nb = ceil_div(n, kScanTile);
    if (nb > scratch_cap) return;  // caller sized the scratch with scan_scratch_words
    hipLaunchKernelGGL(k_scan_reduce, dim3(nb), dim3(kThreads), 0, s, in, n, scratch);
    hipLaunchKernelGGL(k_scan_parts, dim3(1), dim3(kThreads), 0, s, scratch, nb, total_dev);
    hipLaunchKernelGGL(k_scan_down, dim3(nb), dim3(kThreads), 0, s, in, out, n,
                       (const uint32_t *)scratch);
}

uint32_t radix_blocks_for(uint32_t n) { return ceil_div(n ? n : 1, kRadixTile); }

void radix_sort_u64_u32(uint64_t *keys, uint32_t *vals, uint64_t *keys_alt, uint32_t *vals_alt,
                        uint32_t n, int bits, uint32_t *hist, uint32_t hist_blocks,
                        uint32_t *scan_scratch, uint32_t scan_cap, hipStream_t s) {
    if (n <= 1) return;
    const uint32_t nb = radix_blocks_for(n);
    if (nb > hist_blocks) return;
    int passes = (bits + 7) / 8;
    if (passes & 1) ++passes;  // even pass count: the result lands back in keys/vals
    uint64_t *ka = keys, *kb = keys_alt;
    uint32_t *va = vals, *vb = vals_alt;
    for (int p = 0; p < passes; ++p) {
        const int shift = 8 * p;
        hipLaunchKernelGGL(k_radix_hist, dim3(nb), dim3(kThreads), 0, s, ka, n, shift, hist, nb);
        scan_exclusive_u32(hist, hist, 256u * nb, nullptr, scan_scratch, scan_cap, s);
        hipLaunchKernelGGL(k_radix_scatter, dim3(nb), dim3(kThreads), 0, s, ka, va, kb, vb, n,
                           shift, (const uint32_t *)hist, nb);
        uint64_t *tk = ka; ka = kb; kb = tk;
        uint32_t *tv = va; va = vb; vb = tv;
    }
}

}  // namespace nb

// nb_scan.hip -- device-wide exclusive scan and stable LSD radix sort.
//
// Both are written for 64-lane wavefronts: prefix sums use __shfl_up inside a
// wave and LDS across the 4 waves of a 256-thread block; the radix sort ranks
// keys stably with a 9-ballot "match any" on the 9-bit digit, so no
// per-thread histogram is needed, and takes one launch per digit (plus one
// counting launch per sort).  The octree build uses the scan for node
// numbering and the sort for the octant keys (rocPRIM/hipCUB are not used).
#include "nb_internal.h"

namespace nb {
namespace {

constexpr int kThreads = 256;
constexpr int kWaves = kThreads / 64;
constexpr int kScanItems = 8;
constexpr int kScanTile = kThreads * kScanItems;
constexpr int kRadixTile = 4096;       // keys per radix tile
constexpr int kRadixTileBig = 8192;    // keys per tile from kRadixBigMin keys on
constexpr uint32_t kRadixBigMin = 1u << 21;
constexpr int kPassThreads = 512;
constexpr int kPassWaves = kPassThreads / 64;

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

// ---- radix sort: one sweep per digit -----------------------------------------
//
// Each pass is one k_radix_pass launch: a block takes the next tile id from a
// counter, publishes its tile's per-digit counts, ranks its keys in LDS, and
// learns how many keys of each digit the earlier tiles hold from their
// published counts (decoupled look-back: a tile that has finished its
// look-back publishes an inclusive prefix, so later tiles stop there).  Tile
// ids are handed out in the order blocks start, so every tile a block waits on
// belongs to a block that is running and that publishes its counts before it
// waits on anything: every wait ends.  The digit totals (which do not depend
// on the order) come from k_radix_count for pass 0 and from pass p - 1 for
// pass p, which counts the next digit of the keys it has loaded anyway.
//
// Digits are 9 bits wide: 7 passes cover the 63-bit octree keys (an odd
// count, so the result lands in the scratch buffers and the caller swaps).
//
// Control words (u32):
//   [0, 7 * 512)           digit totals of pass p at p * kBins
//   kCtlTiles + p          tile counter of pass p
//   kCtlStatus + plane     two status planes of nb x kBins words, plane p & 1
//                          for pass p; pass p clears its tile's row of the
//                          other plane for pass p + 1, and k_radix_count
//                          clears plane 0 (an odd pass count leaves it used)
// Pass p > 0 zeroes the totals and the tile counter of pass p - 1 (that
// launch has ended), and pass 0 those of the last pass.  Between sorts every
// word is therefore zero except the last pass's totals and tile counter,
// which pass 0 clears before pass passes - 2 adds to them (allocation sets
// everything to zero).  A status word is 0 (not yet published),
// kStatAgg | tile count, or kStatPrefix | count in tiles [0, t].  Counts stay
// below 2^30: octree contexts hold fewer than 2^29 bodies.

constexpr int kDigitBits = 9;
constexpr uint32_t kBins = 1u << kDigitBits;
constexpr int kMaxPasses = 7;
static_assert(kMaxPasses * kDigitBits >= 63, "the passes cover the keys");
static_assert(kBins == (uint32_t)kPassThreads, "thread tid owns digit tid");
constexpr uint32_t kCtlTiles = kMaxPasses * kBins;
constexpr uint32_t kCtlStatus = kCtlTiles + 64;  // planes start 256-byte aligned
constexpr uint32_t kStatAgg = 1u << 30;
constexpr uint32_t kStatPrefix = 2u << 30;
constexpr uint32_t kStatCount = kStatAgg - 1;
constexpr int kLookWindow = 32;               // predecessor tiles read per look-back step
constexpr uint32_t kLookSpinLimit = 1u << 22; // a safety net: the waits end (see above)
constexpr int kMinPasses = 4;  // the control-word rotation above needs 3 or more
constexpr int kCountThreads = 1024;
constexpr uint32_t kCountBlocks = 256;

__device__ inline uint32_t status_load(const uint32_t *p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ inline void status_store(uint32_t *p, uint32_t v) {
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// A sort that cannot run reports it in the caller's status word.
__global__ void k_raise_flag(uint32_t *flags, uint32_t f) {
    if (threadIdx.x == 0) atomicOr(flags, f);
}

// Totals of the lowest digit of the keys: total[digit].  Also clears the
// status plane of the first pass, plane0[0, plane_words): the previous sort's
// last pass used it when the pass count is odd.
__global__ void __launch_bounds__(kCountThreads) k_radix_count(const uint64_t *__restrict__ keys,
                                                               uint32_t n,
                                                               uint32_t *__restrict__ total,
                                                               uint32_t *__restrict__ plane0,
                                                               uint64_t plane_words) {
    __shared__ uint32_t h[kBins];
    if (threadIdx.x < kBins) h[threadIdx.x] = 0;
    __syncthreads();
    const uint64_t stride = (uint64_t)gridDim.x * kCountThreads;
    for (uint64_t i = (uint64_t)blockIdx.x * kCountThreads + threadIdx.x; i < plane_words; i += stride)
        plane0[i] = 0;
#pragma unroll 4
    for (uint64_t i = (uint64_t)blockIdx.x * kCountThreads + threadIdx.x; i < n; i += stride)
        atomicAdd(&h[(uint32_t)keys[i] & (kBins - 1)], 1u);
    __syncthreads();
    if (threadIdx.x < kBins && h[threadIdx.x]) atomicAdd(&total[threadIdx.x], h[threadIdx.x]);
}

// Exclusive prefix over the block's threads of one value per thread (thread
// tid owns digit tid); ws holds kPassWaves words and is not reused.
__device__ inline uint32_t digit_exclusive_scan(uint32_t c, uint32_t *ws) {
    const uint32_t inc = wave_inclusive_scan(c);
    const uint32_t w = threadIdx.x >> 6;
    if (lane_id() == 63) ws[w] = inc;
    __syncthreads();
    uint32_t ex = inc - c;
    for (uint32_t k = 0; k < w; ++k) ex += ws[k];
    return ex;
}

// Keys of digit d in tiles [0, t), t > 0: walk back from tile t - 1,
// kLookWindow tiles per step, adding published counts up to and including the
// first inclusive prefix.  False if a predecessor never published (cannot
// happen, see above; the caller then raises a flag instead of waiting forever).
__device__ inline bool look_back(const uint32_t *stat, uint32_t t, uint32_t d, uint32_t *excl) {
    uint32_t sum = 0, spins = 0;
    int64_t j = (int64_t)t - 1;
    for (;;) {
        uint32_t v[kLookWindow];
#pragma unroll
        for (int k = 0; k < kLookWindow; ++k)
            v[k] = j - k >= 0 ? status_load(stat + (uint64_t)(j - k) * kBins + d) : kStatPrefix;
        bool pending = false;
#pragma unroll
        for (int k = 0; k < kLookWindow; ++k) pending |= v[k] == 0;
        while (pending) {
            if (++spins > kLookSpinLimit) return false;
            __builtin_amdgcn_s_sleep(2);
            pending = false;
#pragma unroll
            for (int k = 0; k < kLookWindow; ++k) {
                if (v[k] == 0) v[k] = status_load(stat + (uint64_t)(j - k) * kBins + d);
                pending |= v[k] == 0;
            }
        }
        bool found = false;
#pragma unroll
        for (int k = 0; k < kLookWindow; ++k) {
            if (!found) {
                sum += v[k] & kStatCount;
                found = v[k] >= kStatPrefix;
            }
        }
        if (found) {
            *excl = sum;
            return true;
        }
        j -= kLookWindow;
    }
}

// One stable pass over digit (key >> 9 pass) & 511 of a tile of kTile
// keys.  Wave w holds the tile's keys [w * kTile / 8, (w + 1) * kTile / 8) and
// ranks them 64 at a time against its own digit counters, with no barrier between
// rounds: a key's place among its digit's keys in the tile is (same digit in
// earlier waves) + (in earlier rounds of its wave) + (in lower lanes of its
// round), so the order is stable.  The tile is reordered in LDS by digit and
// written out digit run by digit run, so consecutive threads write
// consecutive addresses.
template <int kTile>
__global__ void __launch_bounds__(kPassThreads) k_radix_pass(
    const uint64_t *__restrict__ keys, const uint32_t *__restrict__ vals,
    uint64_t *__restrict__ okeys, uint32_t *__restrict__ ovals, uint32_t n, int pass, int passes,
    uint32_t *__restrict__ ctl, uint32_t nb, uint32_t *__restrict__ flags) {
    __shared__ uint32_t wcnt[kPassWaves][kBins];  // digit-d keys of wave w, then their first place
    __shared__ uint32_t next_hist[kBins];          // tile histogram of the next pass's digit
    __shared__ uint32_t dbase[kBins];              // first tile position of digit d
    __shared__ uint32_t gbase[kBins];              // output position of the tile's first digit-d key
    __shared__ uint32_t wsum[2][kPassWaves];
    __shared__ uint64_t lk[kTile];
    __shared__ uint32_t lv[kTile];
    __shared__ uint32_t tile_id, stuck;
    const uint32_t tid = threadIdx.x, lane = lane_id(), w = tid >> 6;
    const bool owner = tid < kBins;  // thread tid owns digit tid
    const int shift = kDigitBits * pass;
    uint32_t *total = ctl + pass * kBins;
    uint32_t *stat = ctl + kCtlStatus + (uint64_t)(pass & 1) * kBins * nb;
    uint32_t *stat_next = ctl + kCtlStatus + (uint64_t)((pass + 1) & 1) * kBins * nb;
    const bool has_next = pass + 1 < passes;
    if (tid == 0) {
        tile_id = atomicAdd(ctl + kCtlTiles + pass, 1u);
        stuck = 0;
    }
    for (uint32_t i = tid; i < kPassWaves * kBins; i += kPassThreads) (&wcnt[0][0])[i] = 0;
    if (owner) next_hist[tid] = 0;
    __syncthreads();
    const uint32_t t = tile_id;
    if (t >= nb) {  // only with corrupt control words: report it and touch nothing
        if (tid == 0) atomicOr(flags, kFlagSortWait);
        return;  // block-uniform, before any further barrier
    }
    if (t == 0 && owner) {
        // the previous pass (or the last pass of the previous sort) has ended
        const int prev = pass > 0 ? pass - 1 : passes - 1;
        ctl[prev * kBins + tid] = 0;
        if (tid == 0) ctl[kCtlTiles + prev] = 0;
    }
    const uint64_t tile = (uint64_t)t * kTile;
    const uint64_t rest = (uint64_t)n - tile;  // > 0: one block per tile
    const uint32_t tn = rest < (uint64_t)kTile ? (uint32_t)rest : (uint32_t)kTile;
    constexpr int kPassItems = kTile / kPassThreads;  // keys per lane
    const uint32_t wbase = w * (kPassItems * 64u);
    uint64_t kk[kPassItems];
    uint32_t vv[kPassItems], dg[kPassItems], rk[kPassItems];
#pragma unroll
    for (int r = 0; r < kPassItems; ++r) {
        const uint32_t j = wbase + (uint32_t)r * 64u + lane;
        kk[r] = j < tn ? keys[tile + j] : ~0ull;
        vv[r] = j < tn ? vals[tile + j] : 0u;
    }
    const uint64_t lt_mask = lane == 0 ? 0ull : (~0ull >> (64u - lane));
#pragma unroll
    for (int r = 0; r < kPassItems; ++r) {
        const bool valid = wbase + (uint32_t)r * 64u + lane < tn;
        const uint32_t d = (uint32_t)(kk[r] >> shift) & (kBins - 1);
        dg[r] = d;
        uint64_t peers = __ballot(valid);
#pragma unroll
        for (int b = 0; b < kDigitBits; ++b) {
            const bool bit = (d >> b) & 1u;
            const uint64_t bal = __ballot(bit);
            peers &= bit ? bal : ~bal;
        }
        const uint32_t rank = (uint32_t)__popcll(peers & lt_mask);
        // a wave's LDS accesses complete in order: the next round reads this write
        const uint32_t before = wcnt[w][d];
        rk[r] = before + rank;
        if (valid && rank == 0) wcnt[w][d] = before + (uint32_t)__popcll(peers);
        if (valid && has_next)
            atomicAdd(&next_hist[(uint32_t)(kk[r] >> (shift + kDigitBits)) & (kBins - 1)], 1u);
    }
    __syncthreads();
    uint32_t c = 0, gtot = 0;
    if (owner) {
        // per-wave counts of digit tid become offsets among the tile's digit-tid keys
#pragma unroll
        for (int k = 0; k < kPassWaves; ++k) {
            const uint32_t cnt = wcnt[k][tid];
            wcnt[k][tid] = c;
            c += cnt;
        }
        status_store(stat + (uint64_t)t * kBins + tid, (t == 0 ? kStatPrefix : kStatAgg) | c);
        stat_next[(uint64_t)t * kBins + tid] = 0;
        if (has_next && next_hist[tid]) atomicAdd(total + kBins + tid, next_hist[tid]);
        gtot = total[tid];
    }
    const uint32_t ex = digit_exclusive_scan(c, wsum[0]);
    const uint32_t goff = digit_exclusive_scan(gtot, wsum[1]);
    if (owner) {
        dbase[tid] = ex;
#pragma unroll
        for (int k = 0; k < kPassWaves; ++k) wcnt[k][tid] += ex;
    }
    __syncthreads();
#pragma unroll
    for (int r = 0; r < kPassItems; ++r) {
        if (wbase + (uint32_t)r * 64u + lane < tn) {
            const uint32_t p = wcnt[w][dg[r]] + rk[r];
            lk[p] = kk[r];
            lv[p] = vv[r];
        }
    }
    if (owner) {
        uint32_t excl = 0;
        if (t > 0) {
            if (!look_back(stat, t, tid, &excl)) stuck = 1;
            // published even when stuck, so that later tiles do not wait as well
            status_store(stat + (uint64_t)t * kBins + tid, kStatPrefix | (excl + c));
        }
        gbase[tid] = goff + excl;
    }
    __syncthreads();
    if (stuck) {
        if (tid == 0) atomicOr(flags, kFlagSortWait);
        return;
    }
    for (uint32_t j = tid; j < tn; j += kPassThreads) {
        const uint64_t key = lk[j];
        const uint32_t d = (uint32_t)(key >> shift) & (kBins - 1);
        const uint32_t pos = gbase[d] + (j - dbase[d]);
        if (pos < n) {
            okeys[pos] = key;
            ovals[pos] = lv[j];
        }
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

// tiles of a sort of n keys at the smaller tile size (the most control words)
uint32_t radix_blocks_for(uint32_t n) { return ceil_div(n ? n : 1, kRadixTile); }

uint32_t radix_ctl_words(uint32_t n) { return kCtlStatus + 2u * kBins * radix_blocks_for(n); }

SortEnd radix_sort_u64_u32(uint64_t *keys, uint32_t *vals, uint64_t *keys_alt,
                           uint32_t *vals_alt, uint32_t n, int bits, uint32_t *ctl,
                           uint32_t *flags, hipStream_t s) {
    if (n >= kStatAgg || bits <= 0 || bits > kDigitBits * kMaxPasses) {
        // counts of 2^30 keys and more do not fit the status words: nothing
        // is sorted, and the caller's status word says so
        hipLaunchKernelGGL(k_raise_flag, dim3(1), dim3(64), 0, s, flags, kFlagSortSize);
        return SortEnd::kRejected;
    }
    if (n <= 1) return SortEnd::kInPlace;
    // larger tiles for large sorts: half the look-back chain and longer digit
    // runs in the scatter, once there are enough tiles to fill the device
    const bool big = n >= kRadixBigMin;
    const uint32_t nb = ceil_div(n, big ? kRadixTileBig : kRadixTile);
    const int passes = std::max((bits + kDigitBits - 1) / kDigitBits, kMinPasses);
    hipLaunchKernelGGL(k_radix_count, dim3(std::min(nb, kCountBlocks)), dim3(kCountThreads), 0, s,
                       (const uint64_t *)keys, n, ctl, ctl + kCtlStatus, (uint64_t)kBins * nb);
    uint64_t *ka = keys, *kb = keys_alt;
    uint32_t *va = vals, *vb = vals_alt;
    for (int p = 0; p < passes; ++p) {
        if (big)
            hipLaunchKernelGGL(k_radix_pass<kRadixTileBig>, dim3(nb), dim3(kPassThreads), 0, s,
                               (const uint64_t *)ka, (const uint32_t *)va, kb, vb, n, p, passes,
                               ctl, nb, flags);
        else
            hipLaunchKernelGGL(k_radix_pass<kRadixTile>, dim3(nb), dim3(kPassThreads), 0, s,
                               (const uint64_t *)ka, (const uint32_t *)va, kb, vb, n, p, passes,
                               ctl, nb, flags);
        uint64_t *tk = ka; ka = kb; kb = tk;
        uint32_t *tv = va; va = vb; vb = tv;
    }
    // an odd count ends in the scratch buffers
    return (passes & 1) ? SortEnd::kInAlt : SortEnd::kInPlace;
}

}  // namespace nb

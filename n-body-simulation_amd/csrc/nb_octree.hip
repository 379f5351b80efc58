// nb_octree.hip -- Barnes-Hut octree build for gfx950.
//
// The reference builds its octree by lock-based top-down insertion
// (ParallelOctreeTopDownSubtrees.cpp:95-813): one work-group inserts every
// body down to level 7, then one work-group per subtree finishes the job.
// The resulting tree is the unique octree in which a cell is split iff it
// holds two or more bodies; only the node ids depend on the insertion race.
// This file builds that same tree without locks:
//
//  1. AABB (FP64 min/max including the origin) and cubify, on the device.
//  2. Per body, the first 21 octant digits with the reference's own FP64
//     classification chain: mid = cellMin + edge/2, digit = 4[y>mid_y] +
//     2[x>mid_x] + [z<mid_z], child min moved to mid on the x/y bits and on a
//     cleared z bit.  Digits are packed most significant first into a 63-bit
//     key, so key order is the reference's in-order body order.
//  3. Stable radix sort of (key, body).
//  4. Runs of bodies sharing all 21 digits get digits 21..41 and are sorted
//     inside the run (rare: bodies closer than edge/2^21).
//  5. lcp[i] = depth of the deepest cell holding sorted bodies i and i+1.
//     The internal cells are exactly the pairs (i, depth l) with
//     lcp[i-1] < l <= lcp[i]: each cell is numbered by the first adjacent pair
//     of its body range.  An exclusive scan of the chain lengths numbers the
//     internal cells k in depth-first pre-order; internal cell k stores its 8
//     children at node ids 1 + 8k + digit.
//  6. Parents come from a galloping search for the first body of the parent
//     cell; children count themselves into their parent.
//  7. Centre of mass bottom-up: the last child to arrive at a cell sums the 8
//     children in digit order 0..7 from 0.0 -- the reference's summation
//     (BarnesHutOctree.cpp:299-311), so every sum is bit-identical.  Cells
//     inside a block of 4096 sorted bodies are finished by that block's
//     workgroup (workgroup-scope atomics: agent-scope ones would write back
//     and invalidate the L2 of an XCD at every arrival); the few cells that
//     cross block boundaries are finished level by level by one workgroup.
//  8. Node records for traversal (COM = numerator / mass as the reference
//     divides per visit, edge = rootEdge * 2^-depth; an FP32 hi/lo record for
//     the group traversal).
#include <algorithm>
#include <cmath>

#include "nb_internal.h"

namespace nb {
namespace {

constexpr int kThreads = 256;

__device__ inline double sycl_min(double a, double b) { return b < a ? b : a; }
__device__ inline double sycl_max(double a, double b) { return a < b ? b : a; }

__device__ inline void raise_flag(TreeHdr *h, uint32_t f) { atomicOr(&h->status, f); }

// ---- 1. AABB ---------------------------------------------------------------
__global__ void __launch_bounds__(kThreads) k_build_begin(TreeHdr *h) {
    if (threadIdx.x == 0 && blockIdx.x == 0) {
        h->internal_count = 0;
        h->max_depth = 0;
        h->status = 0;
        h->deep_runs = 0;
        h->stack_peak = 0;
        h->group_count = 0;
        h->group_next = 0;
        h->cross_count = 0;
    }
}

// Block partial extrema.  Like the reference's partials (BarnesHutOctree.cpp:
// 58-72, value-initialised to 0.0) every partial starts at the origin.
__global__ void __launch_bounds__(kThreads) k_aabb_partial(const double *__restrict__ x,
                                                           const double *__restrict__ y,
                                                           const double *__restrict__ z,
                                                           uint32_t n, double *part,
                                                           TreeHdr *h) {
    double mn[3] = {0.0, 0.0, 0.0}, mx[3] = {0.0, 0.0, 0.0};
    bool bad = false;
    for (uint64_t i = (uint64_t)blockIdx.x * kThreads + threadIdx.x; i < n;
         i += (uint64_t)gridDim.x * kThreads) {
        const double p[3] = {x[i], y[i], z[i]};
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            bad |= !isfinite(p[c]);
            mn[c] = sycl_min(mn[c], p[c]);
            mx[c] = sycl_max(mx[c], p[c]);
        }
    }
    if (bad) raise_flag(h, kFlagAabb);
    // min/max are exact, so any reduction order gives the reference's result
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) {
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            mn[c] = sycl_min(mn[c], __shfl_xor(mn[c], off, 64));
            mx[c] = sycl_max(mx[c], __shfl_xor(mx[c], off, 64));
        }
    }
    __shared__ double sh[kThreads / 64][6];
    const int w = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0) {
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            sh[w][c] = mn[c];
            sh[w][3 + c] = mx[c];
        }
    }
    __syncthreads();
    if (threadIdx.x < 6) {
        double v = sh[0][threadIdx.x];
        for (int k = 1; k < kThreads / 64; ++k)
            v = threadIdx.x < 3 ? sycl_min(v, sh[k][threadIdx.x]) : sycl_max(v, sh[k][threadIdx.x]);
        part[(uint64_t)threadIdx.x * gridDim.x + blockIdx.x] = v;
    }
}

// Cubify the extents ext = {min_x, min_y, min_z, max_x, max_y, max_z}
// (BarnesHutOctree.cpp:161-190) and store the root cube.
__device__ inline void cubify(const double ext[6], TreeHdr *h) {
    double min_x = ext[0], min_y = ext[1], min_z = ext[2];
    double max_x = ext[3], max_y = ext[4], max_z = ext[5];
    const double x_length = fabs(max_x - min_x);
    const double y_length = fabs(max_y - min_y);
    const double z_length = fabs(max_z - min_z);
    // std::max(a, b) == (a < b) ? b : a
    const double yz = y_length < z_length ? z_length : y_length;
    const double maxEdgeLength = x_length < yz ? yz : x_length;
    if (maxEdgeLength == x_length) {
        min_z = min_z - ((maxEdgeLength - z_length) / 2);
        min_y = min_y - ((maxEdgeLength - y_length) / 2);
        max_z = max_z + ((maxEdgeLength - z_length) / 2);
        max_y = max_y + ((maxEdgeLength - y_length) / 2);
    } else if (maxEdgeLength == y_length) {
        min_x = min_x - ((maxEdgeLength - x_length) / 2);
        min_z = min_z - ((maxEdgeLength - z_length) / 2);
        max_x = max_x + ((maxEdgeLength - x_length) / 2);
        max_z = max_z + ((maxEdgeLength - z_length) / 2);
    } else if (maxEdgeLength == z_length) {
        min_x = min_x - ((maxEdgeLength - x_length) / 2);
        min_y = min_y - ((maxEdgeLength - y_length) / 2);
        max_x = max_x + ((maxEdgeLength - x_length) / 2);
        max_y = max_y + ((maxEdgeLength - y_length) / 2);
    } else {
        raise_flag(h, kFlagAabb);
    }
    h->minx = min_x;
    h->miny = min_y;
    h->minz = min_z;
    h->maxx = max_x;
    h->maxy = max_y;
    h->maxz = max_z;
    h->edge = maxEdgeLength;
}

// Final reduction on one thread per component, then one thread cubifies.
__global__ void __launch_bounds__(kThreads) k_aabb_final(const double *part, uint32_t np,
                                                         TreeHdr *h) {
    __shared__ double sh[kThreads / 64][6];
    __shared__ double ext[6];
    // every partial already includes the origin, so 0.0 is a neutral start
    double v[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
    for (uint32_t k = threadIdx.x; k < np; k += kThreads) {
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            v[c] = sycl_min(v[c], part[(uint64_t)c * np + k]);
            v[3 + c] = sycl_max(v[3 + c], part[(uint64_t)(3 + c) * np + k]);
        }
    }
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) {
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            v[c] = sycl_min(v[c], __shfl_xor(v[c], off, 64));
            v[3 + c] = sycl_max(v[3 + c], __shfl_xor(v[3 + c], off, 64));
        }
    }
    if ((threadIdx.x & 63) == 0)
        for (int c = 0; c < 6; ++c) sh[threadIdx.x >> 6][c] = v[c];
    __syncthreads();
    if (threadIdx.x < 6) {
        double r = sh[0][threadIdx.x];
        for (int w = 1; w < kThreads / 64; ++w)
            r = threadIdx.x < 3 ? sycl_min(r, sh[w][threadIdx.x]) : sycl_max(r, sh[w][threadIdx.x]);
        ext[threadIdx.x] = r;
    }
    __syncthreads();
    if (threadIdx.x != 0) return;
    const double e6[6] = {ext[0], ext[1], ext[2], ext[3], ext[4], ext[5]};
    cubify(e6, h);
}

// ---- 2. keys ---------------------------------------------------------------
__device__ inline uint64_t key_from(Cell &c, double px, double py, double pz) {
    uint64_t key = 0;
#pragma unroll
    for (int l = 0; l < kKeyLevels; ++l) key = (key << 3) | descend(c, px, py, pz);
    return key;
}

// Keys, identity permutation, and the packed body record {x, y, z, m} that
// the centre-of-mass and node-record passes gather (one 32-byte load).
__global__ void __launch_bounds__(kThreads) k_keys(const double *__restrict__ x,
                                                   const double *__restrict__ y,
                                                   const double *__restrict__ z,
                                                   const double *__restrict__ m, uint32_t n,
                                                   const TreeHdr *__restrict__ h,
                                                   uint64_t *skey, uint32_t *sidx,
                                                   double4 *body4) {
    const uint32_t i = blockIdx.x * kThreads + threadIdx.x;
    if (i >= n) return;
    Cell c{h->minx, h->miny, h->minz, h->edge};
    const double px = x[i], py = y[i], pz = z[i];
    skey[i] = key_from(c, px, py, pz);
    sidx[i] = i;
    body4[i] = make_double4(px, py, pz, m[i]);
}

// ---- 4./5. LCP and deep runs --------------------------------------------------
__device__ inline int lcp_depth(uint64_t a, uint64_t b) {
    // number of equal leading digits of two 63-bit keys (bit 63 is always 0)
    return (__clzll((long long)(a ^ b)) - 1) / 3;
}

// lcp of adjacent pair i (i + 1 < n); the first pair of a run of equal keys
// extends and sorts the run.  Shared by the one-system build and the
// workgroup-per-system ensemble build.
__device__ inline void lcp_at(uint32_t i, const double *__restrict__ x,
                              const double *__restrict__ y, const double *__restrict__ z,
                              uint32_t n, const Cell &root, TreeHdr *h, uint64_t *skey,
                              uint32_t *sidx, uint64_t *klo, int32_t *lcp) {
    const uint64_t k0 = skey[i], k1 = skey[i + 1];
    if (k0 != k1) {
        lcp[i] = lcp_depth(k0, k1);
        return;
    }
    if (i > 0 && skey[i - 1] == k0) return;  // not the first pair of the run
    // run [a, b] of bodies sharing all 21 digits: extend their keys by 21
    // more digits from the common level-21 cell and sort the run by them.
    const uint32_t a = i;
    uint32_t b = i + 1;
    while (b + 1 < n && skey[b + 1] == k0) ++b;
    atomicAdd(&h->deep_runs, 1u);
    Cell c0 = root;
    {
        const uint32_t q = sidx[a];
        key_from(c0, x[q], y[q], z[q]);  // walk to the shared level-21 cell
    }
    for (uint32_t p = a; p <= b; ++p) {
        Cell c = c0;
        const uint32_t q = sidx[p];
        klo[p] = key_from(c, x[q], y[q], z[q]);
    }
    // insertion sort by the second key word; stable, so equal keys keep body order
    for (uint32_t p = a + 1; p <= b; ++p) {
        const uint64_t kv = klo[p];
        const uint32_t iv = sidx[p];
        uint32_t q = p;
        while (q > a && klo[q - 1] > kv) {
            klo[q] = klo[q - 1];
            sidx[q] = sidx[q - 1];
            --q;
        }
        klo[q] = kv;
        sidx[q] = iv;
    }
    for (uint32_t p = a; p < b; ++p) {
        const uint64_t d = klo[p] ^ klo[p + 1];
        if (d == 0) {
            raise_flag(h, kFlagDepth);  // not separated within kMaxDepth digits
            lcp[p] = kMaxDepth - 1;
        } else {
            lcp[p] = kKeyLevels + lcp_depth(klo[p], klo[p + 1]);
        }
    }
}

__global__ void __launch_bounds__(kThreads) k_lcp(const double *__restrict__ x,
                                                  const double *__restrict__ y,
                                                  const double *__restrict__ z, uint32_t n,
                                                  TreeHdr *h, uint64_t *skey, uint32_t *sidx,
                                                  uint64_t *klo, int32_t *lcp) {
    const uint32_t i = blockIdx.x * kThreads + threadIdx.x;
    if (i + 1 >= n) return;
    const Cell root{h->minx, h->miny, h->minz, h->edge};
    lcp_at(i, x, y, z, n, root, h, skey, sidx, klo, lcp);
}

// chain[i] = number of internal cells owned by pair i = max(0, lcp[i] - lcp[i-1])
__device__ inline void chain_at(uint32_t i, const int32_t *__restrict__ lcp, uint32_t *chain) {
    const int cp = i > 0 ? lcp[i - 1] : -1;
    const int cn = lcp[i];
    chain[i] = cn > cp ? (uint32_t)(cn - cp) : 0u;
}

__global__ void __launch_bounds__(kThreads) k_chain(const int32_t *__restrict__ lcp, uint32_t n,
                                                    uint32_t *chain) {
    const uint32_t i = blockIdx.x * kThreads + threadIdx.x;
    if (i + 1 >= n) return;
    chain_at(i, lcp, chain);
}

// h->max_depth = depth of the deepest leaf = 1 + the deepest cell that two
// sorted neighbours share (max lcp), for n >= 2; 0 for one body.  Run on
// demand (tree queries, cell moments), not in every build: a grid-stride
// maximum per workgroup, and the device-scope RMW only when it raises the
// stored value.
constexpr int kDepthBlocks = 256;
__global__ void __launch_bounds__(kThreads) k_max_depth(const int32_t *__restrict__ lcp, uint32_t n,
                                                        TreeHdr *h) {
    __shared__ int wmax[kThreads / 64];
    int v = -1;
    for (uint32_t i = blockIdx.x * kThreads + threadIdx.x; i + 1 < n; i += gridDim.x * kThreads)
        v = max(v, lcp[i]);
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) v = max(v, __shfl_xor(v, off, 64));
    if ((threadIdx.x & 63u) == 0) wmax[threadIdx.x >> 6] = v;
    __syncthreads();
    if (threadIdx.x == 0) {
        int b = wmax[0];
        for (int w = 1; w < kThreads / 64; ++w) b = max(b, wmax[w]);
        const uint32_t d = (uint32_t)(b + 1);
        if (b >= 0 &&
            d > __hip_atomic_load(&h->max_depth, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT))
            atomicMax(&h->max_depth, d);
    }
}

// ---- 6. topology --------------------------------------------------------------
// The sorted keys as the topology pass reads them.  `win` (optional) holds
// skey[w0, w1) in LDS: the searches of one block stay mostly inside it.
struct KeyView {
    const uint64_t *skey;
    const uint64_t *klo;
    const uint64_t *win = nullptr;
    uint32_t w0 = 0, w1 = 0;
    __device__ inline uint64_t key(uint32_t p) const {
        return p - w0 < w1 - w0 ? win[p - w0] : skey[p];
    }
    __device__ inline uint32_t digit(uint32_t p, int l) const {
        return l < kKeyLevels ? (uint32_t)(key(p) >> (3 * (kKeyLevels - 1 - l))) & 7u
                              : (uint32_t)(klo[p] >> (3 * (2 * kKeyLevels - 1 - l))) & 7u;
    }
    // do sorted bodies p and q share their first v digits?
    __device__ inline bool same(uint32_t p, uint32_t q, int v) const {
        if (v <= 0) return true;
        const uint64_t kp = key(p), kq = key(q);
        if (v <= kKeyLevels) return ((kp ^ kq) >> (63 - 3 * v)) == 0;
        return kp == kq && ((klo[p] ^ klo[q]) >> (63 - 3 * (v - kKeyLevels))) == 0;
    }
    // first sorted position of the depth-v cell that holds sorted body i
    __device__ inline uint32_t cell_start(uint32_t i, int v) const {
        if (v <= 0 || i == 0) return 0;
        int64_t in = i;       // shares the prefix
        int64_t out = -1;     // does not (or -1)
        uint32_t step = 1;
        while (true) {
            const int64_t q = (int64_t)i - step;
            if (q < 0) break;
            if (!same((uint32_t)q, i, v)) {
                out = q;
                break;
            }
            in = q;
            step <<= 1;
        }
        while (in - out > 1) {
            const int64_t mid = out + (in - out) / 2;
            if (same((uint32_t)mid, i, v)) in = mid; else out = mid;
        }
        return (uint32_t)in;
    }
};

__global__ void __launch_bounds__(kThreads) k_fill(TreeHdr *h, uint32_t icap, int32_t *ref,
                                                   uint32_t *spanning, uint32_t spanning_cap) {
    const uint32_t I = min(h->internal_count, icap);
    const uint64_t nodes = 1 + 8ull * I;
    const uint64_t stride = (uint64_t)gridDim.x * kThreads;
    for (uint64_t j = (uint64_t)blockIdx.x * kThreads + threadIdx.x; j < nodes; j += stride)
        ref[j] = kRefEmpty;
    for (uint64_t j = (uint64_t)blockIdx.x * kThreads + threadIdx.x; j < spanning_cap; j += stride)
        spanning[j] = kNoCell;
}

// Internal cells owned by pair i and the leaf of sorted body i.  A cell is
// "local" (int_local 1) when its bodies stay inside one com_block-body block
// of the sorted order; the others go to the table
// spanning[block * kMaxDepth + depth], with int_local 2 if they also cross a
// boundary of the com_super-body super-blocks and 0 if not.
// No atomics: every cell, node slot and table entry has exactly one writer.
__device__ inline void link_at(uint32_t i, uint32_t n, uint32_t icap, TreeHdr *h, KeyView kv,
                               const uint32_t *__restrict__ sidx, const int32_t *__restrict__ lcp,
                               const uint32_t *__restrict__ off, int32_t *ref, uint8_t *int_depth,
                               uint32_t *int_parent, uint32_t *int_node, uint32_t *body_parent,
                               uint8_t *int_local, uint32_t *spanning, uint32_t spanning_cap,
                               uint32_t *int_first, uint32_t com_block, uint32_t com_super) {
    if (n == 1) {  // the root is a leaf holding the only body (max_depth stays 0)
        ref[0] = ref_of_body(sidx[0]);
        body_parent[0] = 0xffffffffu;
        return;
    }
    const int cp = i > 0 ? lcp[i - 1] : -1;
    const int cn = i + 1 < n ? lcp[i] : -1;
    auto owner_index = [&](uint32_t a, int v) -> uint32_t {
        const int ca = a > 0 ? lcp[a - 1] : -1;
        return off[a] + (uint32_t)(v - ca - 1);
    };
    // internal cells owned by pair i: depths cp+1 .. cn
    for (int l = cp + 1; l <= cn; ++l) {
        const uint32_t k = off[i] + (uint32_t)(l - cp - 1);
        if (k >= icap) {
            raise_flag(h, kFlagNodeOverflow);
            break;
        }
        int_depth[k] = (uint8_t)l;
        int_first[k] = i;
        // The cell's bodies start at i; it crosses into the next COM block iff
        // that block's first body shares its l-digit prefix with body i.
        const uint32_t next_block = (i / com_block + 1) * com_block;
        const bool local = !(next_block < n && kv.same(next_block, i, l));
        if (local) {
            int_local[k] = 1;
        } else {
            // the same test at the next super-block boundary
            const uint32_t next_super = (i / com_super + 1) * com_super;
            int_local[k] = next_super < n && kv.same(next_super, i, l) ? 2 : 0;
            // the block of body i owns at most one crossing cell per depth:
            // the depth-l cell holding the block's last and the next block's
            // first body
            const uint64_t slot = (uint64_t)(i / com_block) * kMaxDepth + (uint32_t)l;
            if (slot < spanning_cap) spanning[slot] = k;
            else raise_flag(h, kFlagNodeOverflow);
        }
        if (l == 0) {
            int_node[k] = 0;
            int_parent[k] = 0xffffffffu;
            ref[0] = (int32_t)k;
            continue;
        }
        const uint32_t pk = l > cp + 1 ? k - 1 : owner_index(kv.cell_start(i, cp), cp);
        if (pk >= icap) {
            raise_flag(h, kFlagNodeOverflow);
            break;
        }
        const uint32_t node = 1u + 8u * pk + kv.digit(i, l - 1);
        int_node[k] = node;
        int_parent[k] = pk;
        ref[node] = (int32_t)k;
    }
    // the body's leaf: one level below the deepest cell it shares with a neighbour
    const int D = cp > cn ? cp : cn;
    const uint32_t pk = owner_index(kv.cell_start(i, D), D);
    if (pk >= icap) {
        raise_flag(h, kFlagNodeOverflow);
        body_parent[i] = 0xffffffffu;
        return;
    }
    const uint32_t leaf = 1u + 8u * pk + kv.digit(i, D);
    ref[leaf] = ref_of_body(sidx[i]);
    body_parent[i] = pk;
    // the deepest leaf, 1 + max lcp, is taken from the lcp array (max_depth_of)
}

constexpr uint32_t kLinkBack = 768;  // keys before a link block's own kept in LDS
__global__ void __launch_bounds__(kThreads) k_link(uint32_t n, uint32_t icap, TreeHdr *h,
                                                   KeyView kv, const uint32_t *__restrict__ sidx,
                                                   const int32_t *__restrict__ lcp,
                                                   const uint32_t *__restrict__ off,
                                                   int32_t *ref, uint8_t *int_depth,
                                                   uint32_t *int_parent, uint32_t *int_node,
                                                   uint32_t *body_parent, uint8_t *int_local,
                                                   uint32_t *spanning, uint32_t spanning_cap,
                                                   uint32_t *int_first) {
    // the block's keys and kLinkBack before them in LDS
    __shared__ uint64_t win[kLinkBack + kThreads];
    const uint32_t b0 = blockIdx.x * kThreads;
    const uint32_t w0 = b0 > kLinkBack ? b0 - kLinkBack : 0u;
    const uint32_t w1 = min(n, b0 + (uint32_t)kThreads);
    for (uint32_t p = w0 + threadIdx.x; p < w1; p += kThreads) win[p - w0] = kv.skey[p];
    __syncthreads();
    kv.win = win;
    kv.w0 = w0;
    kv.w1 = w1;
    const uint32_t i = b0 + threadIdx.x;
    if (i >= n) return;
    link_at(i, n, icap, h, kv, sidx, lcp, off, ref, int_depth, int_parent, int_node, body_parent,
            int_local, spanning, spanning_cap, int_first, kComBlock, kComSuper);
}

// ---- 7. centre of mass --------------------------------------------------------
// Sum of the 8 children of internal cell pk in digit order 0..7 from 0.0
// (BarnesHutOctree.cpp:299-311): a body child adds (x m, y m, z m, m), a cell
// child its sums.  Empty children add +0.0, which changes no running sum that
// starts at +0.0, so they are skipped.  All eight children are loaded first,
// so their loads are in flight together; the additions keep digit order.
__device__ inline void add_children(const int32_t (&r)[8], const double4 (&v)[8],
                                    const uint32_t (&c)[8], InternalSums &out, uint32_t &cnt) {
    double sx = 0, sy = 0, sz = 0, sm = 0;
    cnt = 0;
#pragma unroll
    for (int d = 0; d < 8; ++d) {
        if (r[d] != kRefEmpty) {
            sx += v[d].x;
            sy += v[d].y;
            sz += v[d].z;
            sm += v[d].w;
            cnt += c[d];
        }
    }
    out = InternalSums{sx, sy, sz, sm};
}

// the term of a body child: (x m, y m, z m, m)
__device__ inline double4 body_term(const double4 p) {
    return make_double4(p.x * p.w, p.y * p.w, p.z * p.w, p.w);
}

// Traversal records and groups, written by the centre-of-mass kernels: a cell,
// once summed, writes the records of its 8 child slots (their data is at
// hand) and, if it holds more than rule.group_max bodies, marks the traversal
// groups among its children (a child cell with at most rule.group_max
// bodies, or a body).  The kernel that sums the root writes the root's
// record, and its group if the root holds at most rule.group_max bodies.
struct RecOut {
    double G;
    int f64, f32;  // FP64 records (cm, edge); FP32 records and traversal groups
    double4 *cm;
    double *edge;
    NodeF32 *rec;
    uint32_t *group_size, *group_slot;
    GroupRule rule;  // group size and packing limit
};

// Record centre of a child from its term: a cell's sums, or a body's
// (x m, y m, z m, m).  The reference divides numerator by mass at every visit
// (BarnesHutAlgorithm.cpp:351-353); the quotient is the same.
__device__ inline double4 term_com(const double4 v) {
    return make_double4(v.x / v.w, v.y / v.w, v.z / v.w, v.w);
}

__device__ inline void put_record(const RecOut &o, uint64_t j, int32_t r, const double4 c,
                                  double e) {
    if (o.f64) {
        o.cm[j] = c;
        o.edge[j] = e;
    }
    if (o.f32) {
        NodeF32 q;
        q.x = (float)c.x;
        q.y = (float)c.y;
        q.z = (float)c.z;
        q.gm = (float)(o.G * c.w);
        q.ref = r;
        q.lx = (float)(c.x - (double)q.x);
        q.ly = (float)(c.y - (double)q.y);
        q.lz = (float)(c.z - (double)q.z);
        o.rec[j] = q;
    }
}

// a traversal group of `size` bodies starts at sorted position p
__device__ inline void put_group(const RecOut &o, uint32_t p, uint32_t size) {
    if (o.f32) {
        o.group_size[p] = size;
        o.group_slot[p] = 1u;
    }
}

// Traversal groups among the children of a cell that holds more than
// group_max bodies; first: the cell's first sorted position, m[d]: child d's
// body count (0 if empty).  Each child with at most group_max bodies is a
// group; in a cell of at most pack_max bodies, runs of such children are
// packed in digit order into groups of at most group_max bodies instead
// (siblings are consecutive in the sorted order; a child with more bodies
// ends a run).
__device__ inline void put_child_groups(const RecOut &o, uint32_t first, const uint32_t (&m)[8]) {
    const uint32_t gmax = o.rule.group_max;
    uint32_t total = 0;
#pragma unroll
    for (int d = 0; d < 8; ++d) total += m[d];
    const bool pack = total <= o.rule.pack_max;
    uint32_t pos = first, g0 = first, gs = 0;
#pragma unroll
    for (int d = 0; d < 8; ++d) {
        if (m[d] == 0) continue;
        if (gs && (!pack || m[d] > gmax || gs + m[d] > gmax)) {
            put_group(o, g0, gs);
            gs = 0;
        }
        if (m[d] <= gmax) {
            if (gs == 0) g0 = pos;
            gs += m[d];
        }
        pos += m[d];
    }
    if (gs) put_group(o, g0, gs);
}

// Records and groups of the children of cell pk (depth `depth`, first sorted
// position `first`), given their refs, terms and body counts; the root's own
// record and group when pk is the root.
__device__ inline void cell_outputs(const RecOut &o, uint32_t pk, uint32_t depth, uint32_t first,
                                    double root_edge, const int32_t (&r)[8],
                                    const double4 (&v)[8], const uint32_t (&c)[8],
                                    const double4 s, uint32_t cnt) {
    const uint64_t base = 1ull + 8ull * pk;
    const double ce = ldexp(root_edge, -(int)(depth + 1));
#pragma unroll
    for (int d = 0; d < 8; ++d) {
        if (r[d] == kRefEmpty)
            put_record(o, base + d, kRefEmpty, make_double4(0.0, 0.0, 0.0, 0.0), 0.0);
        else
            put_record(o, base + d, r[d], term_com(v[d]), r[d] >= 0 ? ce : 0.0);
    }
    if (cnt > o.rule.group_max) put_child_groups(o, first, c);
    if (depth == 0) {
        put_record(o, 0, (int32_t)pk, term_com(s), root_edge);
        if (cnt <= o.rule.group_max) put_group(o, first, cnt);
    }
}

// The global-memory path: the cell's 8 children summed in digit order from 0.0,
// then the children's records and groups.
__device__ inline void sum_children_out(uint32_t pk, uint32_t depth, uint32_t first,
                                        double root_edge, const double4 *__restrict__ body4,
                                        const int32_t *__restrict__ ref, InternalSums *sums,
                                        uint32_t *counts, TreeHdr *h, const RecOut &o) {
    const uint64_t base = 1ull + 8ull * pk;
    int32_t r[8];
    double4 v[8];
    uint32_t c[8];
#pragma unroll
    for (int d = 0; d < 8; ++d) r[d] = ref[base + d];
#pragma unroll
    for (int d = 0; d < 8; ++d) {
        v[d] = make_double4(0.0, 0.0, 0.0, 0.0);
        c[d] = 0;
        if (r[d] >= 0) {
            const InternalSums q = sums[r[d]];
            v[d] = make_double4(q.x, q.y, q.z, q.m);
            c[d] = counts[r[d]];
        } else if (r[d] != kRefEmpty) {
            v[d] = body_term(body4[body_of_ref(r[d])]);
            c[d] = 1;
        }
    }
    InternalSums q;
    uint32_t cnt;
    add_children(r, v, c, q, cnt);
    sums[pk] = q;
    counts[pk] = cnt;
    if (q.m == 0) raise_flag(h, kFlagZeroMass);
    cell_outputs(o, pk, depth, first, root_edge, r, v, c, make_double4(q.x, q.y, q.z, q.m), cnt);
}

// Centre of mass, level by level from the deepest cells up: a cell's sums
// (its children in digit order) need only deeper cells, so every cell of one
// depth can be summed at once, with a barrier between depths.
constexpr int kComThreads = 1024;

// Cells [kb, ke) bucketed by depth into lvl[kb ..) (cells with keep(k) only),
// then sum(k) for each, deepest depth first.  cnt/cur: kMaxDepth + 1 and
// kMaxDepth LDS words.  Every thread of the workgroup calls it.
template <typename Keep, typename Sum>
__device__ inline void com_levels(uint32_t kb, uint32_t ke, Keep keep,
                                  const uint8_t *__restrict__ int_depth, uint32_t *lvl,
                                  uint32_t *cnt, uint32_t *cur, Sum sum) {
    const uint32_t t = threadIdx.x, nt = blockDim.x;
    for (uint32_t d = t; d <= (uint32_t)kMaxDepth; d += nt) cnt[d] = 0;
    __syncthreads();
    for (uint32_t k = kb + t; k < ke; k += nt)
        if (keep(k)) atomicAdd(&cnt[min((uint32_t)int_depth[k], (uint32_t)kMaxDepth - 1)], 1u);
    __syncthreads();
    if (t == 0) {
        uint32_t run = 0;
        for (int d = 0; d < kMaxDepth; ++d) {
            const uint32_t c = cnt[d];
            cnt[d] = run;
            cur[d] = run;
            run += c;
        }
        cnt[kMaxDepth] = run;
    }
    __syncthreads();
    for (uint32_t k = kb + t; k < ke; k += nt)
        if (keep(k))
            lvl[kb + atomicAdd(&cur[min((uint32_t)int_depth[k], (uint32_t)kMaxDepth - 1)], 1u)] = k;
    __syncthreads();
    for (int d = kMaxDepth - 1; d >= 0; --d) {
        const uint32_t lb = cnt[d], le = cnt[d + 1];
        if (lb == le) continue;  // the same for every thread: LDS, read after a barrier
        for (uint32_t j = lb + t; j < le; j += nt) sum(lvl[kb + j]);
        __syncthreads();
    }
}

// Pass 1: one workgroup per kComBlock consecutive (sorted) bodies sums the
// cells that lie inside its block.  Pair i owns cells [off[i], off[i+1]), so
// the block's cells are one contiguous range [kb, ke).  Up to kComLdsCells
// cells are summed entirely in LDS.  Staged once (all loads in flight
// together): the block's bodies and their input indices in sorted order, and
// per cell its depth, first position and a 16-bit code per child slot
// (empty, body, or the child cell's place in the block).  The children of a
// cell cover consecutive sorted ranges in digit order, so a body child is the
// body at the running position: the cell's first position plus the counts of
// its earlier children; the level pass stores that position in the body's
// code.  The levels touch LDS only (a barrier also waits for global stores).
// A last pass writes the sums, the node records of the cells' child slots and
// the group marks, consecutive threads on consecutive slots.  A larger block
// falls back to the same work in global memory.  The block also clears the
// group marks of its positions before any is set.
constexpr uint32_t kComLdsCells = 576;   // about 490 per block; 537 the most seen (disk, Plummer)
constexpr int kComLocalThreads = 512;
constexpr int kComBodiesPerThread = kComBlock / kComLocalThreads;
constexpr int kComRefsPerThread = 8 * kComLdsCells / kComLocalThreads;  // child slots
static_assert(kComBlock % kComLocalThreads == 0 && (8 * kComLdsCells) % kComLocalThreads == 0,
              "whole staging rounds");
constexpr uint16_t kCodeEmpty = 0xffffu, kCodeBody = 0xfffeu, kCodeFar = 0xfffdu;
constexpr uint16_t kCodeAt = 0x8000u;    // | sorted position - block start: a placed body
static_assert(kComLdsCells < kCodeAt && kComBlock <= 0x7ffdu, "16-bit LDS codes");
// Block blk of one octree (the one-system kernel below, and the ensemble's,
// which runs the blocks of every system in one launch).
__device__ __forceinline__ void com_local_block(
    uint32_t blk, uint32_t n, uint32_t icap, TreeHdr *h, const double4 *__restrict__ body4,
    const uint32_t *__restrict__ sidx, const int32_t *__restrict__ ref,
    const uint32_t *__restrict__ off, const uint8_t *__restrict__ int_depth,
    const uint8_t *__restrict__ int_local, const uint32_t *__restrict__ int_first, uint32_t *lvl,
    InternalSums *sums, uint32_t *counts, const RecOut &o) {
    constexpr uint32_t nt = kComLocalThreads;
    __shared__ uint32_t cnt[kMaxDepth + 1], cur[kMaxDepth];
    __shared__ unsigned long long levels;  // bit d: a local cell of depth d
    __shared__ double4 b_s[kComBlock];     // the block's bodies {x, y, z, m}, sorted order
    __shared__ uint32_t i_s[kComBlock];    // their input indices
    __shared__ double4 c_s[kComLdsCells];  // cell sums {x m, y m, z m, m}
    __shared__ uint32_t n_s[kComLdsCells]; // cell body counts
    __shared__ uint16_t code_s[8 * kComLdsCells];
    __shared__ uint16_t first_s[kComLdsCells];  // first sorted position - block start
    __shared__ uint8_t depth_s[kComLdsCells];   // depth; kMaxDepth if not local
    const uint32_t t = threadIdx.x;
    const uint32_t s = blk * kComBlock;
    const uint32_t e = min(n, s + kComBlock);
    const double root_edge = h->edge;
    {
        // all index loads, then all body loads, in flight together
        uint32_t bi[kComBodiesPerThread];
        double4 b[kComBodiesPerThread];
#pragma unroll
        for (int q = 0; q < kComBodiesPerThread; ++q) {
            const uint32_t i = s + t + q * nt;
            bi[q] = i < e ? sidx[i] : 0u;
        }
#pragma unroll
        for (int q = 0; q < kComBodiesPerThread; ++q)
            if (s + t + q * nt < e) b[q] = body4[bi[q]];
#pragma unroll
        for (int q = 0; q < kComBodiesPerThread; ++q) {
            const uint32_t i = s + t + q * nt;
            if (i >= e) continue;
            if (b[q].w == 0) raise_flag(h, kFlagZeroMass);
            b_s[i - s] = b[q];
            i_s[i - s] = bi[q];
            if (o.f32) {
                o.group_size[i] = 0;
                o.group_slot[i] = 0;
            }
        }
    }
    if (n < 2) {  // a lone body is the root leaf: no internal cell
        __syncthreads();
        if (t == 0) {
            put_record(o, 0, ref_of_body(i_s[0]), term_com(body_term(b_s[0])), 0.0);
            put_group(o, 0, 1);
        }
        return;
    }
    const uint32_t I = min(h->internal_count, icap);
    const uint32_t kb = s < n - 1 ? min(off[s], I) : I;
    const uint32_t ke = e < n - 1 ? min(off[e], I) : I;
    const uint32_t nc = ke - kb;
    const uint32_t nbody = e - s;
    if (nc > kComLdsCells) {  // the same for the whole workgroup
        com_levels(kb, ke, [&](uint32_t k) { return int_local[k] == 1; }, int_depth, lvl, cnt,
                   cur, [&](uint32_t k) {
                       sum_children_out(k, int_depth[k], int_first[k], root_edge, body4, ref,
                                        sums, counts, h, o);
                   });
        return;
    }
    if (t == 0) levels = 0;
    for (uint32_t c = t; c < nc; c += nt) {
        const bool local = int_local[kb + c] == 1;
        depth_s[c] = local ? (uint8_t)min((uint32_t)int_depth[kb + c], (uint32_t)kMaxDepth - 1)
                           : (uint8_t)kMaxDepth;
        first_s[c] = (uint16_t)(int_first[kb + c] - s);
    }
    {
        int32_t r[kComRefsPerThread];
#pragma unroll
        for (int q = 0; q < kComRefsPerThread; ++q) {
            const uint32_t j = t + q * nt;
            r[q] = j < 8 * nc ? ref[1ull + 8ull * kb + j] : kRefEmpty;
        }
#pragma unroll
        for (int q = 0; q < kComRefsPerThread; ++q) {
            const uint32_t j = t + q * nt;
            if (j >= 8 * nc) continue;
            uint16_t code = kCodeEmpty;
            if (r[q] >= 0) {
                const uint32_t rl = (uint32_t)r[q] - kb;
                code = rl < nc ? (uint16_t)rl : kCodeFar;  // far: not for a local cell
            } else if (r[q] != kRefEmpty) {
                code = kCodeBody;
            }
            code_s[j] = code;
        }
    }
    __syncthreads();
    for (uint32_t c = t; c < nc; c += nt)
        if (depth_s[c] < kMaxDepth) atomicOr(&levels, 1ull << depth_s[c]);
    __syncthreads();
    const unsigned long long lv = levels;  // the same for every thread
    for (int d = kMaxDepth - 1; d >= 0; --d) {
        if (!((lv >> d) & 1ull)) continue;
        for (uint32_t c = t; c < nc; c += nt) {
            if (depth_s[c] != (uint32_t)d) continue;
            double sx = 0, sy = 0, sz = 0, sm = 0;
            uint32_t cn = 0;
            uint32_t pos = first_s[c];
            // digit order from 0.0, as sum_children_out
#pragma unroll
            for (int dd = 0; dd < 8; ++dd) {
                const uint16_t code = code_s[8 * c + dd];
                if (code == kCodeEmpty) continue;
                double4 v;
                uint32_t m = 1;
                if (code < kComLdsCells) {
                    v = c_s[code];
                    m = n_s[code];
                } else if (code == kCodeBody && pos < nbody) {
                    v = body_term(b_s[pos]);
                    code_s[8 * c + dd] = (uint16_t)(kCodeAt | pos);
                } else {  // not for a local cell: the child from global memory
                    const int32_t r = ref[1ull + 8ull * (kb + c) + dd];
                    if (r >= 0) {
                        const InternalSums g = sums[r];
                        v = make_double4(g.x, g.y, g.z, g.m);
                        m = counts[r];
                    } else {
                        v = body_term(body4[body_of_ref(r)]);
                    }
                }
                sx += v.x;
                sy += v.y;
                sz += v.z;
                sm += v.w;
                cn += m;
                pos += m;
            }
            c_s[c] = make_double4(sx, sy, sz, sm);
            n_s[c] = cn;
            if (sm == 0) raise_flag(h, kFlagZeroMass);
        }
        __syncthreads();
    }
    // sums, group marks and child records of the local cells
    for (uint32_t c = t; c < nc; c += nt) {
        if (depth_s[c] >= kMaxDepth) continue;
        const double4 g = c_s[c];
        sums[kb + c] = InternalSums{g.x, g.y, g.z, g.w};
        counts[kb + c] = n_s[c];
        if (n_s[c] > o.rule.group_max) {
            uint32_t m[8];
#pragma unroll
            for (int dd = 0; dd < 8; ++dd) {
                const uint16_t code = code_s[8 * c + dd];
                m[dd] = code == kCodeEmpty ? 0u : 1u;
                if (code < kComLdsCells) {
                    m[dd] = n_s[code];
                } else if (code == kCodeFar) {  // not for a local cell: global memory
                    m[dd] = counts[ref[1ull + 8ull * (kb + c) + dd]];
                }
            }
            put_child_groups(o, s + first_s[c], m);
        }
    }
    for (uint32_t j = t; j < 8 * nc; j += nt) {
        const uint32_t c = j >> 3;
        const uint32_t dc = depth_s[c];
        if (dc >= kMaxDepth) continue;
        const uint64_t slot = 1ull + 8ull * kb + j;
        const uint16_t code = code_s[j];
        const uint32_t pc = n_s[c];
        if (code == kCodeEmpty) {
            put_record(o, slot, kRefEmpty, make_double4(0.0, 0.0, 0.0, 0.0), 0.0);
        } else if (code < kComLdsCells) {
            put_record(o, slot, (int32_t)(kb + code), term_com(c_s[code]),
                       ldexp(root_edge, -(int)(dc + 1)));
        } else if ((code & kCodeAt) && code < kCodeFar) {
            const uint32_t pos = code & ~kCodeAt;
            put_record(o, slot, ref_of_body(i_s[pos]), term_com(body_term(b_s[pos])), 0.0);
        } else {  // not for a local cell: the child from global memory
            const int32_t r = ref[slot];
            if (r >= 0) {
                const InternalSums g = sums[r];
                put_record(o, slot, r, term_com(make_double4(g.x, g.y, g.z, g.m)),
                           ldexp(root_edge, -(int)(dc + 1)));
            } else {
                put_record(o, slot, r, term_com(body_term(body4[body_of_ref(r)])), 0.0);
            }
        }
        if (dc == 0 && (j & 7u) == 0) {  // the root (the block holds every body)
            put_record(o, 0, (int32_t)(kb + c), term_com(c_s[c]), root_edge);
            if (pc <= o.rule.group_max) put_group(o, s + first_s[c], pc);
        }
    }
}

__global__ void __launch_bounds__(kComLocalThreads) k_com_local(
    uint32_t n, uint32_t icap, TreeHdr *h, const double4 *__restrict__ body4,
    const uint32_t *__restrict__ sidx, const int32_t *__restrict__ ref,
    const uint32_t *__restrict__ off, const uint8_t *__restrict__ int_depth,
    const uint8_t *__restrict__ int_local, const uint32_t *__restrict__ int_first, uint32_t *lvl,
    InternalSums *sums, uint32_t *counts, RecOut o) {
    com_local_block(blockIdx.x, n, icap, h, body4, sidx, ref, off, int_depth, int_local, int_first,
                    lvl, sums, counts, o);
}

// Pass 2: the cells crossing block boundaries (the table
// spanning[block * kMaxDepth + depth]), in two launches.  want = 0: one
// workgroup per super-block of kComSuperBlocks blocks sums the cells that
// cross block boundaries inside it and lists those crossing its end (at most
// one per depth) in `cross`; want = 2: one workgroup sums the listed cells,
// which cross super-block boundaries.  A workgroup stages up to kSpanCells cells
// in LDS: the term of every child (a body, or a cell summed by an earlier
// pass) is loaded once, all loads in flight together, and its record is
// written then; a child summed by this workgroup is named by its LDS id
// (cid[k], written first).  The levels, deepest first, then read LDS only.
// More cells fall back to the same sums in global memory.
constexpr uint32_t kSpanCells = 448;
constexpr uint32_t kChildEmpty = 0xffffffffu;
constexpr uint32_t kChildHere = 0x80000000u;  // | LDS id of a child summed here
// Super-block (want 0) or cross list (want 2) sb of one octree.
__device__ __forceinline__ void com_span_block(
    uint32_t sb, uint32_t nblocks, uint32_t rows, uint32_t want, TreeHdr *h,
    const double4 *__restrict__ body4, const int32_t *__restrict__ ref,
    const uint32_t *__restrict__ spanning, const uint8_t *__restrict__ int_local,
    const uint32_t *__restrict__ int_first, uint32_t *cid, InternalSums *sums, uint32_t *counts,
    uint32_t *cross, uint32_t cross_cap, const RecOut &o) {
    __shared__ uint32_t ncell, cnt[kMaxDepth + 1], cur[kMaxDepth];
    __shared__ uint32_t nx, xbase, xl[kMaxDepth];  // want 0: this super-block's crossing cells
    __shared__ uint32_t list[kSpanCells];  // cell index by LDS id
    __shared__ uint32_t fst[kSpanCells];   // first sorted position
    __shared__ uint8_t ldep[kSpanCells];
    __shared__ uint16_t order[kSpanCells];  // LDS ids bucketed by depth
    __shared__ double4 cv[8 * kSpanCells];  // child terms
    __shared__ uint32_t cc[8 * kSpanCells]; // child body counts, kChildEmpty, or kChildHere | id
    __shared__ double4 sv[kSpanCells];
    __shared__ uint32_t sn[kSpanCells];
    const uint32_t t = threadIdx.x;
    const uint32_t r0 = sb * rows;
    const uint32_t r1 = min(nblocks, r0 + rows);
    const uint64_t slot0 = (uint64_t)r0 * kMaxDepth, slot1 = (uint64_t)r1 * kMaxDepth;
    const double root_edge = h->edge;
    if (t == 0) {
        ncell = 0;
        nx = 0;
    }
    for (uint32_t d = t; d <= (uint32_t)kMaxDepth; d += kComThreads) cnt[d] = 0;
    __syncthreads();
    auto take = [&](uint32_t k, uint32_t d) {
        const uint32_t id = atomicAdd(&ncell, 1u);
        if (id < kSpanCells) {
            list[id] = k;
            fst[id] = int_first[k];
            ldep[id] = (uint8_t)d;
            cid[k] = id;
            atomicAdd(&cnt[d], 1u);
        }
    };
    // want 0 scans this super-block's table rows and keeps its crossing cells
    // (they all hold the super-block's end: at most one per depth); want 2
    // reads the list of those cells
    const uint32_t ncross = want == 2 ? min(h->cross_count, cross_cap) : 0u;
    if (want == 2) {
        for (uint32_t i = t; i < ncross; i += kComThreads) take(cross[i] >> 6, cross[i] & 63u);
    } else {
        for (uint64_t j = slot0 + t; j < slot1; j += kComThreads) {
            const uint32_t k = spanning[j];
            if (k == kNoCell) continue;
            const uint8_t kind = int_local[k];
            const uint32_t d = (uint32_t)(j % kMaxDepth);
            if (kind == 2) {
                const uint32_t x = atomicAdd(&nx, 1u);
                if (x < (uint32_t)kMaxDepth) xl[x] = (k << 6) | d;
            } else if (kind == want) {
                take(k, d);
            }
        }
    }
    __syncthreads();
    if (want != 2) {
        const uint32_t m = min(nx, (uint32_t)kMaxDepth);
        if (t == 0) xbase = m ? atomicAdd(&h->cross_count, m) : 0u;
        __syncthreads();
        if (t < m) {
            if (xbase + t < cross_cap) cross[xbase + t] = xl[t];
            else raise_flag(h, kFlagNodeOverflow);
        }
        if (t == 0 && nx > (uint32_t)kMaxDepth) raise_flag(h, kFlagNodeOverflow);
    }
    const uint32_t nk = ncell;  // the same for every thread
    if (nk > kSpanCells) {
        for (int d = kMaxDepth - 1; d >= 0; --d) {
            if (want == 2) {
                for (uint32_t i = t; i < ncross; i += kComThreads)
                    if ((cross[i] & 63u) == (uint32_t)d) {
                        const uint32_t k = cross[i] >> 6;
                        sum_children_out(k, (uint32_t)d, int_first[k], root_edge, body4, ref,
                                         sums, counts, h, o);
                    }
            } else {
                for (uint64_t j = slot0 + (uint32_t)d + (uint64_t)t * kMaxDepth; j < slot1;
                     j += (uint64_t)kComThreads * kMaxDepth) {
                    const uint32_t k = spanning[j];
                    if (k != kNoCell && int_local[k] == want)
                        sum_children_out(k, (uint32_t)d, int_first[k], root_edge, body4, ref,
                                         sums, counts, h, o);
                }
            }
            __syncthreads();
        }
        return;
    }
    if (nk == 0) return;
    if (t == 0) {
        uint32_t run = 0;
        for (int d = 0; d < kMaxDepth; ++d) {
            const uint32_t c = cnt[d];
            cnt[d] = run;
            cur[d] = run;
            run += c;
        }
        cnt[kMaxDepth] = run;
    }
    __syncthreads();
    for (uint32_t id = t; id < nk; id += kComThreads) order[atomicAdd(&cur[ldep[id]], 1u)] = (uint16_t)id;
    for (uint32_t p = t; p < 8 * nk; p += kComThreads) {
        const uint32_t id = p >> 3;
        const uint32_t k = list[id];
        const uint64_t j = 1ull + 8ull * k + (p & 7u);
        const int32_t r = ref[j];
        double4 v = make_double4(0.0, 0.0, 0.0, 0.0);
        uint32_t code = kChildEmpty;
        if (r >= 0) {
            if (int_local[r] == want) {
                code = kChildHere | cid[r];  // summed by this workgroup, one level deeper
            } else {
                const InternalSums q = sums[r];
                v = make_double4(q.x, q.y, q.z, q.m);
                code = counts[r];
                put_record(o, j, r, term_com(v), ldexp(root_edge, -(int)(ldep[id] + 1u)));
            }
        } else if (r != kRefEmpty) {
            v = body_term(body4[body_of_ref(r)]);
            code = 1;
            put_record(o, j, r, term_com(v), 0.0);
        } else {
            put_record(o, j, kRefEmpty, v, 0.0);
        }
        cv[p] = v;
        cc[p] = code;
    }
    __syncthreads();
    for (int d = kMaxDepth - 1; d >= 0; --d) {
        const uint32_t lb = cnt[d], le = cnt[d + 1];
        if (lb == le) continue;  // the same for every thread: LDS, read after a barrier
        for (uint32_t j = lb + t; j < le; j += kComThreads) {
            const uint32_t id = order[j];
            double sx = 0, sy = 0, sz = 0, sm = 0;
            uint32_t cn = 0;
#pragma unroll
            for (int dd = 0; dd < 8; ++dd) {
                const uint32_t code = cc[8 * id + dd];
                if (code == kChildEmpty) continue;
                double4 v;
                uint32_t m;
                if (code & kChildHere) {
                    v = sv[code & ~kChildHere];
                    m = sn[code & ~kChildHere];
                } else {
                    v = cv[8 * id + dd];
                    m = code;
                }
                sx += v.x;
                sy += v.y;
                sz += v.z;
                sm += v.w;
                cn += m;
            }
            sv[id] = make_double4(sx, sy, sz, sm);
            sn[id] = cn;
            if (sm == 0) raise_flag(h, kFlagZeroMass);
        }
        __syncthreads();
    }
    // global stores only now (a barrier also waits for them): sums, the
    // records of the children summed here, group marks, the root
    for (uint32_t id = t; id < nk; id += kComThreads) {
        const double4 g = sv[id];
        sums[list[id]] = InternalSums{g.x, g.y, g.z, g.w};
        counts[list[id]] = sn[id];
        if (sn[id] > o.rule.group_max) {
            uint32_t m[8];
#pragma unroll
            for (int dd = 0; dd < 8; ++dd) {
                const uint32_t code = cc[8 * id + dd];
                m[dd] = code == kChildEmpty ? 0u : code & kChildHere ? sn[code & ~kChildHere] : code;
            }
            put_child_groups(o, fst[id], m);
        }
    }
    for (uint32_t p = t; p < 8 * nk; p += kComThreads) {
        const uint32_t id = p >> 3;
        const uint32_t dd = p & 7u;
        const uint32_t code = cc[p];
        const uint32_t pc = sn[id];
        if (code != kChildEmpty && (code & kChildHere))
            put_record(o, 1ull + 8ull * list[id] + dd, (int32_t)list[code & ~kChildHere],
                       term_com(sv[code & ~kChildHere]), ldexp(root_edge, -(int)(ldep[id] + 1u)));
        if (dd == 0 && ldep[id] == 0) {  // the root
            put_record(o, 0, (int32_t)list[id], term_com(sv[id]), root_edge);
            if (pc <= o.rule.group_max) put_group(o, fst[id], pc);
        }
    }
}

__global__ void __launch_bounds__(kComThreads) k_com_span(
    uint32_t nblocks, uint32_t rows, uint32_t want, TreeHdr *h, const double4 *__restrict__ body4,
    const int32_t *__restrict__ ref, const uint32_t *__restrict__ spanning,
    const uint8_t *__restrict__ int_local, const uint32_t *__restrict__ int_first, uint32_t *cid,
    InternalSums *sums, uint32_t *counts, uint32_t *cross, uint32_t cross_cap, RecOut o) {
    com_span_block(blockIdx.x, nblocks, rows, want, h, body4, ref, spanning, int_local, int_first,
                   cid, sums, counts, cross, cross_cap, o);
}

// ---- traversal groups -------------------------------------------------------------
// A group is a cell with at most group_max bodies whose parent holds more, or
// a lone body whose parent holds more.  Groups partition the bodies into
// spatially compact runs of the octree order.  Every sorted position starts
// at most one group: mark the starts, scan, scatter -- the group list comes
// out in octree order, so neighbouring groups share most of their walk.
__global__ void __launch_bounds__(kThreads) k_group_scatter(uint32_t n,
                                                            const uint32_t *__restrict__ group_size,
                                                            const uint32_t *__restrict__ group_slot,
                                                            uint2 *groups) {
    const uint32_t i = blockIdx.x * kThreads + threadIdx.x;
    if (i >= n) return;
    const uint32_t s = group_size[i];
    if (s) groups[group_slot[i]] = make_uint2(i, s);
}

// ---- ensemble: the octrees of many systems ------------------------------------------
constexpr int kForestThreads = 1024;
constexpr int kForestWaves = kForestThreads / 64;
// Per-wave contiguous range [*b, *e) of m items: 16 ranges of whole 64-item
// groups, so a wave reads 64 consecutive items at a time (coalesced) and the
// ranges keep the item order wave by wave.
__device__ inline void wave_range(uint32_t m, uint32_t *b, uint32_t *e) {
    const uint32_t chunk = ((m + kForestWaves - 1) / kForestWaves + 63u) & ~63u;
    const uint32_t w = threadIdx.x >> 6;
    *b = min(m, w * chunk);
    *e = min(m, *b + chunk);
}

// Exclusive scan in place of a[0..m) (global memory, this workgroup only):
// wave totals, their scan, then a running wave scan over each wave's range.
__device__ inline uint32_t wg_scan_array(uint32_t *a, uint32_t m, uint32_t *wsum) {
    const uint32_t t = threadIdx.x, lane = t & 63u, w = t >> 6;
    uint32_t b, e;
    wave_range(m, &b, &e);
    uint32_t s = 0;
    for (uint32_t i = b + lane; i < e; i += 64) s += a[i];
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) s += __shfl_xor(s, off, 64);
    if (lane == 0) wsum[w] = s;
    __syncthreads();
    if (t == 0) {
        uint32_t run = 0;
        for (int k = 0; k < kForestWaves; ++k) {
            const uint32_t c = wsum[k];
            wsum[k] = run;
            run += c;
        }
        wsum[kForestWaves] = run;
    }
    __syncthreads();
    uint32_t carry = wsum[w];
    const uint32_t total = wsum[kForestWaves];
    for (uint32_t g = b; g < e; g += 64) {
        const uint32_t i = g + lane;
        const uint32_t v = i < e ? a[i] : 0u;
        uint32_t inc = v;
#pragma unroll
        for (int off = 1; off < 64; off <<= 1) {
            const uint32_t o = __shfl_up(inc, off, 64);
            if (lane >= (uint32_t)off) inc += o;
        }
        if (i < e) a[i] = carry + inc - v;
        carry += __shfl(inc, 63, 64);
    }
    __syncthreads();
    return total;
}

// 1. Per Barnes-Hut system (one workgroup each): header reset, AABB including
// the origin (every partial starts at 0.0), then cubify into the header.
__global__ void __launch_bounds__(kForestThreads) k_forest_aabb(
    const EnsSystem *__restrict__ sys, const uint32_t *__restrict__ order,
    const double *__restrict__ gx, const double *__restrict__ gy, const double *__restrict__ gz,
    ForestBufs fb) {
    __shared__ double red[kForestWaves][6];
    const uint32_t s = order[blockIdx.x];
    const EnsSystem S = sys[s];
    const uint32_t n = S.n, t = threadIdx.x;
    const double *x = gx + S.body_off, *y = gy + S.body_off, *z = gz + S.body_off;
    TreeHdr *h = fb.hdr + s;
    if (t == 0) {
        h->internal_count = 0;
        h->max_depth = 0;
        h->status = 0;
        h->deep_runs = 0;
        h->stack_peak = 0;
        h->group_count = 0;
        h->group_next = 0;
        h->cross_count = 0;
        h->corr_pairs = 0;
    }
    __syncthreads();  // the reset precedes every flag raised below
    double mn[3] = {0.0, 0.0, 0.0}, mx[3] = {0.0, 0.0, 0.0};
    bool bad = false;
    for (uint32_t i = t; i < n; i += kForestThreads) {
        const double p[3] = {x[i], y[i], z[i]};
        for (int c = 0; c < 3; ++c) {
            bad |= !isfinite(p[c]);
            mn[c] = sycl_min(mn[c], p[c]);
            mx[c] = sycl_max(mx[c], p[c]);
        }
    }
    if (bad) raise_flag(h, kFlagAabb);
    for (int off = 32; off >= 1; off >>= 1)
        for (int c = 0; c < 3; ++c) {
            mn[c] = sycl_min(mn[c], __shfl_xor(mn[c], off, 64));
            mx[c] = sycl_max(mx[c], __shfl_xor(mx[c], off, 64));
        }
    if ((t & 63) == 0)
        for (int c = 0; c < 3; ++c) {
            red[t >> 6][c] = mn[c];
            red[t >> 6][3 + c] = mx[c];
        }
    __syncthreads();
    if (t == 0) {
        double v[6];
        for (int c = 0; c < 6; ++c) v[c] = red[0][c];
        for (int w = 1; w < kForestWaves; ++w)
            for (int c = 0; c < 3; ++c) {
                v[c] = sycl_min(v[c], red[w][c]);
                v[3 + c] = sycl_max(v[3 + c], red[w][3 + c]);
            }
        cubify(v, h);
    }
}

// The largest 63-bit key: the padding between two systems' slots sorts after
// their bodies (ties keep slot order, and padding slots follow the bodies).
constexpr uint64_t kPadKey = ~0ull >> 1;

// 2. Keys of every slot j of the forest's body space (all systems at once):
// a Barnes-Hut system's body i = j - fb_off gets its key in the system's root
// cell and its packed record; padding and direct-sum slots get kPadKey.
__global__ void __launch_bounds__(kThreads) k_forest_keys(
    uint32_t slots, const uint16_t *__restrict__ fb_sys, const EnsSystem *__restrict__ sys,
    const double *__restrict__ gx, const double *__restrict__ gy, const double *__restrict__ gz,
    const double *__restrict__ gm, ForestBufs fb, uint64_t *key, uint32_t *val) {
    const uint32_t j = blockIdx.x * kThreads + threadIdx.x;
    if (j >= slots) return;
    const uint32_t s = fb_sys[j];
    const EnsSystem S = sys[s];
    const uint32_t i = j - S.fb_off;
    uint64_t k = kPadKey;
    if (S.bh && i < S.n) {
        TreeHdr *h = fb.hdr + s;
        Cell c{h->minx, h->miny, h->minz, h->edge};
        const uint32_t b = S.body_off + i;
        const double px = gx[b], py = gy[b], pz = gz[b], pm = gm[b];
        k = key_from(c, px, py, pz);
        fb.body4[j] = make_double4(px, py, pz, pm);
        fb.klo[j] = k;  // the full key by slot, for k_forest_rekey
        if (pm == 0) raise_flag(h, kFlagZeroMass);  // an error of the reference's centre of mass
    }
    key[j] = k;
    val[j] = j;
}

// Low key bits sorted by the first forest-wide sort; the second sorts the
// system above the remaining high bits.  Both sorts then take 4 passes of 9
// bits (for up to 512 systems), the minimum of a sort.
constexpr int kForestLowBits = 36;

// 3b. After the sort by the low key bits: each slot's key becomes (its system,
// its high key bits), for the stable sort that gathers every system's slots
// back into its own range, in key order.  A slot's system is the last whose
// first slot is at or before it: up to kSysrankLds systems the first slots sit
// in LDS and each slot's system is found by bisection there, instead of a
// gathered load of fb_sys at a scattered slot.
constexpr uint32_t kSysrankLds = 8192;
constexpr uint32_t kSysrankBlocks = 2048;
__global__ void __launch_bounds__(kThreads) k_forest_sysrank(uint32_t slots, uint32_t nsys,
                                                             const uint32_t *__restrict__ first,
                                                             const uint16_t *__restrict__ fb_sys,
                                                             uint64_t *key,
                                                             const uint32_t *__restrict__ val) {
    __shared__ uint32_t fl[kSysrankLds];
    const bool lds = nsys <= kSysrankLds;
    if (lds) {
        for (uint32_t k = threadIdx.x; k < nsys; k += kThreads) fl[k] = first[k];
        __syncthreads();
    }
    const uint32_t stride = gridDim.x * kThreads;
    for (uint32_t j = blockIdx.x * kThreads + threadIdx.x; j < slots; j += stride) {
        const uint32_t v = val[j];
        uint32_t sys;
        if (lds) {
            // fl[0] = 0 <= v: keep fl[lo] <= v while halving the range
            uint32_t lo = 0, len = nsys;
            while (len > 1) {
                const uint32_t half = len >> 1;
                if (fl[lo + half] <= v) lo += half;
                len -= half;
            }
            sys = lo;
        } else {
            sys = fb_sys[v];
        }
        key[j] = ((uint64_t)sys << (3 * kKeyLevels - kForestLowBits)) | (key[j] >> kForestLowBits);
    }
}

// After the forest-wide sort, the build alternates stages over every body
// slot of the forest (all systems at once, the whole GPU) with stages of one
// workgroup per system, which hold what needs a barrier across one system:
// the cell-numbering scan, the level-by-level centre of mass and the group
// scan.  Every stage is the one-system build's per-element step; the trees
// are bit-identical to a one-system build.

// A slot's system and its index inside that system's slots.
struct SlotSys {
    uint32_t s, li;
    EnsSystem S;
};
__device__ inline SlotSys slot_sys(uint32_t j, const uint16_t *__restrict__ fb_sys,
                                   const EnsSystem *__restrict__ sys) {
    SlotSys r;
    r.s = fb_sys[j];
    r.S = sys[r.s];
    r.li = j - r.S.fb_off;
    return r;
}

// 3. The full keys of the bodies in sorted order (the second sort kept only
// the high key bits; k_forest_keys left every slot's full key in klo, which
// holds nothing else until the deep runs' second key words), and the
// system-local body indices.  `sorted` may be fb.sidx itself: each slot is
// read, then rewritten, by one thread.
__global__ void __launch_bounds__(kThreads) k_forest_rekey(uint32_t slots,
                                                           const uint16_t *__restrict__ fb_sys,
                                                           const EnsSystem *__restrict__ sys,
                                                           ForestBufs fb, const uint32_t *sorted) {
    const uint32_t j = blockIdx.x * kThreads + threadIdx.x;
    if (j >= slots) return;
    const SlotSys ss = slot_sys(j, fb_sys, sys);
    if (!ss.S.bh || ss.li >= ss.S.n) return;
    TreeHdr *h = fb.hdr + ss.s;
    uint32_t q = sorted[j] - ss.S.fb_off;
    if (q >= ss.S.n) {  // only after a failed sort, which the flag reports
        raise_flag(h, kFlagSortWait);
        q = 0;
    }
    fb.skey[j] = fb.klo[ss.S.fb_off + q];  // one gathered 8-byte key per body
    fb.sidx[j] = q;
}

// 4. lcp of each adjacent pair of a system's sorted bodies; a run of bodies
// sharing all 21 digits is extended and sorted by its first pair's thread.
__global__ void __launch_bounds__(kThreads) k_forest_lcp(
    uint32_t slots, const uint16_t *__restrict__ fb_sys, const EnsSystem *__restrict__ sys,
    const double *__restrict__ gx, const double *__restrict__ gy, const double *__restrict__ gz,
    ForestBufs fb) {
    const uint32_t j = blockIdx.x * kThreads + threadIdx.x;
    if (j >= slots) return;
    const SlotSys ss = slot_sys(j, fb_sys, sys);
    if (!ss.S.bh || ss.li + 1 >= ss.S.n) return;
    TreeHdr *h = fb.hdr + ss.s;
    const Cell root{h->minx, h->miny, h->minz, h->edge};
    const uint32_t fo = ss.S.fb_off, bo = ss.S.body_off;
    lcp_at(ss.li, gx + bo, gy + bo, gz + bo, ss.S.n, root, h, fb.skey + fo, fb.sidx + fo,
           fb.klo + fo, fb.lcp + fo);
}

// 5. chain: the internal cells owned by each pair.  (A system's deepest leaf
// is not needed by the ensemble; the header's max_depth stays 0.)
__global__ void __launch_bounds__(kThreads) k_forest_chain(uint32_t slots,
                                                           const uint16_t *__restrict__ fb_sys,
                                                           const EnsSystem *__restrict__ sys,
                                                           ForestBufs fb) {
    const uint32_t j = blockIdx.x * kThreads + threadIdx.x;
    if (j >= slots) return;
    const SlotSys ss = slot_sys(j, fb_sys, sys);
    if (ss.S.bh && ss.li + 1 < ss.S.n) chain_at(ss.li, fb.lcp + ss.S.fb_off, fb.chain + ss.S.fb_off);
}

// 6. Per system: cell numbering (exclusive scan of the chain), the internal
// cell count, the empty node slots the topology fills in, and an empty table
// of block-crossing cells.
__global__ void __launch_bounds__(kForestThreads) k_forest_number(
    const EnsSystem *__restrict__ sys, const uint32_t *__restrict__ order, ForestBufs fb) {
    __shared__ uint32_t wsum[kForestWaves + 1];
    const uint32_t s = order[blockIdx.x];
    const EnsSystem S = sys[s];
    const uint32_t n = S.n;
    TreeHdr *h = fb.hdr + s;
    const uint32_t I_all = n > 1 ? wg_scan_array(fb.chain + S.fb_off, n - 1, wsum) : 0u;
    if (threadIdx.x == 0) h->internal_count = I_all;
    const uint32_t I = min(I_all, S.icap);
    int32_t *ref = fb.ref + S.node_off;
    for (uint64_t j = threadIdx.x; j < 1 + 8ull * I; j += kForestThreads) ref[j] = kRefEmpty;
    uint32_t *spanning = fb.spanning + S.span_off;
    const uint32_t span_cap = kMaxDepth * ((n + kComBlock - 1) / kComBlock);
    for (uint32_t j = threadIdx.x; j < span_cap; j += kForestThreads) spanning[j] = kNoCell;
}

// 7. Topology of every body of the forest, as k_link: a block stages the keys
// of its slots and kLinkBack before them (whatever systems they belong to;
// a search only reads keys of its own system).  Cells are local to, or cross,
// the system's kComBlock-body blocks as in the one-system build.
__global__ void __launch_bounds__(kThreads) k_forest_link(uint32_t slots,
                                                          const uint16_t *__restrict__ fb_sys,
                                                          const EnsSystem *__restrict__ sys,
                                                          ForestBufs fb) {
    __shared__ uint64_t win[kLinkBack + kThreads];
    const uint32_t b0 = blockIdx.x * kThreads;
    const uint32_t w0 = b0 > kLinkBack ? b0 - kLinkBack : 0u;
    const uint32_t w1 = min(slots, b0 + (uint32_t)kThreads);
    for (uint32_t p = w0 + threadIdx.x; p < w1; p += kThreads) win[p - w0] = fb.skey[p];
    __syncthreads();
    const uint32_t j = b0 + threadIdx.x;
    if (j >= slots) return;
    const SlotSys ss = slot_sys(j, fb_sys, sys);
    const EnsSystem &S = ss.S;
    if (!S.bh || ss.li >= S.n) return;
    const uint32_t fo = S.fb_off, io = S.int_off;
    KeyView kv{fb.skey + fo, fb.klo + fo};
    // window bounds relative to the system's first slot; the unsigned
    // differences in KeyView::key place a position in or out of the window
    // correctly also when the window starts before the system
    kv.win = win;
    kv.w0 = w0 - fo;
    kv.w1 = w1 - fo;
    link_at(ss.li, S.n, S.icap, fb.hdr + ss.s, kv, fb.sidx + fo, fb.lcp + fo, fb.chain + fo,
            fb.ref + S.node_off, fb.int_depth + io, fb.int_parent + io, fb.int_node + io,
            fb.body_parent + fo, fb.int_local + io, fb.spanning + S.span_off,
            kMaxDepth * ((S.n + kComBlock - 1) / kComBlock), fb.int_first + io, kComBlock,
            kComSuper);
}

// The one-system build's view of system s, for the centre-of-mass passes.
struct ForestTree {
    EnsSystem S;
    TreeHdr *h;
    RecOut o;
};
__device__ inline ForestTree forest_tree(const EnsSystem *__restrict__ sys, uint32_t s,
                                         const ForestBufs &fb, double G, GroupRule rule) {
    ForestTree f;
    f.S = sys[s];
    f.h = fb.hdr + s;
    const uint32_t fo = f.S.fb_off;
    f.o = RecOut{G, 0, 1, nullptr, nullptr, fb.rec + f.S.node_off, fb.group_size + fo,
                 fb.group_slot + fo, rule};
    return f;
}

// 8. Centre of mass, node records and group marks, as k_com_local and
// k_com_span, for the blocks of every system in one launch each: block
// (system, b) of fb.blk_tab; then (system, super-block) of fb.span_tab
// (want 0), then the systems of fb.big_tab (want 2: cells crossing
// super-block ends).
__global__ void __launch_bounds__(kComLocalThreads) k_forest_com_local(
    const EnsSystem *__restrict__ sys, ForestBufs fb, double G, GroupRule rule) {
    const uint2 e = fb.blk_tab[blockIdx.x];
    const ForestTree f = forest_tree(sys, e.x, fb, G, rule);
    const EnsSystem &S = f.S;
    const uint32_t fo = S.fb_off, io = S.int_off;
    com_local_block(e.y, S.n, S.icap, f.h, fb.body4 + fo, fb.sidx + fo, fb.ref + S.node_off,
                    fb.chain + fo, fb.int_depth + io, fb.int_local + io, fb.int_first + io,
                    fb.int_lvl + io, fb.int_sums + io, fb.int_count + io, f.o);
}

__global__ void __launch_bounds__(kComThreads) k_forest_com_span(
    const EnsSystem *__restrict__ sys, const uint2 *__restrict__ tab, uint32_t want, ForestBufs fb,
    double G, GroupRule rule) {
    const uint2 e = tab[blockIdx.x];
    const ForestTree f = forest_tree(sys, e.x, fb, G, rule);
    const EnsSystem &S = f.S;
    const uint32_t fo = S.fb_off, io = S.int_off;
    const uint32_t nblocks = (S.n + kComBlock - 1) / kComBlock;
    com_span_block(e.y, nblocks, want == 2 ? nblocks : kComSuperBlocks, want, f.h,
                   fb.body4 + fo, fb.ref + S.node_off, fb.spanning + S.span_off,
                   fb.int_local + io, fb.int_first + io, fb.int_lvl + io, fb.int_sums + io,
                   fb.int_count + io, fb.cross + S.cross_off,
                   kMaxDepth * ((S.n + kComSuper - 1) / kComSuper + 1), f.o);
}

// 9. The traversal groups of every system: after one exclusive scan of the
// group marks over all slots of the forest (padding included, whatever it
// holds: a system's group numbers are differences of the scan, exact modulo
// 2^32), each group start writes its group; a system's first slot writes its
// group count and merges its status flags into *status.
__global__ void __launch_bounds__(kThreads) k_forest_groups(uint32_t slots,
                                                            const uint16_t *__restrict__ fb_sys,
                                                            const EnsSystem *__restrict__ sys,
                                                            ForestBufs fb,
                                                            const uint32_t *__restrict__ total,
                                                            uint32_t *items, uint32_t *status) {
    const uint32_t j = blockIdx.x * kThreads + threadIdx.x;
    if (j >= slots) return;
    const SlotSys ss = slot_sys(j, fb_sys, sys);
    const EnsSystem &S = ss.S;
    if (!S.bh || ss.li >= S.n) return;
    const uint32_t fo = S.fb_off;
    const uint32_t base = fb.group_slot[fo];
    const uint32_t gs = fb.group_size[j];
    if (gs) fb.groups[fo + (fb.group_slot[j] - base)] = make_uint2(ss.li, gs);
    if (ss.li == 0) {
        const uint32_t end = fo + S.n < slots ? fb.group_slot[fo + S.n] : *total;
        const uint32_t ng = end - base;
        TreeHdr *h = fb.hdr + ss.s;
        h->group_count = ng;
        items[ss.s] = ng;
        // flags were raised with device-scope atomics: read the word the same way
        const uint32_t st = __hip_atomic_load(&h->status, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        if (st) atomicOr(status, st);
    }
}

}  // namespace

void launch_forest_build(const EnsSystem *sys, const uint32_t *order, uint32_t nbh,
                         uint32_t nsys, const double *x, const double *y, const double *z,
                         const double *m, double G, const ForestBufs &fb, GroupRule rule,
                         uint32_t *items, uint32_t *status, hipStream_t s) {
    if (nbh == 0) return;
    hipLaunchKernelGGL(k_forest_aabb, dim3(nbh), dim3(kForestThreads), 0, s, sys, order, x, y, z,
                       fb);
    // LSD over (system, key): one radix sort of every system's slots by the
    // low key bits, then a stable one by (system, high key bits): each
    // system's bodies end in its own slot range, in key order with ties in
    // index order (the order of a sort of that system alone)
    const uint32_t slots = fb.slots, blocks = ceil_div(slots, kThreads);
    uint64_t *k = fb.skey, *ka = fb.skey_alt;
    uint32_t *v = fb.sidx, *va = fb.sidx_alt;
    hipLaunchKernelGGL(k_forest_keys, dim3(blocks), dim3(kThreads), 0, s, slots,
                       (const uint16_t *)fb.fb_sys, sys, x, y, z, m, fb, k, v);
    SortEnd end = radix_sort_u64_u32(k, v, ka, va, slots, kForestLowBits, fb.radix_ctl, status, s);
    if (end == SortEnd::kRejected) return;  // reported in *status
    if (end == SortEnd::kInAlt) {
        std::swap(k, ka);
        std::swap(v, va);
    }
    hipLaunchKernelGGL(k_forest_sysrank, dim3(std::min(blocks, kSysrankBlocks)), dim3(kThreads), 0,
                       s, slots, nsys, (const uint32_t *)fb.sys_first, (const uint16_t *)fb.fb_sys,
                       k, (const uint32_t *)v);
    int sys_bits = 1;
    while (sys_bits < 16 && (1u << sys_bits) < nsys) ++sys_bits;
    // its own control words: a control buffer must always see the same pass count
    end = radix_sort_u64_u32(k, v, ka, va, slots, 3 * kKeyLevels - kForestLowBits + sys_bits,
                             fb.radix_ctl_sys, status, s);
    if (end == SortEnd::kRejected) return;
    if (end == SortEnd::kInAlt) {
        std::swap(k, ka);
        std::swap(v, va);
    }
    const uint16_t *fs = fb.fb_sys;
    hipLaunchKernelGGL(k_forest_rekey, dim3(blocks), dim3(kThreads), 0, s, slots, fs, sys, fb,
                       (const uint32_t *)v);
    hipLaunchKernelGGL(k_forest_lcp, dim3(blocks), dim3(kThreads), 0, s, slots, fs, sys, x, y, z,
                       fb);
    hipLaunchKernelGGL(k_forest_chain, dim3(blocks), dim3(kThreads), 0, s, slots, fs, sys, fb);
    hipLaunchKernelGGL(k_forest_number, dim3(nbh), dim3(kForestThreads), 0, s, sys, order, fb);
    hipLaunchKernelGGL(k_forest_link, dim3(blocks), dim3(kThreads), 0, s, slots, fs, sys, fb);
    if (fb.nblk)
        hipLaunchKernelGGL(k_forest_com_local, dim3(fb.nblk), dim3(kComLocalThreads), 0, s, sys, fb,
                           G, rule);
    if (fb.nspan)
        hipLaunchKernelGGL(k_forest_com_span, dim3(fb.nspan), dim3(kComThreads), 0, s, sys,
                           (const uint2 *)fb.span_tab, 0u, fb, G, rule);
    if (fb.nbig)
        hipLaunchKernelGGL(k_forest_com_span, dim3(fb.nbig), dim3(kComThreads), 0, s, sys,
                           (const uint2 *)fb.big_tab, 2u, fb, G, rule);
    uint32_t *total = fb.scan_part + (fb.scan_cap - 1);
    scan_exclusive_u32(fb.group_slot, fb.group_slot, slots, total, fb.scan_part, fb.scan_cap - 1,
                       s);
    hipLaunchKernelGGL(k_forest_groups, dim3(blocks), dim3(kThreads), 0, s, slots, fs, sys, fb,
                       (const uint32_t *)total, items, status);
}

void launch_aabb(const TreeBufs &t, const double *x, const double *y, const double *z,
                 hipStream_t s) {
    hipLaunchKernelGGL(k_build_begin, dim3(1), dim3(kThreads), 0, s, t.hdr);
    const uint32_t nb = std::min<uint32_t>(kAabbBlocks, ceil_div(t.n, 4 * kThreads));
    hipLaunchKernelGGL(k_aabb_partial, dim3(nb), dim3(kThreads), 0, s, x, y, z, t.n, t.aabb_part,
                       t.hdr);
    hipLaunchKernelGGL(k_aabb_final, dim3(1), dim3(kThreads), 0, s, (const double *)t.aabb_part, nb,
                       t.hdr);
}

void launch_keys(TreeBufs &t, const double *x, const double *y, const double *z,
                 const double *m, hipStream_t s) {
    hipLaunchKernelGGL(k_keys, dim3(ceil_div(t.n, kThreads)), dim3(kThreads), 0, s, x, y, z, m,
                       t.n, (const TreeHdr *)t.hdr, t.skey, t.sidx, t.body4);
}

void launch_key_sort(TreeBufs &t, hipStream_t s) {
    // a rejected sort (2^30 bodies or more, which the context's capacity
    // check already excludes) raises kFlagSortSize; the build reports it
    if (radix_sort_u64_u32(t.skey, t.sidx, t.skey_alt, t.sidx_alt, t.n, 3 * kKeyLevels,
                           t.radix_ctl, &t.hdr->status, s) == SortEnd::kInAlt) {
        std::swap(t.skey, t.skey_alt);
        std::swap(t.sidx, t.sidx_alt);
    }
}

void launch_topology(const TreeBufs &t, const double *x, const double *y, const double *z,
                     hipStream_t s) {
    const uint32_t n = t.n;
    const dim3 grid(ceil_div(n, kThreads));
    if (n > 1) {
        hipLaunchKernelGGL(k_lcp, grid, dim3(kThreads), 0, s, x, y, z, n, t.hdr, t.skey, t.sidx,
                           t.klo, t.lcp);
        hipLaunchKernelGGL(k_chain, grid, dim3(kThreads), 0, s, (const int32_t *)t.lcp, n,
                           t.chain);
        scan_exclusive_u32(t.chain, t.chain, n - 1, &t.hdr->internal_count, t.scan_part,
                           t.scan_part_cap, s);
    }
    const uint32_t fill_blocks = std::max<uint32_t>(1, std::min<uint32_t>(4096, ceil_div(8ull * n, kThreads)));
    hipLaunchKernelGGL(k_fill, dim3(fill_blocks), dim3(kThreads), 0, s, t.hdr, t.icap, t.ref,
                       t.spanning, t.spanning_cap);
    KeyView kv{t.skey, t.klo};
    hipLaunchKernelGGL(k_link, grid, dim3(kThreads), 0, s, n, t.icap, t.hdr, kv,
                       (const uint32_t *)t.sidx, (const int32_t *)t.lcp,
                       (const uint32_t *)t.chain, t.ref, t.int_depth, t.int_parent, t.int_node,
                       t.body_parent, t.int_local, t.spanning, t.spanning_cap, t.int_first);
}

void launch_max_depth(const TreeBufs &t, hipStream_t s) {
    if (t.n < 2) return;
    const uint32_t blocks = std::min<uint32_t>(kDepthBlocks, ceil_div(t.n, kThreads));
    hipLaunchKernelGGL(k_max_depth, dim3(blocks), dim3(kThreads), 0, s, (const int32_t *)t.lcp,
                       t.n, t.hdr);
}

void launch_com(const TreeBufs &t, double G, bool f64, bool f32, GroupRule rule, hipStream_t s) {
    const uint32_t nblocks = ceil_div(t.n, kComBlock);
    const RecOut o{G,          (int)f64,      (int)f32,      t.cm,
                   t.edge,     t.rec,         t.group_size,  t.group_slot,
                   rule};
    hipLaunchKernelGGL(k_com_local, dim3(nblocks), dim3(kComLocalThreads), 0, s, t.n, t.icap, t.hdr,
                       (const double4 *)t.body4, (const uint32_t *)t.sidx, (const int32_t *)t.ref,
                       (const uint32_t *)t.chain, (const uint8_t *)t.int_depth,
                       (const uint8_t *)t.int_local, (const uint32_t *)t.int_first, t.int_lvl,
                       t.int_sums, t.int_count, o);
    if (nblocks >= 2) {
        hipLaunchKernelGGL(k_com_span, dim3(ceil_div(nblocks, kComSuperBlocks)), dim3(kComThreads),
                           0, s, nblocks, kComSuperBlocks, 0u, t.hdr, (const double4 *)t.body4,
                           (const int32_t *)t.ref, (const uint32_t *)t.spanning,
                           (const uint8_t *)t.int_local, (const uint32_t *)t.int_first, t.int_lvl,
                           t.int_sums, t.int_count, t.cross, t.cross_cap, o);
        if (nblocks > kComSuperBlocks)
            hipLaunchKernelGGL(k_com_span, dim3(1), dim3(kComThreads), 0, s, nblocks, nblocks, 2u,
                               t.hdr, (const double4 *)t.body4, (const int32_t *)t.ref,
                               (const uint32_t *)t.spanning, (const uint8_t *)t.int_local,
                               (const uint32_t *)t.int_first, t.int_lvl, t.int_sums, t.int_count,
                               t.cross, t.cross_cap, o);
    }
    if (f32) {
        // the group list: slots by an exclusive scan of the marks, then a scatter
        scan_exclusive_u32(t.group_slot, t.group_slot, t.n, &t.hdr->group_count, t.scan_part,
                           t.scan_part_cap, s);
        hipLaunchKernelGGL(k_group_scatter, dim3(ceil_div(t.n, kThreads)), dim3(kThreads), 0, s,
                           t.n, (const uint32_t *)t.group_size, (const uint32_t *)t.group_slot,
                           t.groups);
    }
}

}  // namespace nb

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
struct KeyView {
    const uint64_t *skey;
    const uint64_t *klo;
    __device__ inline uint32_t digit(uint32_t p, int l) const {
        return l < kKeyLevels ? (uint32_t)(skey[p] >> (3 * (kKeyLevels - 1 - l))) & 7u
                              : (uint32_t)(klo[p] >> (3 * (2 * kKeyLevels - 1 - l))) & 7u;
    }
    // do sorted bodies p and q share their first v digits?
    __device__ inline bool same(uint32_t p, uint32_t q, int v) const {
        if (v <= 0) return true;
        if (v <= kKeyLevels) return ((skey[p] ^ skey[q]) >> (63 - 3 * v)) == 0;
        return skey[p] == skey[q] && ((klo[p] ^ klo[q]) >> (63 - 3 * (v - kKeyLevels))) == 0;
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
// "local" when its bodies stay inside one com_block-body block of the sorted
// order; the others go to the table spanning[block * kMaxDepth + depth].
// No atomics: every cell, node slot and table entry has exactly one writer.
__device__ inline void link_at(uint32_t i, uint32_t n, uint32_t icap, TreeHdr *h, KeyView kv,
                               const uint32_t *__restrict__ sidx, const int32_t *__restrict__ lcp,
                               const uint32_t *__restrict__ off, int32_t *ref, uint8_t *int_depth,
                               uint32_t *int_parent, uint32_t *int_node, uint32_t *body_parent,
                               uint8_t *int_local, uint32_t *spanning, uint32_t spanning_cap,
                               uint32_t *int_first, uint32_t com_block) {
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
        int_local[k] = local ? 1 : 0;
        if (!local) {
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

__global__ void __launch_bounds__(kThreads) k_link(uint32_t n, uint32_t icap, TreeHdr *h,
                                                   KeyView kv, const uint32_t *__restrict__ sidx,
                                                   const int32_t *__restrict__ lcp,
                                                   const uint32_t *__restrict__ off,
                                                   int32_t *ref, uint8_t *int_depth,
                                                   uint32_t *int_parent, uint32_t *int_node,
                                                   uint32_t *body_parent, uint8_t *int_local,
                                                   uint32_t *spanning, uint32_t spanning_cap,
                                                   uint32_t *int_first) {
    const uint32_t i = blockIdx.x * kThreads + threadIdx.x;
    if (i >= n) return;
    link_at(i, n, icap, h, kv, sidx, lcp, off, ref, int_depth, int_parent, int_node, body_parent,
            int_local, spanning, spanning_cap, int_first, kComBlock);
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

__device__ inline void sum_children(uint32_t pk, const double4 *__restrict__ body4,
                                    const int32_t *__restrict__ ref, InternalSums *sums,
                                    uint32_t *counts, TreeHdr *h) {
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
            const InternalSums s = sums[r[d]];
            v[d] = make_double4(s.x, s.y, s.z, s.m);
            c[d] = counts[r[d]];
        } else if (r[d] != kRefEmpty) {
            v[d] = body_term(body4[body_of_ref(r[d])]);
            c[d] = 1;
        }
    }
    InternalSums s;
    uint32_t cnt;
    add_children(r, v, c, s, cnt);
    sums[pk] = s;
    counts[pk] = cnt;
    if (s.m == 0) raise_flag(h, kFlagZeroMass);
}

// Centre of mass, level by level from the deepest cells up: a cell's sums
// (its children in digit order) need only deeper cells, so every cell of one
// depth can be summed at once, with a barrier between depths.
constexpr int kComThreads = 1024;

// Cells [kb, ke) bucketed by depth into lvl[kb ..) (cells with keep(k) only),
// then summed deepest depth first.  cnt/cur: kMaxDepth + 1 and kMaxDepth
// LDS words.  Every thread of the workgroup calls it.
template <typename Keep>
__device__ inline void com_levels(uint32_t kb, uint32_t ke, Keep keep,
                                  const uint8_t *__restrict__ int_depth, uint32_t *lvl,
                                  uint32_t *cnt, uint32_t *cur,
                                  const double4 *__restrict__ body4,
                                  const int32_t *__restrict__ ref, InternalSums *sums,
                                  uint32_t *counts, TreeHdr *h) {
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
        for (uint32_t j = lb + t; j < le; j += nt)
            sum_children(lvl[kb + j], body4, ref, sums, counts, h);
        __syncthreads();
    }
}

// Pass 1: one workgroup per kComBlock consecutive (sorted) bodies sums the
// cells that lie inside its block.  Pair i owns cells [off[i], off[i+1]), so
// the block's cells are one contiguous range [kb, ke).  Up to kComLdsCells of
// them are summed in LDS: their child references are staged once
// (coalesced), and a child cell's sums are read from LDS, so a level costs one
// global round trip (the body children) instead of three.  A larger block
// falls back to the same sums in global memory (lvl: depth buckets).
constexpr uint32_t kComLdsCells = 1536;  // a 2048-body block holds about 1,000 cells
__global__ void __launch_bounds__(kComThreads) k_com_local(
    uint32_t n, uint32_t icap, TreeHdr *h, const double4 *__restrict__ body4,
    const uint32_t *__restrict__ sidx, const int32_t *__restrict__ ref,
    const uint32_t *__restrict__ off, const uint8_t *__restrict__ int_depth,
    const uint8_t *__restrict__ int_local, uint32_t *lvl, InternalSums *sums, uint32_t *counts) {
    __shared__ uint32_t cnt[kMaxDepth + 1], cur[kMaxDepth];
    __shared__ uint32_t lvl_s[kComLdsCells];
    __shared__ int32_t ref_s[8 * kComLdsCells];
    __shared__ double sx_s[kComLdsCells], sy_s[kComLdsCells], sz_s[kComLdsCells],
        sm_s[kComLdsCells];
    __shared__ uint32_t n_s[kComLdsCells];
    const uint32_t t = threadIdx.x;
    const uint32_t s = blockIdx.x * kComBlock;
    const uint32_t e = min(n, s + kComBlock);
    for (uint32_t i = s + t; i < e; i += kComThreads)
        if (body4[sidx[i]].w == 0) raise_flag(h, kFlagZeroMass);
    if (n < 2) return;  // a lone body is the root leaf: no internal cell
    const uint32_t I = min(h->internal_count, icap);
    const uint32_t kb = s < n - 1 ? min(off[s], I) : I;
    const uint32_t ke = e < n - 1 ? min(off[e], I) : I;
    const uint32_t nc = ke - kb;
    if (nc > kComLdsCells) {  // the same for the whole workgroup
        com_levels(kb, ke, [&](uint32_t k) { return int_local[k] != 0; }, int_depth, lvl, cnt,
                   cur, body4, ref, sums, counts, h);
        return;
    }
    for (uint32_t j = t; j < 8 * nc; j += kComThreads) ref_s[j] = ref[1ull + 8ull * kb + j];
    for (uint32_t d = t; d <= (uint32_t)kMaxDepth; d += kComThreads) cnt[d] = 0;
    __syncthreads();
    for (uint32_t c = t; c < nc; c += kComThreads)
        if (int_local[kb + c])
            atomicAdd(&cnt[min((uint32_t)int_depth[kb + c], (uint32_t)kMaxDepth - 1)], 1u);
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
    for (uint32_t c = t; c < nc; c += kComThreads)
        if (int_local[kb + c])
            lvl_s[atomicAdd(&cur[min((uint32_t)int_depth[kb + c], (uint32_t)kMaxDepth - 1)], 1u)] = c;
    __syncthreads();
    for (int d = kMaxDepth - 1; d >= 0; --d) {
        const uint32_t lb = cnt[d], le = cnt[d + 1];
        if (lb == le) continue;  // the same for every thread: LDS, read after a barrier
        for (uint32_t j = lb + t; j < le; j += kComThreads) {
            const uint32_t c = lvl_s[j];
            // the children's data first, then the sums in digit order (as sum_children)
            int32_t r[8];
            double4 v[8];
            uint32_t cc[8];
#pragma unroll
            for (int dd = 0; dd < 8; ++dd) r[dd] = ref_s[8 * c + dd];
#pragma unroll
            for (int dd = 0; dd < 8; ++dd) {
                v[dd] = make_double4(0.0, 0.0, 0.0, 0.0);
                cc[dd] = 0;
                if (r[dd] >= 0) {
                    const uint32_t rl = (uint32_t)r[dd] - kb;
                    if (rl < nc) {
                        v[dd] = make_double4(sx_s[rl], sy_s[rl], sz_s[rl], sm_s[rl]);
                        cc[dd] = n_s[rl];
                    } else {  // not in this block: cannot happen for a local cell
                        const InternalSums q = sums[r[dd]];
                        v[dd] = make_double4(q.x, q.y, q.z, q.m);
                        cc[dd] = counts[r[dd]];
                    }
                } else if (r[dd] != kRefEmpty) {
                    v[dd] = body_term(body4[body_of_ref(r[dd])]);
                    cc[dd] = 1;
                }
            }
            InternalSums q;
            uint32_t cn;
            add_children(r, v, cc, q, cn);
            sx_s[c] = q.x;
            sy_s[c] = q.y;
            sz_s[c] = q.z;
            sm_s[c] = q.m;
            n_s[c] = cn;
            sums[kb + c] = q;
            counts[kb + c] = cn;
            if (q.m == 0) raise_flag(h, kFlagZeroMass);
        }
        __syncthreads();
    }
}

// Pass 2: one workgroup sums the cells that cross block boundaries, the
// table spanning[block * kMaxDepth + depth], deepest depth first.
__global__ void __launch_bounds__(kComThreads) k_com_spanning(
    uint32_t nblocks, TreeHdr *h, const double4 *__restrict__ body4,
    const int32_t *__restrict__ ref, const uint32_t *__restrict__ spanning, InternalSums *sums,
    uint32_t *counts) {
    __shared__ uint32_t live[kMaxDepth];
    const uint32_t t = threadIdx.x;
    if (t < (uint32_t)kMaxDepth) live[t] = 0;
    __syncthreads();
    for (uint32_t j = t; j < nblocks * (uint32_t)kMaxDepth; j += kComThreads)
        if (spanning[j] != kNoCell) live[j % kMaxDepth] = 1;
    __syncthreads();
    for (int d = kMaxDepth - 1; d >= 0; --d) {
        if (!live[d]) continue;  // the same for every thread: LDS, read after a barrier
        for (uint32_t b = t; b < nblocks; b += kComThreads) {
            const uint32_t k = spanning[(uint64_t)b * kMaxDepth + d];
            if (k != kNoCell) sum_children(k, body4, ref, sums, counts, h);
        }
        __syncthreads();
    }
}

// ---- traversal groups -------------------------------------------------------------
// A group is a cell with at most kGroupMax bodies whose parent holds more, or
// a lone body whose parent holds more.  Groups partition the bodies into
// spatially compact runs of the octree order.  Every sorted position starts
// at most one group: mark the starts, scan, scatter -- the group list comes
// out in octree order, so neighbouring groups share most of their walk.
// pass 1: every sorted position; lone bodies start a group of one
__device__ inline void group_mark_body(uint32_t i, uint32_t n, const uint32_t *__restrict__ counts,
                                       const uint32_t *__restrict__ body_parent,
                                       uint32_t *group_size, uint32_t *group_slot) {
    const uint32_t pk = body_parent[i];
    const bool lone = pk == 0xffffffffu ? (n == 1) : counts[pk] > kGroupMax;
    group_size[i] = lone ? 1u : 0u;
    group_slot[i] = lone ? 1u : 0u;
}

__global__ void __launch_bounds__(kThreads) k_group_marks_bodies(
    uint32_t n, const uint32_t *__restrict__ counts, const uint32_t *__restrict__ body_parent,
    uint32_t *group_size, uint32_t *group_slot) {
    const uint32_t i = blockIdx.x * kThreads + threadIdx.x;
    if (i >= n) return;
    group_mark_body(i, n, counts, body_parent, group_size, group_slot);
}

// pass 2: group cells mark their first sorted position (never a lone body's)
__device__ inline void group_mark_cell(uint64_t k, uint32_t n, const uint32_t *__restrict__ counts,
                                       const uint32_t *__restrict__ int_parent,
                                       const uint32_t *__restrict__ int_first,
                                       uint32_t *group_size, uint32_t *group_slot) {
    const uint32_t c = counts[k];
    const uint32_t pc = k == 0 ? 0xffffffffu : counts[int_parent[k]];
    if (c <= kGroupMax && pc > kGroupMax) {
        const uint32_t f = int_first[k];
        if (f < n) {
            group_size[f] = c;
            group_slot[f] = 1u;
        }
    }
}

__global__ void __launch_bounds__(kThreads) k_group_marks_cells(uint32_t n, uint32_t icap,
                                                                const TreeHdr *__restrict__ h,
                                                                const uint32_t *__restrict__ counts,
                                                                const uint32_t *__restrict__ int_parent,
                                                                const uint32_t *__restrict__ int_first,
                                                                uint32_t *group_size,
                                                                uint32_t *group_slot) {
    const uint32_t I = min(h->internal_count, icap);
    const uint64_t stride = (uint64_t)gridDim.x * kThreads;
    const uint64_t t0 = (uint64_t)blockIdx.x * kThreads + threadIdx.x;
    for (uint64_t k = t0; k < I; k += stride)
        group_mark_cell(k, n, counts, int_parent, int_first, group_size, group_slot);
}

__global__ void __launch_bounds__(kThreads) k_group_scatter(uint32_t n,
                                                            const uint32_t *__restrict__ group_size,
                                                            const uint32_t *__restrict__ group_slot,
                                                            uint2 *groups) {
    const uint32_t i = blockIdx.x * kThreads + threadIdx.x;
    if (i >= n) return;
    const uint32_t s = group_size[i];
    if (s) groups[group_slot[i]] = make_uint2(i, s);
}

// ---- 8. node records -----------------------------------------------------------
__device__ inline void record_at(uint64_t j, double root_edge, const double4 *__restrict__ body4,
                                 const int32_t *__restrict__ ref,
                                 const InternalSums *__restrict__ sums,
                                 const uint8_t *__restrict__ int_depth, double G, int write_f64,
                                 int write_f32, double4 *cm, double *edge, NodeF32 *rec) {
    const int32_t r = ref[j];
    double4 c = make_double4(0.0, 0.0, 0.0, 0.0);
    double e = 0.0;
    if (r >= 0) {
        // the reference divides numerator by mass at every visit
        // (BarnesHutAlgorithm.cpp:351-353); the quotient is the same
        const InternalSums s = sums[r];
        c = make_double4(s.x / s.m, s.y / s.m, s.z / s.m, s.m);
        e = ldexp(root_edge, -(int)int_depth[r]);
    } else if (r != kRefEmpty) {
        const double4 p = body4[body_of_ref(r)];
        const double mb = p.w;
        c = make_double4((p.x * mb) / mb, (p.y * mb) / mb, (p.z * mb) / mb, mb);
    }
    if (write_f64) {
        cm[j] = c;
        edge[j] = e;
    }
    if (write_f32) {
        NodeF32 o;
        o.x = (float)c.x;
        o.y = (float)c.y;
        o.z = (float)c.z;
        o.gm = (float)(G * c.w);
        o.ref = r;
        o.lx = (float)(c.x - (double)o.x);
        o.ly = (float)(c.y - (double)o.y);
        o.lz = (float)(c.z - (double)o.z);
        rec[j] = o;
    }
}

__global__ void __launch_bounds__(kThreads) k_records(const TreeHdr *__restrict__ h,
                                                      uint32_t icap,
                                                      const double4 *__restrict__ body4,
                                                      const int32_t *__restrict__ ref,
                                                      const InternalSums *__restrict__ sums,
                                                      const uint8_t *__restrict__ int_depth,
                                                      double G, int write_f64, int write_f32,
                                                      double4 *cm, double *edge, NodeF32 *rec) {
    const uint32_t I = min(h->internal_count, icap);
    const uint64_t nodes = 1 + 8ull * I;
    const double root_edge = h->edge;
    const uint64_t stride = (uint64_t)gridDim.x * kThreads;
    for (uint64_t j = (uint64_t)blockIdx.x * kThreads + threadIdx.x; j < nodes; j += stride)
        record_at(j, root_edge, body4, ref, sums, int_depth, G, write_f64, write_f32, cm, edge,
                  rec);
}

// ---- ensemble: one workgroup builds one system's octree ----------------------------
// The same per-element steps as the one-system build, separated by workgroup
// barriers instead of kernel boundaries; scans and the radix sort run inside
// the workgroup, and every cell is local (one centre-of-mass pass).
constexpr int kForestThreads = 1024;
constexpr int kForestWaves = kForestThreads / 64;
constexpr int kSortBits = 8;  // radix digit width: 256 bins x 16 waves of LDS counters
constexpr int kSortBins = 1 << kSortBits;
constexpr int kSortPasses = (63 + kSortBits - 1) / kSortBits;  // 8: the result ends in place
static_assert(kSortPasses % 2 == 0, "the sorted pairs must end in the input buffers");
constexpr int kSortUnroll = 4;  // 64-key groups per iteration of a wave

// lanes below this one, as a ballot mask
__device__ inline uint64_t forest_lanes_below() {
    const uint32_t lane = threadIdx.x & 63u;
    return lane == 0 ? 0ull : (~0ull >> (64u - lane));
}

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

// Lanes of the wave holding the same digit as this lane (valid lanes only).
__device__ inline uint64_t digit_peers(uint32_t dig, bool valid) {
    uint64_t peers = __ballot(valid);
#pragma unroll
    for (int b = 0; b < kSortBits; ++b) {
        const bool bit = (dig >> b) & 1u;
        const uint64_t mask = __ballot(bit);
        peers &= bit ? mask : ~mask;
    }
    return peers;
}

// Stable LSD radix sort of (k, v) pairs by the 63-bit keys, 8-bit digits,
// (k2, v2) the ping-pong buffers.  Wave w owns a contiguous range and reads
// it 64 consecutive pairs at a time; a pair's rank among the equal digits of
// its 64 comes from ballots, and the offsets run digit-major over the waves
// (whist[w * 256 + d]), so equal digits keep their order: stable.
__device__ inline void wg_radix_sort(uint64_t *k, uint32_t *v, uint64_t *k2, uint32_t *v2,
                                     uint32_t n, uint32_t *whist, uint32_t *wsum) {
    const uint32_t t = threadIdx.x, lane = t & 63u, w = t >> 6;
    const uint64_t below = forest_lanes_below();
    uint32_t b, e;
    wave_range(n, &b, &e);
    uint32_t *row = whist + w * kSortBins;
    uint64_t *sk = k, *dk = k2;
    uint32_t *sv = v, *dv = v2;
    for (int p = 0; p < kSortPasses; ++p) {
        const int sh = kSortBits * p;
        for (uint32_t d = lane; d < (uint32_t)kSortBins; d += 64) row[d] = 0;
        // kSortUnroll groups of 64 per iteration: their loads are in flight together
        for (uint32_t g = b; g < e; g += 64 * kSortUnroll) {
            uint32_t dig[kSortUnroll];
            bool val[kSortUnroll];
#pragma unroll
            for (int u = 0; u < kSortUnroll; ++u) {
                const uint32_t i = g + 64u * u + lane;
                val[u] = i < e;
                dig[u] = val[u] ? (uint32_t)(sk[i] >> sh) & (kSortBins - 1) : 0u;
            }
#pragma unroll
            for (int u = 0; u < kSortUnroll; ++u) {
                const uint64_t peers = digit_peers(dig[u], val[u]);
                if (val[u] && (peers & below) == 0) row[dig[u]] += (uint32_t)__popcll(peers);
            }
        }
        __syncthreads();
        // offsets: digit-major exclusive scan over (digit, wave); threads
        // 0..255 (waves 0..3) take one digit each
        uint32_t dig_excl = 0;
        if (t < (uint32_t)kSortBins) {
            uint32_t run = 0;
            for (int ww = 0; ww < kForestWaves; ++ww) {
                const uint32_t c = whist[ww * kSortBins + t];
                whist[ww * kSortBins + t] = run;
                run += c;
            }
            uint32_t inc = run;
#pragma unroll
            for (int off = 1; off < 64; off <<= 1) {
                const uint32_t o = __shfl_up(inc, off, 64);
                if (lane >= (uint32_t)off) inc += o;
            }
            dig_excl = inc - run;
            if (lane == 63) wsum[w] = inc;
        }
        __syncthreads();
        if (t < (uint32_t)kSortBins) {
            for (uint32_t ww = 0; ww < w; ++ww) dig_excl += wsum[ww];
            for (int ww = 0; ww < kForestWaves; ++ww) whist[ww * kSortBins + t] += dig_excl;
        }
        __syncthreads();
        for (uint32_t g = b; g < e; g += 64 * kSortUnroll) {
            uint64_t key[kSortUnroll];
            uint32_t val[kSortUnroll];
            bool ok[kSortUnroll];
#pragma unroll
            for (int u = 0; u < kSortUnroll; ++u) {
                const uint32_t i = g + 64u * u + lane;
                ok[u] = i < e;
                key[u] = ok[u] ? sk[i] : 0ull;
                val[u] = ok[u] ? sv[i] : 0u;
            }
            // the groups in order: group u's ranks follow group u-1's
#pragma unroll
            for (int u = 0; u < kSortUnroll; ++u) {
                const uint32_t dig = (uint32_t)(key[u] >> sh) & (kSortBins - 1);
                const uint64_t peers = digit_peers(dig, ok[u]);
                // every lane reads its digit's offset before the leader advances it
                const uint32_t base = row[dig];
                if (ok[u]) {
                    const uint32_t pos = base + (uint32_t)__popcll(peers & below);
                    dk[pos] = key[u];
                    dv[pos] = val[u];
                    if ((peers & below) == 0) row[dig] = base + (uint32_t)__popcll(peers);
                }
            }
        }
        __syncthreads();
        uint64_t *tk = sk;
        sk = dk;
        dk = tk;
        uint32_t *tv = sv;
        sv = dv;
        dv = tv;
    }
}

__global__ void __launch_bounds__(kForestThreads) k_forest_build(
    const EnsSystem *__restrict__ sys, const uint32_t *__restrict__ order,
    const double *__restrict__ gx, const double *__restrict__ gy, const double *__restrict__ gz,
    const double *__restrict__ gm, double G, ForestBufs fb, uint32_t *items, uint32_t *status) {
    __shared__ uint32_t whist[kSortBins * kForestWaves];
    __shared__ uint32_t wsum[kForestWaves + 1];
    __shared__ uint32_t lvl_n[kMaxDepth + 1];  // cells per depth, then their start offsets
    __shared__ uint32_t lvl_c[kMaxDepth];      // fill cursors
    __shared__ int lcp_max;
    __shared__ double red[kForestWaves][6];
    __shared__ double cube[4];  // root cell {min x, min y, min z, edge}
    const uint32_t s = order[blockIdx.x];
    const EnsSystem S = sys[s];
    const uint32_t n = S.n, t = threadIdx.x;
    const double *x = gx + S.body_off, *y = gy + S.body_off, *z = gz + S.body_off,
                 *m = gm + S.body_off;
    TreeHdr *h = fb.hdr + s;
    const uint32_t fo = S.fb_off;
    uint64_t *skey = fb.skey + fo, *skey_alt = fb.skey_alt + fo, *klo = fb.klo + fo;
    uint32_t *sidx = fb.sidx + fo, *sidx_alt = fb.sidx_alt + fo, *chain = fb.chain + fo,
             *body_parent = fb.body_parent + fo, *group_size = fb.group_size + fo,
             *group_slot = fb.group_slot + fo;
    int32_t *lcp = fb.lcp + fo;
    uint2 *groups = fb.groups + fo;
    double4 *body4 = fb.body4 + fo;
    uint8_t *int_depth = fb.int_depth + S.int_off, *int_local = fb.int_local + S.int_off;
    uint32_t *int_parent = fb.int_parent + S.int_off, *int_node = fb.int_node + S.int_off,
             *lvl = fb.int_lvl + S.int_off, *counts = fb.int_count + S.int_off,
             *int_first = fb.int_first + S.int_off;
    InternalSums *sums = fb.int_sums + S.int_off;
    int32_t *ref = fb.ref + S.node_off;
    NodeF32 *rec = fb.rec + S.node_off;
    const uint32_t icap = S.icap;

    if (t == 0) {
        h->internal_count = 0;
        h->max_depth = 0;
        h->status = 0;
        h->deep_runs = 0;
        h->stack_peak = 0;
        h->group_count = 0;
        h->group_next = 0;
        h->corr_pairs = 0;
        lcp_max = -1;
    }
    __syncthreads();  // the reset precedes every flag raised below
    // 1. AABB including the origin (every partial starts at 0.0), then cubify
    {
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
            cube[0] = h->minx;
            cube[1] = h->miny;
            cube[2] = h->minz;
            cube[3] = h->edge;
        }
        __syncthreads();
    }
    // the workgroup reads the root cell from LDS, never back from the header
    const Cell root{cube[0], cube[1], cube[2], cube[3]};
    // 2. keys, 3. sort
    {
        for (uint32_t i = t; i < n; i += kForestThreads) {
            Cell c = root;
            const double px = x[i], py = y[i], pz = z[i];
            skey[i] = key_from(c, px, py, pz);
            sidx[i] = i;
            body4[i] = make_double4(px, py, pz, m[i]);
        }
        __syncthreads();
        wg_radix_sort(skey, sidx, skey_alt, sidx_alt, n, whist, wsum);
    }
    // 4./5. lcp, deep runs, cell numbering
    for (uint32_t i = t; i + 1 < n; i += kForestThreads)
        lcp_at(i, x, y, z, n, root, h, skey, sidx, klo, lcp);
    __syncthreads();
    {
        int lmax = -1;  // deepest cell shared by two neighbours
        for (uint32_t i = t; i + 1 < n; i += kForestThreads) {
            chain_at(i, lcp, chain);
            lmax = max(lmax, lcp[i]);
        }
#pragma unroll
        for (int off = 32; off >= 1; off >>= 1) lmax = max(lmax, __shfl_xor(lmax, off, 64));
        if ((t & 63u) == 0) atomicMax(&lcp_max, lmax);
    }
    __syncthreads();
    if (t == 0 && lcp_max >= 0) h->max_depth = (uint32_t)lcp_max + 1;  // the deepest leaf
    const uint32_t I_all = n > 1 ? wg_scan_array(chain, n - 1, wsum) : 0u;
    if (t == 0) h->internal_count = I_all;
    const uint32_t I = min(I_all, icap);
    // 6. topology
    for (uint64_t j = t; j < 1 + 8ull * I; j += kForestThreads) ref[j] = kRefEmpty;
    __syncthreads();
    const KeyView kv{skey, klo};
    for (uint32_t i = t; i < n; i += kForestThreads)
        link_at(i, n, icap, h, kv, sidx, lcp, chain, ref, int_depth, int_parent, int_node,
                body_parent, int_local, nullptr, 0u, int_first, n);
    __syncthreads();
    // 7. centre of mass, level by level: every cell is local to this workgroup
    for (uint32_t i = t; i < n; i += kForestThreads)
        if (m[i] == 0) raise_flag(h, kFlagZeroMass);
    com_levels(0u, I, [](uint32_t) { return true; }, int_depth, lvl, lvl_n, lvl_c, body4, ref, sums,
               counts, h);
    // 8. records and traversal groups
    const double root_edge = root.e;
    for (uint64_t j = t; j < 1 + 8ull * I; j += kForestThreads)
        record_at(j, root_edge, body4, ref, sums, int_depth, G, 0, 1, nullptr, nullptr, rec);
    for (uint32_t i = t; i < n; i += kForestThreads)
        group_mark_body(i, n, counts, body_parent, group_size, group_slot);
    __syncthreads();
    for (uint32_t k = t; k < I; k += kForestThreads)
        group_mark_cell(k, n, counts, int_parent, int_first, group_size, group_slot);
    __syncthreads();
    const uint32_t ng = wg_scan_array(group_slot, n, wsum);
    for (uint32_t i = t; i < n; i += kForestThreads) {
        const uint32_t gs = group_size[i];
        if (gs) groups[group_slot[i]] = make_uint2(i, gs);
    }
    __syncthreads();
    if (t == 0) {
        h->group_count = ng;
        items[s] = ng;
        // flags were raised with device-scope atomics: read the word the same way
        const uint32_t st = __hip_atomic_load(&h->status, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        if (st) atomicOr(status, st);
    }
}

}  // namespace

void launch_forest_build(const EnsSystem *sys, const uint32_t *order, uint32_t nbh,
                         const double *x, const double *y, const double *z, const double *m,
                         double G, const ForestBufs &fb, uint32_t *items, uint32_t *status,
                         hipStream_t s) {
    if (nbh == 0) return;
    hipLaunchKernelGGL(k_forest_build, dim3(nbh), dim3(kForestThreads), 0, s, sys, order, x, y, z,
                       m, G, fb, items, status);
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

void launch_keys(const TreeBufs &t, const double *x, const double *y, const double *z,
                 const double *m, hipStream_t s) {
    hipLaunchKernelGGL(k_keys, dim3(ceil_div(t.n, kThreads)), dim3(kThreads), 0, s, x, y, z, m,
                       t.n, (const TreeHdr *)t.hdr, t.skey, t.sidx, t.body4);
    radix_sort_u64_u32(t.skey, t.sidx, t.skey_alt, t.sidx_alt, t.n, 3 * kKeyLevels, t.radix_ctl,
                       &t.hdr->status, s);
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

void launch_com(const TreeBufs &t, hipStream_t s) {
    const uint32_t nblocks = ceil_div(t.n, kComBlock);
    hipLaunchKernelGGL(k_com_local, dim3(nblocks), dim3(kComThreads), 0, s, t.n, t.icap, t.hdr,
                       (const double4 *)t.body4, (const uint32_t *)t.sidx, (const int32_t *)t.ref,
                       (const uint32_t *)t.chain, (const uint8_t *)t.int_depth,
                       (const uint8_t *)t.int_local, t.int_lvl, t.int_sums, t.int_count);
    hipLaunchKernelGGL(k_com_spanning, dim3(1), dim3(kComThreads), 0, s, nblocks, t.hdr,
                       (const double4 *)t.body4, (const int32_t *)t.ref,
                       (const uint32_t *)t.spanning, t.int_sums, t.int_count);
}

void launch_node_records(const TreeBufs &t, double G, bool f64, bool f32, hipStream_t s) {
    const uint32_t blocks = std::max<uint32_t>(1, std::min<uint32_t>(8192, ceil_div(8ull * t.n, kThreads)));
    hipLaunchKernelGGL(k_records, dim3(blocks), dim3(kThreads), 0, s, (const TreeHdr *)t.hdr,
                       t.icap, (const double4 *)t.body4, (const int32_t *)t.ref,
                       (const InternalSums *)t.int_sums, (const uint8_t *)t.int_depth, G,
                       (int)f64, (int)f32, t.cm, t.edge, t.rec);
    if (f32) {
        const dim3 grid(ceil_div(t.n, kThreads));
        hipLaunchKernelGGL(k_group_marks_bodies, grid, dim3(kThreads), 0, s, t.n,
                           (const uint32_t *)t.int_count, (const uint32_t *)t.body_parent,
                           t.group_size, t.group_slot);
        const uint32_t gb = std::max<uint32_t>(1, std::min<uint32_t>(4096, ceil_div(t.n, kThreads)));
        hipLaunchKernelGGL(k_group_marks_cells, dim3(gb), dim3(kThreads), 0, s, t.n, t.icap,
                           (const TreeHdr *)t.hdr, (const uint32_t *)t.int_count,
                           (const uint32_t *)t.int_parent, (const uint32_t *)t.int_first,
                           t.group_size, t.group_slot);
        scan_exclusive_u32(t.group_slot, t.group_slot, t.n, &t.hdr->group_count, t.scan_part,
                           t.scan_part_cap, s);
        hipLaunchKernelGGL(k_group_scatter, grid, dim3(kThreads), 0, s, t.n,
                           (const uint32_t *)t.group_size, (const uint32_t *)t.group_slot,
                           t.groups);
    }
}

}  // namespace nb

// nb_internal.h -- shared declarations of libnbody_hip.so (not installed).
//
// Data layout in HBM (all arrays structure-of-arrays, 256-byte aligned):
//   state     x y z vx vy vz vhx vhy vhz ax ay az m        13 x N doubles
//   octree    per sorted position:  skey (u64), sidx (u32), klo (u64), lcp (i32)
//             per internal node k:  depth, parent, node id, child count,
//                                   arrival counter, sums {x*m, y*m, z*m, m}, count
//             per node id:          ref (i32), cm {com_x, com_y, com_z, m} (double4),
//                                   edge (double), r_open^2 (float)
// A node id is 0 for the root and 1 + 8*k + digit for the child with octant
// digit `digit` of internal node k, so the 8 children of a cell are adjacent.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "nbody_capi.h"

namespace nb {

constexpr int kKeyLevels = 21;  // octant digits held by one 63-bit key word
constexpr int kMaxDepth = 42;   // two key words
constexpr int kAabbBlocks = 1024;

// device status bits, turned into NB_E_* codes by the host
enum : uint32_t {
    kFlagNodeOverflow = 1u,
    kFlagDepth = 2u,
    kFlagZeroMass = 4u,
    kFlagStack = 8u,
    kFlagAabb = 16u,
    kFlagSortWait = 32u,  // radix sort look-back safety net (never expected)
    kFlagSortSize = 64u,  // radix sort asked for 2^30 keys or more: nothing sorted
};

// node reference encoding
constexpr int32_t kRefEmpty = -1;  // empty leaf
constexpr uint32_t kNoCell = 0xffffffffu;  // no internal cell (tables, parents)
__host__ __device__ inline int32_t ref_of_body(uint32_t b) { return -2 - (int32_t)b; }
__host__ __device__ inline uint32_t body_of_ref(int32_t r) { return (uint32_t)(-2 - r); }

// Written by the kernels, read by the host after a build.  One 128-byte line
// per header, so headers of an ensemble's systems never share a cache line.
struct alignas(128) TreeHdr {
    double minx, miny, minz, maxx, maxy, maxz, edge;  // root cube (cubified AABB)
    uint32_t internal_count;                           // internal nodes I; nodes = 1 + 8 I
    uint32_t max_depth;                                // deepest leaf
    uint32_t status;                                   // kFlag* bits
    uint32_t deep_runs;                                // runs of bodies sharing 21 digits
    uint32_t stack_peak;                              // deepest group-traversal stack seen
    uint32_t group_count;                              // traversal groups
    uint32_t group_next;                               // traversal work-queue cursor
    uint32_t cross_count;                              // cells crossing super-block boundaries
    unsigned long long corr_pairs;                     // learned-correction evaluations
    unsigned long long quad_pairs;                     // quadrupole cell interactions
};

// Bodies per traversal group: a group is an octree cell holding at most
// kGroupMax bodies whose parent holds more (or a lone body under such a
// parent), and the sibling groups of a cell with at most kPackMax bodies are
// packed, in digit order, into groups of at most kGroupMax bodies (fewer
// walks for somewhat larger boxes).  A 64-lane wave walks one group (the
// walk's template also has a two-bodies-per-lane form for groups of up to
// 128, compiled only if kGroupMax exceeds 64).
// Measured with the two-pass flush (groups whose size leaves lanes idle
// evaluate each list in two passes), same box, against 32 / 128:
//  - 64 / 256 (this): 4M-body Plummer sphere at its gated opening angle
//    325.3 M against 304.6 M body-steps/s (theta 0.374 against 0.365: the
//    larger boxes are stricter, so the error bound allows a larger theta),
//    traversal 11.66 against 12.53 ms; 1M-body disk 0.977 against 1.022 ms;
//  - 48 / 192: Plummer 316.8 M, disk 0.940 ms (the disk's best); 64 / 128:
//    315.8 M, 0.966 ms; 40 / 160: 313.4 M, 0.967 ms; 64 / 384 and 64 / 512:
//    Plummer 326.6 M and 326.5 M (the limit stays at 4 groups, away from the
//    unbounded case below).
// Before the two-pass flush, groups of 33 to 64 bodies left up to half the
// lanes idle and 32 / 128 was the fastest (64 / 128 and 96 / 192 slower on
// both inputs; one group per cell, no packing: disk +16 %, Plummer +6 %).
// Packing without a limit made the Plummer traversal ten times slower:
// siblings around a dense non-group sibling make a box that contains it, and
// the walk must resolve that cell to its leaves.
constexpr uint32_t kGroupMax = 64;
constexpr uint32_t kPackMax = 4 * kGroupMax;
// The largest group a walk takes (one body per lane of a 64-lane wave); the
// ensemble's autotuner chooses group_max in [1, kGroupLimit].
constexpr uint32_t kGroupLimit = 64;
struct GroupRule {
    uint32_t group_max;  // bodies per traversal group
    uint32_t pack_max;   // sibling groups are packed in cells of at most this many bodies
};

struct InternalSums {
    double x, y, z, m;  // centre-of-mass numerators and mass
};

// An octree cell: minimum corner and edge.
struct Cell {
    double mx, my, mz, e;
};

// One level of the reference's classification chain; returns the digit and
// moves the cell to the child (Subtrees.cpp:256-315, :348-358).
__device__ inline uint32_t descend(Cell &c, double px, double py, double pz) {
    const double h = c.e / 2.0;
    const double midx = c.mx + h, midy = c.my + h, midz = c.mz + h;
    const bool up = py > midy;
    const bool right = px > midx;
    const bool back = pz < midz;
    if (up) c.my = midy;
    if (right) c.mx = midx;
    if (!back) c.mz = midz;
    c.e = h;
    return (uint32_t)up * 4u + (uint32_t)right * 2u + (uint32_t)back;
}

// Traversal record of one node (32 bytes, children of a cell are adjacent):
// FP32 centre of mass split into a high part and the FP32 rounding of the
// remainder (hi + lo carries ~48 bits), G*m in FP32, and the node reference.
struct NodeF32 {
    float x, y, z, gm;
    int32_t ref;
    float lx, ly, lz;
};

// Bodies per block of the first centre-of-mass pass: cells whose body range
// stays inside one block are summed there, level by level in LDS (a disk or
// a Plummer sphere has about 0.48 internal cells per body, so a block's cells
// fit the LDS stage, and two blocks share a CU); the few cells crossing a
// block boundary are finished by the spanning passes.
constexpr uint32_t kComBlock = 1024;
// Blocks per super-block: cells crossing block boundaries are summed by one
// workgroup per super-block, those crossing super-block boundaries by one more.
constexpr uint32_t kComSuperBlocks = 64;
constexpr uint32_t kComSuper = kComBlock * kComSuperBlocks;

// Device buffers of one octree.
struct TreeBufs {
    uint32_t n = 0;
    uint32_t icap = 0;  // internal-node capacity
    uint64_t mcap = 0;  // node capacity = 1 + 8 icap
    TreeHdr *hdr = nullptr;
    double *aabb_part = nullptr;  // 6 * kAabbBlocks
    uint64_t *skey = nullptr, *skey_alt = nullptr, *klo = nullptr;
    uint32_t *sidx = nullptr, *sidx_alt = nullptr;
    int32_t *lcp = nullptr;       // n entries (pair i between sorted i and i+1)
    uint32_t *chain = nullptr;    // n entries: internal nodes owned by pair i, then offsets
    uint32_t *body_parent = nullptr;  // n: internal index of each sorted body's parent
    uint8_t *int_depth = nullptr;
    uint32_t *int_parent = nullptr, *int_node = nullptr, *int_count = nullptr;
    uint32_t *int_lvl = nullptr;    // centre of mass: depth buckets, or LDS ids of spanning cells
    InternalSums *int_sums = nullptr;
    uint8_t *int_local = nullptr;   // 1: bodies in one kComBlock block; 2: crosses a kComSuper
                                    // boundary; 0: crosses block boundaries only
    uint32_t *int_first = nullptr;  // first sorted position of the cell's bodies
    uint2 *groups = nullptr;        // traversal groups {first sorted position, body count}
    uint32_t *group_size = nullptr; // per sorted position: size of the group starting there
    uint32_t *group_slot = nullptr; // per sorted position: 1 if a group starts there, then slots
    uint32_t *spanning = nullptr;   // [block * kMaxDepth + depth]: cell crossing the block's end
    uint32_t *cross = nullptr;      // cells crossing super-block boundaries: k << 6 | depth
    uint32_t cross_cap = 0;
    double4 *body4 = nullptr;       // {x, y, z, m} per body, input order: one load per body
    uint32_t *gq = nullptr;         // traversal queue counters, one per XCD, 128 bytes apart
    static constexpr uint32_t kGroupQueueWords = 8 * 32;
    uint32_t spanning_cap = 0;
    int32_t *ref = nullptr;
    double4 *cm = nullptr;          // FP64 records (reference-semantics traversal)
    double *edge = nullptr;
    NodeF32 *rec = nullptr;         // FP32 records (group traversal)
    // scan / sort scratch
    uint32_t *scan_part = nullptr;
    uint32_t *radix_ctl = nullptr;  // radix sort control words (owned by the sort)
    uint32_t scan_part_cap = 0;
};

// ---- ensemble of independent systems ------------------------------------------
// One concatenated state (bodies of system s at [body_off, body_off + n)); a
// Barnes-Hut system owns per-body forest entries [fb_off, fb_off + n),
// internal cells [int_off, int_off + icap) and nodes [node_off, node_off +
// 1 + 8 icap) of the forest arrays, each range starting on a fresh 64-entry
// boundary so that no cache line holds two systems' data; its body indices in
// sidx and node references are local to the system.
struct EnsSystem {
    uint32_t body_off, n;
    uint32_t int_off, icap;
    uint64_t node_off;
    uint32_t fb_off;
    uint32_t bh;     // 1: Barnes-Hut, 0: direct sum
    // the centre-of-mass passes' tables: kMaxDepth entries per kComBlock-body
    // block (cells crossing its end), and per kComSuper-body super-block + 1
    // (cells crossing super-block ends)
    uint32_t span_off, cross_off;
};
constexpr uint32_t kEnsAlign = 64;

struct ForestBufs {
    TreeHdr *hdr;  // one per system
    // per body (offset body_off)
    uint64_t *skey, *skey_alt, *klo;
    uint32_t *sidx, *sidx_alt, *chain, *body_parent, *group_size, *group_slot;
    int32_t *lcp;
    uint2 *groups;
    double4 *body4;  // {x, y, z, m} in the system's input order
    // per internal cell (offset int_off)
    uint8_t *int_depth, *int_local;
    uint32_t *int_parent, *int_node, *int_lvl, *int_count, *int_first;
    InternalSums *int_sums;
    // per node (offset node_off)
    int32_t *ref;
    NodeF32 *rec;
    // the body slots of all systems (padding included): slot -> system, and
    // the control words of the forest-wide radix sorts, one buffer per sort
    // (by key, then by system): a buffer's rotation assumes one pass count
    uint32_t slots;
    uint16_t *fb_sys;
    uint32_t *sys_first;  // per system: its first slot (ascending)
    uint32_t *radix_ctl, *radix_ctl_sys;
    uint32_t *spanning, *cross;  // tables of every system (EnsSystem::span_off, cross_off)
    // the centre-of-mass launches over the Barnes-Hut systems: (system,
    // block) of every kComBlock-body block, (system, super-block) of the
    // systems of two or more blocks, and the systems of more than one
    // super-block; counts kept on the host
    uint2 *blk_tab, *span_tab, *big_tab;
    uint32_t nblk, nspan, nbig;
    // the group scan over all slots: scan_scratch_words(slots) partials and
    // the total in the last word
    uint32_t *scan_part;
    uint32_t scan_cap;
};

// ---- launchers (each .hip file defines the kernels it launches) -------------
// nb_scan.hip
// Exclusive scan of n u32 values (out may alias in); *total_dev receives the sum.
void scan_exclusive_u32(const uint32_t *in, uint32_t *out, uint32_t n, uint32_t *total_dev,
                        uint32_t *scratch, uint32_t scratch_cap, hipStream_t s);
uint32_t scan_scratch_words(uint32_t n);
// Where a radix sort left its result.
enum class SortEnd {
    kInPlace,   // keys/vals (an even pass count, or n <= 1)
    kInAlt,     // keys_alt/vals_alt (an odd pass count)
    kRejected,  // n >= 2^30 or bits outside [1, 63]: nothing sorted, kFlagSortSize raised
};
// Stable LSD radix sort of (key, val) pairs on key bits [0, bits), bits <= 63,
// n < 2^30.  ctl holds
// radix_ctl_words(n) words, zeroed at allocation and then owned by the sort
// (one ctl per stream: sorts sharing it must not overlap); a failed look-back
// wait raises kFlagSortWait in *flags.
SortEnd radix_sort_u64_u32(uint64_t *keys, uint32_t *vals, uint64_t *keys_alt,
                           uint32_t *vals_alt, uint32_t n, int bits, uint32_t *ctl,
                           uint32_t *flags, hipStream_t s);
uint32_t radix_blocks_for(uint32_t n);
uint32_t radix_ctl_words(uint32_t n);

// nb_octree.hip
void launch_aabb(const TreeBufs &t, const double *x, const double *y, const double *z,
                 hipStream_t s);
// keys and t.body4 = {x, y, z, m} of every body (input order)
void launch_keys(TreeBufs &t, const double *x, const double *y, const double *z,
                 const double *m, hipStream_t s);
// radix sort of the keys (the sorted keys and body indices end in
// t.skey/t.sidx: the buffers swap with their scratch copies when the sort
// ends there)
void launch_key_sort(TreeBufs &t, hipStream_t s);
void launch_topology(const TreeBufs &t, const double *x, const double *y, const double *z,
                     hipStream_t s);
// Centre of mass with the node records and traversal groups.  f64: FP64
// records (cm, edge) for the reference-semantics traversal; f32: packed FP32
// records and traversal groups for the group traversal.
void launch_com(const TreeBufs &t, double G, bool f64, bool f32, GroupRule rule, hipStream_t s);
// hdr->max_depth of the last build (not maintained by the build itself)
void launch_max_depth(const TreeBufs &t, hipStream_t s);

// Octree of every Barnes-Hut system of an ensemble: the one-system build's
// stages, each over the slots or blocks of all systems (order: the BH
// systems' indices, largest first, for the per-system stages; fb's block
// tables list their blocks); items[s] receives the system's traversal-group
// count.
void launch_forest_build(const EnsSystem *sys, const uint32_t *order, uint32_t nbh,
                         uint32_t nsys, const double *x, const double *y, const double *z,
                         const double *m, double G, const ForestBufs &fb, GroupRule rule,
                         uint32_t *items, uint32_t *status, hipStream_t s);

// nb_forces.hip
void launch_direct_f64(uint32_t n, const double *x, const double *y, const double *z,
                       const double *m, double G, double eps2, double *ax, double *ay,
                       double *az, hipStream_t s);
void launch_direct_targets_f64(uint32_t n, const double *x, const double *y, const double *z,
                               const double *m, double G, double eps2, uint32_t nt,
                               const uint32_t *targets, double *ax, double *ay, double *az,
                               hipStream_t s);
// Split direct sum (FP32 pairs with FP64 accumulation, or fp64: FP64 pairs
// with a corrected v_rsq_f64 estimate), deterministic; partial holds
// direct_split_scratch_bytes(n) bytes (none needed with one split).
void launch_direct_split(bool fp64, uint32_t n, const double *x, const double *y,
                         const double *z, const double *m, double G, double eps2, double *ax,
                         double *ay, double *az, double *partial, hipStream_t s);
uint32_t direct_split_count(uint32_t n);
uint64_t direct_split_scratch_bytes(uint32_t n);
void launch_bh_exact(const TreeBufs &t, const double *x, const double *y, const double *z,
                     double G, double eps2, double theta, int sort_bodies, uint32_t *stack,
                     uint32_t stack_cap, double *ax, double *ay, double *az,
                     uint32_t *interactions, hipStream_t s);
// gquad: G*Q records of the internal cells (two float4 each) for the
// quadrupole walk, or null for monopole cells
void launch_bh_group(const TreeBufs &t, const double *x, const double *y, const double *z,
                     const double *m, double eps2, double theta, const float4 *gquad, double *ax,
                     double *ay, double *az, uint32_t *interactions, hipStream_t s);

// per-body kinetic and potential energy terms (reference arithmetic)
void launch_energy_terms(uint32_t n, const double *x, const double *y, const double *z,
                         const double *vx, const double *vy, const double *vz, const double *m,
                         double G, double *ekin, double *epot, hipStream_t s);

// nb_integrate.hip
void launch_kick_drift(uint32_t n, double dt, const double *ax, const double *ay,
                       const double *az, const double *vx, const double *vy, const double *vz,
                       double *vhx, double *vhy, double *vhz, double *x, double *y, double *z,
                       hipStream_t s);
void launch_kick(uint32_t n, double dt, const double *ax, const double *ay, const double *az,
                 const double *vhx, const double *vhy, const double *vhz, double *vx,
                 double *vy, double *vz, hipStream_t s);
// kick of step k and kick-drift of step k+1 in one pass; bit-identical to
// launch_kick followed by launch_kick_drift, without storing v
void launch_kick_kick_drift(uint32_t n, double dt, const double *ax, const double *ay,
                            const double *az, double *vhx, double *vhy, double *vhz, double *x,
                            double *y, double *z, hipStream_t s);
// ensemble: body i advances by dt_sys[sys_of[i]]
void launch_ens_kick_kick_drift(uint32_t n, const double *dt_sys, const uint16_t *sys_of,
                                const double *ax, const double *ay, const double *az, double *vhx,
                                double *vhy, double *vhz, double *x, double *y, double *z,
                                hipStream_t s);
void launch_ens_kick_drift(uint32_t n, const double *dt_sys, const uint16_t *sys_of,
                           const double *ax, const double *ay, const double *az, const double *vx,
                           const double *vy, const double *vz, double *vhx, double *vhy,
                           double *vhz, double *x, double *y, double *z, hipStream_t s);
void launch_ens_kick(uint32_t n, const double *dt_sys, const uint16_t *sys_of, const double *ax,
                     const double *ay, const double *az, const double *vhx, const double *vhy,
                     const double *vhz, double *vx, double *vy, double *vz, hipStream_t s);

// ensemble forces: one persistent launch over a queue of work items -- the
// traversal groups of Barnes-Hut systems and 64-body direct-sum tiles of the
// others -- in queue order qorder (item_off[q] = first item of position q).
struct EnsForceArgs {
    const EnsSystem *sys;
    const uint32_t *qorder;
    const uint32_t *item_off;  // nq + 1 entries
    uint32_t nq;
    const double *x, *y, *z, *m;
    double *ax, *ay, *az;
    ForestBufs fb;
    double G, theta;
    float eps2;
    uint32_t *cursor;  // work-queue cursor (reset by launch_ens_items)
    uint32_t *status;  // kFlag* bits
    uint32_t *inter;   // per body: accepted interactions (Barnes-Hut) or sources (direct)
    uint32_t direct_tpl;  // direct-sum targets per lane: 1 or 2 (items of 64 or 128 targets)
};
constexpr uint32_t kEnsTile = 64;  // targets per direct-sum item and lane
// item_off = exclusive prefix of items[qorder[q]]; resets the cursor and status
void launch_ens_items(uint32_t nq, const uint32_t *qorder, const uint32_t *items,
                      uint32_t *item_off, uint32_t *cursor, uint32_t *status, hipStream_t s);
void launch_ens_forces(const EnsForceArgs &a, uint32_t waves, hipStream_t s);

// ---------------------------------------------------------------------------
// Correctly rounded FP64 square root and division for the reference-order
// kernels, within ranges where the compiler's IEEE sequences need none of
// their range handling.  The compiler's square root takes a v_rsq_f64
// estimate, one Goldschmidt step for the root and its half-reciprocal and two
// residual corrections, scaling inputs below 2^-767 and passing zero and
// infinity through; its division takes v_rcp_f64 of the divisor, two Newton
// steps and one residual correction of the quotient, with operand scaling
// (v_div_scale_f64) and a special-value fix-up (v_div_fixup_f64) that change
// nothing while dividend, divisor and quotient are normal and their exponents
// differ by less than 768.  The cores below run the same operations without
// that handling, so inside the guarded ranges they give the same bits; the
// GPU tests compare every pair of the direct sum, the traversal and the
// energy with the oracle's IEEE operations.  Outside: the plain operators.
__device__ inline double sqrt_core(double v) {  // v in [2^-767, 2^1023)
    const double g = __builtin_amdgcn_rsq(v);
    double s = v * g;
    double h = g * 0.5;
    const double r = fma(-h, s, 0.5);
    s = fma(s, r, s);
    h = fma(h, r, h);
    double d = fma(-s, s, v);
    s = fma(d, h, s);
    d = fma(-s, s, v);
    return fma(d, h, s);
}
__device__ inline double div_core(double a, double b) {  // |a|, |b| in [2^-300, 2^300]
    double y = __builtin_amdgcn_rcp(b);
    double e = fma(-b, y, 1.0);
    y = fma(y, e, y);
    e = fma(-b, y, 1.0);
    y = fma(y, e, y);
    const double q = a * y;
    e = fma(-b, q, a);
    return fma(e, y, q);
}
// 1.0 / sqrt(v), both operations correctly rounded (the reference's
// inverse square root); the core for v in [2^-512, 2^512]
__device__ inline double inv_sqrt_ieee(double v) {
    if (!(v >= 0x1p-512 && v <= 0x1p512)) return 1.0 / sqrt(v);
    const double s = sqrt_core(v);
    double y = __builtin_amdgcn_rcp(s);
    double e = fma(-s, y, 1.0);
    y = fma(y, e, y);
    e = fma(-s, y, 1.0);
    y = fma(y, e, y);
    e = fma(-s, y, 1.0);  // the quotient 1.0 * y is y
    return fma(e, y, y);
}
// a / sqrt(v), both operations correctly rounded; the cores for v in
// [2^-600, 2^600] and |a| in [2^-300, 2^300]
__device__ inline double div_sqrt_ieee(double a, double v) {
    const double fa = fabs(a);
    if (!(v >= 0x1p-600 && v <= 0x1p600 && fa >= 0x1p-300 && fa <= 0x1p300)) return a / sqrt(v);
    return div_core(a, sqrt_core(v));
}

inline uint32_t ceil_div(uint64_t a, uint64_t b) { return (uint32_t)((a + b - 1) / b); }

// Host: canonical depth-first dump of an octree (children in digit order),
// from host copies of its node references, cell sums and counts and of the
// bodies.  Returns the record count, 0 if cap is too small.
inline uint64_t canonical_dump(const int32_t *ref, const InternalSums *sums, const uint32_t *cnt,
                               const double *hx, const double *hy, const double *hz,
                               const double *hm, uint32_t n, uint64_t cap, uint8_t *depth,
                               uint8_t *kind, uint32_t *body, double *mass, double *com_x,
                               double *com_y, double *com_z, uint32_t *count) {
    struct Item {
        uint64_t node;
        uint32_t d;
    };
    Item stack[1 + 7 * (kMaxDepth + 1)];
    int sp = 0;
    stack[sp++] = {0, 0};
    uint64_t k = 0;
    while (sp > 0) {
        const Item it = stack[--sp];
        if (k >= cap) return 0;
        const int32_t r = ref[it.node];
        depth[k] = (uint8_t)it.d;
        if (r >= 0) {
            kind[k] = 2;
            body[k] = n;
            mass[k] = sums[r].m;
            com_x[k] = sums[r].x;
            com_y[k] = sums[r].y;
            com_z[k] = sums[r].z;
            count[k] = cnt[r];
            if (sp + 8 > (int)(sizeof(stack) / sizeof(stack[0]))) return 0;  // deeper than 42 levels
            for (int dg = 7; dg >= 0; --dg) stack[sp++] = {1ull + 8ull * (uint64_t)r + dg, it.d + 1};
        } else if (r != kRefEmpty) {
            const uint32_t b = body_of_ref(r);
            kind[k] = 1;
            body[k] = b;
            mass[k] = hm[b];
            com_x[k] = hx[b] * hm[b];
            com_y[k] = hy[b] * hm[b];
            com_z[k] = hz[b] * hm[b];
            count[k] = 1;
        } else {
            kind[k] = 0;
            body[k] = n;
            mass[k] = com_x[k] = com_y[k] = com_z[k] = 0.0;
            count[k] = 0;
        }
        ++k;
    }
    return k;
}

}  // namespace nb

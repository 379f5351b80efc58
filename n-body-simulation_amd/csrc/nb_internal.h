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
};

// node reference encoding
constexpr int32_t kRefEmpty = -1;  // empty leaf
__host__ __device__ inline int32_t ref_of_body(uint32_t b) { return -2 - (int32_t)b; }
__host__ __device__ inline uint32_t body_of_ref(int32_t r) { return (uint32_t)(-2 - r); }

// Written by the kernels, read by the host after a build.
struct TreeHdr {
    double minx, miny, minz, maxx, maxy, maxz, edge;  // root cube (cubified AABB)
    uint32_t internal_count;                           // internal nodes I; nodes = 1 + 8 I
    uint32_t max_depth;                                // deepest leaf
    uint32_t status;                                   // kFlag* bits
    uint32_t deep_runs;                                // runs of bodies sharing 21 digits
    uint32_t spanning_count;                           // cells crossing a COM block boundary
    uint32_t stack_peak;                               // deepest group-traversal stack seen
    uint32_t group_count;                              // traversal groups
    uint32_t group_next;                               // traversal work-queue cursor
    unsigned long long corr_pairs;                     // learned-correction evaluations
};

// Bodies per traversal group: a group is an octree cell holding at most
// kGroupMax bodies whose parent holds more (or a lone body under such a parent).
constexpr uint32_t kGroupMax = 64;

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
// stays inside one block are summed there with workgroup-scope atomics; the
// few cells crossing a block boundary are finished by one workgroup.
constexpr uint32_t kComBlock = 4096;

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
    uint32_t *int_parent = nullptr, *int_node = nullptr, *int_nchild = nullptr,
             *int_arrive = nullptr, *int_count = nullptr;
    InternalSums *int_sums = nullptr;
    uint8_t *int_local = nullptr;   // 1 if the cell's bodies lie in one kComBlock block
    uint32_t *int_first = nullptr;  // first sorted position of the cell's bodies
    uint2 *groups = nullptr;        // traversal groups {first sorted position, body count}
    uint32_t *group_size = nullptr; // per sorted position: size of the group starting there
    uint32_t *group_slot = nullptr; // per sorted position: 1 if a group starts there, then slots
    uint32_t *spanning = nullptr;   // cells crossing a block boundary
    uint32_t spanning_cap = 0;
    int32_t *ref = nullptr;
    double4 *cm = nullptr;          // FP64 records (reference-semantics traversal)
    double *edge = nullptr;
    NodeF32 *rec = nullptr;         // FP32 records (group traversal)
    // scan / sort scratch
    uint32_t *scan_part = nullptr;
    uint32_t *radix_hist = nullptr;
    uint32_t radix_blocks = 0;
    uint32_t scan_part_cap = 0;
};

// ---- launchers (each .hip file defines the kernels it launches) -------------
// nb_scan.hip
// Exclusive scan of n u32 values (out may alias in); *total_dev receives the sum.
void scan_exclusive_u32(const uint32_t *in, uint32_t *out, uint32_t n, uint32_t *total_dev,
                        uint32_t *scratch, uint32_t scratch_cap, hipStream_t s);
uint32_t scan_scratch_words(uint32_t n);
// Stable LSD radix sort of (key, val) pairs on key bits [0, bits).  Result in
// keys/vals (the alt buffers are scratch).
void radix_sort_u64_u32(uint64_t *keys, uint32_t *vals, uint64_t *keys_alt, uint32_t *vals_alt,
                        uint32_t n, int bits, uint32_t *hist, uint32_t hist_blocks,
                        uint32_t *scan_scratch, uint32_t scan_cap, hipStream_t s);
uint32_t radix_blocks_for(uint32_t n);

// nb_octree.hip
void launch_aabb(const TreeBufs &t, const double *x, const double *y, const double *z,
                 hipStream_t s);
void launch_keys(const TreeBufs &t, const double *x, const double *y, const double *z,
                 hipStream_t s);
void launch_topology(const TreeBufs &t, const double *x, const double *y, const double *z,
                     hipStream_t s);
void launch_com(const TreeBufs &t, const double *x, const double *y, const double *z,
                const double *m, hipStream_t s);
// f64: FP64 records (cm, edge) for the reference-semantics traversal;
// f32: packed FP32 records and traversal groups for the group traversal.
void launch_node_records(const TreeBufs &t, const double *x, const double *y, const double *z,
                         const double *m, double G, bool f64, bool f32, hipStream_t s);

// nb_forces.hip
void launch_direct_f64(uint32_t n, const double *x, const double *y, const double *z,
                       const double *m, double G, double eps2, double *ax, double *ay,
                       double *az, hipStream_t s);
void launch_direct_targets_f64(uint32_t n, const double *x, const double *y, const double *z,
                               const double *m, double G, double eps2, uint32_t nt,
                               const uint32_t *targets, double *ax, double *ay, double *az,
                               hipStream_t s);
void launch_direct_f32c(uint32_t n, const double *x, const double *y, const double *z,
                        const double *m, double G, double eps2, double *ax, double *ay,
                        double *az, double *partial, uint32_t splits, hipStream_t s);
uint32_t direct_f32c_splits(uint32_t n);
void launch_bh_exact(const TreeBufs &t, const double *x, const double *y, const double *z,
                     double G, double eps2, double theta, int sort_bodies, uint32_t *stack,
                     uint32_t stack_cap, double *ax, double *ay, double *az,
                     uint32_t *interactions, hipStream_t s);
void launch_bh_group(const TreeBufs &t, const double *x, const double *y, const double *z,
                     const double *m, double eps2, double theta, double *ax, double *ay,
                     double *az, uint32_t *interactions, hipStream_t s);

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

inline uint32_t ceil_div(uint64_t a, uint64_t b) { return (uint32_t)((a + b - 1) / b); }

}  // namespace nb

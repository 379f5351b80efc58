/*
 * nbody_capi.h -- C ABI of libnbody_hip.so, the MI355X (gfx950) n-body hot path.
 *
 * The reference (TimThuering/N-Body-Simulation) has no C boundary: its hot
 * path sits behind the C++ class nBodyAlgorithm and its two subclasses
 * (src/simulationBackend/nBodyAlgorithm.hpp:19-187, NaiveAlgorithm.hpp,
 * BarnesHutAlgorithm.hpp), which own SYCL buffers and run every kernel with
 * queue.submit(...).wait().  This header is the replacement seam: the
 * re-created host classes (n-body-simulation_amd/host/) call these entry
 * points, and so can any FFI (ctypes, cgo, JNI).  Each entry point names the
 * reference function it replaces.
 *
 * Conventions
 *  - Plain pointers and sizes; all pointers are HOST pointers unless the name
 *    says _device.
 *  - Every call is synchronous (it returns after its GPU work finished), the
 *    same contract as the reference's .wait() after each submit.
 *  - Every call returns NB_OK (0) or an NB_E_* code; nb_last_error() gives a
 *    message.  Device-side failures (octree storage exhausted, bodies that
 *    cannot be separated, zero-mass cells, traversal stack exhausted) are
 *    reported as codes instead of the reference's silent corruption or hang.
 *  - One host thread per context.  A context owns one HIP stream on the
 *    device that was current when it was created.
 */
#ifndef NBODY_CAPI_H
#define NBODY_CAPI_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes ------------------------------------------------------ */
enum {
    NB_OK = 0,
    NB_E_INVALID_ARG = 1,
    NB_E_HIP = 2,             /* a HIP runtime call failed */
    NB_E_NODE_OVERFLOW = 3,   /* octree needs more nodes than storage_size_param allows */
    NB_E_STACK_OVERFLOW = 4,  /* per-body traversal stack exhausted */
    NB_E_DEPTH = 5,           /* two bodies share all 42 octree digits (coincident bodies) */
    NB_E_ZERO_MASS = 6,       /* a non-empty cell has zero mass (reference: endless COM loop) */
    NB_E_NOT_BUILT = 7,       /* forces requested before an octree was built */
    NB_E_AABB = 8,            /* bounding box has no longest edge (non-finite positions) */
    NB_E_NO_DEVICE = 10
};

/* ---- parameters ---------------------------------------------------------- */
enum {
    /* Reference arithmetic: FP64 pair forces, the reference's summation order
     * and per-body opening test. */
    NB_PREC_FP64 = 0,
    /* FP32 pair arithmetic with FP64 accumulation of per-tile FP32 partials;
     * Barnes-Hut opens a cell when any body of a 64-body group needs it. */
    NB_PREC_FP32C = 1,
    /* Direct sum in FP64 pair arithmetic with the inverse square root from
     * the hardware estimate and a second-order correction, fused
     * multiply-adds, and the
     * sources split over many waves whose partial sums are added in a fixed
     * order: deterministic, within a stated relative tolerance of the
     * reference's sum rather than bit-identical.  Barnes-Hut runs as NB_PREC_FP64. */
    NB_PREC_FP64_FAST = 2
};

enum { NB_ALGO_NAIVE = 0, NB_ALGO_BARNES_HUT = 1 };

typedef struct nb_params {
    double G;             /* nBodyAlgorithm.hpp:55-61 (AU^3 kg^-1 d^-2 by default) */
    double eps2;          /* configuration::epsilon2, Configuration.cpp:6 */
    double theta;         /* configuration::barnes_hut_algorithm::theta, Configuration.cpp:10 */
    uint32_t precision;   /* NB_PREC_* for forces (integrator, AABB, COM are always FP64) */
    uint32_t sort_bodies; /* configuration::barnes_hut_algorithm::sortBodies, Configuration.cpp:15 */
    uint32_t storage_size_param; /* main.cpp:122-127; node capacity = param * N (default 16) */
    uint32_t stack_size_param;   /* main.cpp:129-134; traversal stack rule (default 16) */
    /* No counterpart in the reference (which uses monopole cells,
     * BarnesHutAlgorithm.cpp:361-367): nonzero adds each accepted cell's
     * traceless quadrupole moment to its monopole force in the NB_PREC_FP32C
     * group traversal, so a larger theta reaches the same accuracy.  The
     * FP64 modes ignore it.  Default 0. */
    uint32_t quadrupole;
    /* No counterpart in the reference: the NB_PREC_FP32C traversal walks the
     * octree once per group of nearby bodies, a cell of at most group_max
     * bodies (1..64); sibling groups inside a cell of at most pack_max bodies
     * are packed into groups of up to group_max.  0 selects the defaults
     * (64 and 256). */
    uint32_t group_max;
    uint32_t pack_max;
} nb_params;

/* Fill *p with the reference defaults (G, eps2 = 1e-22, theta = 1.05,
 * NB_PREC_FP64, sorting on, storage 16, stack 16, monopole cells;
 * Configuration.cpp:5-33). */
void nb_default_params(nb_params *p);

/* ---- context ------------------------------------------------------------- */
typedef struct nb_ctx nb_ctx;

/* Allocates device state for n bodies (the reference allocates in the
 * nBodyAlgorithm/BarnesHutOctree constructors, BarnesHutOctree.cpp:5-43,
 * BarnesHutAlgorithm.cpp:8-15). */
int nb_create(nb_ctx **out, uint32_t n, const nb_params *p);
int nb_destroy(nb_ctx *ctx);
int nb_set_params(nb_ctx *ctx, const nb_params *p);
int nb_get_params(const nb_ctx *ctx, nb_params *p);
uint32_t nb_body_count(const nb_ctx *ctx);
const char *nb_last_error(const nb_ctx *ctx);
const char *nb_strerror(int code);
/* Name of the device the library runs on (info::device::name in the reference). */
int nb_device_name(char *buf, size_t len);

/* Host -> device copy of the initial state (NaiveAlgorithm.cpp:27-66). */
int nb_set_state(nb_ctx *ctx, const double *x, const double *y, const double *z,
                 const double *vx, const double *vy, const double *vz, const double *m);
/* Device -> host copies (the reference's host_accessor reads,
 * NaiveAlgorithm.cpp:172-180, :233-242).  Any pointer may be NULL. */
int nb_get_state(nb_ctx *ctx, double *x, double *y, double *z, double *vx, double *vy,
                 double *vz);
int nb_get_accel(nb_ctx *ctx, double *ax, double *ay, double *az);
/* Asynchronous copy-out of a visualised step, replacing the synchronous
 * snapshot reads of NaiveAlgorithm.cpp:170-181, :227-243 (and the
 * acceleration read of storeAccelerations, nBodyAlgorithm.cpp:88-102).
 * nb_snapshot_begin enqueues, behind the work already issued on the context,
 * a device-side copy of x, y, z, vx, vy, vz, ax, ay, az into a staging buffer
 * and a DMA of it into pinned host memory on a second stream, and returns
 * without waiting: the next nb_steps runs while the DMA is in flight.
 * nb_snapshot_wait blocks until the DMA of the last begin is done and stores
 * nine pointers (x, y, z, vx, vy, vz, ax, ay, az; n doubles each) into
 * `arrays`.  They point into one of two pinned host slots that successive
 * begins use in turn, hold the state as it was at that begin, and stay valid
 * until the second nb_snapshot_begin after it (or nb_destroy): a caller may
 * read one snapshot on its own threads while the next one is in flight.
 * Waiting with no snapshot begun is NB_E_INVALID_ARG.  Like every call on a
 * context, begin and wait are made from one thread at a time. */
int nb_snapshot_begin(nb_ctx *ctx);
int nb_snapshot_wait(nb_ctx *ctx, const double **arrays);
/* Overwrite positions only (used by tests to evaluate forces on given points). */
int nb_set_positions(nb_ctx *ctx, const double *x, const double *y, const double *z);

/* ---- hot path ------------------------------------------------------------ */
/* Direct-sum accelerations at the current positions
 * (NaiveAlgorithm::computeAccelerations_opt_{0,1,2}, NaiveAlgorithm.cpp:262-482). */
int nb_forces_direct(nb_ctx *ctx);
/* FP64 direct sum for the listed targets only, in reference arithmetic; the
 * FP64 error reference on a sample of bodies.  Outputs have nt entries. */
int nb_forces_direct_targets(nb_ctx *ctx, uint32_t nt, const uint32_t *targets, double *ax,
                             double *ay, double *az);
/* Octree for the current positions: AABB, octant keys, radix sort, topology,
 * centre of mass, body order (BarnesHutOctree::buildOctree,
 * BarnesHutOctree.hpp:102-103; ParallelOctreeTopDownSubtrees.cpp:15-93). */
int nb_bh_build(nb_ctx *ctx);
/* Barnes-Hut accelerations from the last octree
 * (BarnesHutAlgorithm::computeAccelerations, BarnesHutAlgorithm.cpp:280-401). */
int nb_bh_forces(nb_ctx *ctx);
/* Leapfrog part 1: v_half = v + a*dt/2; x += v_half*dt (NaiveAlgorithm.cpp:140-164). */
int nb_kick_drift(nb_ctx *ctx, double dt);
/* Leapfrog part 2: v = v_half + a*dt/2 (NaiveAlgorithm.cpp:206-222). */
int nb_kick(nb_ctx *ctx, double dt);
/* `count` iterations of the reference's loop body (NaiveAlgorithm.cpp:131-259,
 * BarnesHutAlgorithm.cpp:149-277 without the visualisation copies):
 * kick-drift, [octree build], forces, kick.  One host synchronisation at the
 * end.  algorithm = NB_ALGO_*. */
int nb_steps(nb_ctx *ctx, int algorithm, double dt, uint32_t count);

/* Direct sum without any allocation.  No counterpart in the reference: for
 * callers that cannot allocate device memory (code under HIP graph capture,
 * serving loops, harnesses that forbid hipMalloc).  It runs the kernels of
 * nb_forces_direct (precision NB_PREC_FP32C or NB_PREC_FP64_FAST) on DEVICE
 * arrays x, y, z, m (n doubles each) into DEVICE arrays ax, ay, az, scaled by
 * G with softening eps2, enqueued on `stream` (a hipStream_t; NULL is the
 * null stream), and returns without waiting.  The partial sums of the
 * source splits (3 * 8 * n * splits bytes, at most 24 MiB for every n) live in
 * a library-owned device array of NB_NOALLOC_SCRATCH_BYTES per GPU, held until
 * the process ends; n above NB_NOALLOC_MAX_BODIES is NB_E_INVALID_ARG.  All
 * calls on one GPU share that array: calls that overlap on one GPU race and
 * give undefined results.  The first call on a device looks the array up
 * (hipGetSymbolAddress), which allocates nothing. */
#define NB_NOALLOC_SCRATCH_BYTES (24ull * 1024 * 1024)
#define NB_NOALLOC_MAX_BODIES (1u << 23)
int nb_forces_direct_noalloc(uint32_t n, const double *x, const double *y, const double *z,
                             const double *m, double G, double eps2, uint32_t precision,
                             double *ax, double *ay, double *az, void *stream);

/* System energies at the current state (nBodyAlgorithm::computeEnergy,
 * nBodyAlgorithm.cpp:11-86): kinetic sum of 0.5*m*v^2, potential
 * -sum_{i<j} G m_i m_j / r without softening, per-body terms summed in body
 * order on the host as the reference does.  Any output pointer may be NULL. */
int nb_energy(nb_ctx *ctx, double *kinetic, double *potential, double *total,
              double *virial);

/* ---- timing (TimeMeasurement keys, BarnesHutAlgorithm.cpp:81-100) --------- */
enum {
    NB_PHASE_AABB = 0,        /* "AABB creation" */
    NB_PHASE_KEYS_SORT = 1,   /* keys + radix sort ("Build octree to level" + "Sort bodies for subtrees") */
    NB_PHASE_TOPOLOGY = 2,    /* LCP, node numbering, linking ("Build subtrees") */
    NB_PHASE_COM = 3,         /* "Compute center of mass" */
    NB_PHASE_OCTREE = 4,      /* "Octree creation" (sum of the above + node records) */
    NB_PHASE_ACCEL = 5,       /* "Acceleration Kernel Time" */
    NB_PHASE_LF1 = 6,         /* "Leapfrog Part 1" */
    NB_PHASE_LF2 = 7,         /* "Leapfrog Part 2" */
    /* the two parts of NB_PHASE_KEYS_SORT */
    NB_PHASE_KEYS = 8,        /* octant keys of every body ("Build octree to level") */
    NB_PHASE_SORT = 9,        /* radix sort of the keys ("Sort bodies for subtrees") */
    NB_PHASE_COUNT = 10
};
/* GPU time (HIP events) of each phase summed over the last nb_* call;
 * ms must hold NB_PHASE_COUNT doubles. */
int nb_phase_ms(const nb_ctx *ctx, double *ms);

/* ---- octree inspection (parity checks) --------------------------------- */
/* aabb = {min_x, min_y, min_z, max_x, max_y, max_z, edge} (BarnesHutOctree.hpp:22-29). */
int nb_tree_info(nb_ctx *ctx, uint64_t *node_count, uint32_t *internal_count,
                 uint32_t *max_depth, double aabb[7]);
/* sortedBodiesInOrder (BarnesHutOctree.hpp:93-95): body ids in octree in-order. */
int nb_tree_sorted(nb_ctx *ctx, uint32_t *sorted);
/* Canonical form of the octree: depth-first pre-order from the root with
 * children in octant-digit order 0..7, one record per node.  kind: 0 empty
 * leaf, 1 body leaf, 2 internal; body = N unless kind == 1; mass and com_* are
 * the reference's SUM_MASSES and centre-of-mass numerators, count its
 * bodyCountNode.  Returns the record count (0 if cap is too small). */
uint64_t nb_tree_canonical(nb_ctx *ctx, uint64_t cap, uint8_t *depth, uint8_t *kind,
                           uint32_t *body, double *mass, double *com_x, double *com_y,
                           double *com_z, uint32_t *count);
/* Accepted interactions (cells + bodies) per body in the last nb_bh_forces,
 * indexed by body id. */
int nb_bh_interactions(nb_ctx *ctx, uint32_t *count);

/* ---- learned far-field correction ---------------------------------------
 * No counterpart in the reference.  A network of 13 inputs -> 64 -> 64 -> 64
 * (SiLU) -> 3 outputs, with BF16 weights on the matrix cores and FP32
 * accumulation, predicts for an accepted octree cell and a target body the
 * residual of the monopole approximation:
 *   correction = net(f) * G M e^2 / d^4
 * Here r = com - target, d = |r|, M is the cell mass and e the cell edge. The
 * features f are:
 *   f[0..2]  = r / d
 *   f[3]     = e / d
 *   f[4]     = ln(M / M_root) / 16
 *   f[5..10] = Q / (M e^2), Q = sum m (3 s s^T - |s|^2 I) about the cell's
 *              centre of mass, as {xx, yy, zz, xy, xz, yz}
 *   f[11]    = o . r/d
 *   f[12]    = |o - (o . r/d) r/d|, with o = (geometric centre - com) / e.
 * Training targets are (a_exact - a_monopole) * d^4 / (G M e^2), with a_exact
 * summed over the cell's bodies in FP64.
 *
 * Parameters, 9411 floats, row-major with the input index first:
 *   W1[13][64] b1[64] W2[64][64] b2[64] W3[64][64] b3[64] W4[64][3] b4[3]
 */
typedef struct nb_mlp nb_mlp;

typedef struct nb_mlp_train_opts {
    uint32_t epochs;  /* passes over the data set */
    uint32_t batch;   /* samples per Adam step */
    float lr;         /* initial learning rate (cosine decay to lr_final) */
    float lr_final;
    float beta1, beta2, eps;  /* Adam */
    uint32_t augment; /* nonzero: rotate every sample by a fresh random rotation
                         each epoch (direction, quadrupole and target together;
                         the physics is rotation invariant) */
    uint64_t seed;    /* data-order permutations and rotations */
} nb_mlp_train_opts;

void nb_mlp_default_train_opts(nb_mlp_train_opts *o);
/* Weights uniform in +-sqrt(6/fan_in) from a splitmix64 stream, zero biases. */
int nb_mlp_create(nb_mlp **out, uint64_t seed);
int nb_mlp_destroy(nb_mlp *mlp);
const char *nb_mlp_last_error(const nb_mlp *mlp);
int nb_mlp_set_params(nb_mlp *mlp, const float *params);  /* also resets Adam state */
int nb_mlp_get_params(nb_mlp *mlp, float *params);
/* out[count][3] = net(features[count][13]) on the GPU. */
int nb_mlp_forward(nb_mlp *mlp, uint64_t count, const float *features, float *out);
/* Gradient of L = mean over samples and components of (net(f) - t)^2 (the
 * training loss) for all 9411 parameters, computed by the training kernels. */
int nb_mlp_grad(nb_mlp *mlp, uint32_t count, const float *features, const float *targets,
                float *grad, float *loss);
/* One Adam step (beta1 0.9, beta2 0.999, eps 1e-8, step number t >= 1) from a given gradient. */
int nb_mlp_adam_step(nb_mlp *mlp, const float *grad, float lr, uint32_t t);
/* Train on the GPU; epoch_loss (may be NULL) receives the mean batch loss of each epoch. */
int nb_mlp_train(nb_mlp *mlp, uint64_t count, const float *features, const float *targets,
                 const nb_mlp_train_opts *opts, float *epoch_loss);

/* Training pairs from the context's current octree (after nb_bh_build): for
 * each listed target body, a depth-first walk accepts internal cells with
 * e / d < theta (d to the cell's centre of mass) and opens the others.  For
 * up to cap pairs it writes the cell (internal-cell index = its rank among
 * internal cells in depth-first pre-order with children in digit order), the
 * body, the 13 features and the 3 normalised residuals.  *count receives the
 * number of pairs found; only min(*count, cap) are written. */
int nb_mlp_pairs(nb_ctx *ctx, double theta, uint32_t nt, const uint32_t *targets, uint64_t cap,
                 uint32_t *cell, uint32_t *body, float *features, float *residuals,
                 uint64_t *count);
/* Attach a trained network (NULL detaches) to the FP32C Barnes-Hut traversal.
 * Every accepted cell with e / d_group > tau (d_group is the distance to the
 * group's box) then also gets the learned correction for each body.  The
 * network must outlive the attachment. */
int nb_set_correction(nb_ctx *ctx, nb_mlp *mlp, double tau);
/* Network evaluations ((cell, body) pairs) of the last corrected nb_bh_forces
 * (18,432 flop each, forward). */
int nb_correction_pairs(nb_ctx *ctx, uint64_t *pairs);
/* (cell, body) interactions evaluated with the quadrupole term in the last
 * nb_bh_forces (0 unless nb_params.quadrupole); included in the counts of
 * nb_bh_interactions.  No counterpart in the reference. */
int nb_quadrupole_pairs(nb_ctx *ctx, uint64_t *pairs);
/* Quadrupole moments Q = sum m (3 s s^T - |s|^2 I) of every node, about its
 * centre of mass, as {xx, yy, zz, xy, xz, yz}, in the canonical order of
 * nb_tree_canonical (zero for leaves).  q has 6 * cap doubles.  Returns the
 * record count (0 if cap is too small). */
uint64_t nb_tree_quadrupoles(nb_ctx *ctx, uint64_t cap, double *q);

/* ---- ensembles of independent systems -------------------------------------
 * No counterpart in the reference, which runs one system per process.  Many
 * systems of uneven size advance together: the states are concatenated
 * (system s holds bodies [sum of sizes[0..s), + sizes[s]) of every array),
 * each system has its own time step, and one loop iteration of the whole
 * ensemble is one kick-drift, one octree build of every Barnes-Hut system
 * (one workgroup per system), one persistent force launch over a work queue
 * of traversal groups and direct-sum tiles, and one kick -- the reference's
 * loop body (NaiveAlgorithm.cpp:131-259, BarnesHutAlgorithm.cpp:149-277)
 * for every system at once.  Forces are FP32 pair arithmetic with FP64
 * accumulation (NB_PREC_FP32C); octree, integrator and sums are FP64.
 * Systems with fewer than bh_threshold bodies use the direct sum. */
typedef struct nb_ensemble nb_ensemble;

typedef struct nb_ensemble_tuning {
    uint32_t bh_threshold; /* systems with n >= bh_threshold use Barnes-Hut */
    uint32_t waves;        /* persistent wavefronts of the force launch */
    uint32_t group_max;    /* bodies per Barnes-Hut traversal group, 1..64 (default 64) */
    uint32_t pack_max;     /* sibling groups are packed inside cells of at most this many
                              bodies, >= group_max (default 256) */
    uint32_t direct_tile;  /* targets per direct-sum work item: 64 or 128 (default 64) */
} nb_ensemble_tuning;

/* p->precision must be NB_PREC_FP32C; G, eps2, theta and storage_size_param
 * apply to every system.  At most 65535 systems, 2^32 bodies in all. */
int nb_ensemble_create(nb_ensemble **out, uint32_t nsys, const uint32_t *sizes,
                       const nb_params *p);
int nb_ensemble_destroy(nb_ensemble *e);
const char *nb_ensemble_last_error(const nb_ensemble *e);
uint64_t nb_ensemble_body_count(const nb_ensemble *e);
/* Concatenated host arrays; dt has one entry per system. */
int nb_ensemble_set_state(nb_ensemble *e, const double *x, const double *y, const double *z,
                          const double *vx, const double *vy, const double *vz, const double *m,
                          const double *dt);
int nb_ensemble_get_state(nb_ensemble *e, double *x, double *y, double *z, double *vx, double *vy,
                          double *vz);
int nb_ensemble_get_accel(nb_ensemble *e, double *ax, double *ay, double *az);
/* Forces at the current positions (the t = 0 evaluation). */
int nb_ensemble_forces(nb_ensemble *e);
/* `count` loop iterations of every system, one host synchronisation. */
int nb_ensemble_steps(nb_ensemble *e, uint32_t count);
/* GPU time per phase of the last call (NB_PHASE_OCTREE, _ACCEL, _LF1, _LF2). */
int nb_ensemble_phase_ms(const nb_ensemble *e, double *ms);
/* Per body, last force evaluation: accepted cells and bodies (Barnes-Hut
 * systems) or source bodies (direct-sum systems). */
int nb_ensemble_interactions(nb_ensemble *e, uint32_t *count);
int nb_ensemble_set_tuning(nb_ensemble *e, const nb_ensemble_tuning *t);
int nb_ensemble_get_tuning(const nb_ensemble *e, nb_ensemble_tuning *t);
/* flags[s] = 1 if system s uses Barnes-Hut under the current tuning. */
int nb_ensemble_uses_bh(const nb_ensemble *e, uint8_t *flags);
/* Octree of Barnes-Hut system s from the last force evaluation: node count,
 * and the canonical dump in the format of nb_tree_canonical (body indices
 * local to the system).  0 if s is not a Barnes-Hut system or cap is small. */
uint64_t nb_ensemble_node_count(nb_ensemble *e, uint32_t s);
uint64_t nb_ensemble_tree_canonical(nb_ensemble *e, uint32_t s, uint64_t cap, uint8_t *depth,
                                    uint8_t *kind, uint32_t *body, double *mass, double *com_x,
                                    double *com_y, double *com_z, uint32_t *count);
/* Autotuner: times `steps` loop iterations from the current state for each
 * candidate, one parameter at a time from the current best: the
 * direct-sum/Barnes-Hut threshold, the traversal group size, the packing
 * limit, the direct-sum tile and the persistent grid size; keeps the fastest
 * and restores the state.  *best and *best_step_ms (either may be NULL)
 * receive the choice and its time. */
int nb_ensemble_autotune(nb_ensemble *e, uint32_t steps, nb_ensemble_tuning *best,
                         double *best_step_ms);

/* ---- synthetic inputs (seeded; same bits on every machine) -------------- */
/* One star of 1 Msun at rest at the origin plus n-1 bodies on Kepler orbits:
 * a ~ U[2.1, 3.3] AU, e ~ U[0, 0.3], i ~ U[0, 20 deg], angles ~ U[0, 2pi),
 * m = 10^U[-20,-12] Msun; Kepler -> Cartesian as in the reference's dataset
 * converter (dataset_converter/src/conversion.cpp:13-89).  Units kg, AU, AU/d. */
int nb_gen_star_disk(uint32_t n, uint64_t seed, double *x, double *y, double *z, double *vx,
                     double *vy, double *vz, double *m);
/* Plummer sphere (Aarseth, Henon & Wielen 1974), scale radius 1, total mass 1
 * in units where G = 1, equal masses, radius truncated at 100, centre of mass
 * at rest at the origin.  Masses are written as (1/n)/G_ref kg so that the
 * reference's fixed G reproduces G = 1 in these units. */
int nb_gen_plummer(uint32_t n, uint64_t seed, double *x, double *y, double *z, double *vx,
                   double *vy, double *vz, double *m);

#ifdef __cplusplus
}
#endif

#endif

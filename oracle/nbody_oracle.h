/*
 * nbody_oracle.h -- CPU restatement of TimThuering/N-Body-Simulation's hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker: tests/, the
 * __graft_entry__ smoke() check and bench.py's cpu_baseline leg are the only
 * code allowed to load it.  The product (libnbody_hip.so and the host
 * application) never links or calls it.
 *
 * Every function restates one piece of the reference and cites the file:line
 * it follows (paths relative to the reference repository root).  All
 * arithmetic is FP64 and the library is built with -ffp-contract=off so that
 * each `a*b+c` in the reference is a multiply followed by an add, never a
 * fused multiply-add.  sycl::rsqrt(x) is restated as 1.0/std::sqrt(x) (two
 * correctly rounded IEEE operations); the SYCL implementations leave its
 * rounding unspecified.
 *
 * Error codes are the same numbers as the product's NB_E_* codes.
 */
#ifndef NBODY_ORACLE_H
#define NBODY_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
    ORC_OK = 0,
    ORC_E_INVALID_ARG = 1,
    ORC_E_NODE_OVERFLOW = 3,   /* octree storage (storageSizeParameter) exhausted */
    ORC_E_STACK_OVERFLOW = 4,  /* traversal stack (stackSize) exhausted */
    ORC_E_ZERO_MASS = 6,       /* a non-empty cell has zero mass: the reference COM loop never ends */
    ORC_E_AABB = 8             /* cubify found no longest edge (NaN input) */
};

/* G in AU^3 kg^-1 d^-2 exactly as nBodyAlgorithm.hpp:55-61 evaluates it. */
double orc_G_ref(void);
/* epsilon^2 = pow(10,-22), Configuration.cpp:6. */
double orc_eps2_ref(void);

/* Number of OpenMP threads the oracle uses (0 = leave the runtime default). */
void orc_set_threads(int threads);
int orc_get_threads(void);

/* ---- naive direct sum (NaiveAlgorithm.cpp:262-482) ------------------------ */
/* a_i = G * sum_{j=0}^{N-1} m_j * (r_ij * rsqrt((|r_ij|^2+eps2)^3)), j in index
 * order, self term exactly 0 (opt_0 :384-411 == opt_2 :299-352). */
void orc_naive_accel(uint32_t n, const double *x, const double *y, const double *z,
                     const double *m, double G, double eps2,
                     double *ax, double *ay, double *az);
/* Same sum, only for the listed targets (out arrays have nt entries).  Used as
 * the FP64 direct-sum error reference on sampled targets. */
void orc_naive_accel_targets(uint32_t n, const double *x, const double *y, const double *z,
                             const double *m, double G, double eps2,
                             uint32_t nt, const uint32_t *targets,
                             double *ax, double *ay, double *az);

/* ---- leapfrog (NaiveAlgorithm.cpp:140-164 and :206-222) ----------------- */
/* v_half = v + a*(dt/2); x = x + v_half*dt */
void orc_kick_drift(uint32_t n, double dt,
                    const double *ax, const double *ay, const double *az,
                    const double *vx, const double *vy, const double *vz,
                    double *vhx, double *vhy, double *vhz,
                    double *x, double *y, double *z);
/* v = v_half + a*(dt/2) */
void orc_kick(uint32_t n, double dt,
              const double *ax, const double *ay, const double *az,
              const double *vhx, const double *vhy, const double *vhz,
              double *vx, double *vy, double *vz);

/* ---- energy (nBodyAlgorithm.cpp:11-86) ----------------------------------- */
/* out = {kinetic, potential, total, virial 2K/|U|}; no softening. */
void orc_energy(uint32_t n, const double *x, const double *y, const double *z,
                const double *vx, const double *vy, const double *vz, const double *m,
                double G, double out[4]);

/* ---- AABB (BarnesHutOctree.cpp:45-191) ---------------------------------- */
/* out = {min_x, min_y, min_z, max_x, max_y, max_z, edge} after cubify. */
int orc_aabb(uint32_t n, int work_items, const double *x, const double *y, const double *z,
             double out[7]);

/* ---- octree (BarnesHutOctree + ParallelOctreeTopDown{Subtrees,Synchronized}) */
typedef struct orc_tree orc_tree;

/* storage = storageSizeParameter (nodes per array; reference default 16*N). */
orc_tree *orc_tree_create(uint32_t n, uint32_t storage);
void orc_tree_destroy(orc_tree *t);

enum { ORC_BUILDER_SUBTREES = 0, ORC_BUILDER_SYNCHRONIZED = 1 };

/* Build the octree for positions/masses: AABB -> insertion -> COM -> sortBodies.
 * builder: ORC_BUILDER_SUBTREES (default build, Subtrees.cpp:15-93) or
 * ORC_BUILDER_SYNCHRONIZED (Synchronized.cpp:10-388).
 * phase_ms (may be NULL) receives 8 wall times in ms:
 *   [0] AABB creation, [1] build octree to level (or whole build),
 *   [2] prepare subtrees, [3] sort bodies for subtrees, [4] build subtrees,
 *   [5] compute center of mass, [6] sort bodies, [7] total. */
int orc_tree_build(orc_tree *t, int builder, const double *x, const double *y, const double *z,
                   const double *m, int max_build_level, int aabb_work_items,
                   int top_work_items, int subtree_work_items, int sort_bodies,
                   double *phase_ms);

uint32_t orc_tree_node_count(const orc_tree *t);
void orc_tree_get_aabb(const orc_tree *t, double out[7]);

/* Copy out raw reference-layout arrays (any pointer may be NULL).
 * octants has 8*storage entries, octants[d*storage + node] = child with digit d.
 * Node arrays have storage entries; sorted has n entries. */
void orc_tree_export(const orc_tree *t, uint32_t *octants, uint32_t *body_of_node,
                     int32_t *node_is_leaf, double *sum_masses,
                     double *com_x, double *com_y, double *com_z,
                     uint32_t *body_count, double *edge,
                     double *min_x, double *min_y, double *min_z, uint32_t *sorted);

/* Canonical form of the tree: depth-first pre-order from the root, children in
 * digit order 0..7, one record per node (internal, body leaf and empty leaf).
 * kind: 0 empty leaf, 1 body leaf, 2 internal.  body = N for non-body nodes.
 * mass / com_* are the reference's SUM_MASSES and centre-of-mass numerators.
 * Returns the number of records (== node count), or 0 if cap is too small. */
uint64_t orc_tree_canonical(const orc_tree *t, uint64_t cap, uint8_t *depth, uint8_t *kind,
                            uint32_t *body, double *mass, double *com_x, double *com_y,
                            double *com_z, uint32_t *count);

/* ---- Barnes-Hut traversal (BarnesHutAlgorithm.cpp:280-401) -------------- */
/* Forces for all bodies from a built tree; per-body stack of stack_size
 * entries (BarnesHutAlgorithm.cpp:8-15, :340-342).  Returns
 * ORC_E_STACK_OVERFLOW if a body needs a deeper stack.
 * interactions (may be NULL) receives the accepted-node count per body. */
int orc_bh_accel(const orc_tree *t, const double *x, const double *y, const double *z,
                 double theta, double G, double eps2, uint32_t stack_size, int sort_bodies,
                 double *ax, double *ay, double *az, uint32_t *interactions);

/* stackSize rule of Configuration.cpp:24-33 */
uint32_t orc_stack_size(uint32_t n, int stack_size_param);

/* ---- whole simulation loop (NaiveAlgorithm.cpp:15-260, BarnesHutAlgorithm.cpp:18-278) */
/* algorithm: 0 naive, 1 Barnes-Hut (Subtrees builder).  x..vz are updated in
 * place; ax..az receive the last accelerations.  Visualisation is not
 * modelled (the reference's visualised steps only copy data out).
 * loop_ms (may be NULL) = wall time of the while-loop (t=0 force excluded).
 * Returns an ORC_* code; *steps_done receives the number of loop iterations. */
int orc_simulate(int algorithm, uint32_t n, double *x, double *y, double *z,
                 double *vx, double *vy, double *vz, const double *m,
                 double dt, double t_end, double theta, int storage_size_param,
                 int stack_size_param, int max_build_level,
                 double *ax, double *ay, double *az, int32_t *steps_done, double *loop_ms,
                 double *phase_ms_sum);

#ifdef __cplusplus
}
#endif

#endif

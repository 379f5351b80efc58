// nbody_oracle.cpp -- CPU restatement of the reference's n-body hot path.
//
// TEST INFRASTRUCTURE ONLY (the checker).  Loaded by tests/, by the smoke()
// check of __graft_entry__.py and by bench.py's cpu_baseline leg; the product
// never links or calls it.
//
// The restatement keeps the reference's arithmetic order and its work
// decomposition.  The reference runs on the CPU through a SYCL OpenMP backend
// that maps one work-group to one OpenMP task and runs the work-items of a
// group one after another; this file does the same:
//   * kernels launched over range(N) / nd_range(N, wg) run as an OpenMP loop
//     over work-groups of 64 bodies;
//   * kernels launched as ONE work-group (top-of-octree build,
//     Subtrees.cpp:199-200; centre of mass, BarnesHutOctree.cpp:424; the
//     synchronized build, Synchronized.cpp:129-131) run on one core, with the
//     work-items in id order, so all locks are uncontended;
//   * buildSubtrees runs its work-groups (one per subtree) in parallel.
//
// Built with -ffp-contract=off (see oracle/Makefile).
#include "nbody_oracle.h"

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <cstring>
#include <limits>
#include <vector>

#include <omp.h>

namespace {

int g_threads = 0;

inline int nthreads() { return g_threads > 0 ? g_threads : omp_get_max_threads(); }

// sycl::rsqrt restated as two correctly rounded IEEE operations.
inline double rsqrt_ref(double v) { return 1.0 / std::sqrt(v); }

// SYCL common min/max: min(x, y) = y < x ? y : x; max(x, y) = x < y ? y : x.
inline double sycl_min(double a, double b) { return b < a ? b : a; }
inline double sycl_max(double a, double b) { return a < b ? b : a; }

using clk = std::chrono::steady_clock;
inline double ms_between(clk::time_point a, clk::time_point b) {
    return std::chrono::duration<double, std::milli>(b - a).count();
}

// Octant digit of a point inside a cell (ParallelOctreeTopDownSubtrees.cpp:348-358,
// :400-406; BarnesHutOctree.cpp:586-591): 4*[y > mid_y] + 2*[x > mid_x] + [z < mid_z]
// with mid = cellMin + edge/2 evaluated in FP64.
inline uint32_t octant_digit(double px, double py, double pz, double mnx, double mny,
                             double mnz, double e) {
    bool upperPart = py > mny + (e / 2.0);
    bool rightPart = px > mnx + (e / 2.0);
    bool backPart = pz < mnz + (e / 2.0);
    return ((uint32_t)upperPart) * 4u + ((uint32_t)rightPart) * 2u + ((uint32_t)backPart);
}

}  // namespace

struct orc_tree {
    uint32_t N = 0;
    uint32_t S = 0;  // storageSizeParameter
    std::vector<uint32_t> octants;  // 8*S, octants[d*S + node]
    std::vector<double> edge, minx, miny, minz;
    std::vector<uint32_t> bodyOfNode;  // N == "no body"
    std::vector<int32_t> isLeaf;
    std::vector<double> mass, cx, cy, cz;  // SUM_MASSES and centre-of-mass numerators
    std::vector<uint32_t> count;           // bodyCountNode
    std::vector<uint32_t> sorted;          // sortedBodiesInOrder
    std::vector<uint32_t> subtreeOfBody;   // Subtrees builder
    std::vector<uint32_t> sortedBodies;    // Subtrees builder
    std::atomic<uint32_t> nextFree{0};
    std::atomic<int> overflow{0};
    double aabb[7] = {0, 0, 0, 0, 0, 0, 0};
};

extern "C" {

double orc_G_ref(void) {
    // nBodyAlgorithm.hpp:55-61
    double G = 6.67428 * std::pow(10, -11);
    double meter_AU = 1.0 / (1.49597870691 * std::pow(10, 11));
    double second_Days = 1.0 / 86400;
    G = G * (std::pow(meter_AU, 3) / std::pow(second_Days, 2));
    return G;
}

double orc_eps2_ref(void) { return std::pow(10, -22); }  // Configuration.cpp:6

void orc_set_threads(int threads) { g_threads = threads; }
int orc_get_threads(void) { return nthreads(); }

// ---------------------------------------------------------------------------
// naive direct sum: NaiveAlgorithm.cpp:384-411 (opt_0).  opt_2 (:299-352)
// performs the same operations in the same j order (LDS tiles only change
// where m_j, x_j come from), so one restatement covers all stages.
static inline void naive_one(uint32_t n, const double *x, const double *y, const double *z,
                             const double *m, double G, double eps2, uint32_t i,
                             double *oax, double *oay, double *oaz) {
    double acc_x = 0, acc_y = 0, acc_z = 0;
    const double pos_x = x[i], pos_y = y[i], pos_z = z[i];
    for (uint32_t j = 0; j < n; ++j) {
        double r_x = x[j] - pos_x;
        double r_y = y[j] - pos_y;
        double r_z = z[j] - pos_z;
        double denominator = (r_x * r_x) + (r_y * r_y) + (r_z * r_z) + eps2;
        denominator = denominator * denominator * denominator;
        denominator = rsqrt_ref(denominator);
        acc_x += m[j] * (r_x * denominator);
        acc_y += m[j] * (r_y * denominator);
        acc_z += m[j] * (r_z * denominator);
    }
    *oax = acc_x * G;
    *oay = acc_y * G;
    *oaz = acc_z * G;
}

void orc_naive_accel(uint32_t n, const double *x, const double *y, const double *z,
                     const double *m, double G, double eps2, double *ax, double *ay,
                     double *az) {
#pragma omp parallel for schedule(static, 64) num_threads(nthreads())
    for (int64_t i = 0; i < (int64_t)n; ++i)
        naive_one(n, x, y, z, m, G, eps2, (uint32_t)i, &ax[i], &ay[i], &az[i]);
}

void orc_naive_accel_targets(uint32_t n, const double *x, const double *y, const double *z,
                             const double *m, double G, double eps2, uint32_t nt,
                             const uint32_t *targets, double *ax, double *ay, double *az) {
#pragma omp parallel for schedule(dynamic, 16) num_threads(nthreads())
    for (int64_t t = 0; t < (int64_t)nt; ++t)
        naive_one(n, x, y, z, m, G, eps2, targets[t], &ax[t], &ay[t], &az[t]);
}

// ---------------------------------------------------------------------------
// leapfrog, NaiveAlgorithm.cpp:154-163 (== BarnesHutAlgorithm.cpp:171-181)
void orc_kick_drift(uint32_t n, double dt, const double *ax, const double *ay,
                    const double *az, const double *vx, const double *vy, const double *vz,
                    double *vhx, double *vhy, double *vhz, double *x, double *y, double *z) {
    const double delta_t = dt;
#pragma omp parallel for schedule(static) num_threads(nthreads())
    for (int64_t i = 0; i < (int64_t)n; ++i) {
        vhx[i] = vx[i] + ax[i] * (delta_t / 2.0);
        vhy[i] = vy[i] + ay[i] * (delta_t / 2.0);
        vhz[i] = vz[i] + az[i] * (delta_t / 2.0);
        x[i] = x[i] + vhx[i] * delta_t;
        y[i] = y[i] + vhy[i] * delta_t;
        z[i] = z[i] + vhz[i] * delta_t;
    }
}

// NaiveAlgorithm.cpp:217-221 (== BarnesHutAlgorithm.cpp:234-238)
void orc_kick(uint32_t n, double dt, const double *ax, const double *ay, const double *az,
              const double *vhx, const double *vhy, const double *vhz, double *vx, double *vy,
              double *vz) {
    const double delta_t = dt;
#pragma omp parallel for schedule(static) num_threads(nthreads())
    for (int64_t i = 0; i < (int64_t)n; ++i) {
        vx[i] = vhx[i] + ax[i] * (delta_t / 2.0);
        vy[i] = vhy[i] + ay[i] * (delta_t / 2.0);
        vz[i] = vhz[i] + az[i] * (delta_t / 2.0);
    }
}

// ---------------------------------------------------------------------------
// energy: nBodyAlgorithm.cpp:48-85 (per-body terms in parallel, host sums in
// body order)
void orc_energy(uint32_t n, const double *x, const double *y, const double *z,
                const double *vx, const double *vy, const double *vz, const double *m,
                double G, double out[4]) {
    std::vector<double> E_KIN(n), E_POT(n);
#pragma omp parallel for schedule(dynamic, 64) num_threads(nthreads())
    for (int64_t j = 0; j < (int64_t)n; ++j) {
        double v = vx[j] * vx[j] + vy[j] * vy[j] + vz[j] * vz[j];
        E_KIN[j] = 0.5 * m[j] * v;
        E_POT[j] = 0;
        for (int64_t i = 0; i < j; ++i) {
            double r_x = x[j] - x[i];
            double r_y = y[j] - y[i];
            double r_z = z[j] - z[i];
            double r = std::sqrt(r_x * r_x + r_y * r_y + r_z * r_z);
            E_POT[j] += G * m[i] * m[j] / r;
        }
    }
    double E_kin_result = 0, E_pot_result = 0;
    for (uint32_t i = 0; i < n; ++i) {
        E_kin_result += E_KIN[i];
        E_pot_result += E_POT[i];
    }
    E_pot_result *= -1;
    out[0] = E_kin_result;
    out[1] = E_pot_result;
    out[2] = E_kin_result + E_pot_result;
    out[3] = (2.0 * E_kin_result) / std::abs(E_pot_result);
}

// ---------------------------------------------------------------------------
// AABB: BarnesHutOctree.cpp:45-191.  Partial extrema start at 0.0 (the
// vectors at :58-72 are value-initialised), so the box always contains the
// origin; then the box is made a cube by growing the shorter edges
// symmetrically, testing x, then y, then z (:161-188).
int orc_aabb(uint32_t n, int work_items, const double *x, const double *y, const double *z,
             double out[7]) {
    if (work_items <= 0) return ORC_E_INVALID_ARG;
    double min_x = std::numeric_limits<double>::infinity();
    double min_y = std::numeric_limits<double>::infinity();
    double min_z = std::numeric_limits<double>::infinity();
    double max_x = std::numeric_limits<double>::lowest();
    double max_y = std::numeric_limits<double>::lowest();
    double max_z = std::numeric_limits<double>::lowest();

    const int threadCount = work_items;
    const int bodiesPerThread = (int)std::ceil((double)n / (double)threadCount);
    std::vector<double> lmax_x(threadCount), lmax_y(threadCount), lmax_z(threadCount);
    std::vector<double> lmin_x(threadCount), lmin_y(threadCount), lmin_z(threadCount);

#pragma omp parallel for schedule(static) num_threads(nthreads())
    for (int idx = 0; idx < threadCount; ++idx) {
        for (int64_t i = (int64_t)bodiesPerThread * idx;
             i < (int64_t)bodiesPerThread * idx + bodiesPerThread; ++i) {
            if (i < (int64_t)n) {
                lmin_x[idx] = sycl_min(lmin_x[idx], x[i]);
                lmin_y[idx] = sycl_min(lmin_y[idx], y[i]);
                lmin_z[idx] = sycl_min(lmin_z[idx], z[i]);
                lmax_x[idx] = sycl_max(lmax_x[idx], x[i]);
                lmax_y[idx] = sycl_max(lmax_y[idx], y[i]);
                lmax_z[idx] = sycl_max(lmax_z[idx], z[i]);
            }
        }
    }
    for (int i = 0; i < threadCount; ++i) {
        if ((uint32_t)i < n) {
            if (lmin_x[i] < min_x) min_x = lmin_x[i];
            if (lmin_y[i] < min_y) min_y = lmin_y[i];
            if (lmin_z[i] < min_z) min_z = lmin_z[i];
            if (lmax_x[i] > max_x) max_x = lmax_x[i];
            if (lmax_y[i] > max_y) max_y = lmax_y[i];
            if (lmax_z[i] > max_z) max_z = lmax_z[i];
        }
    }
    double x_length = std::abs(max_x - min_x);
    double y_length = std::abs(max_y - min_y);
    double z_length = std::abs(max_z - min_z);
    double maxEdgeLength = std::max(x_length, std::max(y_length, z_length));
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
        return ORC_E_AABB;  // reference throws std::invalid_argument (:187)
    }
    out[0] = min_x; out[1] = min_y; out[2] = min_z;
    out[3] = max_x; out[4] = max_y; out[5] = max_z;
    out[6] = maxEdgeLength;
    return ORC_OK;
}

// ---------------------------------------------------------------------------
// octree

orc_tree *orc_tree_create(uint32_t n, uint32_t storage) {
    if (n == 0 || storage == 0) return nullptr;
    orc_tree *t = new orc_tree();
    t->N = n;
    t->S = storage;
    // BarnesHutOctree.cpp:5-43 and Subtrees.cpp:4-13
    t->octants.assign((size_t)8 * storage, 0);
    t->edge.assign(storage, 0.0);
    t->minx.assign(storage, 0.0);
    t->miny.assign(storage, 0.0);
    t->minz.assign(storage, 0.0);
    t->bodyOfNode.assign(storage, n);
    t->isLeaf.assign(storage, 1);
    t->mass.assign(storage, 0.0);
    t->cx.assign(storage, 0.0);
    t->cy.assign(storage, 0.0);
    t->cz.assign(storage, 0.0);
    t->count.assign(storage, 0);
    t->sorted.assign(n, n);
    t->subtreeOfBody.assign(n, 0);
    t->sortedBodies.assign(n, 0);
    return t;
}

void orc_tree_destroy(orc_tree *t) { delete t; }

namespace {

// Reserve 8 children for `cur` and initialise them exactly as
// Subtrees.cpp:252-339 does.  Children are stored at slots first..first+7 in
// the slot order of the reference: digit {5,7,4,6,1,3,0,2}.  Child minimum:
// x += e_c if the x-bit is set, y += e_c if the y-bit is set, z += e_c if the
// z-bit is NOT set (the z-bit means "z below the midpoint").
bool split_node(orc_tree *t, uint32_t cur) {
    const uint32_t S = t->S;
    uint32_t first = t->nextFree.fetch_add(8);
    if ((uint64_t)first + 8 > S) {
        t->overflow.store(1);
        return false;
    }
    static const uint32_t slot_digit[8] = {5, 7, 4, 6, 1, 3, 0, 2};
    const double childEdgeLength = t->edge[cur] / 2;
    const double pmx = t->minx[cur], pmy = t->miny[cur], pmz = t->minz[cur];
    for (uint32_t s = 0; s < 8; ++s) {
        const uint32_t idx = first + s;
        const uint32_t d = slot_digit[s];
        t->edge[idx] = childEdgeLength;
        t->octants[(size_t)d * S + cur] = idx;
        t->minx[idx] = (d & 2u) ? pmx + childEdgeLength : pmx;
        t->miny[idx] = (d & 4u) ? pmy + childEdgeLength : pmy;
        t->minz[idx] = (d & 1u) ? pmz : pmz + childEdgeLength;
        for (uint32_t k = 0; k < 8; ++k) t->octants[(size_t)k * S + idx] = 0;
        t->bodyOfNode[idx] = t->N;
        t->count[idx] = 0;
        t->isLeaf[idx] = 1;
        t->cx[idx] = 0;
        t->cy[idx] = 0;
        t->cz[idx] = 0;
        t->mass[idx] = 0;
    }
    return true;
}

inline uint32_t child_for(const orc_tree *t, uint32_t node, double px, double py, double pz) {
    uint32_t d = octant_digit(px, py, pz, t->minx[node], t->miny[node], t->minz[node],
                              t->edge[node]);
    return t->octants[(size_t)d * t->S + node];
}

void init_root(orc_tree *t) {
    // Subtrees.cpp:135-163
    const uint32_t S = t->S;
    t->edge[0] = t->aabb[6];
    t->minx[0] = t->aabb[0];
    t->miny[0] = t->aabb[1];
    t->minz[0] = t->aabb[2];
    for (uint32_t k = 0; k < 8; ++k) t->octants[(size_t)k * S] = 0;
    t->nextFree.store(1);
    t->isLeaf[0] = 1;
    t->count[0] = 0;
    t->mass[0] = 0;
    t->cx[0] = t->cy[0] = t->cz[0] = 0;
    t->bodyOfNode[0] = t->N;
    t->overflow.store(0);
}

// Unbounded top-down insertion of body i starting at `node`
// (Subtrees.cpp:625-808; Synchronized.cpp insertion loop).
bool insert_from(orc_tree *t, uint32_t i, uint32_t node, const double *x, const double *y,
                 const double *z) {
    const uint32_t N = t->N;
    while (true) {
        if (t->isLeaf[node] == 1) {
            if (t->bodyOfNode[node] == N) {
                t->bodyOfNode[node] = i;
                return true;
            }
            const uint32_t old = t->bodyOfNode[node];
            if (!split_node(t, node)) return false;
            const uint32_t c = child_for(t, node, x[old], y[old], z[old]);
            t->bodyOfNode[c] = old;
            t->bodyOfNode[node] = N;
            t->isLeaf[node] = 0;
        } else {
            node = child_for(t, node, x[i], y[i], z[i]);
        }
    }
}

// Top of the tree to maxLevel (Subtrees.cpp:207-430), executed as the single
// work-group it is launched as: work-items in id order, each inserting its
// contiguous chunk of bodies, i.e. bodies in index order.
bool build_top(orc_tree *t, int maxLevel, const double *x, const double *y, const double *z) {
    const uint32_t N = t->N;
    for (uint32_t i = 0; i < N; ++i) {
        t->subtreeOfBody[i] = 0;
        int currentDepth = 0;
        uint32_t node = 0;
        while (true) {
            if (currentDepth < maxLevel) {
                if (t->isLeaf[node] == 1) {
                    if (t->bodyOfNode[node] == N) {
                        t->bodyOfNode[node] = i;
                        break;
                    }
                    const uint32_t old = t->bodyOfNode[node];
                    if (!split_node(t, node)) return false;
                    const uint32_t c = child_for(t, node, x[old], y[old], z[old]);
                    t->bodyOfNode[c] = old;
                    t->bodyOfNode[node] = N;
                    if (currentDepth == maxLevel - 1) {
                        // the old body's children sit on maxLevel: it becomes a
                        // pending member of that subtree (:371-376)
                        t->bodyOfNode[c] = N;
                        t->subtreeOfBody[old] = c;
                    }
                    t->isLeaf[node] = 0;
                } else {
                    node = child_for(t, node, x[i], y[i], z[i]);
                    currentDepth += 1;
                }
            } else {
                t->subtreeOfBody[i] = node;  // :420-427
                break;
            }
        }
    }
    return true;
}

// prepareCenterOfMass + computeCenterOfMass_CPU (BarnesHutOctree.cpp:193-227,
// :391-548).  Leaves with a body get (x*m, y*m, z*m, m, 1); an internal node
// gets the sum over its children in digit order 0..7 starting from 0.0.  The
// reference launches this as one work-group that waits until all children are
// "ready" (SUM_MASSES != 0); children always have larger ids than their parent,
// so visiting ids from the highest down is an order in which every wait is
// already satisfied.  A zero-mass cell never becomes ready in the reference
// (the loop at :487-544 would not end): reported as ORC_E_ZERO_MASS.
int compute_com(orc_tree *t, const double *x, const double *y, const double *z,
                const double *m) {
    const uint32_t S = t->S, N = t->N;
    const uint32_t nodes = t->nextFree.load();
#pragma omp parallel for schedule(static) num_threads(nthreads())
    for (int64_t i = 0; i < (int64_t)nodes; ++i) {
        const uint32_t b = t->bodyOfNode[i];
        if (t->isLeaf[i] == 1 && b != N) {
            t->cx[i] = x[b] * m[b];
            t->cy[i] = y[b] * m[b];
            t->cz[i] = z[b] * m[b];
            t->mass[i] = m[b];
            t->count[i] = 1;
        }
    }
    int zero = 0;
    for (int64_t idx = (int64_t)nodes - 1; idx >= 0; --idx) {
        if (t->isLeaf[idx] == 1) {
            if (t->bodyOfNode[idx] != N && t->mass[idx] == 0) zero = 1;
            continue;
        }
        double sumMasses = 0, centerMassX = 0, centerMassY = 0, centerMassZ = 0;
        uint32_t bodyCount = 0;
        for (uint32_t octant = 0; octant < 8; ++octant) {
            const uint32_t c = t->octants[idx + (size_t)octant * S];
            centerMassX += t->cx[c];
            centerMassY += t->cy[c];
            centerMassZ += t->cz[c];
            sumMasses += t->mass[c];
            bodyCount += t->count[c];
        }
        t->count[idx] = bodyCount;
        t->cx[idx] = centerMassX;
        t->cy[idx] = centerMassY;
        t->cz[idx] = centerMassZ;
        t->mass[idx] = sumMasses;
        if (sumMasses == 0) zero = 1;
    }
    return zero ? ORC_E_ZERO_MASS : ORC_OK;
}

// sortBodies: BarnesHutOctree.cpp:550-613
void sort_bodies(orc_tree *t, const double *x, const double *y, const double *z) {
    const uint32_t S = t->S, N = t->N;
#pragma omp parallel for schedule(static, 64) num_threads(nthreads())
    for (int64_t i = 0; i < (int64_t)N; ++i) {
        uint32_t insertionIndex = 0;
        uint32_t currentNode = 0;
        while (!t->isLeaf[currentNode]) {
            const uint32_t d = octant_digit(x[i], y[i], z[i], t->minx[currentNode],
                                            t->miny[currentNode], t->minz[currentNode],
                                            t->edge[currentNode]);
            const uint32_t octantID = t->octants[(size_t)d * S + currentNode];
            for (uint32_t j = 0; j < d; ++j)
                insertionIndex += t->count[t->octants[(size_t)j * S + currentNode]];
            currentNode = octantID;
        }
        t->sorted[insertionIndex] = (uint32_t)i;
    }
}

}  // namespace

int orc_tree_build(orc_tree *t, int builder, const double *x, const double *y, const double *z,
                   const double *m, int max_build_level, int aabb_work_items,
                   int top_work_items, int subtree_work_items, int sort_bodies_flag,
                   double *phase_ms) {
    if (!t || max_build_level < 1 || top_work_items <= 0 || subtree_work_items <= 0)
        return ORC_E_INVALID_ARG;
    (void)top_work_items;  // serial within the single work-group: order == index order
    double ph[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    const auto t0 = clk::now();
    int rc = orc_aabb(t->N, aabb_work_items, x, y, z, t->aabb);
    if (rc != ORC_OK) return rc;
    const auto t1 = clk::now();
    ph[0] = ms_between(t0, t1);
    init_root(t);
    const uint32_t N = t->N;

    if (builder == ORC_BUILDER_SYNCHRONIZED) {
        // Synchronized.cpp:100-356: one work-group inserts every body from the root.
        for (uint32_t i = 0; i < N; ++i)
            if (!insert_from(t, i, 0, x, y, z)) return ORC_E_NODE_OVERFLOW;
        const auto t2 = clk::now();
        ph[1] = ms_between(t1, t2);
    } else {
        if (!build_top(t, max_build_level, x, y, z)) return ORC_E_NODE_OVERFLOW;
        const auto t2 = clk::now();
        ph[1] = ms_between(t1, t2);

        // prepareSubtrees (Subtrees.cpp:436-476)
        const uint32_t nodeCount = t->nextFree.load();
        std::vector<std::atomic<uint32_t>> cnt(nodeCount);
        for (auto &c : cnt) c.store(0, std::memory_order_relaxed);
#pragma omp parallel for schedule(static, 64) num_threads(nthreads())
        for (int64_t i = 0; i < (int64_t)N; ++i)
            cnt[t->subtreeOfBody[i]].fetch_add(1, std::memory_order_relaxed);
        std::vector<uint32_t> bodyCountSubtree(nodeCount);
        for (uint32_t i = 0; i < nodeCount; ++i) bodyCountSubtree[i] = cnt[i].load();
        std::vector<uint32_t> subtrees;
        subtrees.reserve(nodeCount);
        for (uint32_t i = 1; i < nodeCount; ++i)  // single_task compaction, root skipped
            if (bodyCountSubtree[i] > 0) subtrees.push_back(i);
        const int64_t numberOfSubtrees = (int64_t)subtrees.size();
        const auto t3 = clk::now();
        ph[2] = ms_between(t2, t3);

        // sortBodiesForSubtrees (Subtrees.cpp:478-534): O(S^2) start indices and
        // an O(N*S) linear search per body, as written.
        std::vector<uint32_t> startIndex(numberOfSubtrees);
        std::vector<std::atomic<uint32_t>> nextPos(numberOfSubtrees);
#pragma omp parallel for schedule(dynamic, 64) num_threads(nthreads())
        for (int64_t i = 0; i < numberOfSubtrees; ++i) {
            uint32_t firstIndex = 0;
            for (int64_t j = i - 1; j >= 0; j--) firstIndex += bodyCountSubtree[subtrees[j]];
            startIndex[i] = firstIndex;
            nextPos[i].store(firstIndex, std::memory_order_relaxed);
        }
#pragma omp parallel for schedule(dynamic, 256) num_threads(nthreads())
        for (int64_t i = 0; i < (int64_t)N; ++i) {
            const uint32_t sub = t->subtreeOfBody[i];
            int64_t subtreeIndex = -1;
            for (int64_t j = 0; j < numberOfSubtrees; ++j) {
                if (subtrees[j] == sub) {
                    subtreeIndex = j;
                    break;
                }
            }
            if (subtreeIndex != -1) {
                const uint32_t pos = nextPos[subtreeIndex].fetch_add(1);
                t->sortedBodies[pos] = (uint32_t)i;
            }
        }
        const auto t4 = clk::now();
        ph[3] = ms_between(t3, t4);

        // buildSubtrees (Subtrees.cpp:537-813): one work-group per subtree;
        // inside a group the work-items run in id order over their chunks.
        const int64_t localSize = subtree_work_items;
#pragma omp parallel for schedule(dynamic, 1) num_threads(nthreads())
        for (int64_t g = 0; g < numberOfSubtrees; ++g) {
            if (t->overflow.load()) continue;
            const uint32_t root = subtrees[g];
            const int64_t cntg = bodyCountSubtree[root];
            const int64_t start = startIndex[g];
            const int64_t bpw =
                cntg > localSize ? (int64_t)std::ceil((double)cntg / (double)localSize) : 1;
            for (int64_t lid = 0; lid < localSize; ++lid) {
                for (int64_t index = start + lid * bpw; index < start + lid * bpw + bpw; ++index) {
                    if (index < start + cntg) {
                        if (!insert_from(t, t->sortedBodies[index], root, x, y, z)) break;
                    }
                }
            }
        }
        if (t->overflow.load()) return ORC_E_NODE_OVERFLOW;
        const auto t5 = clk::now();
        ph[4] = ms_between(t4, t5);
    }
    const auto tc0 = clk::now();
    rc = compute_com(t, x, y, z, m);
    const auto tc1 = clk::now();
    ph[5] = ms_between(tc0, tc1);
    if (rc != ORC_OK) return rc;
    if (sort_bodies_flag) sort_bodies(t, x, y, z);
    const auto tc2 = clk::now();
    ph[6] = ms_between(tc1, tc2);
    ph[7] = ms_between(t0, tc2);
    if (phase_ms)
        for (int k = 0; k < 8; ++k) phase_ms[k] = ph[k];
    return ORC_OK;
}

uint32_t orc_tree_node_count(const orc_tree *t) { return t ? t->nextFree.load() : 0; }

void orc_tree_get_aabb(const orc_tree *t, double out[7]) {
    for (int k = 0; k < 7; ++k) out[k] = t->aabb[k];
}

void orc_tree_export(const orc_tree *t, uint32_t *octants, uint32_t *body_of_node,
                     int32_t *node_is_leaf, double *sum_masses, double *com_x, double *com_y,
                     double *com_z, uint32_t *body_count, double *edge, double *min_x,
                     double *min_y, double *min_z, uint32_t *sorted) {
    const size_t S = t->S;
    if (octants) std::memcpy(octants, t->octants.data(), 8 * S * sizeof(uint32_t));
    if (body_of_node) std::memcpy(body_of_node, t->bodyOfNode.data(), S * sizeof(uint32_t));
    if (node_is_leaf) std::memcpy(node_is_leaf, t->isLeaf.data(), S * sizeof(int32_t));
    if (sum_masses) std::memcpy(sum_masses, t->mass.data(), S * sizeof(double));
    if (com_x) std::memcpy(com_x, t->cx.data(), S * sizeof(double));
    if (com_y) std::memcpy(com_y, t->cy.data(), S * sizeof(double));
    if (com_z) std::memcpy(com_z, t->cz.data(), S * sizeof(double));
    if (body_count) std::memcpy(body_count, t->count.data(), S * sizeof(uint32_t));
    if (edge) std::memcpy(edge, t->edge.data(), S * sizeof(double));
    if (min_x) std::memcpy(min_x, t->minx.data(), S * sizeof(double));
    if (min_y) std::memcpy(min_y, t->miny.data(), S * sizeof(double));
    if (min_z) std::memcpy(min_z, t->minz.data(), S * sizeof(double));
    if (sorted) std::memcpy(sorted, t->sorted.data(), t->N * sizeof(uint32_t));
}

uint64_t orc_tree_canonical(const orc_tree *t, uint64_t cap, uint8_t *depth, uint8_t *kind,
                            uint32_t *body, double *mass, double *com_x, double *com_y,
                            double *com_z, uint32_t *count) {
    const uint32_t nodes = t->nextFree.load();
    if (cap < nodes) return 0;
    const size_t S = t->S;
    std::vector<std::pair<uint32_t, uint32_t>> stack;  // (node, depth)
    stack.push_back({0u, 0u});
    uint64_t k = 0;
    while (!stack.empty()) {
        auto [node, d] = stack.back();
        stack.pop_back();
        if (k >= cap) return 0;
        const bool internal = t->isLeaf[node] == 0;
        depth[k] = (uint8_t)d;
        kind[k] = internal ? 2 : (t->bodyOfNode[node] != t->N ? 1 : 0);
        body[k] = internal ? t->N : t->bodyOfNode[node];
        mass[k] = t->mass[node];
        com_x[k] = t->cx[node];
        com_y[k] = t->cy[node];
        com_z[k] = t->cz[node];
        count[k] = t->count[node];
        ++k;
        if (internal)
            for (int dg = 7; dg >= 0; --dg) stack.push_back({t->octants[dg * S + node], d + 1});
    }
    return k;
}

// ---------------------------------------------------------------------------
// traversal: BarnesHutAlgorithm.cpp:319-393
int orc_bh_accel(const orc_tree *t, const double *x, const double *y, const double *z,
                 double theta, double G, double eps2, uint32_t stack_size, int sort_bodies_flag,
                 double *ax, double *ay, double *az, uint32_t *interactions) {
    const uint32_t N = t->N;
    const size_t S = t->S;
    const double THETA = theta;
    std::atomic<int> overflow{0};
#pragma omp parallel num_threads(nthreads())
    {
        std::vector<uint32_t> stack(stack_size + 9);
#pragma omp for schedule(static, 64)
        for (int64_t id = 0; id < (int64_t)N; ++id) {
            const uint32_t i = sort_bodies_flag ? t->sorted[id] : (uint32_t)id;
            double acc_x = 0, acc_y = 0, acc_z = 0;
            uint32_t inter = 0;
            const double pos_x_i = x[i], pos_y_i = y[i], pos_z_i = z[i];
            // local index k mirrors NODES_ON_STACK[stack_size*i + k]; k = 0 is
            // the empty-stack sentinel, the root goes to k = 1.
            uint32_t k = 1;
            stack[k] = 0;
            bool ovf = false;
            while (k != 0) {
                const uint32_t node = stack[k];
                k -= 1;
                if (t->mass[node] != 0 && t->bodyOfNode[node] != i) {
                    const double d_x = (t->cx[node] / t->mass[node]) - pos_x_i;
                    const double d_y = (t->cy[node] / t->mass[node]) - pos_y_i;
                    const double d_z = (t->cz[node] / t->mass[node]) - pos_z_i;
                    const double d = rsqrt_ref(d_x * d_x + d_y * d_y + d_z * d_z);
                    const double currentTheta = t->edge[node] * d;
                    if ((currentTheta < THETA) || t->bodyOfNode[node] != N) {
                        double denominator = (d_x * d_x) + (d_y * d_y) + (d_z * d_z) + eps2;
                        denominator = denominator * denominator * denominator;
                        denominator = rsqrt_ref(denominator);
                        acc_x += t->mass[node] * (d_x * denominator);
                        acc_y += t->mass[node] * (d_y * denominator);
                        acc_z += t->mass[node] * (d_z * denominator);
                        ++inter;
                    } else {
                        if (k + 8 >= stack_size) {
                            ovf = true;
                            break;
                        }
                        static const uint32_t push_order[8] = {5, 7, 4, 6, 1, 3, 0, 2};
                        for (int p = 0; p < 8; ++p)
                            stack[++k] = t->octants[push_order[p] * S + node];
                    }
                }
            }
            if (ovf) overflow.store(1);
            ax[i] = acc_x * G;
            ay[i] = acc_y * G;
            az[i] = acc_z * G;
            if (interactions) interactions[i] = inter;
        }
    }
    return overflow.load() ? ORC_E_STACK_OVERFLOW : ORC_OK;
}

uint32_t orc_stack_size(uint32_t n, int stack_size_param) {
    // Configuration.cpp:24-33
    if (n < 15000)
        return (uint32_t)stack_size_param * (uint32_t)std::ceil(std::log2(n)) + 500;
    return (uint32_t)stack_size_param * (uint32_t)std::ceil(std::log2(n));
}

// ---------------------------------------------------------------------------
// The simulation loop: NaiveAlgorithm.cpp:15-260 / BarnesHutAlgorithm.cpp:18-278
// with visualisation switched off (it only copies state out).
int orc_simulate(int algorithm, uint32_t n, double *x, double *y, double *z, double *vx,
                 double *vy, double *vz, const double *m, double dt, double t_end, double theta,
                 int storage_size_param, int stack_size_param, int max_build_level, double *ax,
                 double *ay, double *az, int32_t *steps_done, double *loop_ms,
                 double *phase_ms_sum) {
    if (n == 0 || (algorithm != 0 && algorithm != 1)) return ORC_E_INVALID_ARG;
    const double G = orc_G_ref();
    const double eps2 = orc_eps2_ref();
    std::vector<double> vhx(n), vhy(n), vhz(n);
    orc_tree *tree = nullptr;
    uint32_t stack = 0;
    double ph[4] = {0, 0, 0, 0};  // LF1, accel, LF2, octree (loop only)
    auto forces = [&](bool in_loop) -> int {
        if (algorithm == 0) {
            const auto a = clk::now();
            orc_naive_accel(n, x, y, z, m, G, eps2, ax, ay, az);
            if (in_loop) ph[1] += ms_between(a, clk::now());
            return ORC_OK;
        }
        const auto a = clk::now();
        int rc = orc_tree_build(tree, ORC_BUILDER_SUBTREES, x, y, z, m, max_build_level, 1024,
                                1024, 640, 1, nullptr);
        const auto b = clk::now();
        if (rc != ORC_OK) return rc;
        rc = orc_bh_accel(tree, x, y, z, theta, G, eps2, stack, 1, ax, ay, az, nullptr);
        const auto c = clk::now();
        if (in_loop) {
            ph[3] += ms_between(a, b);
            ph[1] += ms_between(b, c);
        }
        return rc;
    };
    if (algorithm == 1) {
        tree = orc_tree_create(n, (uint32_t)storage_size_param * n);
        if (!tree) return ORC_E_INVALID_ARG;
        stack = orc_stack_size(n, stack_size_param);
    }
    int rc = forces(false);
    int32_t steps = 0;
    double time = 0.0;
    time += dt;
    const auto loop0 = clk::now();
    while (rc == ORC_OK && time <= t_end + 0.000001) {
        auto a = clk::now();
        orc_kick_drift(n, dt, ax, ay, az, vx, vy, vz, vhx.data(), vhy.data(), vhz.data(), x, y, z);
        ph[0] += ms_between(a, clk::now());
        rc = forces(true);
        if (rc != ORC_OK) break;
        a = clk::now();
        orc_kick(n, dt, ax, ay, az, vhx.data(), vhy.data(), vhz.data(), vx, vy, vz);
        ph[2] += ms_between(a, clk::now());
        ++steps;
        time += dt;
    }
    const auto loop1 = clk::now();
    if (loop_ms) *loop_ms = ms_between(loop0, loop1);
    if (steps_done) *steps_done = steps;
    if (phase_ms_sum)
        for (int k = 0; k < 4; ++k) phase_ms_sum[k] = ph[k];
    orc_tree_destroy(tree);
    return rc;
}

}  // extern "C"

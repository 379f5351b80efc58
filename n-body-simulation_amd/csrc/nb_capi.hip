// nb_capi.hip -- the extern "C" boundary of libnbody_hip.so (include/nbody_capi.h).
//
// A context owns the device state of one system and one HIP stream.  Each
// entry point enqueues its kernels, synchronises once, converts the device
// status word into an NB_E_* code and returns -- the contract of the
// reference's queue.submit(...).wait().
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "nb_mlp.h"

using nb::TreeBufs;
using nb::TreeHdr;

struct nb_ctx {
    uint32_t n = 0;
    nb_params p{};
    int device = 0;
    hipStream_t stream = nullptr;
    double *x = nullptr, *y = nullptr, *z = nullptr;
    double *vx = nullptr, *vy = nullptr, *vz = nullptr;
    double *vhx = nullptr, *vhy = nullptr, *vhz = nullptr;
    double *ax = nullptr, *ay = nullptr, *az = nullptr;
    double *m = nullptr;
    uint32_t *inter = nullptr;
    TreeBufs tree;
    bool tree_valid = false;
    TreeHdr *hdr_host = nullptr;  // pinned copy of tree.hdr
    uint32_t *exact_stack = nullptr;
    uint32_t exact_cap = 0;
    double *split_scratch = nullptr;  // partial sums of the split direct sums
    uint64_t split_scratch_bytes = 0;
    uint32_t *targets_dev = nullptr;
    double *targets_out = nullptr;
    uint32_t targets_cap = 0;
    double *energy_terms = nullptr;
    // learned far-field correction
    nb_mlp *corr = nullptr;
    float corr_tau2 = 0.f;
    double *quad = nullptr;          // 6 per internal cell
    nb::MlpCell *mcell = nullptr;    // per internal cell
    float4 *gquad = nullptr;         // 2 per internal cell: G*Q for the quadrupole walk
    bool quad_valid = false;         // quadrupoles match the current octree
    // asynchronous snapshots: device staging (9 n doubles), two pinned host
    // slots used in turn, the DMA stream and the event that ends each DMA
    double *snap_dev = nullptr;
    double *snap_host[2] = {nullptr, nullptr};
    int snap_slot = 1;  // slot of the last begin
    hipStream_t copy_stream = nullptr;
    hipEvent_t snap_staged = nullptr, snap_done = nullptr;
    bool snap_begun = false;
    std::vector<hipEvent_t> events;
    struct Mark {
        int phase;
        size_t begin, end;  // event indices (end: kOpen until closed)
    };
    std::vector<Mark> marks;
    std::vector<std::pair<int, bool>> pending;  // (phase, begin?) not recorded yet
    size_t next_event = 0;
    double phase[NB_PHASE_COUNT] = {0};
    std::vector<void *> allocations;
    std::string err;
};

namespace {

int fail(nb_ctx *c, int code, const std::string &msg) {
    if (c) c->err = msg;
    return code;
}

#define NB_HIP(ctx, call)                                                              \
    do {                                                                               \
        hipError_t e_ = (call);                                                        \
        if (e_ != hipSuccess)                                                          \
            return fail(ctx, NB_E_HIP, std::string(#call) + ": " + hipGetErrorString(e_)); \
    } while (0)

template <typename T>
int dev_alloc(nb_ctx *c, T **ptr, uint64_t count) {
    void *p = nullptr;
    const uint64_t bytes = std::max<uint64_t>(count, 1) * sizeof(T);
    hipError_t e = hipMalloc(&p, bytes);
    if (e != hipSuccess)
        return fail(c, NB_E_HIP, "hipMalloc(" + std::to_string(bytes) + " B): " + hipGetErrorString(e));
    c->allocations.push_back(p);
    *ptr = static_cast<T *>(p);
    return NB_OK;
}

void dev_free(nb_ctx *c, void *p) {
    if (!p) return;
    auto it = std::find(c->allocations.begin(), c->allocations.end(), p);
    if (it != c->allocations.end()) c->allocations.erase(it);
    (void)hipFree(p);
}

// Phase timing with HIP events on the context stream.  Phases may nest
// (the octree phases inside "Octree creation"); an end closes the most
// recent open mark of its phase.  A mark is only requested by mark_begin /
// mark_end: flush_marks, called before each launch and at the end of a call,
// records one event for every mark requested since the last launch, so
// phase boundaries that coincide share one event (each record idles the GPU
// for a few microseconds).  The events only time: they are read after the
// stream has been synchronised, so they skip the system-scope fence (a cache
// write-back at every record, about 10 us of idle GPU each).
constexpr size_t kOpen = ~size_t(0);

void mark_begin(nb_ctx *c, int phase) { c->pending.push_back({phase, true}); }
void mark_end(nb_ctx *c, int phase) { c->pending.push_back({phase, false}); }

void flush_marks(nb_ctx *c) {
    if (c->pending.empty()) return;
    const size_t idx = c->next_event;
    if (c->events.size() <= idx) {
        hipEvent_t e;
        if (hipEventCreateWithFlags(&e, hipEventDisableSystemFence) != hipSuccess) {
            c->pending.clear();  // timing is lost, nothing else
            return;
        }
        c->events.push_back(e);
    }
    (void)hipEventRecord(c->events[idx], c->stream);
    c->next_event = idx + 1;
    for (const auto &pm : c->pending) {
        if (pm.second) {
            c->marks.push_back({pm.first, idx, kOpen});
            continue;
        }
        for (size_t k = c->marks.size(); k-- > 0;) {
            if (c->marks[k].phase == pm.first && c->marks[k].end == kOpen) {
                c->marks[k].end = idx;
                break;
            }
        }
    }
    c->pending.clear();
}

void timing_reset(nb_ctx *c) {
    c->marks.clear();
    c->pending.clear();
    c->next_event = 0;
    for (double &v : c->phase) v = 0;
}

void timing_collect(nb_ctx *c) {
    for (const auto &mk : c->marks) {
        float ms = 0.f;
        if (mk.end != kOpen &&
            hipEventElapsedTime(&ms, c->events[mk.begin], c->events[mk.end]) == hipSuccess)
            c->phase[mk.phase] += ms;
    }
    c->marks.clear();
    c->next_event = 0;
}

int status_to_code(nb_ctx *c, uint32_t st) {
    if (st & nb::kFlagAabb) return fail(c, NB_E_AABB, "bounding box: non-finite positions");
    if (st & nb::kFlagDepth)
        return fail(c, NB_E_DEPTH, "two bodies share all 42 octant digits (coincident bodies)");
    if (st & nb::kFlagNodeOverflow)
        return fail(c, NB_E_NODE_OVERFLOW, "octree needs more nodes than storage_size_param * N");
    if (st & nb::kFlagZeroMass)
        return fail(c, NB_E_ZERO_MASS, "a non-empty octree cell has zero mass");
    if (st & nb::kFlagStack) return fail(c, NB_E_STACK_OVERFLOW, "traversal stack exhausted");
    if (st & nb::kFlagSortWait)
        return fail(c, NB_E_HIP, "radix sort: a look-back wait did not end");
    if (st & nb::kFlagSortSize)
        return fail(c, NB_E_INVALID_ARG, "radix sort: 2^30 bodies or more, nothing sorted");
    return NB_OK;
}

// Synchronise the stream, fetch the octree header and turn flags into a code.
int finish(nb_ctx *c, bool read_tree_status) {
    flush_marks(c);
    if (read_tree_status)
        NB_HIP(c, hipMemcpyAsync(c->hdr_host, c->tree.hdr, sizeof(TreeHdr), hipMemcpyDeviceToHost,
                                 c->stream));
    NB_HIP(c, hipStreamSynchronize(c->stream));
    timing_collect(c);
    if (read_tree_status) return status_to_code(c, c->hdr_host->status);
    return NB_OK;
}

uint32_t exact_stack_cap(const nb_ctx *c) {
    // the reference's stackSize (Configuration.cpp:24-33), raised to the
    // largest depth-first stack a depth-42 octree can need (1 + 7*42)
    const uint32_t n = c->n;
    const uint32_t lg = (uint32_t)std::ceil(std::log2((double)std::max<uint32_t>(n, 2)));
    uint32_t ref = c->p.stack_size_param * lg + (n < 15000 ? 500u : 0u);
    return std::max<uint32_t>(ref, 1 + 7 * nb::kMaxDepth + 8);
}

// Internal-cell capacity for storage_size_param * n nodes, in 64 bits; the
// group traversal packs an internal-cell index into 26 bits.
bool tree_capacity(uint32_t storage_size_param, uint32_t n, uint32_t *icap) {
    const uint64_t storage = (uint64_t)std::max<uint32_t>(storage_size_param, 1) * n;
    const uint64_t cells = std::max<uint64_t>(1, storage / 8);
    if (cells >= (1u << 26)) return false;
    *icap = (uint32_t)cells;
    return true;
}

// Every device buffer of an octree (null pointers are skipped).
void tree_free(nb_ctx *c, TreeBufs &t) {
    void *bufs[] = {t.hdr, t.aabb_part, t.skey, t.skey_alt, t.klo, t.sidx, t.sidx_alt,
                    t.lcp, t.chain, t.body_parent, t.int_depth, t.int_parent, t.int_node,
                    t.int_lvl, t.int_count, t.int_sums, t.int_local, t.int_first, t.groups,
                    t.group_size, t.group_slot, t.body4, t.gq, t.spanning, t.cross, t.ref,
                    t.cm, t.edge, t.rec, t.radix_ctl, t.scan_part};
    for (void *b : bufs) dev_free(c, b);
    t = TreeBufs();
}

// Allocates the octree buffers of storage_size_param * n nodes into t (which
// the caller passes empty); on failure whatever was allocated is freed again
// and t is left empty.
int alloc_tree_into(nb_ctx *c, TreeBufs &t, uint32_t storage_size_param);

int alloc_tree(nb_ctx *c) { return alloc_tree_into(c, c->tree, c->p.storage_size_param); }

int alloc_tree_into(nb_ctx *c, TreeBufs &t, uint32_t storage_size_param) {
    const uint32_t n = c->n;
    uint32_t icap = 0;
    if (!tree_capacity(storage_size_param, n, &icap))
        return fail(c, NB_E_INVALID_ARG, "storage_size_param * N exceeds 2^29 octree nodes");
    t.n = n;
    t.icap = icap;
    t.mcap = 1 + 8ull * t.icap;
    int rc = NB_OK;
    auto undo = [&](int code) {
        tree_free(c, t);
        return code;
    };
#define NB_TREE_ALLOC(ptr, count) \
    if ((rc = dev_alloc(c, &(ptr), (count)))) return undo(rc)
    NB_TREE_ALLOC(t.hdr, 1);
    NB_TREE_ALLOC(t.aabb_part, 6ull * nb::kAabbBlocks);
    NB_TREE_ALLOC(t.skey, n);
    NB_TREE_ALLOC(t.skey_alt, n);
    NB_TREE_ALLOC(t.klo, n);
    NB_TREE_ALLOC(t.sidx, n);
    NB_TREE_ALLOC(t.sidx_alt, n);
    NB_TREE_ALLOC(t.lcp, n);
    NB_TREE_ALLOC(t.chain, n);
    NB_TREE_ALLOC(t.body_parent, n);
    NB_TREE_ALLOC(t.int_depth, t.icap);
    NB_TREE_ALLOC(t.int_parent, t.icap);
    NB_TREE_ALLOC(t.int_node, t.icap);
    NB_TREE_ALLOC(t.int_lvl, t.icap);
    NB_TREE_ALLOC(t.int_count, t.icap);
    NB_TREE_ALLOC(t.int_sums, t.icap);
    NB_TREE_ALLOC(t.int_local, t.icap);
    NB_TREE_ALLOC(t.int_first, t.icap);
    NB_TREE_ALLOC(t.groups, n);
    NB_TREE_ALLOC(t.group_size, n);
    NB_TREE_ALLOC(t.group_slot, n);
    NB_TREE_ALLOC(t.body4, n);
    NB_TREE_ALLOC(t.gq, TreeBufs::kGroupQueueWords);
    // a cell crossing a COM block boundary holds that boundary's two bodies:
    // at most one per depth (kMaxDepth) and block
    t.spanning_cap = (uint32_t)nb::kMaxDepth * (nb::ceil_div(n, nb::kComBlock) + 1);
    NB_TREE_ALLOC(t.spanning, t.spanning_cap);
    // a super-block owns at most one crossing cell per depth (the cells holding its end)
    t.cross_cap = (uint32_t)nb::kMaxDepth * (nb::ceil_div(n, nb::kComSuper) + 1);
    NB_TREE_ALLOC(t.cross, t.cross_cap);
    NB_TREE_ALLOC(t.ref, t.mcap);
    NB_TREE_ALLOC(t.cm, t.mcap);
    NB_TREE_ALLOC(t.edge, t.mcap);
    NB_TREE_ALLOC(t.rec, t.mcap);
    const uint32_t ctl_words = nb::radix_ctl_words(n);
    NB_TREE_ALLOC(t.radix_ctl, ctl_words);
    t.scan_part_cap = nb::scan_scratch_words(n);
    NB_TREE_ALLOC(t.scan_part, t.scan_part_cap);
#undef NB_TREE_ALLOC
    const hipError_t e[] = {
        hipMemsetAsync(t.hdr, 0, sizeof(TreeHdr), c->stream),
        hipMemsetAsync(t.radix_ctl, 0, 4ull * ctl_words, c->stream),
        // sort scratch starts as valid data (body index 0), so that nothing
        // the sort could leave behind is out of range
        hipMemsetAsync(t.skey_alt, 0, 8ull * n, c->stream),
        hipMemsetAsync(t.sidx_alt, 0, 4ull * n, c->stream)};
    for (hipError_t err : e)
        if (err != hipSuccess)
            return undo(fail(c, NB_E_HIP, std::string("octree buffers: ") + hipGetErrorString(err)));
    return NB_OK;
}

bool params_ok(const nb_params *p) {
    return p &&
           (p->precision == NB_PREC_FP64 || p->precision == NB_PREC_FP32C ||
            p->precision == NB_PREC_FP64_FAST) &&
           p->storage_size_param >= 1 && p->stack_size_param >= 1 && std::isfinite(p->G) &&
           std::isfinite(p->eps2) && p->eps2 >= 0 && std::isfinite(p->theta) && p->theta >= 0 &&
           p->group_max <= nb::kGroupLimit &&
           (p->pack_max == 0 || p->pack_max >= (p->group_max ? p->group_max : nb::kGroupMax));
}

// the traversal group rule of the parameters (0: the defaults)
nb::GroupRule group_rule(const nb_params &p) {
    const uint32_t g = p.group_max ? p.group_max : nb::kGroupMax;
    return nb::GroupRule{g, p.pack_max ? p.pack_max : std::max(nb::kPackMax, g)};
}

int build_tree(nb_ctx *c) {
    TreeBufs &t = c->tree;
    mark_begin(c, NB_PHASE_AABB);
    flush_marks(c);
    nb::launch_aabb(t, c->x, c->y, c->z, c->stream);
    mark_end(c, NB_PHASE_AABB);
    mark_begin(c, NB_PHASE_KEYS_SORT);
    mark_begin(c, NB_PHASE_KEYS);
    flush_marks(c);
    nb::launch_keys(t, c->x, c->y, c->z, c->m, c->stream);
    mark_end(c, NB_PHASE_KEYS);
    mark_begin(c, NB_PHASE_SORT);
    flush_marks(c);
    nb::launch_key_sort(t, c->stream);
    mark_end(c, NB_PHASE_SORT);
    mark_end(c, NB_PHASE_KEYS_SORT);
    mark_begin(c, NB_PHASE_TOPOLOGY);
    flush_marks(c);
    nb::launch_topology(t, c->x, c->y, c->z, c->stream);
    mark_end(c, NB_PHASE_TOPOLOGY);
    mark_begin(c, NB_PHASE_COM);
    const bool fp64 = c->p.precision != NB_PREC_FP32C;  // FP64 and FP64_FAST
    flush_marks(c);
    nb::launch_com(t, c->p.G, fp64, !fp64, group_rule(c->p), c->stream);
    c->quad_valid = false;
    if (!fp64 && (c->corr || c->p.quadrupole)) {
        flush_marks(c);
        nb::launch_quadrupoles(t, c->x, c->y, c->z, c->m, c->quad, c->corr ? c->mcell : nullptr,
                               c->p.quadrupole ? c->gquad : nullptr, c->p.G, c->stream);
        c->quad_valid = true;
    }
    mark_end(c, NB_PHASE_COM);
    return NB_OK;
}

// the moment buffers of the learned correction and the quadrupole walk
int ensure_quad_buffers(nb_ctx *c) {
    int rc;
    if (!c->quad && (rc = dev_alloc(c, &c->quad, 6ull * c->tree.icap))) return rc;
    if (!c->mcell && (rc = dev_alloc(c, &c->mcell, c->tree.icap))) return rc;
    if (!c->gquad && (rc = dev_alloc(c, &c->gquad, 2ull * c->tree.icap))) return rc;
    return NB_OK;
}

int bh_forces_enqueue(nb_ctx *c) {
    mark_begin(c, NB_PHASE_ACCEL);
    flush_marks(c);
    if (c->p.precision != NB_PREC_FP32C) {  // FP64 and FP64_FAST: the reference traversal
        const uint32_t cap = exact_stack_cap(c);
        if (!c->exact_stack || c->exact_cap != cap) {
            dev_free(c, c->exact_stack);
            c->exact_stack = nullptr;
            int rc = dev_alloc(c, &c->exact_stack, (uint64_t)cap * c->n);
            if (rc) return rc;
            c->exact_cap = cap;
        }
        nb::launch_bh_exact(c->tree, c->x, c->y, c->z, c->p.G, c->p.eps2, c->p.theta,
                            c->p.sort_bodies, c->exact_stack, c->exact_cap, c->ax, c->ay, c->az,
                            c->inter, c->stream);
    } else if (c->corr && c->quad_valid) {
        const nb::MlpWeights w = nb::mlp_weights(c->corr);
        const nb::CorrArgs ca{w.pack, w.master, c->mcell, c->corr_tau2};
        nb::launch_bh_group_corr(c->tree, c->x, c->y, c->z, c->m, c->p.eps2, c->p.theta, ca,
                                 c->p.quadrupole ? c->gquad : nullptr, c->ax, c->ay, c->az,
                                 c->inter, c->stream);
    } else {
        const bool quad = c->p.quadrupole && c->quad_valid;
        nb::launch_bh_group(c->tree, c->x, c->y, c->z, c->m, c->p.eps2, c->p.theta,
                            quad ? c->gquad : nullptr, c->ax, c->ay, c->az, c->inter, c->stream);
    }
    mark_end(c, NB_PHASE_ACCEL);
    return NB_OK;
}

int direct_enqueue(nb_ctx *c) {
    mark_begin(c, NB_PHASE_ACCEL);
    flush_marks(c);
    if (c->p.precision == NB_PREC_FP64) {
        nb::launch_direct_f64(c->n, c->x, c->y, c->z, c->m, c->p.G, c->p.eps2, c->ax, c->ay, c->az,
                              c->stream);
    } else {
        const uint64_t bytes = nb::direct_split_scratch_bytes(c->n);
        if (bytes && (!c->split_scratch || c->split_scratch_bytes < bytes)) {
            dev_free(c, c->split_scratch);
            c->split_scratch = nullptr;
            int rc = dev_alloc(c, &c->split_scratch, bytes / sizeof(double));
            if (rc) return rc;
            c->split_scratch_bytes = bytes;
        }
        nb::launch_direct_split(c->p.precision == NB_PREC_FP64_FAST, c->n, c->x, c->y, c->z,
                                c->m, c->p.G, c->p.eps2, c->ax, c->ay, c->az, c->split_scratch,
                                c->stream);
    }
    mark_end(c, NB_PHASE_ACCEL);
    return NB_OK;
}

}  // namespace

extern "C" {

void nb_default_params(nb_params *p) {
    if (!p) return;
    // nBodyAlgorithm.hpp:55-61
    double G = 6.67428 * std::pow(10, -11);
    const double meter_AU = 1.0 / (1.49597870691 * std::pow(10, 11));
    const double second_Days = 1.0 / 86400;
    G = G * (std::pow(meter_AU, 3) / std::pow(second_Days, 2));
    p->G = G;
    p->eps2 = std::pow(10, -22);  // Configuration.cpp:6
    p->theta = 1.05;              // Configuration.cpp:10
    p->precision = NB_PREC_FP64;
    p->sort_bodies = 1;
    p->storage_size_param = 16;
    p->stack_size_param = 16;
    p->quadrupole = 0;
    p->group_max = 0;
    p->pack_max = 0;
}

const char *nb_strerror(int code) {
    switch (code) {
        case NB_OK: return "ok";
        case NB_E_INVALID_ARG: return "invalid argument";
        case NB_E_HIP: return "HIP runtime error";
        case NB_E_NODE_OVERFLOW: return "octree storage exhausted";
        case NB_E_STACK_OVERFLOW: return "traversal stack exhausted";
        case NB_E_DEPTH: return "bodies not separable within 42 octree levels";
        case NB_E_ZERO_MASS: return "zero-mass octree cell";
        case NB_E_NOT_BUILT: return "octree not built for the current positions";
        case NB_E_AABB: return "bounding box undefined (non-finite positions)";
        case NB_E_NO_DEVICE: return "no HIP device";
        default: return "unknown error";
    }
}

const char *nb_last_error(const nb_ctx *c) { return c ? c->err.c_str() : "null context"; }

int nb_device_name(char *buf, size_t len) {
    if (!buf || len == 0) return NB_E_INVALID_ARG;
    int dev = 0, count = 0;
    if (hipGetDeviceCount(&count) != hipSuccess || count == 0) return NB_E_NO_DEVICE;
    if (hipGetDevice(&dev) != hipSuccess) return NB_E_HIP;
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, dev) != hipSuccess) return NB_E_HIP;
    std::snprintf(buf, len, "%s (%s)", prop.name, prop.gcnArchName);
    return NB_OK;
}

int nb_create(nb_ctx **out, uint32_t n, const nb_params *p) {
    if (!out || n == 0 || !params_ok(p)) return NB_E_INVALID_ARG;
    *out = nullptr;
    int count = 0;
    if (hipGetDeviceCount(&count) != hipSuccess || count == 0) return NB_E_NO_DEVICE;
    nb_ctx *c = new nb_ctx();
    c->n = n;
    c->p = *p;
    int rc = NB_OK;
    auto bail = [&](int code) {
        nb_destroy(c);
        return code;
    };
    if (hipGetDevice(&c->device) != hipSuccess) return bail(NB_E_HIP);
    if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) return bail(NB_E_HIP);
    double **state[] = {&c->x, &c->y, &c->z, &c->vx, &c->vy, &c->vz, &c->vhx,
                        &c->vhy, &c->vhz, &c->ax, &c->ay, &c->az, &c->m};
    for (double **s : state)
        if ((rc = dev_alloc(c, s, n))) return bail(rc);
    if ((rc = dev_alloc(c, &c->inter, n))) return bail(rc);
    if (hipHostMalloc((void **)&c->hdr_host, sizeof(TreeHdr), hipHostMallocDefault) != hipSuccess)
        return bail(NB_E_HIP);
    std::memset(c->hdr_host, 0, sizeof(TreeHdr));
    if ((rc = alloc_tree(c))) return bail(rc);
    if (p->quadrupole && (rc = ensure_quad_buffers(c))) return bail(rc);
    for (double **s : state)
        if (hipMemsetAsync(*s, 0, (size_t)n * sizeof(double), c->stream) != hipSuccess)
            return bail(NB_E_HIP);
    if (hipMemsetAsync(c->inter, 0, (size_t)n * sizeof(uint32_t), c->stream) != hipSuccess)
        return bail(NB_E_HIP);
    if (hipStreamSynchronize(c->stream) != hipSuccess) return bail(NB_E_HIP);
    *out = c;
    return NB_OK;
}

int nb_destroy(nb_ctx *c) {
    if (!c) return NB_E_INVALID_ARG;
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    if (c->copy_stream) (void)hipStreamSynchronize(c->copy_stream);
    for (void *p : c->allocations) (void)hipFree(p);
    for (double *h : c->snap_host)
        if (h) (void)hipHostFree(h);
    if (c->snap_staged) (void)hipEventDestroy(c->snap_staged);
    if (c->snap_done) (void)hipEventDestroy(c->snap_done);
    if (c->copy_stream) (void)hipStreamDestroy(c->copy_stream);
    for (hipEvent_t e : c->events) (void)hipEventDestroy(e);
    if (c->hdr_host) (void)hipHostFree(c->hdr_host);
    if (c->stream) (void)hipStreamDestroy(c->stream);
    delete c;
    return NB_OK;
}

int nb_set_params(nb_ctx *c, const nb_params *p) {
    if (!c || !params_ok(p)) return fail(c, NB_E_INVALID_ARG, "invalid parameters");
    if (p->storage_size_param != c->p.storage_size_param) {
        // The new octree storage is allocated first; on any failure (capacity
        // beyond 2^29 nodes, out of memory) the old parameters and buffers
        // stay as they were and the context remains usable.
        TreeBufs fresh;
        int rc = alloc_tree_into(c, fresh, p->storage_size_param);
        if (rc) return rc;
        double *quad = nullptr;
        nb::MlpCell *mcell = nullptr;
        float4 *gquad = nullptr;
        if (c->corr || p->quadrupole) {
            if ((rc = dev_alloc(c, &quad, 6ull * fresh.icap)) ||
                (rc = dev_alloc(c, &mcell, fresh.icap)) ||
                (rc = dev_alloc(c, &gquad, 2ull * fresh.icap))) {
                dev_free(c, quad);
                dev_free(c, mcell);
                tree_free(c, fresh);
                return rc;
            }
        }
        if (hipStreamSynchronize(c->stream) != hipSuccess) {
            dev_free(c, quad);
            dev_free(c, mcell);
            dev_free(c, gquad);
            tree_free(c, fresh);
            return fail(c, NB_E_HIP, "nb_set_params: stream synchronisation failed");
        }
        tree_free(c, c->tree);
        dev_free(c, c->quad);
        dev_free(c, c->mcell);
        dev_free(c, c->gquad);
        c->tree = fresh;
        c->quad = quad;
        c->mcell = mcell;
        c->gquad = gquad;
        c->quad_valid = false;
    } else if (p->quadrupole && !c->p.quadrupole) {
        int rc = ensure_quad_buffers(c);
        if (rc) return rc;
    }
    c->p = *p;
    c->tree_valid = false;  // opening radii depend on theta
    return NB_OK;
}

int nb_get_params(const nb_ctx *c, nb_params *p) {
    if (!c || !p) return NB_E_INVALID_ARG;
    *p = c->p;
    return NB_OK;
}

uint32_t nb_body_count(const nb_ctx *c) { return c ? c->n : 0; }

int nb_set_state(nb_ctx *c, const double *x, const double *y, const double *z, const double *vx,
                 const double *vy, const double *vz, const double *m) {
    if (!c || !x || !y || !z || !vx || !vy || !vz || !m)
        return fail(c, NB_E_INVALID_ARG, "nb_set_state: null pointer");
    const size_t bytes = (size_t)c->n * sizeof(double);
    const double *src[] = {x, y, z, vx, vy, vz, m};
    double *dst[] = {c->x, c->y, c->z, c->vx, c->vy, c->vz, c->m};
    for (int k = 0; k < 7; ++k)
        NB_HIP(c, hipMemcpyAsync(dst[k], src[k], bytes, hipMemcpyHostToDevice, c->stream));
    NB_HIP(c, hipMemsetAsync(c->ax, 0, bytes, c->stream));
    NB_HIP(c, hipMemsetAsync(c->ay, 0, bytes, c->stream));
    NB_HIP(c, hipMemsetAsync(c->az, 0, bytes, c->stream));
    c->tree_valid = false;
    NB_HIP(c, hipStreamSynchronize(c->stream));
    return NB_OK;
}

int nb_set_positions(nb_ctx *c, const double *x, const double *y, const double *z) {
    if (!c || !x || !y || !z) return fail(c, NB_E_INVALID_ARG, "nb_set_positions: null pointer");
    const size_t bytes = (size_t)c->n * sizeof(double);
    NB_HIP(c, hipMemcpyAsync(c->x, x, bytes, hipMemcpyHostToDevice, c->stream));
    NB_HIP(c, hipMemcpyAsync(c->y, y, bytes, hipMemcpyHostToDevice, c->stream));
    NB_HIP(c, hipMemcpyAsync(c->z, z, bytes, hipMemcpyHostToDevice, c->stream));
    c->tree_valid = false;
    NB_HIP(c, hipStreamSynchronize(c->stream));
    return NB_OK;
}

int nb_get_state(nb_ctx *c, double *x, double *y, double *z, double *vx, double *vy,
                 double *vz) {
    if (!c) return NB_E_INVALID_ARG;
    const size_t bytes = (size_t)c->n * sizeof(double);
    double *dst[] = {x, y, z, vx, vy, vz};
    const double *src[] = {c->x, c->y, c->z, c->vx, c->vy, c->vz};
    for (int k = 0; k < 6; ++k)
        if (dst[k]) NB_HIP(c, hipMemcpyAsync(dst[k], src[k], bytes, hipMemcpyDeviceToHost, c->stream));
    NB_HIP(c, hipStreamSynchronize(c->stream));
    return NB_OK;
}

int nb_get_accel(nb_ctx *c, double *ax, double *ay, double *az) {
    if (!c) return NB_E_INVALID_ARG;
    const size_t bytes = (size_t)c->n * sizeof(double);
    double *dst[] = {ax, ay, az};
    const double *src[] = {c->ax, c->ay, c->az};
    for (int k = 0; k < 3; ++k)
        if (dst[k]) NB_HIP(c, hipMemcpyAsync(dst[k], src[k], bytes, hipMemcpyDeviceToHost, c->stream));
    NB_HIP(c, hipStreamSynchronize(c->stream));
    return NB_OK;
}

int nb_snapshot_begin(nb_ctx *c) {
    if (!c) return NB_E_INVALID_ARG;
    const size_t n = c->n, bytes = n * sizeof(double);
    if (!c->snap_done) {
        // first use; a failure part-way leaves what was made for the next try
        if (!c->snap_dev) {
            int rc = dev_alloc(c, &c->snap_dev, 9ull * n);
            if (rc) return rc;
        }
        for (double *&h : c->snap_host)
            if (!h) NB_HIP(c, hipHostMalloc((void **)&h, 9 * bytes, hipHostMallocDefault));
        if (!c->copy_stream)
            NB_HIP(c, hipStreamCreateWithFlags(&c->copy_stream, hipStreamNonBlocking));
        if (!c->snap_staged)
            NB_HIP(c, hipEventCreateWithFlags(&c->snap_staged, hipEventDisableTiming));
        NB_HIP(c, hipEventCreateWithFlags(&c->snap_done, hipEventDisableTiming));
    } else {
        // the staging buffer is reused: the previous DMA must have read it
        NB_HIP(c, hipStreamWaitEvent(c->stream, c->snap_done, 0));
    }
    const double *src[9] = {c->x, c->y, c->z, c->vx, c->vy, c->vz, c->ax, c->ay, c->az};
    for (int k = 0; k < 9; ++k)
        NB_HIP(c, hipMemcpyAsync(c->snap_dev + k * n, src[k], bytes, hipMemcpyDeviceToDevice,
                                 c->stream));
    NB_HIP(c, hipEventRecord(c->snap_staged, c->stream));
    NB_HIP(c, hipStreamWaitEvent(c->copy_stream, c->snap_staged, 0));
    c->snap_slot ^= 1;
    NB_HIP(c, hipMemcpyAsync(c->snap_host[c->snap_slot], c->snap_dev, 9 * bytes,
                             hipMemcpyDeviceToHost, c->copy_stream));
    NB_HIP(c, hipEventRecord(c->snap_done, c->copy_stream));
    c->snap_begun = true;
    return NB_OK;
}

int nb_snapshot_wait(nb_ctx *c, const double **arrays) {
    if (!c || !arrays) return NB_E_INVALID_ARG;
    if (!c->snap_begun) return fail(c, NB_E_INVALID_ARG, "nb_snapshot_wait: no snapshot begun");
    NB_HIP(c, hipStreamSynchronize(c->copy_stream));
    for (int k = 0; k < 9; ++k) arrays[k] = c->snap_host[c->snap_slot] + (size_t)k * c->n;
    return NB_OK;
}

int nb_forces_direct(nb_ctx *c) {
    if (!c) return NB_E_INVALID_ARG;
    timing_reset(c);
    int rc = direct_enqueue(c);
    if (rc) return rc;
    return finish(c, false);
}

int nb_forces_direct_targets(nb_ctx *c, uint32_t nt, const uint32_t *targets, double *ax,
                             double *ay, double *az) {
    if (!c || (nt > 0 && (!targets || !ax || !ay || !az)))
        return fail(c, NB_E_INVALID_ARG, "nb_forces_direct_targets: bad arguments");
    if (nt == 0) return NB_OK;
    for (uint32_t t = 0; t < nt; ++t)
        if (targets[t] >= c->n) return fail(c, NB_E_INVALID_ARG, "target index out of range");
    if (nt > c->targets_cap) {
        dev_free(c, c->targets_dev);
        dev_free(c, c->targets_out);
        c->targets_dev = nullptr;
        c->targets_out = nullptr;
        int rc;
        if ((rc = dev_alloc(c, &c->targets_dev, nt))) return rc;
        if ((rc = dev_alloc(c, &c->targets_out, 3ull * nt))) return rc;
        c->targets_cap = nt;
    }
    timing_reset(c);
    NB_HIP(c, hipMemcpyAsync(c->targets_dev, targets, (size_t)nt * 4, hipMemcpyHostToDevice, c->stream));
    mark_begin(c, NB_PHASE_ACCEL);
    flush_marks(c);
    nb::launch_direct_targets_f64(c->n, c->x, c->y, c->z, c->m, c->p.G, c->p.eps2, nt,
                                  c->targets_dev, c->targets_out, c->targets_out + nt,
                                  c->targets_out + 2ull * nt, c->stream);
    mark_end(c, NB_PHASE_ACCEL);
    flush_marks(c);  // before the copies
    double *dst[] = {ax, ay, az};
    for (int k = 0; k < 3; ++k)
        NB_HIP(c, hipMemcpyAsync(dst[k], c->targets_out + (uint64_t)k * nt, (size_t)nt * 8,
                                 hipMemcpyDeviceToHost, c->stream));
    return finish(c, false);
}

int nb_bh_build(nb_ctx *c) {
    if (!c) return NB_E_INVALID_ARG;
    timing_reset(c);
    mark_begin(c, NB_PHASE_OCTREE);
    build_tree(c);
    mark_end(c, NB_PHASE_OCTREE);
    int rc = finish(c, true);
    c->tree_valid = rc == NB_OK;
    return rc;
}

int nb_bh_forces(nb_ctx *c) {
    if (!c) return NB_E_INVALID_ARG;
    if (!c->tree_valid) return fail(c, NB_E_NOT_BUILT, "nb_bh_forces: build the octree first");
    timing_reset(c);
    int rc = bh_forces_enqueue(c);
    if (rc) return rc;
    return finish(c, true);
}

int nb_kick_drift(nb_ctx *c, double dt) {
    if (!c) return NB_E_INVALID_ARG;
    timing_reset(c);
    mark_begin(c, NB_PHASE_LF1);
    flush_marks(c);
    nb::launch_kick_drift(c->n, dt, c->ax, c->ay, c->az, c->vx, c->vy, c->vz, c->vhx, c->vhy,
                          c->vhz, c->x, c->y, c->z, c->stream);
    mark_end(c, NB_PHASE_LF1);
    c->tree_valid = false;
    return finish(c, false);
}

int nb_kick(nb_ctx *c, double dt) {
    if (!c) return NB_E_INVALID_ARG;
    timing_reset(c);
    mark_begin(c, NB_PHASE_LF2);
    flush_marks(c);
    nb::launch_kick(c->n, dt, c->ax, c->ay, c->az, c->vhx, c->vhy, c->vhz, c->vx, c->vy, c->vz,
                    c->stream);
    mark_end(c, NB_PHASE_LF2);
    return finish(c, false);
}

int nb_steps(nb_ctx *c, int algorithm, double dt, uint32_t count) {
    if (!c || (algorithm != NB_ALGO_NAIVE && algorithm != NB_ALGO_BARNES_HUT))
        return fail(c, NB_E_INVALID_ARG, "nb_steps: bad algorithm");
    timing_reset(c);
    // Between two steps of the run nothing reads v, so the kick that ends
    // step s - 1 and the kick-drift that starts step s run as one pass
    // (timed as leapfrog part 1); the last step's kick writes v.
    for (uint32_t s = 0; s < count; ++s) {
        mark_begin(c, NB_PHASE_LF1);
        flush_marks(c);
        if (s == 0)
            nb::launch_kick_drift(c->n, dt, c->ax, c->ay, c->az, c->vx, c->vy, c->vz, c->vhx,
                                  c->vhy, c->vhz, c->x, c->y, c->z, c->stream);
        else
            nb::launch_kick_kick_drift(c->n, dt, c->ax, c->ay, c->az, c->vhx, c->vhy, c->vhz,
                                       c->x, c->y, c->z, c->stream);
        mark_end(c, NB_PHASE_LF1);
        int rc;
        if (algorithm == NB_ALGO_BARNES_HUT) {
            mark_begin(c, NB_PHASE_OCTREE);
            build_tree(c);
            mark_end(c, NB_PHASE_OCTREE);
            rc = bh_forces_enqueue(c);
        } else {
            rc = direct_enqueue(c);
        }
        if (rc) return rc;
    }
    if (count > 0) {
        mark_begin(c, NB_PHASE_LF2);
        flush_marks(c);
        nb::launch_kick(c->n, dt, c->ax, c->ay, c->az, c->vhx, c->vhy, c->vhz, c->vx, c->vy,
                        c->vz, c->stream);
        mark_end(c, NB_PHASE_LF2);
    }
    const bool bh = algorithm == NB_ALGO_BARNES_HUT && count > 0;
    int rc = finish(c, bh);
    c->tree_valid = false;  // positions moved after the last build
    return rc;
}

int nb_energy(nb_ctx *c, double *kinetic, double *potential, double *total, double *virial) {
    if (!c) return NB_E_INVALID_ARG;
    if (!c->energy_terms) {
        int rc = dev_alloc(c, &c->energy_terms, 2ull * c->n);
        if (rc) return rc;
    }
    const uint32_t n = c->n;
    double *ekin = c->energy_terms, *epot = c->energy_terms + n;
    nb::launch_energy_terms(n, c->x, c->y, c->z, c->vx, c->vy, c->vz, c->m, c->p.G, ekin, epot,
                            c->stream);
    std::vector<double> hk(n), hp(n);
    NB_HIP(c, hipMemcpyAsync(hk.data(), ekin, (size_t)n * 8, hipMemcpyDeviceToHost, c->stream));
    NB_HIP(c, hipMemcpyAsync(hp.data(), epot, (size_t)n * 8, hipMemcpyDeviceToHost, c->stream));
    NB_HIP(c, hipStreamSynchronize(c->stream));
    // nBodyAlgorithm.cpp:71-85
    double E_kin_result = 0, E_pot_result = 0;
    for (uint32_t i = 0; i < n; ++i) {
        E_kin_result += hk[i];
        E_pot_result += hp[i];
    }
    E_pot_result *= -1;
    if (kinetic) *kinetic = E_kin_result;
    if (potential) *potential = E_pot_result;
    if (total) *total = E_kin_result + E_pot_result;
    if (virial) *virial = (2.0 * E_kin_result) / std::abs(E_pot_result);
    return NB_OK;
}

int nb_phase_ms(const nb_ctx *c, double *ms) {
    if (!c || !ms) return NB_E_INVALID_ARG;
    for (int k = 0; k < NB_PHASE_COUNT; ++k) ms[k] = c->phase[k];
    return NB_OK;
}

int nb_tree_info(nb_ctx *c, uint64_t *node_count, uint32_t *internal_count, uint32_t *max_depth,
                 double aabb[7]) {
    if (!c) return NB_E_INVALID_ARG;
    if (c->tree_valid) nb::launch_max_depth(c->tree, c->stream);  // computed on demand
    NB_HIP(c, hipMemcpyAsync(c->hdr_host, c->tree.hdr, sizeof(TreeHdr), hipMemcpyDeviceToHost,
                             c->stream));
    NB_HIP(c, hipStreamSynchronize(c->stream));
    const TreeHdr &h = *c->hdr_host;
    const uint32_t I = std::min(h.internal_count, c->tree.icap);
    if (node_count) *node_count = 1 + 8ull * I;
    if (internal_count) *internal_count = h.internal_count;
    if (max_depth) *max_depth = h.max_depth;
    if (aabb) {
        const double v[7] = {h.minx, h.miny, h.minz, h.maxx, h.maxy, h.maxz, h.edge};
        for (int k = 0; k < 7; ++k) aabb[k] = v[k];
    }
    return NB_OK;
}

int nb_tree_sorted(nb_ctx *c, uint32_t *sorted) {
    if (!c || !sorted) return NB_E_INVALID_ARG;
    NB_HIP(c, hipMemcpyAsync(sorted, c->tree.sidx, (size_t)c->n * 4, hipMemcpyDeviceToHost, c->stream));
    NB_HIP(c, hipStreamSynchronize(c->stream));
    return NB_OK;
}

int nb_bh_interactions(nb_ctx *c, uint32_t *count) {
    if (!c || !count) return NB_E_INVALID_ARG;
    NB_HIP(c, hipMemcpyAsync(count, c->inter, (size_t)c->n * 4, hipMemcpyDeviceToHost, c->stream));
    NB_HIP(c, hipStreamSynchronize(c->stream));
    return NB_OK;
}

uint64_t nb_tree_canonical(nb_ctx *c, uint64_t cap, uint8_t *depth, uint8_t *kind,
                           uint32_t *body, double *mass, double *com_x, double *com_y,
                           double *com_z, uint32_t *count) {
    if (!c || !depth || !kind || !body || !mass || !com_x || !com_y || !com_z || !count) return 0;
    if (hipMemcpyAsync(c->hdr_host, c->tree.hdr, sizeof(TreeHdr), hipMemcpyDeviceToHost,
                       c->stream) != hipSuccess ||
        hipStreamSynchronize(c->stream) != hipSuccess)
        return 0;
    const uint32_t I = std::min(c->hdr_host->internal_count, c->tree.icap);
    const uint64_t nodes = 1 + 8ull * I;
    if (cap < nodes) return 0;
    const uint32_t n = c->n;
    std::vector<int32_t> ref(nodes);
    std::vector<nb::InternalSums> sums(I);
    std::vector<uint32_t> cnt(I);
    std::vector<double> hx(n), hy(n), hz(n), hm(n);
    bool ok = hipMemcpy(ref.data(), c->tree.ref, nodes * 4, hipMemcpyDeviceToHost) == hipSuccess;
    if (I) {
        ok = ok && hipMemcpy(sums.data(), c->tree.int_sums, (size_t)I * sizeof(nb::InternalSums),
                             hipMemcpyDeviceToHost) == hipSuccess;
        ok = ok && hipMemcpy(cnt.data(), c->tree.int_count, (size_t)I * 4, hipMemcpyDeviceToHost) ==
                       hipSuccess;
    }
    ok = ok && hipMemcpy(hx.data(), c->x, (size_t)n * 8, hipMemcpyDeviceToHost) == hipSuccess;
    ok = ok && hipMemcpy(hy.data(), c->y, (size_t)n * 8, hipMemcpyDeviceToHost) == hipSuccess;
    ok = ok && hipMemcpy(hz.data(), c->z, (size_t)n * 8, hipMemcpyDeviceToHost) == hipSuccess;
    ok = ok && hipMemcpy(hm.data(), c->m, (size_t)n * 8, hipMemcpyDeviceToHost) == hipSuccess;
    if (!ok) return 0;
    return nb::canonical_dump(ref.data(), sums.data(), cnt.data(), hx.data(), hy.data(), hz.data(),
                              hm.data(), n, cap, depth, kind, body, mass, com_x, com_y, com_z,
                              count);
}

// ---- learned far-field correction: context side ---------------------------------

int nb_set_correction(nb_ctx *c, nb_mlp *mlp, double tau) {
    if (!c || !std::isfinite(tau) || tau < 0)
        return fail(c, NB_E_INVALID_ARG, "nb_set_correction: bad arguments");
    if (mlp) {
        int rc = ensure_quad_buffers(c);
        if (rc) return rc;
    }
    c->corr = mlp;
    c->corr_tau2 = (float)(tau * tau);
    c->tree_valid = false;  // the next build adds the quadrupoles
    c->quad_valid = false;
    return NB_OK;
}

int nb_quadrupole_pairs(nb_ctx *c, uint64_t *pairs) {
    if (!c || !pairs) return NB_E_INVALID_ARG;
    NB_HIP(c, hipMemcpyAsync(c->hdr_host, c->tree.hdr, sizeof(TreeHdr), hipMemcpyDeviceToHost,
                             c->stream));
    NB_HIP(c, hipStreamSynchronize(c->stream));
    *pairs = c->hdr_host->quad_pairs;
    return NB_OK;
}

int nb_correction_pairs(nb_ctx *c, uint64_t *pairs) {
    if (!c || !pairs) return NB_E_INVALID_ARG;
    NB_HIP(c, hipMemcpyAsync(c->hdr_host, c->tree.hdr, sizeof(TreeHdr), hipMemcpyDeviceToHost,
                             c->stream));
    NB_HIP(c, hipStreamSynchronize(c->stream));
    *pairs = c->hdr_host->corr_pairs;
    return NB_OK;
}

}  // extern "C"

namespace {

// quadrupoles of the current octree, computed now if the build did not
int quad_now(nb_ctx *c) {
    if (!c->tree_valid) return fail(c, NB_E_NOT_BUILT, "build the octree first");
    if (c->quad_valid) return NB_OK;
    int rc = ensure_quad_buffers(c);
    if (rc) return rc;
    nb::launch_quadrupoles(c->tree, c->x, c->y, c->z, c->m, c->quad, c->mcell,
                           c->p.quadrupole ? c->gquad : nullptr, c->p.G, c->stream);
    NB_HIP(c, hipGetLastError());
    NB_HIP(c, hipStreamSynchronize(c->stream));
    c->quad_valid = true;
    return NB_OK;
}

}  // namespace

extern "C" {

uint64_t nb_tree_quadrupoles(nb_ctx *c, uint64_t cap, double *q) {
    if (!c || !q || quad_now(c)) return 0;
    if (hipMemcpyAsync(c->hdr_host, c->tree.hdr, sizeof(TreeHdr), hipMemcpyDeviceToHost,
                       c->stream) != hipSuccess ||
        hipStreamSynchronize(c->stream) != hipSuccess)
        return 0;
    const uint32_t I = std::min(c->hdr_host->internal_count, c->tree.icap);
    const uint64_t nodes = 1 + 8ull * I;
    if (cap < nodes) return 0;
    std::vector<int32_t> ref(nodes);
    std::vector<double> quad(6ull * I);
    bool ok = hipMemcpy(ref.data(), c->tree.ref, nodes * 4, hipMemcpyDeviceToHost) == hipSuccess;
    if (I) ok = ok && hipMemcpy(quad.data(), c->quad, quad.size() * 8, hipMemcpyDeviceToHost) == hipSuccess;
    if (!ok) return 0;
    std::vector<uint64_t> stack{0};
    uint64_t k = 0;
    while (!stack.empty()) {
        const uint64_t node = stack.back();
        stack.pop_back();
        const int32_t r = ref[node];
        for (int u = 0; u < 6; ++u) q[6 * k + u] = r >= 0 ? quad[6ull * r + u] : 0.0;
        if (r >= 0)
            for (int dg = 7; dg >= 0; --dg) stack.push_back(1ull + 8ull * (uint64_t)r + dg);
        ++k;
    }
    return k;
}

int nb_mlp_pairs(nb_ctx *c, double theta, uint32_t nt, const uint32_t *targets, uint64_t cap,
                 uint32_t *cell, uint32_t *body, float *features, float *residuals,
                 uint64_t *count) {
    if (!c || !count || !(theta > 0) || (nt && !targets) ||
        (cap && (!cell || !body || !features || !residuals)))
        return fail(c, NB_E_INVALID_ARG, "nb_mlp_pairs: bad arguments");
    *count = 0;
    for (uint32_t t = 0; t < nt; ++t)
        if (targets[t] >= c->n) return fail(c, NB_E_INVALID_ARG, "target index out of range");
    int rc = quad_now(c);
    if (rc) return rc;
    if (nt == 0) return NB_OK;
    const uint32_t stack_cap = 1 + 7 * nb::kMaxDepth + 8;
    // scratch of this call only
    struct Scratch {
        nb_ctx *c;
        std::vector<void *> p;
        ~Scratch() {
            for (void *q : p) dev_free(c, q);
        }
    } sc{c, {}};
    uint32_t *tdev = nullptr, *stack = nullptr, *pcell = nullptr, *pbody = nullptr;
    unsigned long long *counter = nullptr;
    float *F = nullptr, *T = nullptr;
    const uint64_t cap1 = std::max<uint64_t>(cap, 1);
    if ((rc = dev_alloc(c, &tdev, nt))) return rc;
    sc.p.push_back(tdev);
    if ((rc = dev_alloc(c, &stack, (uint64_t)nt * stack_cap))) return rc;
    sc.p.push_back(stack);
    if ((rc = dev_alloc(c, &counter, 1))) return rc;
    sc.p.push_back(counter);
    if ((rc = dev_alloc(c, &pcell, cap1))) return rc;
    sc.p.push_back(pcell);
    if ((rc = dev_alloc(c, &pbody, cap1))) return rc;
    sc.p.push_back(pbody);
    if ((rc = dev_alloc(c, &F, cap1 * nb::kMlpIn))) return rc;
    sc.p.push_back(F);
    if ((rc = dev_alloc(c, &T, cap1 * nb::kMlpOut))) return rc;
    sc.p.push_back(T);
    NB_HIP(c, hipMemcpyAsync(tdev, targets, (size_t)nt * 4, hipMemcpyHostToDevice, c->stream));
    NB_HIP(c, hipMemsetAsync(counter, 0, sizeof(unsigned long long), c->stream));
    nb::launch_pair_walk(c->tree, nt, tdev, c->x, c->y, c->z, theta, stack, stack_cap, cap, pcell,
                         pbody, counter, c->stream);
    unsigned long long found = 0;
    NB_HIP(c, hipMemcpyAsync(&found, counter, sizeof(found), hipMemcpyDeviceToHost, c->stream));
    NB_HIP(c, hipStreamSynchronize(c->stream));
    const uint64_t m = std::min<uint64_t>(found, cap);
    nb::launch_pair_targets(c->tree, m, pcell, pbody, c->x, c->y, c->z, c->m, c->quad, c->mcell,
                            c->p.G, c->p.eps2, F, T, c->stream);
    NB_HIP(c, hipGetLastError());
    if (m) {
        NB_HIP(c, hipMemcpyAsync(cell, pcell, m * 4, hipMemcpyDeviceToHost, c->stream));
        NB_HIP(c, hipMemcpyAsync(body, pbody, m * 4, hipMemcpyDeviceToHost, c->stream));
        NB_HIP(c, hipMemcpyAsync(features, F, m * nb::kMlpIn * 4, hipMemcpyDeviceToHost, c->stream));
        NB_HIP(c, hipMemcpyAsync(residuals, T, m * nb::kMlpOut * 4, hipMemcpyDeviceToHost, c->stream));
    }
    NB_HIP(c, hipStreamSynchronize(c->stream));
    *count = found;
    return NB_OK;
}

}  // extern "C"

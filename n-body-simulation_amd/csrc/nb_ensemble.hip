// nb_ensemble.hip -- many independent systems advanced together.
//
// The systems' states are concatenated in one structure-of-arrays.  One
// loop iteration for the whole ensemble is a fixed sequence of launches,
// whatever the number of systems:
//   kick-drift (every body, its system's dt)
//   octree build of every Barnes-Hut system (the one-system build's stages,
//           each over the slots or blocks of all systems at once)
//   forces: one persistent launch over a work queue holding the traversal
//           groups of the Barnes-Hut systems and 64-body direct-sum tiles of
//           the small systems, largest systems first
//   kick
// Systems below a size threshold use the direct sum.  The threshold and the
// persistent grid size are tuning parameters; nb_ensemble_autotune measures
// candidates on the device and keeps the fastest.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <numeric>
#include <string>
#include <vector>

#include "nb_internal.h"

using nb::EnsSystem;
using nb::ForestBufs;
using nb::TreeHdr;

enum { kEnsPhases = NB_PHASE_COUNT };

struct nb_ensemble {
    uint32_t nsys = 0;
    uint64_t nbody = 0;
    nb_params p{};
    int device = 0;
    hipStream_t stream = nullptr;
    std::vector<uint32_t> sizes;
    std::vector<EnsSystem> sys_host;
    nb_ensemble_tuning tune{};
    uint32_t nbh = 0;
    // device
    EnsSystem *sys = nullptr;
    double *x = nullptr, *y = nullptr, *z = nullptr, *vx = nullptr, *vy = nullptr, *vz = nullptr;
    double *vhx = nullptr, *vhy = nullptr, *vhz = nullptr, *ax = nullptr, *ay = nullptr,
           *az = nullptr, *m = nullptr, *dt_sys = nullptr;
    uint16_t *sys_of = nullptr;
    ForestBufs fb{};
    uint32_t *items = nullptr, *item_off = nullptr, *qorder = nullptr, *bh_order = nullptr;
    uint32_t *cursor = nullptr, *status = nullptr, *inter = nullptr;
    uint32_t *status_host = nullptr;
    // phase timing
    std::vector<hipEvent_t> events;
    std::vector<std::pair<int, size_t>> marks;
    double phase[kEnsPhases] = {0};
    std::vector<void *> allocations;
    std::string err;
};

namespace {

int efail(nb_ensemble *e, int code, const std::string &msg) {
    if (e) e->err = msg;
    return code;
}

#define ENS_HIP(e, call)                                                                   \
    do {                                                                                   \
        hipError_t e_ = (call);                                                            \
        if (e_ != hipSuccess)                                                              \
            return efail(e, NB_E_HIP, std::string(#call) + ": " + hipGetErrorString(e_)); \
    } while (0)

template <typename T>
int ealloc(nb_ensemble *e, T **ptr, uint64_t count) {
    void *q = nullptr;
    const uint64_t bytes = std::max<uint64_t>(count, 1) * sizeof(T);
    hipError_t r = hipMalloc(&q, bytes);
    if (r != hipSuccess)
        return efail(e, NB_E_HIP, "hipMalloc(" + std::to_string(bytes) + " B): " + hipGetErrorString(r));
    e->allocations.push_back(q);
    *ptr = static_cast<T *>(q);
    return NB_OK;
}

void mark(nb_ensemble *e, int phase, bool begin) {
    if (begin) {
        const size_t idx = e->marks.size() * 2;
        while (e->events.size() < idx + 2) {
            hipEvent_t ev;
            // timing only, read after a stream synchronisation: no system-scope fence
            if (hipEventCreateWithFlags(&ev, hipEventDisableSystemFence) != hipSuccess) return;
            e->events.push_back(ev);
        }
        e->marks.push_back({phase, idx});
        (void)hipEventRecord(e->events[idx], e->stream);
    } else {
        for (size_t k = e->marks.size(); k-- > 0;)
            if (e->marks[k].first == phase) {
                (void)hipEventRecord(e->events[e->marks[k].second + 1], e->stream);
                return;
            }
    }
}

int finish(nb_ensemble *e) {
    ENS_HIP(e, hipMemcpyAsync(e->status_host, e->status, sizeof(uint32_t), hipMemcpyDeviceToHost,
                              e->stream));
    ENS_HIP(e, hipStreamSynchronize(e->stream));
    for (auto &mk : e->marks) {
        float ms = 0.f;
        if (hipEventElapsedTime(&ms, e->events[mk.second], e->events[mk.second + 1]) == hipSuccess)
            e->phase[mk.first] += ms;
    }
    e->marks.clear();
    const uint32_t st = *e->status_host;
    if (!st) return NB_OK;
    // name the first system whose octree build raised a flag
    std::string where;
    std::vector<TreeHdr> hdr(e->nsys);
    if (hipMemcpy(hdr.data(), e->fb.hdr, e->nsys * sizeof(TreeHdr), hipMemcpyDeviceToHost) ==
        hipSuccess) {
        for (uint32_t s = 0; s < e->nsys; ++s)
            if (e->sys_host[s].bh && hdr[s].status) {
                where = " (system " + std::to_string(s) + ", n = " +
                        std::to_string(e->sys_host[s].n) + ", status " +
                        std::to_string(hdr[s].status) + ", deep runs " +
                        std::to_string(hdr[s].deep_runs) + ")";
                break;
            }
    }
    if (st & nb::kFlagAabb) return efail(e, NB_E_AABB, "bounding box: non-finite positions" + where);
    if (st & nb::kFlagDepth) return efail(e, NB_E_DEPTH, "coincident bodies in a system" + where);
    if (st & nb::kFlagNodeOverflow)
        return efail(e, NB_E_NODE_OVERFLOW, "octree storage exhausted" + where);
    if (st & nb::kFlagZeroMass) return efail(e, NB_E_ZERO_MASS, "zero-mass octree cell" + where);
    if (st & nb::kFlagStack) return efail(e, NB_E_STACK_OVERFLOW, "traversal stack exhausted" + where);
    if (st & nb::kFlagSortWait) return efail(e, NB_E_HIP, "radix sort: a look-back wait did not end");
    if (st & nb::kFlagSortSize)
        return efail(e, NB_E_INVALID_ARG, "radix sort: 2^30 body slots or more, nothing sorted");
    return efail(e, NB_E_HIP, "unknown device status " + std::to_string(st));
}

void timing_reset(nb_ensemble *e) {
    e->marks.clear();
    for (double &v : e->phase) v = 0;
}

// direct-sum or Barnes-Hut per system, queue order, direct-sum item counts
int apply_tuning(nb_ensemble *e) {
    const uint32_t S = e->nsys;
    std::vector<uint32_t> bh_order, qorder(S), items(S, 0);
    std::vector<double> cost(S);
    for (uint32_t s = 0; s < S; ++s) {
        EnsSystem &es = e->sys_host[s];
        es.bh = es.n >= e->tune.bh_threshold ? 1u : 0u;
        if (es.bh) bh_order.push_back(s);
        else items[s] = nb::ceil_div(es.n, e->tune.direct_tile);
        // rough cost model: pair evaluations
        cost[s] = es.bh ? 1500.0 * es.n : (double)es.n * (double)es.n;
    }
    std::stable_sort(bh_order.begin(), bh_order.end(),
                     [&](uint32_t a, uint32_t b) { return e->sys_host[a].n > e->sys_host[b].n; });
    std::iota(qorder.begin(), qorder.end(), 0u);
    std::stable_sort(qorder.begin(), qorder.end(), [&](uint32_t a, uint32_t b) { return cost[a] > cost[b]; });
    e->nbh = (uint32_t)bh_order.size();
    // the centre-of-mass launches: every block of every Barnes-Hut system,
    // the super-blocks of systems of two or more blocks, the systems of more
    // than one super-block
    std::vector<uint2> blk, span, big;
    for (uint32_t s : bh_order) {
        const uint32_t blocks = nb::ceil_div(e->sys_host[s].n, nb::kComBlock);
        for (uint32_t b = 0; b < blocks; ++b) blk.push_back(make_uint2(s, b));
        if (blocks >= 2)
            for (uint32_t b = 0; b < nb::ceil_div(blocks, nb::kComSuperBlocks); ++b)
                span.push_back(make_uint2(s, b));
        if (blocks > nb::kComSuperBlocks) big.push_back(make_uint2(s, 0u));
    }
    e->fb.nblk = (uint32_t)blk.size();
    e->fb.nspan = (uint32_t)span.size();
    e->fb.nbig = (uint32_t)big.size();
    if (!blk.empty())
        ENS_HIP(e, hipMemcpyAsync(e->fb.blk_tab, blk.data(), blk.size() * sizeof(uint2),
                                  hipMemcpyHostToDevice, e->stream));
    if (!span.empty())
        ENS_HIP(e, hipMemcpyAsync(e->fb.span_tab, span.data(), span.size() * sizeof(uint2),
                                  hipMemcpyHostToDevice, e->stream));
    if (!big.empty())
        ENS_HIP(e, hipMemcpyAsync(e->fb.big_tab, big.data(), big.size() * sizeof(uint2),
                                  hipMemcpyHostToDevice, e->stream));
    ENS_HIP(e, hipMemcpyAsync(e->sys, e->sys_host.data(), S * sizeof(EnsSystem),
                              hipMemcpyHostToDevice, e->stream));
    if (e->nbh)
        ENS_HIP(e, hipMemcpyAsync(e->bh_order, bh_order.data(), e->nbh * 4, hipMemcpyHostToDevice,
                                  e->stream));
    ENS_HIP(e, hipMemcpyAsync(e->qorder, qorder.data(), S * 4, hipMemcpyHostToDevice, e->stream));
    ENS_HIP(e, hipMemcpyAsync(e->items, items.data(), S * 4, hipMemcpyHostToDevice, e->stream));
    ENS_HIP(e, hipStreamSynchronize(e->stream));
    return NB_OK;
}

void enqueue_forces(nb_ensemble *e) {
    mark(e, NB_PHASE_OCTREE, true);
    nb::launch_forest_build(e->sys, e->bh_order, e->nbh, e->nsys, e->x, e->y, e->z, e->m, e->p.G,
                            e->fb, nb::GroupRule{e->tune.group_max, e->tune.pack_max}, e->items,
                            e->status, e->stream);
    mark(e, NB_PHASE_OCTREE, false);
    mark(e, NB_PHASE_ACCEL, true);
    nb::launch_ens_items(e->nsys, e->qorder, e->items, e->item_off, e->cursor, e->status,
                         e->stream);
    nb::EnsForceArgs a{e->sys,    e->qorder, e->item_off, e->nsys,     e->x,
                       e->y,      e->z,      e->m,        e->ax,       e->ay,
                       e->az,     e->fb,     e->p.G,      e->p.theta,  (float)e->p.eps2,
                       e->cursor, e->status, e->inter,    e->tune.direct_tile / nb::kEnsTile};
    nb::launch_ens_forces(a, e->tune.waves, e->stream);
    mark(e, NB_PHASE_ACCEL, false);
}

void enqueue_steps(nb_ensemble *e, uint32_t count) {
    const uint32_t n = (uint32_t)e->nbody;
    // the kick ending step k - 1 and the kick-drift starting step k run as one
    // pass (nothing reads v between them); the last step's kick writes v
    for (uint32_t k = 0; k < count; ++k) {
        mark(e, NB_PHASE_LF1, true);
        if (k == 0)
            nb::launch_ens_kick_drift(n, e->dt_sys, e->sys_of, e->ax, e->ay, e->az, e->vx, e->vy,
                                      e->vz, e->vhx, e->vhy, e->vhz, e->x, e->y, e->z, e->stream);
        else
            nb::launch_ens_kick_kick_drift(n, e->dt_sys, e->sys_of, e->ax, e->ay, e->az, e->vhx,
                                           e->vhy, e->vhz, e->x, e->y, e->z, e->stream);
        mark(e, NB_PHASE_LF1, false);
        enqueue_forces(e);
    }
    if (count > 0) {
        mark(e, NB_PHASE_LF2, true);
        nb::launch_ens_kick(n, e->dt_sys, e->sys_of, e->ax, e->ay, e->az, e->vhx, e->vhy, e->vhz,
                            e->vx, e->vy, e->vz, e->stream);
        mark(e, NB_PHASE_LF2, false);
    }
}

int reset_status(nb_ensemble *e) {
    ENS_HIP(e, hipMemsetAsync(e->status, 0, sizeof(uint32_t), e->stream));
    return NB_OK;
}

}  // namespace

extern "C" {

int nb_ensemble_create(nb_ensemble **out, uint32_t nsys, const uint32_t *sizes,
                       const nb_params *p) {
    if (!out || nsys == 0 || nsys > 65535 || !sizes || !p || p->precision != NB_PREC_FP32C ||
        !(p->theta >= 0) || !(p->eps2 >= 0) || p->storage_size_param < 1)
        return NB_E_INVALID_ARG;
    *out = nullptr;
    uint64_t total = 0;
    for (uint32_t s = 0; s < nsys; ++s) {
        if (sizes[s] == 0) return NB_E_INVALID_ARG;
        total += sizes[s];
    }
    if (total >= (1ull << 32)) return NB_E_INVALID_ARG;
    int count = 0;
    if (hipGetDeviceCount(&count) != hipSuccess || count == 0) return NB_E_NO_DEVICE;
    nb_ensemble *e = new nb_ensemble();
    auto bail = [&](int code) {
        nb_ensemble_destroy(e);
        return code;
    };
    e->nsys = nsys;
    e->nbody = total;
    e->p = *p;
    e->sizes.assign(sizes, sizes + nsys);
    if (hipGetDevice(&e->device) != hipSuccess) return bail(NB_E_HIP);
    if (hipStreamCreateWithFlags(&e->stream, hipStreamNonBlocking) != hipSuccess) return bail(NB_E_HIP);
    // storage layout
    e->sys_host.resize(nsys);
    auto up = [](uint64_t v) { return (v + nb::kEnsAlign - 1) / nb::kEnsAlign * nb::kEnsAlign; };
    uint64_t bo = 0, fo = 0, io = 0, no = 0, span = 0, xo = 0, nblk = 0, nsup = 0;
    for (uint32_t s = 0; s < nsys; ++s) {
        EnsSystem &es = e->sys_host[s];
        es.body_off = (uint32_t)bo;
        es.n = sizes[s];
        // checked in 64 bits before the narrowing
        const uint64_t icap = std::max<uint64_t>(1, (uint64_t)p->storage_size_param * sizes[s] / 8);
        if (icap >= (1u << 26)) return bail(NB_E_INVALID_ARG);
        es.icap = (uint32_t)icap;
        es.fb_off = (uint32_t)fo;
        es.int_off = (uint32_t)io;
        es.node_off = no;
        es.bh = 0;
        es.span_off = (uint32_t)span;
        es.cross_off = (uint32_t)xo;
        const uint64_t blocks = nb::ceil_div(sizes[s], nb::kComBlock);
        nblk += blocks;
        nsup += nb::ceil_div(blocks, nb::kComSuperBlocks);
        span += (uint64_t)nb::kMaxDepth * blocks;
        xo += (uint64_t)nb::kMaxDepth * (nb::ceil_div(sizes[s], nb::kComSuper) + 1);
        bo += sizes[s];
        fo = up(fo + sizes[s]);
        io = up(io + es.icap);
        no = up(no + 1 + 8ull * es.icap);
    }
    // the forest-wide radix sorts take fewer than 2^30 body slots
    if (io >= (1ull << 32) || fo >= (1ull << 30) || span >= (1ull << 32) || xo >= (1ull << 32))
        return bail(NB_E_INVALID_ARG);
    int rc;
    double **st[] = {&e->x, &e->y, &e->z, &e->vx, &e->vy, &e->vz, &e->vhx,
                     &e->vhy, &e->vhz, &e->ax, &e->ay, &e->az, &e->m};
    for (double **a : st)
        if ((rc = ealloc(e, a, total))) return bail(rc);
    ForestBufs &f = e->fb;
    if ((rc = ealloc(e, &e->sys, nsys)) || (rc = ealloc(e, &e->dt_sys, nsys)) ||
        (rc = ealloc(e, &e->sys_of, total)) || (rc = ealloc(e, &f.hdr, nsys)) ||
        (rc = ealloc(e, &f.skey, fo)) || (rc = ealloc(e, &f.skey_alt, fo)) ||
        (rc = ealloc(e, &f.klo, fo)) || (rc = ealloc(e, &f.sidx, fo)) ||
        (rc = ealloc(e, &f.sidx_alt, fo)) || (rc = ealloc(e, &f.chain, fo)) ||
        (rc = ealloc(e, &f.body_parent, fo)) || (rc = ealloc(e, &f.group_size, fo)) ||
        (rc = ealloc(e, &f.group_slot, fo)) || (rc = ealloc(e, &f.lcp, fo)) ||
        (rc = ealloc(e, &f.groups, fo)) || (rc = ealloc(e, &f.body4, fo)) ||
        (rc = ealloc(e, &f.int_depth, io)) ||
        (rc = ealloc(e, &f.int_local, io)) || (rc = ealloc(e, &f.int_parent, io)) ||
        (rc = ealloc(e, &f.int_node, io)) || (rc = ealloc(e, &f.int_lvl, io)) ||
        (rc = ealloc(e, &f.int_count, io)) ||
        (rc = ealloc(e, &f.int_first, io)) || (rc = ealloc(e, &f.int_sums, io)) ||
        (rc = ealloc(e, &f.ref, no)) || (rc = ealloc(e, &f.rec, no)) ||
        (rc = ealloc(e, &e->items, nsys)) || (rc = ealloc(e, &e->item_off, nsys + 1ull)) ||
        (rc = ealloc(e, &e->qorder, nsys)) || (rc = ealloc(e, &e->bh_order, nsys)) ||
        (rc = ealloc(e, &e->cursor, 1)) || (rc = ealloc(e, &e->status, 1)) ||
        (rc = ealloc(e, &e->inter, total)) || (rc = ealloc(e, &f.fb_sys, fo)) || (rc = ealloc(e, &f.spanning, span)) || (rc = ealloc(e, &f.cross, xo)) || (rc = ealloc(e, &f.blk_tab, nblk)) || (rc = ealloc(e, &f.span_tab, nsup)) || (rc = ealloc(e, &f.big_tab, nsys)) || (rc = ealloc(e, &f.sys_first, nsys)) ||
        (rc = ealloc(e, &f.scan_part, nb::scan_scratch_words((uint32_t)fo) + 1ull)) ||
        (rc = ealloc(e, &f.radix_ctl, nb::radix_ctl_words((uint32_t)fo))) ||
        (rc = ealloc(e, &f.radix_ctl_sys, nb::radix_ctl_words((uint32_t)fo))))
        return bail(rc);
    f.slots = (uint32_t)fo;
    f.scan_cap = nb::scan_scratch_words((uint32_t)fo) + 1;
    if (hipHostMalloc((void **)&e->status_host, sizeof(uint32_t), hipHostMallocDefault) != hipSuccess)
        return bail(NB_E_HIP);
    *e->status_host = 0;
    // body -> system map
    std::vector<uint16_t> so(total);
    for (uint32_t s = 0; s < nsys; ++s)
        std::fill(so.begin() + e->sys_host[s].body_off,
                  so.begin() + e->sys_host[s].body_off + sizes[s], (uint16_t)s);
    if (hipMemcpy(e->sys_of, so.data(), total * sizeof(uint16_t), hipMemcpyHostToDevice) != hipSuccess)
        return bail(NB_E_HIP);
    // slot -> system over the forest's slots; padding belongs to the system before it
    std::vector<uint16_t> fs(fo);
    for (uint32_t s = 0; s < nsys; ++s) {
        const uint64_t end = s + 1 < nsys ? e->sys_host[s + 1].fb_off : fo;
        std::fill(fs.begin() + e->sys_host[s].fb_off, fs.begin() + end, (uint16_t)s);
    }
    std::vector<uint32_t> first(nsys);
    for (uint32_t s = 0; s < nsys; ++s) first[s] = e->sys_host[s].fb_off;
    if (hipMemcpy(f.fb_sys, fs.data(), fo * sizeof(uint16_t), hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(f.sys_first, first.data(), nsys * sizeof(uint32_t), hipMemcpyHostToDevice) !=
            hipSuccess ||
        hipMemset(f.radix_ctl, 0, (size_t)nb::radix_ctl_words((uint32_t)fo) * 4) != hipSuccess ||
        hipMemset(f.radix_ctl_sys, 0, (size_t)nb::radix_ctl_words((uint32_t)fo) * 4) != hipSuccess)
        return bail(NB_E_HIP);
    for (double **a : st)
        if (hipMemsetAsync(*a, 0, total * sizeof(double), e->stream) != hipSuccess) return bail(NB_E_HIP);
    if (hipMemsetAsync(f.hdr, 0, nsys * sizeof(TreeHdr), e->stream) != hipSuccess) return bail(NB_E_HIP);
    e->tune.bh_threshold = 4096;
    e->tune.waves = 8192;
    e->tune.group_max = nb::kGroupMax;
    e->tune.pack_max = nb::kPackMax;
    e->tune.direct_tile = nb::kEnsTile;
    if ((rc = apply_tuning(e))) return bail(rc);
    *out = e;
    return NB_OK;
}

int nb_ensemble_destroy(nb_ensemble *e) {
    if (!e) return NB_OK;
    if (e->stream) (void)hipStreamSynchronize(e->stream);
    for (void *q : e->allocations) (void)hipFree(q);
    for (hipEvent_t ev : e->events) (void)hipEventDestroy(ev);
    if (e->status_host) (void)hipHostFree(e->status_host);
    if (e->stream) (void)hipStreamDestroy(e->stream);
    delete e;
    return NB_OK;
}

const char *nb_ensemble_last_error(const nb_ensemble *e) { return e ? e->err.c_str() : "null ensemble"; }

uint64_t nb_ensemble_body_count(const nb_ensemble *e) { return e ? e->nbody : 0; }

int nb_ensemble_set_state(nb_ensemble *e, const double *x, const double *y, const double *z,
                          const double *vx, const double *vy, const double *vz, const double *m,
                          const double *dt) {
    if (!e || !x || !y || !z || !vx || !vy || !vz || !m || !dt)
        return efail(e, NB_E_INVALID_ARG, "nb_ensemble_set_state: null pointer");
    const size_t bytes = e->nbody * sizeof(double);
    const double *src[] = {x, y, z, vx, vy, vz, m};
    double *dst[] = {e->x, e->y, e->z, e->vx, e->vy, e->vz, e->m};
    for (int k = 0; k < 7; ++k)
        ENS_HIP(e, hipMemcpyAsync(dst[k], src[k], bytes, hipMemcpyHostToDevice, e->stream));
    ENS_HIP(e, hipMemcpyAsync(e->dt_sys, dt, e->nsys * sizeof(double), hipMemcpyHostToDevice,
                              e->stream));
    ENS_HIP(e, hipMemsetAsync(e->ax, 0, bytes, e->stream));
    ENS_HIP(e, hipMemsetAsync(e->ay, 0, bytes, e->stream));
    ENS_HIP(e, hipMemsetAsync(e->az, 0, bytes, e->stream));
    ENS_HIP(e, hipStreamSynchronize(e->stream));
    return NB_OK;
}

int nb_ensemble_get_state(nb_ensemble *e, double *x, double *y, double *z, double *vx, double *vy,
                          double *vz) {
    if (!e) return NB_E_INVALID_ARG;
    const size_t bytes = e->nbody * sizeof(double);
    double *dst[] = {x, y, z, vx, vy, vz};
    const double *src[] = {e->x, e->y, e->z, e->vx, e->vy, e->vz};
    for (int k = 0; k < 6; ++k)
        if (dst[k]) ENS_HIP(e, hipMemcpyAsync(dst[k], src[k], bytes, hipMemcpyDeviceToHost, e->stream));
    ENS_HIP(e, hipStreamSynchronize(e->stream));
    return NB_OK;
}

int nb_ensemble_get_accel(nb_ensemble *e, double *ax, double *ay, double *az) {
    if (!e) return NB_E_INVALID_ARG;
    const size_t bytes = e->nbody * sizeof(double);
    double *dst[] = {ax, ay, az};
    const double *src[] = {e->ax, e->ay, e->az};
    for (int k = 0; k < 3; ++k)
        if (dst[k]) ENS_HIP(e, hipMemcpyAsync(dst[k], src[k], bytes, hipMemcpyDeviceToHost, e->stream));
    ENS_HIP(e, hipStreamSynchronize(e->stream));
    return NB_OK;
}

int nb_ensemble_forces(nb_ensemble *e) {
    if (!e) return NB_E_INVALID_ARG;
    timing_reset(e);
    int rc = reset_status(e);
    if (rc) return rc;
    enqueue_forces(e);
    ENS_HIP(e, hipGetLastError());
    return finish(e);
}

int nb_ensemble_steps(nb_ensemble *e, uint32_t count) {
    if (!e) return NB_E_INVALID_ARG;
    timing_reset(e);
    int rc = reset_status(e);
    if (rc) return rc;
    enqueue_steps(e, count);
    ENS_HIP(e, hipGetLastError());
    return finish(e);
}

uint64_t nb_ensemble_tree_canonical(nb_ensemble *e, uint32_t s, uint64_t cap, uint8_t *depth,
                                    uint8_t *kind, uint32_t *body, double *mass, double *com_x,
                                    double *com_y, double *com_z, uint32_t *count) {
    if (!e || s >= e->nsys || !e->sys_host[s].bh || !depth || !kind || !body || !mass ||
        !com_x || !com_y || !com_z || !count)
        return 0;
    const EnsSystem &S = e->sys_host[s];
    TreeHdr h;
    if (hipStreamSynchronize(e->stream) != hipSuccess ||
        hipMemcpy(&h, e->fb.hdr + s, sizeof(TreeHdr), hipMemcpyDeviceToHost) != hipSuccess)
        return 0;
    const uint32_t I = std::min(h.internal_count, S.icap);
    const uint64_t nodes = 1 + 8ull * I;
    if (cap < nodes) return 0;
    std::vector<int32_t> ref(nodes);
    std::vector<nb::InternalSums> sums(I);
    std::vector<uint32_t> cnt(I);
    std::vector<double> hx(S.n), hy(S.n), hz(S.n), hm(S.n);
    bool ok = hipMemcpy(ref.data(), e->fb.ref + S.node_off, nodes * 4, hipMemcpyDeviceToHost) ==
              hipSuccess;
    if (I) {
        ok = ok && hipMemcpy(sums.data(), e->fb.int_sums + S.int_off,
                             (size_t)I * sizeof(nb::InternalSums), hipMemcpyDeviceToHost) ==
                       hipSuccess;
        ok = ok && hipMemcpy(cnt.data(), e->fb.int_count + S.int_off, (size_t)I * 4,
                             hipMemcpyDeviceToHost) == hipSuccess;
    }
    const double *src[] = {e->x, e->y, e->z, e->m};
    double *dst[] = {hx.data(), hy.data(), hz.data(), hm.data()};
    for (int k = 0; k < 4; ++k)
        ok = ok && hipMemcpy(dst[k], src[k] + S.body_off, (size_t)S.n * 8, hipMemcpyDeviceToHost) ==
                       hipSuccess;
    if (!ok) return 0;
    return nb::canonical_dump(ref.data(), sums.data(), cnt.data(), hx.data(), hy.data(), hz.data(),
                              hm.data(), S.n, cap, depth, kind, body, mass, com_x, com_y, com_z,
                              count);
}

uint64_t nb_ensemble_node_count(nb_ensemble *e, uint32_t s) {
    if (!e || s >= e->nsys || !e->sys_host[s].bh) return 0;
    TreeHdr h;
    if (hipStreamSynchronize(e->stream) != hipSuccess ||
        hipMemcpy(&h, e->fb.hdr + s, sizeof(TreeHdr), hipMemcpyDeviceToHost) != hipSuccess)
        return 0;
    return 1 + 8ull * std::min(h.internal_count, e->sys_host[s].icap);
}

int nb_ensemble_interactions(nb_ensemble *e, uint32_t *count) {
    if (!e || !count) return NB_E_INVALID_ARG;
    ENS_HIP(e, hipMemcpyAsync(count, e->inter, e->nbody * 4, hipMemcpyDeviceToHost, e->stream));
    ENS_HIP(e, hipStreamSynchronize(e->stream));
    return NB_OK;
}

int nb_ensemble_phase_ms(const nb_ensemble *e, double *ms) {
    if (!e || !ms) return NB_E_INVALID_ARG;
    for (int k = 0; k < kEnsPhases; ++k) ms[k] = e->phase[k];
    return NB_OK;
}

bool tuning_ok(const nb_ensemble_tuning *t) {
    return t && t->waves >= 1 && t->waves <= (1u << 20) && t->group_max >= 1 &&
           t->group_max <= nb::kGroupLimit && t->pack_max >= t->group_max &&
           (t->direct_tile == nb::kEnsTile || t->direct_tile == 2 * nb::kEnsTile);
}

int nb_ensemble_set_tuning(nb_ensemble *e, const nb_ensemble_tuning *t) {
    if (!e || !tuning_ok(t))
        return efail(e, NB_E_INVALID_ARG, "nb_ensemble_set_tuning: bad tuning");
    e->tune = *t;
    return apply_tuning(e);
}

int nb_ensemble_get_tuning(const nb_ensemble *e, nb_ensemble_tuning *t) {
    if (!e || !t) return NB_E_INVALID_ARG;
    *t = e->tune;
    return NB_OK;
}

int nb_ensemble_uses_bh(const nb_ensemble *e, uint8_t *flags) {
    if (!e || !flags) return NB_E_INVALID_ARG;
    for (uint32_t s = 0; s < e->nsys; ++s) flags[s] = (uint8_t)e->sys_host[s].bh;
    return NB_OK;
}

int nb_ensemble_autotune(nb_ensemble *e, uint32_t steps, nb_ensemble_tuning *best,
                         double *best_step_ms) {
    if (!e || steps == 0) return NB_E_INVALID_ARG;
    // save the state: every candidate starts from it
    const uint64_t n = e->nbody;
    std::vector<double *> saved(10, nullptr);
    double *live[] = {e->x, e->y, e->z, e->vx, e->vy, e->vz, e->ax, e->ay, e->az, e->vhx};
    struct Free {
        std::vector<double *> &v;
        ~Free() {
            for (double *q : v) (void)hipFree(q);
        }
    } guard{saved};
    for (int k = 0; k < 10; ++k) {
        ENS_HIP(e, hipMalloc((void **)&saved[k], n * sizeof(double)));
        ENS_HIP(e, hipMemcpyAsync(saved[k], live[k], n * sizeof(double), hipMemcpyDeviceToDevice,
                                  e->stream));
    }
    auto restore = [&]() -> int {
        for (int k = 0; k < 10; ++k)
            ENS_HIP(e, hipMemcpyAsync(live[k], saved[k], n * sizeof(double),
                                      hipMemcpyDeviceToDevice, e->stream));
        return NB_OK;
    };
    auto trial = [&](const nb_ensemble_tuning &t, double *ms) -> int {
        int rc;
        e->tune = t;
        if ((rc = apply_tuning(e)) || (rc = restore())) return rc;
        timing_reset(e);
        if ((rc = reset_status(e))) return rc;
        enqueue_forces(e);  // warm-up: t = 0 forces of this configuration
        if ((rc = finish(e))) return rc;
        hipEvent_t a, b;
        ENS_HIP(e, hipEventCreateWithFlags(&a, hipEventDisableSystemFence));
        ENS_HIP(e, hipEventCreateWithFlags(&b, hipEventDisableSystemFence));
        ENS_HIP(e, hipEventRecord(a, e->stream));
        enqueue_steps(e, steps);
        ENS_HIP(e, hipEventRecord(b, e->stream));
        rc = finish(e);
        float f = 0.f;
        if (!rc) ENS_HIP(e, hipEventElapsedTime(&f, a, b));
        (void)hipEventDestroy(a);
        (void)hipEventDestroy(b);
        *ms = f / steps;
        return rc;
    };
    // candidate thresholds: powers of two and their 1.5 multiples within the sizes present
    const uint32_t nmax = *std::max_element(e->sizes.begin(), e->sizes.end());
    std::vector<uint32_t> thr{0};
    for (uint32_t t = 512; t <= std::min<uint32_t>(nmax, 16384); t *= 2) {
        thr.push_back(t);
        if (t * 3 / 2 <= std::min<uint32_t>(nmax, 16384)) thr.push_back(t * 3 / 2);
    }
    nb_ensemble_tuning top{4096, 8192, nb::kGroupMax, nb::kPackMax, nb::kEnsTile};
    double top_ms = 1e300;
    // one parameter at a time, each candidate from the best so far
    auto sweep = [&](auto set, std::initializer_list<uint32_t> values) -> int {
        const nb_ensemble_tuning base = top;
        for (uint32_t v : values) {
            nb_ensemble_tuning cur = base;
            set(cur, v);
            if (!tuning_ok(&cur)) continue;
            double ms;
            int rc = trial(cur, &ms);
            if (rc) return rc;
            if (ms < top_ms) {
                top_ms = ms;
                top = cur;
            }
        }
        return NB_OK;
    };
    int rc = NB_OK;
    {
        const nb_ensemble_tuning base = top;
        for (uint32_t t : thr) {
            nb_ensemble_tuning cur = base;
            cur.bh_threshold = t;
            double ms;
            if ((rc = trial(cur, &ms))) return rc;
            if (ms < top_ms) {
                top_ms = ms;
                top = cur;
            }
        }
    }
    // traversal group size, with the packing limit at 4 groups
    if ((rc = sweep([](nb_ensemble_tuning &c, uint32_t v) { c.group_max = v; c.pack_max = 4 * v; },
                    {16u, 24u, 32u, 48u})))
        return rc;
    if ((rc = sweep([](nb_ensemble_tuning &c, uint32_t v) { c.pack_max = v * c.group_max; },
                    {1u, 2u, 8u})))
        return rc;
    if ((rc = sweep([](nb_ensemble_tuning &c, uint32_t v) { c.direct_tile = v; },
                    {2u * nb::kEnsTile})))
        return rc;
    if ((rc = sweep([](nb_ensemble_tuning &c, uint32_t v) { c.waves = v; },
                    {1024u, 2048u, 4096u, 16384u, 32768u})))
        return rc;
    e->tune = top;
    if ((rc = apply_tuning(e))) return rc;
    if ((rc = restore())) return rc;
    ENS_HIP(e, hipStreamSynchronize(e->stream));
    timing_reset(e);
    if (best) *best = top;
    if (best_step_ms) *best_step_ms = top_ms;
    return NB_OK;
}

}  // extern "C"

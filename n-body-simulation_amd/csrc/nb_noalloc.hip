// nb_noalloc.hip -- the direct sum for callers that cannot allocate device
// memory (code under HIP graph capture, serving loops, harnesses that forbid
// hipMalloc).  No counterpart in the reference.
//
// nb_forces_direct_noalloc runs the split direct-sum kernels of
// nb_forces_direct on device arrays the caller owns, with the partial sums in
// `scratch`, a file-scope device array of fixed size.  Each
// device holds its own copy of the array at its own address: the host looks
// that address up once per device with hipGetSymbolAddress and passes it to
// the kernels (the array's name in host code is not a device address).  All
// calls on one device share the array, so calls that overlap on one device
// race.
#include <cmath>
#include <mutex>

#include "nb_internal.h"

namespace {

constexpr uint64_t kScratchBytes = NB_NOALLOC_SCRATCH_BYTES;
static_assert(kScratchBytes % 256 == 0, "whole 256-byte lines");

__device__ __align__(256) uint8_t scratch[kScratchBytes];

constexpr int kMaxDevices = 64;
std::mutex g_lock;
void *g_scratch[kMaxDevices] = {};

// this device's address of `scratch`, looked up on the first call on it
int scratch_for_current_device(void **out) {
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) return NB_E_NO_DEVICE;
    if (dev < 0 || dev >= kMaxDevices) return NB_E_INVALID_ARG;
    std::lock_guard<std::mutex> g(g_lock);
    if (!g_scratch[dev]) {
        void *ptr = nullptr;
        if (hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)) != hipSuccess || !ptr) return NB_E_HIP;
        g_scratch[dev] = ptr;
    }
    *out = g_scratch[dev];
    return NB_OK;
}

}  // namespace

extern "C" {

int nb_forces_direct_noalloc(uint32_t n, const double *x, const double *y, const double *z,
                             const double *m, double G, double eps2, uint32_t precision,
                             double *ax, double *ay, double *az, void *stream) {
    if (n == 0 || !x || !y || !z || !m || !ax || !ay || !az) return NB_E_INVALID_ARG;
    if (precision != NB_PREC_FP32C && precision != NB_PREC_FP64_FAST) return NB_E_INVALID_ARG;
    if (!(eps2 >= 0) || !std::isfinite(G)) return NB_E_INVALID_ARG;
    const uint64_t need = nb::direct_split_scratch_bytes(n);
    if (n > NB_NOALLOC_MAX_BODIES || need > kScratchBytes) return NB_E_INVALID_ARG;
    void *buf = nullptr;
    int rc = scratch_for_current_device(&buf);
    if (rc) return rc;
    hipStream_t s = static_cast<hipStream_t>(stream);
    // every partial sum is written before it is read in the same call, so
    // nothing in the array needs clearing
    nb::launch_direct_split(precision == NB_PREC_FP64_FAST, n, x, y, z, m, G, eps2, ax, ay, az,
                            static_cast<double *>(buf), s);
    if (hipGetLastError() != hipSuccess) return NB_E_HIP;
    return NB_OK;
}

}  // extern "C"

// nb_integrate.hip -- FP64 kick-drift-kick leapfrog.
//
// Same operations as NaiveAlgorithm.cpp:154-163 and :217-221 (identical in
// BarnesHutAlgorithm.cpp:171-181, :234-238); compiled with -ffp-contract=off
// so `v + a*(dt/2)` stays a rounded multiply followed by a rounded add.
// HBM-bound: 120 bytes of algorithmic traffic per body-step (read x, v, a;
// write x, v).  Run as written, the two parts move 192 bytes (v-half is
// written by the first and read back by the second); inside a run of steps
// with nobody reading v in between, k_kick_kick_drift does the kick of step k
// and the kick-drift of step k+1 in one pass over 120 bytes, with the same
// roundings in the same order.
#include "nb_internal.h"

namespace nb {
namespace {

constexpr int kThreads = 256;

// dt_sys != nullptr: an ensemble, body i advances by dt_sys[sys_of[i]]
__global__ void __launch_bounds__(kThreads) k_kick_drift(
    uint32_t n, double dt, const double *__restrict__ dt_sys, const uint16_t *__restrict__ sys_of,
    const double *__restrict__ ax, const double *__restrict__ ay,
    const double *__restrict__ az, const double *__restrict__ vx, const double *__restrict__ vy,
    const double *__restrict__ vz, double *__restrict__ vhx, double *__restrict__ vhy,
    double *__restrict__ vhz, double *__restrict__ x, double *__restrict__ y,
    double *__restrict__ z) {
    const uint32_t i = blockIdx.x * kThreads + threadIdx.x;
    if (i >= n) return;
    const double delta_t = dt_sys ? dt_sys[sys_of[i]] : dt;
    const double hx = vx[i] + ax[i] * (delta_t / 2.0);
    const double hy = vy[i] + ay[i] * (delta_t / 2.0);
    const double hz = vz[i] + az[i] * (delta_t / 2.0);
    vhx[i] = hx;
    vhy[i] = hy;
    vhz[i] = hz;
    x[i] = x[i] + hx * delta_t;
    y[i] = y[i] + hy * delta_t;
    z[i] = z[i] + hz * delta_t;
}

__global__ void __launch_bounds__(kThreads) k_kick(
    uint32_t n, double dt, const double *__restrict__ dt_sys, const uint16_t *__restrict__ sys_of,
    const double *__restrict__ ax, const double *__restrict__ ay,
    const double *__restrict__ az, const double *__restrict__ vhx,
    const double *__restrict__ vhy, const double *__restrict__ vhz, double *__restrict__ vx,
    double *__restrict__ vy, double *__restrict__ vz) {
    const uint32_t i = blockIdx.x * kThreads + threadIdx.x;
    if (i >= n) return;
    const double delta_t = dt_sys ? dt_sys[sys_of[i]] : dt;
    vx[i] = vhx[i] + ax[i] * (delta_t / 2.0);
    vy[i] = vhy[i] + ay[i] * (delta_t / 2.0);
    vz[i] = vhz[i] + az[i] * (delta_t / 2.0);
}

// The kick of one step and the kick-drift of the next, without storing the
// full-step velocity in between: v = vh + a*(dt/2), vh' = v + a*(dt/2),
// x' = x + vh'*dt, each a rounded multiply and a rounded add as in k_kick
// followed by k_kick_drift, so the results are bit-identical to the two
// kernels.  The caller runs k_kick after the last step to produce v.
__global__ void __launch_bounds__(kThreads) k_kick_kick_drift(
    uint32_t n, double dt, const double *__restrict__ dt_sys, const uint16_t *__restrict__ sys_of,
    const double *__restrict__ ax, const double *__restrict__ ay,
    const double *__restrict__ az, double *__restrict__ vhx, double *__restrict__ vhy,
    double *__restrict__ vhz, double *__restrict__ x, double *__restrict__ y,
    double *__restrict__ z) {
    const uint32_t i = blockIdx.x * kThreads + threadIdx.x;
    if (i >= n) return;
    const double delta_t = dt_sys ? dt_sys[sys_of[i]] : dt;
    const double kx = ax[i] * (delta_t / 2.0);
    const double ky = ay[i] * (delta_t / 2.0);
    const double kz = az[i] * (delta_t / 2.0);
    const double hx = (vhx[i] + kx) + kx;
    const double hy = (vhy[i] + ky) + ky;
    const double hz = (vhz[i] + kz) + kz;
    vhx[i] = hx;
    vhy[i] = hy;
    vhz[i] = hz;
    x[i] = x[i] + hx * delta_t;
    y[i] = y[i] + hy * delta_t;
    z[i] = z[i] + hz * delta_t;
}

}  // namespace

void launch_kick_kick_drift(uint32_t n, double dt, const double *ax, const double *ay,
                            const double *az, double *vhx, double *vhy, double *vhz, double *x,
                            double *y, double *z, hipStream_t s) {
    hipLaunchKernelGGL(k_kick_kick_drift, dim3(ceil_div(n, kThreads)), dim3(kThreads), 0, s, n,
                       dt, (const double *)nullptr, (const uint16_t *)nullptr, ax, ay, az, vhx,
                       vhy, vhz, x, y, z);
}

void launch_ens_kick_kick_drift(uint32_t n, const double *dt_sys, const uint16_t *sys_of,
                                const double *ax, const double *ay, const double *az, double *vhx,
                                double *vhy, double *vhz, double *x, double *y, double *z,
                                hipStream_t s) {
    hipLaunchKernelGGL(k_kick_kick_drift, dim3(ceil_div(n, kThreads)), dim3(kThreads), 0, s, n,
                       0.0, dt_sys, sys_of, ax, ay, az, vhx, vhy, vhz, x, y, z);
}

void launch_kick_drift(uint32_t n, double dt, const double *ax, const double *ay,
                       const double *az, const double *vx, const double *vy, const double *vz,
                       double *vhx, double *vhy, double *vhz, double *x, double *y, double *z,
                       hipStream_t s) {
    hipLaunchKernelGGL(k_kick_drift, dim3(ceil_div(n, kThreads)), dim3(kThreads), 0, s, n, dt,
                       (const double *)nullptr, (const uint16_t *)nullptr, ax, ay, az, vx, vy, vz,
                       vhx, vhy, vhz, x, y, z);
}

void launch_kick(uint32_t n, double dt, const double *ax, const double *ay, const double *az,
                 const double *vhx, const double *vhy, const double *vhz, double *vx,
                 double *vy, double *vz, hipStream_t s) {
    hipLaunchKernelGGL(k_kick, dim3(ceil_div(n, kThreads)), dim3(kThreads), 0, s, n, dt,
                       (const double *)nullptr, (const uint16_t *)nullptr, ax, ay, az, vhx, vhy,
                       vhz, vx, vy, vz);
}

void launch_ens_kick_drift(uint32_t n, const double *dt_sys, const uint16_t *sys_of,
                           const double *ax, const double *ay, const double *az, const double *vx,
                           const double *vy, const double *vz, double *vhx, double *vhy,
                           double *vhz, double *x, double *y, double *z, hipStream_t s) {
    hipLaunchKernelGGL(k_kick_drift, dim3(ceil_div(n, kThreads)), dim3(kThreads), 0, s, n, 0.0,
                       dt_sys, sys_of, ax, ay, az, vx, vy, vz, vhx, vhy, vhz, x, y, z);
}

void launch_ens_kick(uint32_t n, const double *dt_sys, const uint16_t *sys_of, const double *ax,
                     const double *ay, const double *az, const double *vhx, const double *vhy,
                     const double *vhz, double *vx, double *vy, double *vz, hipStream_t s) {
    hipLaunchKernelGGL(k_kick, dim3(ceil_div(n, kThreads)), dim3(kThreads), 0, s, n, 0.0, dt_sys,
                       sys_of, ax, ay, az, vhx, vhy, vhz, vx, vy, vz);
}

}  // namespace nb

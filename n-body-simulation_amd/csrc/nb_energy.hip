// nb_energy.hip -- per-body energy terms for the --energy diagnostics.
//
// Same arithmetic as nBodyAlgorithm.cpp:48-64 (FP64, no softening, i < j
// pairs in index order; compiled with -ffp-contract=off): body j gets
//   E_kin[j] = 0.5 * m_j * (vx^2 + vy^2 + vz^2)
//   E_pot[j] = sum_{i<j} G * m_i * m_j / |p_j - p_i|
// The host adds the terms in body order (nBodyAlgorithm.cpp:71-74).
// O(N^2/2) pair work staged through LDS tiles of 64 bodies.
#include "nb_internal.h"

namespace nb {
namespace {

constexpr int kTile = 64;

__global__ void __launch_bounds__(kTile) k_energy_terms(
    uint32_t n, const double *__restrict__ x, const double *__restrict__ y,
    const double *__restrict__ z, const double *__restrict__ vx, const double *__restrict__ vy,
    const double *__restrict__ vz, const double *__restrict__ m, double G, double *ekin,
    double *epot) {
    __shared__ double sx[kTile], sy[kTile], sz[kTile], sm[kTile];
    const uint32_t j = blockIdx.x * kTile + threadIdx.x;
    const bool valid = j < n;
    const uint32_t jj = valid ? j : 0;
    const double pxj = x[jj], pyj = y[jj], pzj = z[jj], mj = m[jj];
    if (valid) {
        const double v = vx[j] * vx[j] + vy[j] * vy[j] + vz[j] * vz[j];
        ekin[j] = 0.5 * mj * v;
    }
    double ep = 0;
    const uint32_t last = min(n, (blockIdx.x + 1) * kTile);  // bodies i < j of this block
    for (uint32_t k = 0; k < last; k += kTile) {
        const uint32_t i = k + threadIdx.x;
        if (i < n) {
            sx[threadIdx.x] = x[i];
            sy[threadIdx.x] = y[i];
            sz[threadIdx.x] = z[i];
            sm[threadIdx.x] = m[i];
        }
        __syncthreads();
        if (valid) {
            for (uint32_t ii = 0; ii < (uint32_t)kTile && k + ii < j; ++ii) {
                const double r_x = pxj - sx[ii];
                const double r_y = pyj - sy[ii];
                const double r_z = pzj - sz[ii];
                const double r = sqrt(r_x * r_x + r_y * r_y + r_z * r_z);
                ep += G * sm[ii] * mj / r;
            }
        }
        __syncthreads();
    }
    if (valid) epot[j] = ep;
}

}  // namespace

void launch_energy_terms(uint32_t n, const double *x, const double *y, const double *z,
                         const double *vx, const double *vy, const double *vz, const double *m,
                         double G, double *ekin, double *epot, hipStream_t s) {
    hipLaunchKernelGGL(k_energy_terms, dim3(ceil_div(n, kTile)), dim3(kTile), 0, s, n, x, y, z, vx,
                       vy, vz, m, G, ekin, epot);
}

}  // namespace nb

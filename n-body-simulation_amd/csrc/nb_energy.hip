// nb_energy.hip -- per-body energy terms for the --energy diagnostics.
//
// Same arithmetic as nBodyAlgorithm.cpp:48-64 (FP64, no softening, i < j
// pairs in index order; compiled with -ffp-contract=off): body j gets
//   E_kin[j] = 0.5 * m_j * (vx^2 + vy^2 + vz^2)
//   E_pot[j] = sum_{i<j} G * m_i * m_j / |p_j - p_i|
// The host adds the terms in body order (nBodyAlgorithm.cpp:71-74).
// O(N^2/2) pair work staged through LDS tiles of 256 bodies; every body's sum
// runs over i = 0, 1, ..., j - 1 in order, so the terms are bit-identical to
// the reference's.
#include "nb_internal.h"

namespace nb {
namespace {

constexpr int kThreads = 256;  // bodies j per block, and sources i per LDS stage

// Block b holds bodies j of tile nb - 1 - b: the triangle's longest sums
// (the last bodies, i < j over almost all of the system) start first, and the
// short ones fill in behind them, so the launch does not end on a few long
// blocks.  Sources are staged kThreads at a time for all four waves.
__global__ void __launch_bounds__(kThreads) k_energy_terms(
    uint32_t n, const double *__restrict__ x, const double *__restrict__ y,
    const double *__restrict__ z, const double *__restrict__ vx, const double *__restrict__ vy,
    const double *__restrict__ vz, const double *__restrict__ m, double G, double *ekin,
    double *epot) {
    __shared__ double4 src[kThreads];
    const uint32_t tile = gridDim.x - 1 - blockIdx.x;
    const uint32_t j = tile * kThreads + threadIdx.x;
    const bool valid = j < n;
    const uint32_t jj = valid ? j : 0;
    const double pxj = x[jj], pyj = y[jj], pzj = z[jj], mj = m[jj];
    if (valid) {
        const double v = vx[j] * vx[j] + vy[j] * vy[j] + vz[j] * vz[j];
        ekin[j] = 0.5 * mj * v;
    }
    double ep = 0;
    // The numerator G m_i m_j is (G m_i) m_j, so G m_i is staged with the
    // source.  When G m_i and m_j both lie in [2^-150, 2^150] (the whole tile
    // and this lane), the numerator lies in div_core's range and only the
    // squared distance needs checking.
    const bool mj_ok = fabs(mj) >= 0x1p-150 && fabs(mj) <= 0x1p150;
    const uint32_t last = min(n, (tile + 1) * kThreads);  // bodies i < j of this block
    for (uint32_t k = 0; k < last; k += kThreads) {
        const uint32_t i = k + threadIdx.x;
        bool gm_ok = true;
        if (i < n) {
            const double gm = G * m[i];
            gm_ok = fabs(gm) >= 0x1p-150 && fabs(gm) <= 0x1p150;
            src[threadIdx.x] = make_double4(x[i], y[i], z[i], gm);
        }
        const bool tile_ok = __syncthreads_and(gm_ok);
        if (valid && j > k) {
            const uint32_t cnt = min((uint32_t)kThreads, j - k);  // sources i = k + ii < j
            if (tile_ok && mj_ok) {
                for (uint32_t ii = 0; ii < cnt; ++ii) {
                    const double4 si = src[ii];
                    const double r_x = pxj - si.x;
                    const double r_y = pyj - si.y;
                    const double r_z = pzj - si.z;
                    const double a = si.w * mj;
                    const double v = r_x * r_x + r_y * r_y + r_z * r_z;
                    ep += (v >= 0x1p-600 && v <= 0x1p600) ? div_core(a, sqrt_core(v))
                                                          : a / sqrt(v);
                }
            } else {
                for (uint32_t ii = 0; ii < cnt; ++ii) {
                    const double4 si = src[ii];
                    const double r_x = pxj - si.x;
                    const double r_y = pyj - si.y;
                    const double r_z = pzj - si.z;
                    ep += div_sqrt_ieee(si.w * mj, r_x * r_x + r_y * r_y + r_z * r_z);
                }
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
    if (n == 0) return;
    hipLaunchKernelGGL(k_energy_terms, dim3(ceil_div(n, kThreads)), dim3(kThreads), 0, s, n, x, y,
                       z, vx, vy, vz, m, G, ekin, epot);
}

}  // namespace nb

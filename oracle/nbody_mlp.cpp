// nbody_mlp.cpp -- CPU restatement of the learned far-field correction
// (test infrastructure only; see nbody_mlp_oracle.h for the definitions).
#include "nbody_mlp_oracle.h"

#include <cmath>
#include <cstring>
#include <vector>

namespace {

constexpr int kIn = ORC_MLP_IN, kH = ORC_MLP_HIDDEN, kOut = ORC_MLP_OUT;
constexpr int oW1 = 0, oB1 = oW1 + kIn * kH, oW2 = oB1 + kH, oB2 = oW2 + kH * kH,
              oW3 = oB2 + kH, oB3 = oW3 + kH * kH, oW4 = oB3 + kH, oB4 = oW4 + kH * kOut;
static_assert(oB4 + kOut == ORC_MLP_PARAMS, "parameter layout");

// round-to-nearest-even FP32 -> BF16 -> FP32
double bf(double v, int on) {
    if (!on) return v;
    float f = (float)v;
    uint32_t u;
    std::memcpy(&u, &f, 4);
    u += 0x7fffu + ((u >> 16) & 1u);
    u &= 0xffff0000u;
    std::memcpy(&f, &u, 4);
    return (double)f;
}

double sigmoid(double z) { return 1.0 / (1.0 + std::exp(-z)); }
double silu(double z) { return z * sigmoid(z); }
double dsilu(double z) {
    const double s = sigmoid(z);
    return s + z * s * (1.0 - s);
}

struct Net {
    std::vector<double> w;  // parameters, BF16-rounded weights when emulating
    explicit Net(const float *params, int b16) : w(ORC_MLP_PARAMS) {
        for (int i = 0; i < ORC_MLP_PARAMS; ++i) {
            const bool is_bias = (i >= oB1 && i < oW2) || (i >= oB2 && i < oW3) ||
                                 (i >= oB3 && i < oW4) || i >= oB4;
            w[i] = is_bias ? (double)params[i] : bf((double)params[i], b16);
        }
    }
};

// z[out] = b[out] + sum_i W[i][out] * a[i]
void dense(const double *W, const double *b, int nin, int nout, const double *a, double *z) {
    for (int o = 0; o < nout; ++o) {
        double s = 0.0;
        for (int i = 0; i < nin; ++i) s += W[i * nout + o] * a[i];
        z[o] = s + b[o];
    }
}

struct Acts {
    double x[kIn], z1[kH], h1[kH], z2[kH], h2[kH], z3[kH], h3[kH], y[kOut];
};

void forward_one(const Net &n, const float *X, int b16, Acts &a) {
    for (int i = 0; i < kIn; ++i) a.x[i] = bf((double)X[i], b16);
    const double *w = n.w.data();
    dense(w + oW1, w + oB1, kIn, kH, a.x, a.z1);
    for (int i = 0; i < kH; ++i) a.h1[i] = bf(silu(a.z1[i]), b16);
    dense(w + oW2, w + oB2, kH, kH, a.h1, a.z2);
    for (int i = 0; i < kH; ++i) a.h2[i] = bf(silu(a.z2[i]), b16);
    dense(w + oW3, w + oB3, kH, kH, a.h2, a.z3);
    for (int i = 0; i < kH; ++i) a.h3[i] = bf(silu(a.z3[i]), b16);
    dense(w + oW4, w + oB4, kH, kOut, a.h3, a.y);
}

}  // namespace

extern "C" {

void orc_mlp_forward(const float *params, uint32_t count, const float *X, float *Y, int bf16) {
    const Net n(params, bf16);
#pragma omp parallel for schedule(static)
    for (int64_t s = 0; s < (int64_t)count; ++s) {
        Acts a;
        forward_one(n, X + s * kIn, bf16, a);
        for (int c = 0; c < kOut; ++c) Y[s * kOut + c] = (float)a.y[c];
    }
}

void orc_mlp_grad(const float *params, uint32_t count, const float *X, const float *T,
                  double *grad, double *loss, int bf16) {
    const Net n(params, bf16);
    const double *w = n.w.data();
    const double coef = 2.0 / (3.0 * (double)count);
    std::vector<double> g(ORC_MLP_PARAMS, 0.0);
    double L = 0.0;
    // serial: a fixed summation order keeps the oracle deterministic
    for (uint32_t s = 0; s < count; ++s) {
        Acts a;
        forward_one(n, X + (uint64_t)s * kIn, bf16, a);
        double dy[kOut], dz3[kH], dz2[kH], dz1[kH];
        for (int c = 0; c < kOut; ++c) {
            const double r = a.y[c] - (double)T[(uint64_t)s * kOut + c];
            L += r * r;
            dy[c] = bf(coef * r, bf16);
        }
        for (int i = 0; i < kH; ++i) {  // dh3 = W4 dy
            double t = 0.0;
            for (int c = 0; c < kOut; ++c) t += w[oW4 + i * kOut + c] * dy[c];
            dz3[i] = bf(t * dsilu(a.z3[i]), bf16);
        }
        for (int i = 0; i < kH; ++i) {  // dh2 = W3 dz3
            double t = 0.0;
            for (int o = 0; o < kH; ++o) t += w[oW3 + i * kH + o] * dz3[o];
            dz2[i] = bf(t * dsilu(a.z2[i]), bf16);
        }
        for (int i = 0; i < kH; ++i) {  // dh1 = W2 dz2
            double t = 0.0;
            for (int o = 0; o < kH; ++o) t += w[oW2 + i * kH + o] * dz2[o];
            dz1[i] = bf(t * dsilu(a.z1[i]), bf16);
        }
        for (int i = 0; i < kIn; ++i)
            for (int o = 0; o < kH; ++o) g[oW1 + i * kH + o] += a.x[i] * dz1[o];
        for (int o = 0; o < kH; ++o) g[oB1 + o] += dz1[o];
        for (int i = 0; i < kH; ++i)
            for (int o = 0; o < kH; ++o) g[oW2 + i * kH + o] += a.h1[i] * dz2[o];
        for (int o = 0; o < kH; ++o) g[oB2 + o] += dz2[o];
        for (int i = 0; i < kH; ++i)
            for (int o = 0; o < kH; ++o) g[oW3 + i * kH + o] += a.h2[i] * dz3[o];
        for (int o = 0; o < kH; ++o) g[oB3 + o] += dz3[o];
        for (int i = 0; i < kH; ++i)
            for (int c = 0; c < kOut; ++c) g[oW4 + i * kOut + c] += a.h3[i] * dy[c];
        for (int c = 0; c < kOut; ++c) g[oB4 + c] += dy[c];
    }
    for (int i = 0; i < ORC_MLP_PARAMS; ++i) grad[i] = g[i];
    *loss = L / (3.0 * (double)count);
}

void orc_mlp_adam(uint32_t count, float *w, float *m, float *v, const double *grad, float lr,
                  float beta1, float beta2, float eps, uint32_t t) {
    const double c1 = 1.0 - std::pow((double)beta1, (double)t);
    const double c2 = 1.0 - std::pow((double)beta2, (double)t);
    for (uint32_t i = 0; i < count; ++i) {
        const double g = grad[i];
        const double mi = (double)beta1 * m[i] + (1.0 - beta1) * g;
        const double vi = (double)beta2 * v[i] + (1.0 - beta2) * g * g;
        m[i] = (float)mi;
        v[i] = (float)vi;
        w[i] = (float)((double)w[i] - (double)lr * (mi / c1) / (std::sqrt(vi / c2) + eps));
    }
}

void orc_mlp_pair(uint32_t nb, const double *x, const double *y, const double *z,
                  const double *m, const double centre[3], double edge, double m_root, double G,
                  double eps2, const double p[3], float feat[13], double target[3]) {
    double M = 0, cx = 0, cy = 0, cz = 0;
    for (uint32_t j = 0; j < nb; ++j) {
        M += m[j];
        cx += m[j] * x[j];
        cy += m[j] * y[j];
        cz += m[j] * z[j];
    }
    cx /= M;
    cy /= M;
    cz /= M;
    double Q[6] = {0, 0, 0, 0, 0, 0};  // xx yy zz xy xz yz
    double ax = 0, ay = 0, az = 0;
    for (uint32_t j = 0; j < nb; ++j) {
        const double sx = x[j] - cx, sy = y[j] - cy, sz = z[j] - cz;
        const double s2 = sx * sx + sy * sy + sz * sz;
        Q[0] += m[j] * (3 * sx * sx - s2);
        Q[1] += m[j] * (3 * sy * sy - s2);
        Q[2] += m[j] * (3 * sz * sz - s2);
        Q[3] += m[j] * 3 * sx * sy;
        Q[4] += m[j] * 3 * sx * sz;
        Q[5] += m[j] * 3 * sy * sz;
        const double rx = x[j] - p[0], ry = y[j] - p[1], rz = z[j] - p[2];
        double den = rx * rx + ry * ry + rz * rz + eps2;
        den = 1.0 / std::sqrt(den * den * den);
        ax += G * m[j] * rx * den;
        ay += G * m[j] * ry * den;
        az += G * m[j] * rz * den;
    }
    const double rx = cx - p[0], ry = cy - p[1], rz = cz - p[2];
    const double d2 = rx * rx + ry * ry + rz * rz;
    const double d = std::sqrt(d2);
    double den = d2 + eps2;
    den = 1.0 / std::sqrt(den * den * den);
    const double mx = G * M * rx * den, my = G * M * ry * den, mz = G * M * rz * den;
    const double ux = rx / d, uy = ry / d, uz = rz / d;
    const double e2 = edge * edge;
    feat[0] = (float)ux;
    feat[1] = (float)uy;
    feat[2] = (float)uz;
    feat[3] = (float)(edge / d);
    feat[4] = (float)(std::log(M / m_root) / 16.0);
    for (int k = 0; k < 6; ++k) feat[5 + k] = (float)(Q[k] / (M * e2));
    const double ox = (centre[0] - cx) / edge, oy = (centre[1] - cy) / edge,
                 oz = (centre[2] - cz) / edge;
    const double par = ox * ux + oy * uy + oz * uz;
    const double perp2 = ox * ox + oy * oy + oz * oz - par * par;
    feat[11] = (float)par;
    feat[12] = (float)std::sqrt(perp2 > 0 ? perp2 : 0.0);
    const double scale = d2 * d2 / (G * M * e2);
    target[0] = (ax - mx) * scale;
    target[1] = (ay - my) * scale;
    target[2] = (az - mz) * scale;
}

}  // extern "C"

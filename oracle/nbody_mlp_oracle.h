/*
 * nbody_mlp_oracle.h -- CPU restatement of the learned far-field correction.
 *
 * TEST INFRASTRUCTURE ONLY (the same rule as nbody_oracle.h: tests/, the
 * smoke check and bench.py's CPU baseline are the only users).
 *
 * The correction has no counterpart in the reference.  What is restated here
 * is the definition the GPU code implements:
 *   - the network: 13 inputs -> 64 -> 64 -> 64 (SiLU) -> 3 outputs, evaluated
 *     in FP64, optionally rounding to BF16 at exactly the points where the GPU
 *     feeds a matrix-core operand (inputs, weights, hidden activations, and in
 *     the backward pass the output and pre-activation gradients);
 *   - the loss: mean over samples and 3 components of (y - t)^2, and its
 *     gradient with respect to every parameter;
 *   - the features and the normalised monopole residual of one (cell, target)
 *     pair, computed by brute force from the cell's bodies.
 *
 * Parameter layout (9411 floats, row-major, input index first):
 *   W1[13][64] b1[64] W2[64][64] b2[64] W3[64][64] b3[64] W4[64][3] b4[3]
 */
#ifndef NBODY_MLP_ORACLE_H
#define NBODY_MLP_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum { ORC_MLP_IN = 13, ORC_MLP_HIDDEN = 64, ORC_MLP_OUT = 3, ORC_MLP_PARAMS = 9411 };

/* Y[count][3] = net(X[count][13]).  bf16 != 0 rounds operands as the GPU does. */
void orc_mlp_forward(const float *params, uint32_t count, const float *X, float *Y, int bf16);

/* Gradient of L = (1/(3 count)) sum (y - t)^2 over all parameters (grad has
 * ORC_MLP_PARAMS entries, same layout as params); *loss receives L. */
void orc_mlp_grad(const float *params, uint32_t count, const float *X, const float *T,
                  double *grad, double *loss, int bf16);

/* One Adam step on FP32 master weights (t = 1-based step number), the update
 * the GPU applies after every batch. */
void orc_mlp_adam(uint32_t count, float *w, float *m, float *v, const double *grad, float lr,
                  float beta1, float beta2, float eps, uint32_t t);

/* Features and target of one (cell, target point) pair by brute force over the
 * cell's nb bodies (FP64):
 *   com = sum m x / M; r = com - p; d = |r|; e = cell edge
 *   f[0..2] = r / d, f[3] = e / d, f[4] = ln(M / m_root) / 16,
 *   f[5..10] = {Qxx, Qyy, Qzz, Qxy, Qxz, Qyz} / (M e^2) with
 *              Q = sum m (3 s s^T - |s|^2 I), s = x - com,
 *   f[11] = o . r^, f[12] = |o - (o . r^) r^|, o = (centre - com) / e,
 *   target = (a_exact - a_mono) * d^4 / (G M e^2), with
 *            a_exact = sum G m_j s_j / (|s_j|^2 + eps2)^(3/2), s_j = x_j - p,
 *            a_mono = G M r / (d^2 + eps2)^(3/2).
 * The correction applied to a body is net(f) * G M e^2 / d^4. */
void orc_mlp_pair(uint32_t nb, const double *x, const double *y, const double *z,
                  const double *m, const double centre[3], double edge, double m_root, double G,
                  double eps2, const double p[3], float feat[13], double target[3]);

#ifdef __cplusplus
}
#endif

#endif

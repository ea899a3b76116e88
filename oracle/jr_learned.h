/*
 * CPU oracle (test infrastructure only) for the learned part of the walk:
 * the message-passing potential used as the expensive exact judge and the
 * radial-descriptor MLP surrogate.  FP64 restatements of the device kernels'
 * mathematics; the device runs FP32 (judge) and BF16/FP32 (surrogate), so
 * parity with these functions is to a stated tolerance, not bit-exact.
 * Nothing in the product path links or calls this library.
 */
#ifndef JR_LEARNED_H
#define JR_LEARNED_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Invariant message-passing potential (seeded random weights, FP32 storage).
 *   basis   b_n(r) = sqrt(2/rc) sin(n pi r/rc) / r * u(r/rc), n = 1..NB,
 *           u(x) = 1 - 28 x^6 + 48 x^7 - 21 x^8 (x < 1), 0 beyond
 *   layer 1 S[i][n][s] = sum_{j in N(i), t_j = s} b_n(r_ij)
 *           m1[i][c] = sum_n W1[n][c] sum_s emb[s][c] S[i][n][s]
 *           h1[i] = emb[t_i] + silu(m1[i] U1 + c1)
 *   layer 2 m2[i][c] = sum_n W2[n][c] sum_{j in N(i)} b_n(r_ij) h1[j][c]
 *           h2[i] = h1[i] + silu(m2[i] U2 + c2)
 *   E = sum_i wo . h2[i] + 1/2 sum_{i, j in N(i), r < core_rc}
 *           core_eps ((core_sig/r)^12 - (core_sig/core_rc)^12)
 * N(i): every periodic image of every atom within rc of atom i (excluding
 * i itself at zero shift); cubic cells only (cell[0] = cell[4] = cell[8]). */
typedef struct {
  int S, C, NB, pad_;
  double rc, core_eps, core_sig, core_rc;
  const float* emb; /* [S][C] */
  const float* W1;  /* [NB][C] */
  const float* U1;  /* [C][C]  z[d] = sum_c m[c] U[c][d] + b[d] */
  const float* c1;  /* [C] */
  const float* W2;  /* [NB][C] */
  const float* U2;  /* [C][C] */
  const float* c2;  /* [C] */
  const float* wo;  /* [C] */
} or_mpnn;

/* Energy of one configuration; F [N][3] = -dE/dx when F != NULL.  Returns
 * the number of directed edges (neighbour pairs within rc). */
long or_mpnn_energy_forces(const or_mpnn* net, int N, const double* pos,
                           const int32_t* types, const double* cell,
                           double* E, double* F);

/* Radial-descriptor MLP surrogate.
 *   descriptor  G_i[s][k] = 1/4 sum_{j in N_d(i), t_j = s}
 *                   exp(-(r_ij - mu_k)^2 / (2 w^2)) u(r_ij / rd),
 *               mu_k = 0.8 + 1.7 k / 31 (k < 32), w = 0.1, rd = 3.0, u as above
 *   input       x_i = [G_i[0][0..31], G_i[1][0..31], onehot(t_i), 0...] (80)
 *   network     h1 = silu(x W1 + b1), h2 = silu(h1 W2 + b2), e_i = h2 . w3 + b3
 *   E = sum_i e_i.
 * Parameter vector (FP32): W1 [80][128], b1 [128], W2 [128][128], b2 [128],
 * w3 [128], b3 [1].  bf16 != 0 emulates the device: x, W1, W2, h1 and the
 * backward activations rounded to bfloat16 (nearest-even) where the device
 * feeds them to the matrix cores. */
enum { OR_SUR_D = 80, OR_SUR_H = 128, OR_SUR_K = 32 };
enum {
  OR_SUR_W1 = 0,
  OR_SUR_B1 = OR_SUR_W1 + OR_SUR_D * OR_SUR_H,
  OR_SUR_W2 = OR_SUR_B1 + OR_SUR_H,
  OR_SUR_B2 = OR_SUR_W2 + OR_SUR_H * OR_SUR_H,
  OR_SUR_W3 = OR_SUR_B2 + OR_SUR_H,
  OR_SUR_B3 = OR_SUR_W3 + OR_SUR_H,
  OR_SUR_NP = OR_SUR_B3 + 1
};

/* Descriptor rows x [N][80] (FP64, unrounded) of one cubic configuration. */
void or_sur_descriptors(int N, const double* pos, const int32_t* types,
                        const double* cell, int S, double* x);

/* Energy (and forces when F != NULL) of one configuration.  base =
 * {eps, sig, rc} adds the pair baseline eps ((sig/r)^12 - (sig/rc)^12) for
 * r < rc <= 3 (NULL: none). */
double or_sur_energy_forces(const float* params, int bf16, int N,
                            const double* pos, const int32_t* types,
                            const double* cell, int S, const double* base, double* F);

/* Training loss over n configurations given their descriptor rows
 * (x [sum N][80]), atom counts and exact energies:
 *   loss = 1/n sum_b ((E_hat_b - E_b) / N_b)^2,
 * with its gradient with respect to the parameter vector (grad [OR_SUR_NP]). */
double or_sur_loss_grad(const float* params, int bf16, int n, const int* n_atoms,
                        const double* x, const double* E, double* grad);

/* One Adam step (beta1 0.9, beta2 0.999, eps 1e-8) on FP32 parameters with
 * FP32 moments; t is the 1-based step count. */
void or_adam_step(int np, float* p, float* m, float* v, const float* g,
                  double lr, int t);

#ifdef __cplusplus
}
#endif
#endif

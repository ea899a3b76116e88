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

#ifdef __cplusplus
}
#endif
#endif

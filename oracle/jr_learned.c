/*
 * CPU oracle (test infrastructure only) for the learned part: FP64
 * restatements of the message-passing judge and of the radial-descriptor MLP
 * surrogate (forward, forces, training gradient, Adam).  See jr_learned.h.
 */
#include "jr_learned.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

static inline double silu(double z) { return z / (1.0 + exp(-z)); }
static inline double silu_d(double z) {
  double s = 1.0 / (1.0 + exp(-z));
  return s * (1.0 + z * (1.0 - s));
}

/* Envelope u(x) = 1 - 28x^6 + 48x^7 - 21x^8 and du/dx. */
static inline void envelope(double x, double* u, double* du) {
  double x2 = x * x, x5 = x2 * x2 * x, x6 = x5 * x, x7 = x6 * x;
  *u = 1.0 - 28.0 * x6 + 48.0 * x7 - 21.0 * x7 * x;
  *du = -168.0 * x5 + 336.0 * x6 - 168.0 * x7;
}

typedef struct {
  int i, j;           /* receiver, sender */
  double d[3], r;     /* d = x_j + shift - x_i */
} edge_t;

long or_mpnn_energy_forces(const or_mpnn* net, int N, const double* pos,
                           const int32_t* types, const double* cell,
                           double* E_out, double* F) {
  const int S = net->S, C = net->C, NB = net->NB;
  const double L = cell[0], rc = net->rc, rc2 = rc * rc;
  /* wrapped fractional coordinates */
  double* u = (double*)malloc(sizeof(double) * 3 * (N > 0 ? N : 1));
  for (int a = 0; a < 3 * N; ++a) {
    double f = pos[a] / L;
    u[a] = f - floor(f);
  }
  const int nmax = (int)ceil(rc / L) + 1;
  long cap = 1024, ne = 0;
  edge_t* e = (edge_t*)malloc(sizeof(edge_t) * cap);
  for (int i = 0; i < N; ++i)
    for (int j = 0; j < N; ++j)
      for (int sx = -nmax; sx <= nmax; ++sx)
        for (int sy = -nmax; sy <= nmax; ++sy)
          for (int sz = -nmax; sz <= nmax; ++sz) {
            if (i == j && sx == 0 && sy == 0 && sz == 0) continue;
            double d0 = (u[3 * j] - u[3 * i] + sx) * L;
            double d1 = (u[3 * j + 1] - u[3 * i + 1] + sy) * L;
            double d2 = (u[3 * j + 2] - u[3 * i + 2] + sz) * L;
            double r2 = d0 * d0 + d1 * d1 + d2 * d2;
            if (!(r2 < rc2)) continue;
            if (ne == cap) {
              cap *= 2;
              e = (edge_t*)realloc(e, sizeof(edge_t) * cap);
            }
            e[ne].i = i;
            e[ne].j = j;
            e[ne].d[0] = d0;
            e[ne].d[1] = d1;
            e[ne].d[2] = d2;
            e[ne].r = sqrt(r2);
            ++ne;
          }
  /* basis and derivative per edge */
  double* b = (double*)malloc(sizeof(double) * NB * (ne > 0 ? ne : 1));
  double* db = (double*)malloc(sizeof(double) * NB * (ne > 0 ? ne : 1));
  const double pref = sqrt(2.0 / rc);
  for (long k = 0; k < ne; ++k) {
    double r = e[k].r, x = r / rc, uu, du;
    envelope(x, &uu, &du);
    for (int n = 1; n <= NB; ++n) {
      double a = n * M_PI / rc;
      double s = sin(a * r), c = cos(a * r);
      double f = pref * s / r, fd = pref * (a * c / r - s / (r * r));
      b[k * NB + n - 1] = f * uu;
      db[k * NB + n - 1] = fd * uu + f * du / rc;
    }
  }
#define W(P, n, c) ((double)net->P[(n) * C + (c)])
  double* Sv = (double*)calloc((size_t)N * NB * S + 1, sizeof(double));
  double* z1 = (double*)malloc(sizeof(double) * N * C + 1);
  double* h1 = (double*)malloc(sizeof(double) * N * C + 1);
  double* A2 = (double*)calloc((size_t)N * NB * C + 1, sizeof(double));
  double* z2 = (double*)malloc(sizeof(double) * N * C + 1);
  double* m = (double*)malloc(sizeof(double) * C);
  for (long k = 0; k < ne; ++k)
    for (int n = 0; n < NB; ++n)
      Sv[((size_t)e[k].i * NB + n) * S + types[e[k].j]] += b[k * NB + n];
  for (int i = 0; i < N; ++i) {
    for (int c = 0; c < C; ++c) {
      double acc = 0.0;
      for (int n = 0; n < NB; ++n) {
        double t = 0.0;
        for (int s = 0; s < S; ++s)
          t += W(emb, s, c) * Sv[((size_t)i * NB + n) * S + s];
        acc += W(W1, n, c) * t;
      }
      m[c] = acc;
    }
    for (int d = 0; d < C; ++d) {
      double z = W(c1, 0, d);
      for (int c = 0; c < C; ++c) z += m[c] * W(U1, c, d);
      z1[i * C + d] = z;
      h1[i * C + d] = W(emb, types[i], d) + silu(z);
    }
  }
  for (long k = 0; k < ne; ++k)
    for (int n = 0; n < NB; ++n)
      for (int c = 0; c < C; ++c)
        A2[((size_t)e[k].i * NB + n) * C + c] +=
            b[k * NB + n] * h1[e[k].j * C + c];
  double E = 0.0;
  for (int i = 0; i < N; ++i) {
    for (int c = 0; c < C; ++c) {
      double acc = 0.0;
      for (int n = 0; n < NB; ++n)
        acc += W(W2, n, c) * A2[((size_t)i * NB + n) * C + c];
      m[c] = acc;
    }
    for (int d = 0; d < C; ++d) {
      double z = W(c2, 0, d);
      for (int c = 0; c < C; ++c) z += m[c] * W(U2, c, d);
      z2[i * C + d] = z;
      E += W(wo, 0, d) * (h1[i * C + d] + silu(z));
    }
  }
  const double sc = net->core_sig, rcc = net->core_rc;
  const double shift12 = pow(sc / rcc, 12);
  for (long k = 0; k < ne; ++k)
    if (e[k].r < rcc)
      E += 0.5 * net->core_eps * (pow(sc / e[k].r, 12) - shift12);
  *E_out = E;

  if (F != NULL) {
    double* gm2 = (double*)malloc(sizeof(double) * N * C + 1);
    double* gh1 = (double*)malloc(sizeof(double) * N * C + 1);
    double* gS = (double*)malloc(sizeof(double) * (size_t)N * NB * S + 1);
    double* gb = (double*)calloc((size_t)NB * (ne > 0 ? ne : 1), sizeof(double));
    for (int i = 0; i < N; ++i) {
      for (int c = 0; c < C; ++c) {
        double acc = 0.0;
        for (int d = 0; d < C; ++d)
          acc += W(U2, c, d) * W(wo, 0, d) * silu_d(z2[i * C + d]);
        gm2[i * C + c] = acc;
        gh1[i * C + c] = W(wo, 0, c);
      }
    }
    /* through A2: h1 of the sender, and the basis */
    for (long k = 0; k < ne; ++k) {
      const int i = e[k].i, j = e[k].j;
      for (int n = 0; n < NB; ++n) {
        double acc = 0.0;
        for (int c = 0; c < C; ++c) {
          double gA = W(W2, n, c) * gm2[i * C + c];
          acc += gA * h1[j * C + c];
          gh1[j * C + c] += b[k * NB + n] * gA;
        }
        gb[k * NB + n] += acc;
      }
    }
    for (int i = 0; i < N; ++i) {
      for (int c = 0; c < C; ++c) {
        double acc = 0.0;
        for (int d = 0; d < C; ++d)
          acc += W(U1, c, d) * gh1[i * C + d] * silu_d(z1[i * C + d]);
        m[c] = acc; /* gm1 */
      }
      for (int n = 0; n < NB; ++n)
        for (int s = 0; s < S; ++s) {
          double acc = 0.0;
          for (int c = 0; c < C; ++c) acc += m[c] * W(W1, n, c) * W(emb, s, c);
          gS[((size_t)i * NB + n) * S + s] = acc;
        }
    }
    for (int a = 0; a < 3 * N; ++a) F[a] = 0.0;
    for (long k = 0; k < ne; ++k) {
      const int i = e[k].i, j = e[k].j;
      double dEdr = 0.0;
      for (int n = 0; n < NB; ++n)
        dEdr += (gb[k * NB + n] + gS[((size_t)i * NB + n) * S + types[j]]) *
                db[k * NB + n];
      const double r = e[k].r;
      if (r < rcc) dEdr += 0.5 * net->core_eps * (-12.0) * pow(sc / r, 12) / r;
      for (int q = 0; q < 3; ++q) {
        double g = dEdr * e[k].d[q] / r; /* dE/dx_j; dE/dx_i = -g */
        F[3 * j + q] -= g;
        F[3 * i + q] += g;
      }
    }
    free(gm2);
    free(gh1);
    free(gS);
    free(gb);
  }
#undef W
  free(u);
  free(e);
  free(b);
  free(db);
  free(Sv);
  free(z1);
  free(h1);
  free(A2);
  free(z2);
  free(m);
  return ne;
}

/* ---------------------------------------------------------------------- */
/* Radial-descriptor MLP surrogate                                        */
/* ---------------------------------------------------------------------- */

#include <stdint.h>

#define SUR_RD 3.0
#define SUR_W 0.1
#define SUR_SCALE 0.25

/* float -> bfloat16 (round to nearest even) -> value */
static inline double bf16r(double v) {
  float f = (float)v;
  uint32_t u;
  memcpy(&u, &f, 4);
  if ((u & 0x7f800000u) != 0x7f800000u) {
    u += 0x7fffu + ((u >> 16) & 1u);
    u &= 0xffff0000u;
  }
  memcpy(&f, &u, 4);
  return (double)f;
}

static inline double sur_mu(int k) { return 0.8 + 1.7 * k / 31.0; }

/* g_k(r) and dg_k/dr for all k */
static void sur_gauss(double r, double* g, double* dg) {
  double uu, du;
  envelope(r / SUR_RD, &uu, &du);
  du /= SUR_RD;
  for (int k = 0; k < OR_SUR_K; ++k) {
    double t = r - sur_mu(k);
    double e = exp(-t * t / (2.0 * SUR_W * SUR_W));
    g[k] = e * uu;
    if (dg) dg[k] = e * (-t / (SUR_W * SUR_W)) * uu + e * du;
  }
}

typedef struct {
  int i, j;
  double d[3], r;
} sedge_t;

static long sur_edges(int N, const double* pos, const double* cell, sedge_t** out) {
  const double L = cell[0], rd2 = SUR_RD * SUR_RD;
  double* u = (double*)malloc(sizeof(double) * 3 * (N > 0 ? N : 1));
  for (int a = 0; a < 3 * N; ++a) {
    double f = pos[a] / L;
    u[a] = f - floor(f);
  }
  const int nmax = (int)ceil(SUR_RD / L) + 1;
  long cap = 1024, ne = 0;
  sedge_t* e = (sedge_t*)malloc(sizeof(sedge_t) * cap);
  for (int i = 0; i < N; ++i)
    for (int j = 0; j < N; ++j)
      for (int sx = -nmax; sx <= nmax; ++sx)
        for (int sy = -nmax; sy <= nmax; ++sy)
          for (int sz = -nmax; sz <= nmax; ++sz) {
            if (i == j && sx == 0 && sy == 0 && sz == 0) continue;
            double d0 = (u[3 * j] - u[3 * i] + sx) * L;
            double d1 = (u[3 * j + 1] - u[3 * i + 1] + sy) * L;
            double d2 = (u[3 * j + 2] - u[3 * i + 2] + sz) * L;
            double r2 = d0 * d0 + d1 * d1 + d2 * d2;
            if (!(r2 < rd2)) continue;
            if (ne == cap) {
              cap *= 2;
              e = (sedge_t*)realloc(e, sizeof(sedge_t) * cap);
            }
            e[ne].i = i;
            e[ne].j = j;
            e[ne].d[0] = d0;
            e[ne].d[1] = d1;
            e[ne].d[2] = d2;
            e[ne].r = sqrt(r2);
            ++ne;
          }
  free(u);
  *out = e;
  return ne;
}

void or_sur_descriptors(int N, const double* pos, const int32_t* types,
                        const double* cell, int S, double* x) {
  sedge_t* e;
  long ne = sur_edges(N, pos, cell, &e);
  memset(x, 0, sizeof(double) * N * OR_SUR_D);
  double g[OR_SUR_K];
  for (long k = 0; k < ne; ++k) {
    sur_gauss(e[k].r, g, NULL);
    double* xi = x + (size_t)e[k].i * OR_SUR_D + OR_SUR_K * types[e[k].j];
    for (int q = 0; q < OR_SUR_K; ++q) xi[q] += SUR_SCALE * g[q];
  }
  for (int i = 0; i < N; ++i) x[(size_t)i * OR_SUR_D + OR_SUR_K * S + types[i]] = 1.0;
  free(e);
}

/* Forward of rows x [n][80] -> per-row energies e; keeps z1, h1, z2 when
 * the pointers are set (row-major [n][128]). */
static void sur_forward(const float* P, int bf16, int n, const double* x, double* e,
                        double* z1, double* h1, double* z2) {
  const int D = OR_SUR_D, H = OR_SUR_H;
  double* a = (double*)malloc(sizeof(double) * D);
  double* hb = (double*)malloc(sizeof(double) * H);
  for (int r = 0; r < n; ++r) {
    for (int c = 0; c < D; ++c) a[c] = bf16 ? bf16r(x[(size_t)r * D + c]) : x[(size_t)r * D + c];
    double* Z1 = z1 ? z1 + (size_t)r * H : NULL;
    double* H1 = h1 ? h1 + (size_t)r * H : NULL;
    double* Z2 = z2 ? z2 + (size_t)r * H : NULL;
    double t1[OR_SUR_H], t2[OR_SUR_H];
    for (int h = 0; h < H; ++h) {
      double acc = 0.0;
      for (int c = 0; c < D; ++c) {
        double w = P[OR_SUR_W1 + c * H + h];
        acc += a[c] * (bf16 ? bf16r(w) : w);
      }
      t1[h] = acc + P[OR_SUR_B1 + h];
      double hv = silu(t1[h]);
      hb[h] = bf16 ? bf16r(hv) : hv;
      if (Z1) Z1[h] = t1[h];
      if (H1) H1[h] = hb[h];
    }
    double er = P[OR_SUR_B3];
    for (int h = 0; h < H; ++h) {
      double acc = 0.0;
      for (int c = 0; c < H; ++c) {
        double w = P[OR_SUR_W2 + c * H + h];
        acc += hb[c] * (bf16 ? bf16r(w) : w);
      }
      t2[h] = acc + P[OR_SUR_B2 + h];
      if (Z2) Z2[h] = t2[h];
      er += P[OR_SUR_W3 + h] * silu(t2[h]);
    }
    e[r] = er;
  }
  free(a);
  free(hb);
}

/* Backward of rows given dE/de per row (c): dX [n][80] and, when grad is
 * set, parameter gradients accumulated into grad. */
static void sur_backward(const float* P, int bf16, int n, const double* x,
                         const double* c, const double* z1, const double* h1,
                         const double* z2, double* dX, double* grad) {
  const int D = OR_SUR_D, H = OR_SUR_H;
  for (int r = 0; r < n; ++r) {
    double g2[OR_SUR_H], g2b[OR_SUR_H], g1[OR_SUR_H], g1b[OR_SUR_H];
    const double* Z1 = z1 + (size_t)r * H;
    const double* H1 = h1 + (size_t)r * H;
    const double* Z2 = z2 + (size_t)r * H;
    for (int h = 0; h < H; ++h) {
      g2[h] = c[r] * P[OR_SUR_W3 + h] * silu_d(Z2[h]);
      g2b[h] = bf16 ? bf16r(g2[h]) : g2[h];
    }
    for (int h = 0; h < H; ++h) {
      double acc = 0.0;
      for (int k = 0; k < H; ++k) {
        double w = P[OR_SUR_W2 + h * H + k];
        acc += g2b[k] * (bf16 ? bf16r(w) : w);
      }
      /* the forces path holds silu'(z1) as a BF16 operand on the device */
      const double s1 = (bf16 && dX && !grad) ? bf16r(silu_d(Z1[h])) : silu_d(Z1[h]);
      g1[h] = acc * s1;
      g1b[h] = bf16 ? bf16r(g1[h]) : g1[h];
    }
    if (dX) {
      for (int q = 0; q < D; ++q) {
        double acc = 0.0;
        for (int h = 0; h < H; ++h) {
          double w = P[OR_SUR_W1 + q * H + h];
          acc += g1b[h] * (bf16 ? bf16r(w) : w);
        }
        dX[(size_t)r * D + q] = acc;
      }
    }
    if (grad) {
      for (int h = 0; h < H; ++h) {
        grad[OR_SUR_W3 + h] += c[r] * silu(Z2[h]);
        grad[OR_SUR_B2 + h] += g2[h];
        grad[OR_SUR_B1 + h] += g1[h];
        for (int k = 0; k < H; ++k) grad[OR_SUR_W2 + k * H + h] += H1[k] * g2b[h];
        for (int q = 0; q < D; ++q) {
          double xv = x[(size_t)r * D + q];
          grad[OR_SUR_W1 + q * H + h] += (bf16 ? bf16r(xv) : xv) * g1b[h];
        }
      }
      grad[OR_SUR_B3] += c[r];
    }
  }
}

double or_sur_energy_forces(const float* params, int bf16, int N,
                            const double* pos, const int32_t* types,
                            const double* cell, int S, const double* base, double* F) {
  const int D = OR_SUR_D, H = OR_SUR_H;
  size_t nn = (size_t)(N > 0 ? N : 1);
  double* x = (double*)malloc(sizeof(double) * nn * D);
  double* e = (double*)malloc(sizeof(double) * nn);
  double* z1 = (double*)malloc(sizeof(double) * nn * H);
  double* h1 = (double*)malloc(sizeof(double) * nn * H);
  double* z2 = (double*)malloc(sizeof(double) * nn * H);
  or_sur_descriptors(N, pos, types, cell, S, x);
  sur_forward(params, bf16, N, x, e, z1, h1, z2);
  double E = 0.0;
  for (int i = 0; i < N; ++i) E += e[i];
  const int has_base = base && base[0] != 0.0;
  if (has_base) {
    /* every directed edge carries half of its pair's energy */
    sedge_t* ed;
    long ne = sur_edges(N, pos, cell, &ed);
    const double shift = pow(base[1] / base[2], 12.0);
    for (long k = 0; k < ne; ++k)
      if (ed[k].r < base[2]) E += 0.5 * base[0] * (pow(base[1] / ed[k].r, 12.0) - shift);
    free(ed);
  }
  if (F) {
    double* one = (double*)calloc(nn, sizeof(double));
    double* dX = (double*)malloc(sizeof(double) * nn * D);
    for (int i = 0; i < N; ++i) one[i] = 1.0;
    sur_backward(params, bf16, N, x, one, z1, h1, z2, dX, NULL);
    sedge_t* ed;
    long ne = sur_edges(N, pos, cell, &ed);
    for (int a = 0; a < 3 * N; ++a) F[a] = 0.0;
    double g[OR_SUR_K], dg[OR_SUR_K];
    for (long k = 0; k < ne; ++k) {
      const int i = ed[k].i, j = ed[k].j;
      sur_gauss(ed[k].r, g, dg);
      double ce = 0.0;
      for (int q = 0; q < OR_SUR_K; ++q)
        ce += dX[(size_t)i * D + OR_SUR_K * types[j] + q] * dg[q];
      ce *= SUR_SCALE;
      if (has_base && ed[k].r < base[2])  /* half of d phi / dr per edge */
        ce -= 0.5 * 12.0 * base[0] * pow(base[1] / ed[k].r, 12.0) / ed[k].r;
      for (int q = 0; q < 3; ++q) {
        double gq = ce * ed[k].d[q] / ed[k].r;
        F[3 * j + q] -= gq;
        F[3 * i + q] += gq;
      }
    }
    free(ed);
    free(one);
    free(dX);
  }
  free(x);
  free(e);
  free(z1);
  free(h1);
  free(z2);
  return E;
}

double or_sur_loss_grad(const float* params, int bf16, int n, const int* n_atoms,
                        const double* x, const double* E, double* grad) {
  const int H = OR_SUR_H;
  long rows = 0;
  for (int b = 0; b < n; ++b) rows += n_atoms[b];
  double* e = (double*)malloc(sizeof(double) * (rows > 0 ? rows : 1));
  double* z1 = (double*)malloc(sizeof(double) * (rows > 0 ? rows : 1) * H);
  double* h1 = (double*)malloc(sizeof(double) * (rows > 0 ? rows : 1) * H);
  double* z2 = (double*)malloc(sizeof(double) * (rows > 0 ? rows : 1) * H);
  double* c = (double*)malloc(sizeof(double) * (rows > 0 ? rows : 1));
  sur_forward(params, bf16, (int)rows, x, e, z1, h1, z2);
  double loss = 0.0;
  long r0 = 0;
  for (int b = 0; b < n; ++b) {
    double Eh = 0.0;
    for (int i = 0; i < n_atoms[b]; ++i) Eh += e[r0 + i];
    double d = (Eh - E[b]) / n_atoms[b];
    loss += d * d / n;
    double cb = 2.0 * d / ((double)n_atoms[b] * n);
    for (int i = 0; i < n_atoms[b]; ++i) c[r0 + i] = cb;
    r0 += n_atoms[b];
  }
  if (grad) {
    memset(grad, 0, sizeof(double) * OR_SUR_NP);
    sur_backward(params, bf16, (int)rows, x, c, z1, h1, z2, NULL, grad);
  }
  free(e);
  free(z1);
  free(h1);
  free(z2);
  free(c);
  return loss;
}

void or_adam_step(int np, float* p, float* m, float* v, const float* g,
                  double lr, int t) {
  const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
  const float c1 = (float)(1.0 - pow(0.9, t)), c2 = (float)(1.0 - pow(0.999, t));
  for (int k = 0; k < np; ++k) {
    m[k] = b1 * m[k] + (1.0f - b1) * g[k];
    v[k] = b2 * v[k] + (1.0f - b2) * g[k] * g[k];
    float mh = m[k] / c1, vh = v[k] / c2;
    p[k] = p[k] - (float)lr * mh / (sqrtf(vh) + eps);
  }
}

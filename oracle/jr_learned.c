/*
 * CPU oracle (test infrastructure only) for the learned part: FP64
 * restatement of the message-passing judge.  See jr_learned.h.
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

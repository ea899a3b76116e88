/*
 * CPU oracle (test infrastructure only) -- FP64 restatement of the reference
 * nested-sampling inner walk.  See jr_oracle.h.
 *
 * Canonical evaluation order.  The reference leaves summation order to XLA;
 * this file fixes one order and the device kernels follow the same order so
 * that the two can be compared bit-for-bit:
 *   - "lane partial" of a per-atom quantity q: lane p (0..63) holds
 *     q[p] + q[p+64] + q[p+128] + ... (ascending, atoms < N only);
 *   - "wave sum": butterfly over the 64 lane partials, v[l] += v[l ^ m] for
 *     m = 1, 2, 4, 8, 16, 32 (lanes >= N hold +0.0);
 *   - dot3(a, b) = fma(a2, b2, fma(a1, b1, a0*b0)).
 * Compiled with -ffp-contract=off: only the fma() calls written here fuse.
 */
#include "jr_oracle.h"

#include <assert.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define WAVE 64

static inline double dot3(double a0, double a1, double a2, double b0,
                          double b1, double b2) {
  return fma(a2, b2, fma(a1, b1, a0 * b0));
}

/* Butterfly sum of lane partials built from per-atom values q[0..N). */
static double wave_sum_atoms(const double* q, int N) {
  double v[WAVE];
  for (int l = 0; l < WAVE; ++l) {
    double acc = 0.0;
    int first = 1;
    for (int a = l; a < N; a += WAVE) {
      if (first) {
        acc = q[a];
        first = 0;
      } else {
        acc = acc + q[a];
      }
    }
    v[l] = acc;
  }
  for (int m = 1; m < WAVE; m <<= 1) {
    double t[WAVE];
    for (int l = 0; l < WAVE; ++l) t[l] = v[l] + v[l ^ m];
    memcpy(v, t, sizeof(v));
  }
  return v[0];
}

/* |det| of a 3x3 row-major cell; utils/cell.py:12-14 (get_volume). */
double or_det3(const double* c) {
  double m0 = c[4] * c[8] - c[5] * c[7];
  double m1 = c[3] * c[8] - c[5] * c[6];
  double m2 = c[3] * c[7] - c[4] * c[6];
  return c[0] * m0 - c[1] * m1 + c[2] * m2;
}

/* Adjugate-over-determinant inverse (row-vector convention). */
static void inv3(const double* c, double det, double* inv) {
  double idet = 1.0 / det;
  inv[0] = (c[4] * c[8] - c[5] * c[7]) * idet;
  inv[1] = (c[2] * c[7] - c[1] * c[8]) * idet;
  inv[2] = (c[1] * c[5] - c[2] * c[4]) * idet;
  inv[3] = (c[5] * c[6] - c[3] * c[8]) * idet;
  inv[4] = (c[0] * c[8] - c[2] * c[6]) * idet;
  inv[5] = (c[2] * c[3] - c[0] * c[5]) * idet;
  inv[6] = (c[3] * c[7] - c[4] * c[6]) * idet;
  inv[7] = (c[1] * c[6] - c[0] * c[7]) * idet;
  inv[8] = (c[0] * c[4] - c[1] * c[3]) * idet;
}

/* Cube root by an integer initial guess on the IEEE bits and five Newton
 * steps y <- (2y + x/y^2)/3; used for the isotropic scale (V'/V)^(1/3) of
 * sampling/moves/volume.py:60 and V^(1/3) of utils/cell.py:93. */
double or_cbrt(double x) {
  if (!(x > 0.0) || isinf(x)) return x; /* 0, negative (unused), nan, inf */
  const int64_t bias = (int64_t)1023 << 52;
  int64_t b;
  memcpy(&b, &x, 8);
  int64_t g = (b - bias) / 3 + bias;
  double y;
  memcpy(&y, &g, 8);
  for (int it = 0; it < 5; ++it) {
    double t = y * y;
    y = (2.0 * y + x / t) / 3.0;
  }
  return y;
}

/* exp(x) for the stretch factors exp(+-rv) of sampling/moves/stretch.py:
 * 58-60, from IEEE +, *, rint and exponent arithmetic only (Cody-Waite
 * reduction): k = rint(x log2 e), r = (x - k ln2_hi) - k ln2_lo with ln2_hi
 * the first 32 significant bits of ln 2 (k ln2_hi exact), exp(r) by the
 * degree-13 Taylor polynomial sum r^n / n! in Horner form, times 2^k as
 * 2^(k/2) 2^(k - k/2).  Coefficients are 1/n! rounded to double. */
static double pow2i(int k) {
  int64_t b = (int64_t)(k + 1023) << 52;
  double y;
  memcpy(&y, &b, 8);
  return y;
}

double or_exp(double x) {
  static const double inv_fact[14] = {
      1.0, 1.0, 0.5, 0x1.5555555555555p-3, 0x1.5555555555555p-5,
      0x1.1111111111111p-7, 0x1.6c16c16c16c17p-10, 0x1.a01a01a01a01ap-13,
      0x1.a01a01a01a01ap-16, 0x1.71de3a556c734p-19, 0x1.27e4fb7789f5cp-22,
      0x1.ae64567f544e4p-26, 0x1.1eed8eff8d898p-29, 0x1.6124613a86d09p-33};
  if (isnan(x)) return x;
  if (x > 0x1.62e42fefa39efp+9) return INFINITY; /* ln(DBL_MAX) */
  if (x < -0x1.74910d52d3052p+9) return 0.0;     /* below the least subnormal */
  double k = nearbyint(x * 0x1.71547652b82fep+0);
  double r = (x - k * 0x1.62e42fee00000p-1) - k * 0x1.a39ef35793c76p-33;
  double p = inv_fact[13];
  for (int n = 12; n >= 0; --n) p = p * r + inv_fact[n];
  int ki = (int)k;
  int k1 = ki / 2;
  return (p * pow2i(k1)) * pow2i(ki - k1);
}

/* x**n for integer n >= 0 by binary exponentiation (volume.py:92). */
double or_ipow(double x, int n) {
  double r = 1.0, b = x;
  while (n > 0) {
    if (n & 1) r = r * b;
    b = b * b;
    n >>= 1;
  }
  return r;
}

/* utils/cell.py:80-125 (min_aspect_ratio, check_cell_shape). */
int or_check_cell_shape(const double* c, int n_atoms, double max_vpa,
                        double min_vpa, double min_aspect) {
  double vol = fabs(or_det3(c));
  double vpa = vol / (double)n_atoms;
  double vc = or_cbrt(vol);
  static const int pairs[3][3] = {{0, 1, 2}, {1, 2, 0}, {0, 2, 1}};
  double aspect = 0.0;
  for (int p = 0; p < 3; ++p) {
    const double* a = c + 3 * pairs[p][0];
    const double* b = c + 3 * pairs[p][1];
    const double* k = c + 3 * pairs[p][2];
    double cx = a[1] * b[2] - a[2] * b[1];
    double cy = a[2] * b[0] - a[0] * b[2];
    double cz = a[0] * b[1] - a[1] * b[0];
    double nrm = sqrt(dot3(cx, cy, cz, cx, cy, cz));
    double h = fabs(dot3(cx / nrm, cy / nrm, cz / nrm, k[0], k[1], k[2]));
    double ratio = h / vc;
    aspect = (p == 0) ? ratio : fmin(aspect, ratio);
  }
  return (vpa <= max_vpa) && (vpa >= min_vpa) && (aspect >= min_aspect);
}

/* P*V - mu.counts; backends/ensemble.py:34-59. */
double or_ensemble_energy(int N, const int32_t* types, const double* cell,
                          double pressure, const double* mu, int S) {
  double e = pressure * fabs(or_det3(cell));
  if (mu != NULL && S > 0) {
    int counts[16] = {0};
    for (int a = 0; a < N; ++a) counts[types[a]] += 1;
    double acc = 0.0;
    for (int s = 0; s < S; ++s) acc = acc + mu[s] * (double)counts[s];
    e = e - acc;
  }
  return e;
}

/* One pair, dr = MIC(x_j - x_i) and r^2; backends/lj.py:141-162.
 * cubic != 0: xi, xj are fractional coordinates u = x * inv[0] and
 * dr = du - rint(du) (r^2 in units of L^2). */
typedef struct {
  double dx, dy, dz, r2;
  double eps4, eps24, sig2;
  int ok; /* r2 < rc2 */
} pair_t;

static inline pair_t pair_geom(const double* xi, const double* xj,
                               const double* cell, const double* inv, int ti,
                               int tj, const or_pot* pot, double rc2,
                               int cubic) {
  pair_t r;
  double dx = xj[0] - xi[0], dy = xj[1] - xi[1], dz = xj[2] - xi[2];
  if (cubic) {
    r.dx = dx - nearbyint(dx);
    r.dy = dy - nearbyint(dy);
    r.dz = dz - nearbyint(dz);
  } else {
    double f0 = fma(dz, inv[6], fma(dy, inv[3], dx * inv[0]));
    double f1 = fma(dz, inv[7], fma(dy, inv[4], dx * inv[1]));
    double f2 = fma(dz, inv[8], fma(dy, inv[5], dx * inv[2]));
    double n0 = nearbyint(f0), n1 = nearbyint(f1), n2 = nearbyint(f2);
    r.dx = fma(-n2, cell[6], fma(-n1, cell[3], fma(-n0, cell[0], dx)));
    r.dy = fma(-n2, cell[7], fma(-n1, cell[4], fma(-n0, cell[1], dy)));
    r.dz = fma(-n2, cell[8], fma(-n1, cell[5], fma(-n0, cell[2], dz)));
  }
  r.r2 = dot3(r.dx, r.dy, r.dz, r.dx, r.dy, r.dz);
  r.ok = r.r2 < rc2;
  int ab = ti * pot->n_species + tj;
  r.eps4 = pot->eps4[ab];
  r.eps24 = pot->eps24[ab];
  r.sig2 = pot->sig2[ab];
  return r;
}

/* LJ terms from q = sigma^2/r^2 (one species; eps applied per evaluation)
 * or q = 1/r^2 (species table); backends/lj.py:175-197. */
static inline void pair_lj(const pair_t* p, double q, int multi, double* e,
                           double* fs) {
  double s2 = multi ? p->sig2 * q : q;
  double s6 = (s2 * s2) * s2;
  double el = fma(s6, s6, -s6); /* s^12 - s^6 */
  if (multi) {
    *e = p->eps4 * el;
    *fs = (p->eps24 * fma(s6, s6, el)) * q; /* 2 s^12 - s^6 */
  } else {
    *e = el;
    *fs = fma(s6, s6, el) * q;
  }
}

static inline void acc_pair(int a, int j, const pair_t* p, double q, int multi,
                            int reaction, double escale, double* e,
                            double* own, double* trav) {
  double ep, fs;
  pair_lj(p, q, multi, &ep, &fs);
  e[a] = e[a] + (escale == 1.0 ? ep : escale * ep);
  own[3 * a + 0] = fma(-fs, p->dx, own[3 * a + 0]);
  own[3 * a + 1] = fma(-fs, p->dy, own[3 * a + 1]);
  own[3 * a + 2] = fma(-fs, p->dz, own[3 * a + 2]);
  if (reaction) {
    trav[3 * j + 0] = fma(fs, p->dx, trav[3 * j + 0]);
    trav[3 * j + 1] = fma(fs, p->dy, trav[3 * j + 1]);
    trav[3 * j + 2] = fma(fs, p->dz, trav[3 * j + 2]);
  }
}

/* Lattice translation i0 c0 + i1 c1 + i2 c2 (cell rows). */
static void image_shift(const double* cell, int i0, int i1, int i2,
                        double* t) {
  double a = (double)i0, b = (double)i1, c = (double)i2;
  for (int d = 0; d < 3; ++d)
    t[d] = fma(c, cell[6 + d], fma(b, cell[3 + d], a * cell[d]));
}

/* The pair sums with supercell images, backends/lj.py:152-197:
 *   U = 1/2 sum_T sum_{i,j, not (i = j and T = 0)} phi(|MIC(x_j - x_i) + T|)
 * over the translations T of image_shift with |i_d| <= img[d].  The mirror
 * term (j, i, -T) equals (i, j, T), so each unordered pair is taken once,
 * k = 1..(N-1)/2 (atom a with (a+k) mod N, every T per pair, i0, i1, i2
 * ascending with i2 fastest, one division per term, reaction on the
 * partner), then for even N the k = N/2 pairs from both ends with half the
 * energy and no reaction, then the self terms phi(|T|), T != 0, with half
 * weight and no force (their position derivative is exactly zero). */
static void image_pairs(int N, const double* pos, const int32_t* types,
                        const double* cell, const double* inv,
                        const or_pot* pot, double num, double rc2, int multi,
                        double* e, double* own, double* trav) {
  const int* h = pot->img;
  int kmax = (N - 1) / 2;
  for (int k = 1; k <= kmax + ((N >= 2 && N % 2 == 0) ? 1 : 0); ++k) {
    const int lone = k > kmax; /* the k = N/2 pairs */
    for (int a = 0; a < N; ++a) {
      const int j = (a + k) % N;
      pair_t p0 = pair_geom(pos + 3 * a, pos + 3 * j, cell, inv, types[a],
                            types[j], pot, rc2, 0);
      for (int i0 = -h[0]; i0 <= h[0]; ++i0)
        for (int i1 = -h[1]; i1 <= h[1]; ++i1)
          for (int i2 = -h[2]; i2 <= h[2]; ++i2) {
            double t[3];
            image_shift(cell, i0, i1, i2, t);
            pair_t p = p0;
            p.dx = p0.dx + t[0];
            p.dy = p0.dy + t[1];
            p.dz = p0.dz + t[2];
            p.r2 = dot3(p.dx, p.dy, p.dz, p.dx, p.dy, p.dz);
            p.ok = p.r2 < rc2;
            double qv = p.ok ? num / p.r2 : 0.0;
            acc_pair(a, j, &p, qv, multi, !lone, lone ? 0.5 : 1.0, e, own,
                     trav);
          }
    }
  }
  for (int a = 0; a < N; ++a) {
    pair_t p;
    int ab = types[a] * pot->n_species + types[a];
    p.eps4 = pot->eps4[ab];
    p.eps24 = pot->eps24[ab];
    p.sig2 = pot->sig2[ab];
    for (int i0 = -h[0]; i0 <= h[0]; ++i0)
      for (int i1 = -h[1]; i1 <= h[1]; ++i1)
        for (int i2 = -h[2]; i2 <= h[2]; ++i2) {
          if (i0 == 0 && i1 == 0 && i2 == 0) continue;
          double t[3];
          image_shift(cell, i0, i1, i2, t);
          double r2 = dot3(t[0], t[1], t[2], t[0], t[1], t[2]);
          double ep, fs;
          pair_lj(&p, r2 < rc2 ? num / r2 : 0.0, multi, &ep, &fs);
          e[a] = e[a] + 0.5 * ep;
        }
  }
}

/* Model energy U (and forces F = -dU/dx when F != NULL), then
 * H = U + ensemble term; backends/hamiltonian.py:112-153.
 * Pair schedule: for k = 1..(N-1)/2 atom a pairs with (a+k) mod N; the
 * force on a is accumulated in k order into own[a] and the reaction on
 * j = a+k into trav[j] (also in k order); for even N the k = N/2 pairs are
 * taken from both ends with half the energy and no reaction.  Steps come in
 * groups of G = 4 while they last, then one group of 2, then a lone step; a
 * group shares one division:
 *   G = 4: p01 = r2_0 r2_1, p23 = r2_2 r2_3, d = num/(p01 p23),
 *          d01 = p23 d, d23 = p01 d, q0 = r2_1 d01, q1 = r2_0 d01,
 *          q2 = r2_3 d23, q3 = r2_2 d23;
 *   G = 2: d = num/(r2_0 r2_1), q0 = r2_1 d, q1 = r2_0 d;
 *   G = 1 and k = N/2: q = num/r2
 * (num = sigma^2 for one species, else 1).  One species: U = 4 eps * sum
 * and F = (24 eps / sigma^2) * (own + trav).  A cubic cell (diagonal, equal
 * lengths L) is evaluated in fractional coordinates u = x * inv[0]:
 * dr = du - rint(du), num and rc^2 divided by L^2, and F = cf * (own + trav)
 * with cf = 24 eps / sigma^2 * L (species table: cf = L).
 *
 * Evaluation chunks (pot->chunks = C > 1, minimum image only): the ng =
 * floor(kmax / 4) groups of four steps are split into C consecutive chunks,
 * chunk c holding groups [ceil(c ng / C), ceil((c+1) ng / C)); the last
 * chunk also takes the tail (the group of 2, the lone step, the k = N/2
 * pairs).  Each chunk keeps its own running sums e_c[a], own_c[a],
 * trav_c[a] (k order within the chunk, each starting at 0), and the totals
 * are ((x_0 + x_1) + x_2) + ... in chunk order.  C = 1 is the single running
 * sum above.  A team of C wavefronts evaluates one configuration in this
 * order, one chunk each. */
static void lj_steps(int N, const double* pos, const int32_t* types,
                     const double* cell, const double* inv, const or_pot* pot,
                     double num, double rc2, int multi, int cubic, int k,
                     int kend, int tail, double* e, double* own, double* trav,
                     pair_t* P, double* q) {
  const int kmax = (N - 1) / 2;
  const int klim = tail ? kmax : kend - 1;
  while (k <= klim) {
    const int G = (k + 3 <= kmax) ? 4 : ((k + 1 <= kmax) ? 2 : 1);
    for (int a = 0; a < N; ++a) {
      pair_t* p = P + a; /* p[i * N] is step k + i */
      double* qa = q + a;
      for (int i = 0; i < G; ++i) {
        const int j = (a + k + i) % N;
        p[i * N] = pair_geom(pos + 3 * a, pos + 3 * j, cell, inv, types[a],
                             types[j], pot, rc2, cubic);
      }
      if (G == 4) {
        const double p01 = p[0].r2 * p[N].r2;
        const double p23 = p[2 * N].r2 * p[3 * N].r2;
        const double d = num / (p01 * p23);
        const double d01 = p23 * d, d23 = p01 * d;
        qa[0] = p[0].ok ? p[N].r2 * d01 : 0.0;
        qa[N] = p[N].ok ? p[0].r2 * d01 : 0.0;
        qa[2 * N] = p[2 * N].ok ? p[3 * N].r2 * d23 : 0.0;
        qa[3 * N] = p[3 * N].ok ? p[2 * N].r2 * d23 : 0.0;
      } else if (G == 2) {
        const double d = num / (p[0].r2 * p[N].r2);
        qa[0] = p[0].ok ? p[N].r2 * d : 0.0;
        qa[N] = p[N].ok ? p[0].r2 * d : 0.0;
      } else {
        qa[0] = p[0].ok ? num / p[0].r2 : 0.0;
      }
    }
    for (int i = 0; i < G; ++i)
      for (int a = 0; a < N; ++a)
        acc_pair(a, (a + k + i) % N, &P[i * N + a], q[i * N + a], multi, 1,
                 1.0, e, own, trav);
    k += G;
  }
  if (tail && N >= 2 && N % 2 == 0) {
    int h = N / 2;
    for (int a = 0; a < N; ++a) {
      int j = (a + h) % N;
      pair_t p = pair_geom(pos + 3 * a, pos + 3 * j, cell, inv, types[a],
                           types[j], pot, rc2, cubic);
      double qh = p.ok ? num / p.r2 : 0.0;
      acc_pair(a, j, &p, qh, multi, 0, 0.5, e, own, trav);
    }
  }
}

void or_energy_forces(int N, const double* pos, const int32_t* types,
                      const double* cell_in, const or_pot* pot,
                      double pressure, const double* mu, double* U_out,
                      double* H_out, double* F) {
  double det = or_det3(cell_in);
  double cell[9];
  if (det == 0.0) { /* lj.py:138-139: non-periodic stand-in */
    memset(cell, 0, sizeof(cell));
    cell[0] = cell[4] = cell[8] = 1.0e10;
    det = or_det3(cell);
  } else {
    memcpy(cell, cell_in, sizeof(cell));
  }
  double inv[9];
  inv3(cell, det, inv);
  const int multi = pot->n_species > 1;
  const int diag = cell[1] == 0.0 && cell[2] == 0.0 && cell[3] == 0.0 &&
                   cell[5] == 0.0 && cell[6] == 0.0 && cell[7] == 0.0;
  const int images = pot->img[0] > 0 || pot->img[1] > 0 || pot->img[2] > 0;
  const int cubic = !images && diag && cell[0] == cell[4] && cell[0] == cell[8];
  const int C = pot->chunks > 1 ? pot->chunks : 1;
  assert(!(images && C > 1));
  double num = multi ? 1.0 : pot->sig2[0];
  double rc2 = pot->rc2;
  double cf = multi ? 1.0 : pot->eps24[0] / pot->sig2[0];

  size_t nn = (size_t)(N > 0 ? N : 1);
  double* e = (double*)calloc(nn, sizeof(double));
  double* own = (double*)calloc(3 * nn, sizeof(double));
  double* trav = (double*)calloc(3 * nn, sizeof(double));
  double* frac = NULL;
  if (cubic) { /* fractional coordinates; lengths folded into the constants */
    const double len = cell[0], len2 = len * len;
    num = num / len2;
    rc2 = rc2 / len2;
    cf = cf * len;
    frac = (double*)malloc(3 * nn * sizeof(double));
    for (int i = 0; i < 3 * N; ++i) frac[i] = pos[i] * inv[0];
    pos = frac;
  }
  pair_t* P = (pair_t*)malloc(4 * nn * sizeof(pair_t)); /* [4][N] */
  double* q = (double*)malloc(4 * nn * sizeof(double));
  const int kmax = (N - 1) / 2;
  if (images) {
    image_pairs(N, pos, types, cell, inv, pot, num, rc2, multi, e, own, trav);
  } else if (C == 1) {
    lj_steps(N, pos, types, cell, inv, pot, num, rc2, multi, cubic, 1, 0, 1,
             e, own, trav, P, q);
  } else {
    double* ec = (double*)malloc(nn * sizeof(double));
    double* oc = (double*)malloc(3 * nn * sizeof(double));
    double* tc = (double*)malloc(3 * nn * sizeof(double));
    const int ng = kmax / 4;
    for (int c = 0; c < C; ++c) {
      const int g0 = (c * ng + C - 1) / C, g1 = ((c + 1) * ng + C - 1) / C;
      memset(ec, 0, nn * sizeof(double));
      memset(oc, 0, 3 * nn * sizeof(double));
      memset(tc, 0, 3 * nn * sizeof(double));
      lj_steps(N, pos, types, cell, inv, pot, num, rc2, multi, cubic,
               1 + 4 * g0, 1 + 4 * g1, c == C - 1, ec, oc, tc, P, q);
      for (int a = 0; a < N; ++a) {
        e[a] = c == 0 ? ec[a] : e[a] + ec[a];
        for (int d = 0; d < 3; ++d) {
          own[3 * a + d] = c == 0 ? oc[3 * a + d] : own[3 * a + d] + oc[3 * a + d];
          trav[3 * a + d] = c == 0 ? tc[3 * a + d] : trav[3 * a + d] + tc[3 * a + d];
        }
      }
    }
    free(ec);
    free(oc);
    free(tc);
  }
  double U = wave_sum_atoms(e, N);
  if (!multi) U = pot->eps4[0] * U;
  if (F != NULL) {
    for (int a = 0; a < 3 * N; ++a) {
      F[a] = own[a] + trav[a];
      if (cubic || !multi) F[a] = cf * F[a];
    }
  }
  free(frac);
  free(e);
  free(own);
  free(trav);
  free(P);
  free(q);
  *U_out = U;
  if (H_out != NULL)
    *H_out = U + or_ensemble_energy(N, types, cell_in, pressure, mu,
                                    pot->n_species);
}

/* ---------------------------------------------------------------------- */
/* Moves                                                                  */
/* ---------------------------------------------------------------------- */

typedef struct {
  int N, Nmax, S;
  const or_pot* pot;
  double pressure;
  const double* mu;
  double emax;
} chain_ctx;

static double wave_sum_dot(const double* a, const double* b, int N,
                           double* scratch) {
  for (int i = 0; i < N; ++i)
    scratch[i] = dot3(a[3 * i], a[3 * i + 1], a[3 * i + 2], b[3 * i],
                      b[3 * i + 1], b[3 * i + 2]);
  return wave_sum_atoms(scratch, N);
}

/* sampling/moves/galilean.py:142-300 (use_forces=True).  A species scope
 * (mv->species_mask != 0) multiplies the direction, before each
 * normalisation, and the reflection force by the moving-atom mask m_a = 1
 * for types in the scope, else 0 (galilean.py:58-72,171-180,233-240);
 * refresh_direction draws a fresh direction on every call.  Returns
 * accepted. */
static int gmc_move(const chain_ctx* c, double* x, const int32_t* types,
                    const double* cell, double* E, double* d, double s,
                    const or_move* mv, const double* xi) {
  int N = c->N, Nmax = c->Nmax;
  const int scoped = mv->species_mask != 0;
  double* scr = (double*)malloc(sizeof(double) * (size_t)(N + 1));
  double* v = (double*)malloc(sizeof(double) * 3 * (size_t)(N + 1));
  double* x0 = (double*)malloc(sizeof(double) * 3 * (size_t)(N + 1));
  double* F = (double*)malloc(sizeof(double) * 3 * (size_t)(N + 1));
  double* fq = (double*)malloc(sizeof(double) * (size_t)(N + 1));
  double* fd = (double*)malloc(sizeof(double) * (size_t)(N + 1));

  /* direction: fresh if |d_prev| < 1e-8, else perturbed (galilean.py:160-168) */
  double dn = sqrt(wave_sum_dot(d, d, N, scr));
  int fresh = dn < 1e-8 || mv->refresh_direction;
  for (int a = 0; a < N; ++a) {
    const double m = ((mv->species_mask >> types[a]) & 1) ? 1.0 : 0.0;
    for (int q = 0; q < 3; ++q) {
      double z = xi[q * Nmax + a];
      v[3 * a + q] = fresh ? z : d[3 * a + q] + mv->perturb_angle * z;
      if (scoped) v[3 * a + q] = v[3 * a + q] * m;
    }
  }
  double vn = sqrt(wave_sum_dot(v, v, N, scr));
  double den = fmax(vn, 1e-10);
  for (int i = 0; i < 3 * N; ++i) d[i] = v[i] / den;
  memcpy(x0, x, sizeof(double) * 3 * (size_t)N);
  memcpy(v, d, sizeof(double) * 3 * (size_t)N); /* v := initial direction */

  double H = *E;
  for (int r = 0; r < mv->n_reflect; ++r) {
    for (int i = 0; i < 3 * N; ++i) x[i] = x[i] + s * d[i];
    double U, Hn;
    or_energy_forces(N, x, types, cell, c->pot, c->pressure, c->mu, &U, &Hn,
                     F);
    if (isnan(Hn)) Hn = INFINITY;
    if (scoped)
      for (int a = 0; a < N; ++a) {
        const double m = ((mv->species_mask >> types[a]) & 1) ? 1.0 : 0.0;
        for (int q = 0; q < 3; ++q) F[3 * a + q] = F[3 * a + q] * m;
      }
    for (int a = 0; a < N; ++a) {
      fq[a] = dot3(F[3 * a], F[3 * a + 1], F[3 * a + 2], F[3 * a],
                   F[3 * a + 1], F[3 * a + 2]);
      fd[a] = dot3(F[3 * a], F[3 * a + 1], F[3 * a + 2], d[3 * a],
                   d[3 * a + 1], d[3 * a + 2]);
    }
    double fsq = wave_sum_atoms(fq, N);
    double fdot = wave_sum_atoms(fd, N);
    if (Hn >= c->emax) {
      double fn = fmax(sqrt(fsq), 1e-10);
      double t = 2.0 * (fdot / fn);
      for (int i = 0; i < 3 * N; ++i) d[i] = fma(-t, F[i] / fn, d[i]);
    }
    H = Hn;
  }
  int accepted = H < c->emax;
  if (accepted) {
    *E = H;
  } else {
    memcpy(x, x0, sizeof(double) * 3 * (size_t)N);
    for (int i = 0; i < 3 * N; ++i) d[i] = -v[i];
  }
  free(scr);
  free(v);
  free(x0);
  free(F);
  free(fq);
  free(fd);
  return accepted;
}

/* sampling/moves/volume.py:49-109 + _common.py:16-93.  Returns reason. */
static int volume_move(const chain_ctx* c, double* x, const int32_t* types,
                       double* cell, double* E, double s, const or_move* mv,
                       double xi, double u) {
  int N = c->N;
  double oldV = fabs(or_det3(cell));
  double dV = (s * (double)N) * xi;
  double newV = fabs(oldV + dV);
  double ratio = newV / oldV;
  double scale = or_cbrt(ratio);
  double nc[9];
  for (int i = 0; i < 9; ++i) nc[i] = cell[i] * scale;
  double* nx = (double*)malloc(sizeof(double) * 3 * (size_t)(N + 1));
  for (int i = 0; i < 3 * N; ++i) nx[i] = x[i] * scale;
  double U, Hn;
  or_energy_forces(N, nx, types, nc, c->pot, c->pressure, c->mu, &U, &Hn,
                   NULL);
  int cell_ok = or_check_cell_shape(nc, N, mv->max_vpa, mv->min_vpa,
                                    mv->min_aspect);
  double p_acc = mv->flat_v_prior ? 1.0 : fmin(1.0, or_ipow(ratio, N));
  int energy_ok = Hn < c->emax;
  int prior_ok = u < p_acc;
  int reason;
  if (energy_ok && cell_ok && prior_ok) {
    reason = 0;
    memcpy(x, nx, sizeof(double) * 3 * (size_t)N);
    memcpy(cell, nc, sizeof(nc));
    *E = Hn;
  } else {
    reason = !energy_ok ? 1 : (!cell_ok ? 2 : 3);
  }
  free(nx);
  return reason;
}

/* _common.py:16-93 without the prior factor, for the volume-preserving
 * cell moves: positions x' = x T, one energy evaluation, accepted iff
 * H' < Emax and the cell passes the geometry guard; reason energy > cell. */
static int cell_shape_move(const chain_ctx* c, double* x,
                           const int32_t* types, double* cell, double* E,
                           const or_move* mv, const double* nc,
                           const double* T) {
  int N = c->N;
  double* nx = (double*)malloc(sizeof(double) * 3 * (size_t)(N + 1));
  for (int a = 0; a < N; ++a)
    for (int k = 0; k < 3; ++k)
      nx[3 * a + k] = dot3(x[3 * a], x[3 * a + 1], x[3 * a + 2], T[k],
                           T[3 + k], T[6 + k]);
  double U, Hn;
  or_energy_forces(N, nx, types, nc, c->pot, c->pressure, c->mu, &U, &Hn,
                   NULL);
  int cell_ok = or_check_cell_shape(nc, N, mv->max_vpa, mv->min_vpa,
                                    mv->min_aspect);
  int energy_ok = Hn < c->emax;
  int reason = !energy_ok ? 1 : (!cell_ok ? 2 : 0);
  if (reason == 0) {
    memcpy(x, nx, sizeof(double) * 3 * (size_t)N);
    memcpy(cell, nc, 9 * sizeof(double));
    *E = Hn;
  }
  free(nx);
  return reason;
}

/* sampling/moves/shear.py:19-120: cell vector idx = floor(3 u) gains
 * rv1 e1 + rv2 e2 (rv = step * normal), e1 the unit vector of the first
 * other cell vector v1, e2 that of the second one's part orthogonal to e1;
 * positions follow by T = old^-1 new (utils/cell.py:128-152). */
static int shear_move(const chain_ctx* c, double* x, const int32_t* types,
                      double* cell, double* E, double s, const or_move* mv,
                      double xi0, double xi1, double u) {
  int idx = (int)(u * 3.0);
  if (idx > 2) idx = 2;
  const double* v1 = cell + 3 * (idx == 0 ? 1 : 0);
  const double* v2 = cell + 3 * (idx == 2 ? 1 : 2);
  double n1 = sqrt(dot3(v1[0], v1[1], v1[2], v1[0], v1[1], v1[2]));
  double e1[3], e2[3], w[3];
  for (int d = 0; d < 3; ++d) e1[d] = v1[d] / n1;
  double proj = dot3(v2[0], v2[1], v2[2], e1[0], e1[1], e1[2]);
  for (int d = 0; d < 3; ++d) w[d] = v2[d] - proj * e1[d];
  double n2 = sqrt(dot3(w[0], w[1], w[2], w[0], w[1], w[2]));
  for (int d = 0; d < 3; ++d) e2[d] = w[d] / n2;
  double rv1 = s * xi0, rv2 = s * xi1;
  double nc[9], inv[9], T[9];
  memcpy(nc, cell, sizeof(nc));
  for (int d = 0; d < 3; ++d)
    nc[3 * idx + d] = (cell[3 * idx + d] + rv1 * e1[d]) + rv2 * e2[d];
  inv3(cell, or_det3(cell), inv);
  for (int i = 0; i < 3; ++i)
    for (int k = 0; k < 3; ++k)
      T[3 * i + k] = dot3(inv[3 * i], inv[3 * i + 1], inv[3 * i + 2], nc[k],
                          nc[3 + k], nc[6 + k]);
  return cell_shape_move(c, x, types, cell, E, mv, nc, T);
}

/* sampling/moves/stretch.py:37-97: axis pair floor(3 u) of (0,1), (0,2),
 * (1,2); D = diag with exp(rv) on the first axis and exp(-rv) on the second
 * (rv = step * normal); cell' = cell D, x' = x D. */
static int stretch_move(const chain_ctx* c, double* x, const int32_t* types,
                        double* cell, double* E, double s, const or_move* mv,
                        double xi0, double u) {
  static const int axes[3][2] = {{0, 1}, {0, 2}, {1, 2}};
  int pr = (int)(u * 3.0);
  if (pr > 2) pr = 2;
  double rv = s * xi0;
  double d[3] = {1.0, 1.0, 1.0};
  d[axes[pr][0]] = or_exp(rv);
  d[axes[pr][1]] = or_exp(-rv);
  double nc[9], T[9];
  for (int r = 0; r < 3; ++r)
    for (int k = 0; k < 3; ++k) {
      nc[3 * r + k] = cell[3 * r + k] * d[k];
      T[3 * r + k] = r == k ? d[k] : 0.0;
    }
  return cell_shape_move(c, x, types, cell, E, mv, nc, T);
}

/* sampling/moves/hmc.py:51-152.  Forces at a position are evaluated once and
 * reused by the next half kick (bit-identical to re-evaluating). */
static int hmc_move(const chain_ctx* c, double* x, const int32_t* types,
                    const double* cell, double* E, double s,
                    const or_move* mv, const double* xi) {
  int N = c->N, Nmax = c->Nmax;
  double* p = (double*)malloc(sizeof(double) * 3 * (size_t)(N + 1));
  double* nx = (double*)malloc(sizeof(double) * 3 * (size_t)(N + 1));
  double* F = (double*)malloc(sizeof(double) * 3 * (size_t)(N + 1));
  double* scr = (double*)malloc(sizeof(double) * (size_t)(N + 1));
  for (int a = 0; a < N; ++a)
    for (int q = 0; q < 3; ++q) p[3 * a + q] = xi[q * Nmax + a];
  double K0 = 0.5 * wave_sum_dot(p, p, N, scr);
  memcpy(nx, x, sizeof(double) * 3 * (size_t)N);
  double U, Hn;
  double h = 0.5 * s;
  or_energy_forces(N, nx, types, cell, c->pot, c->pressure, c->mu, &U, &Hn,
                   F);
  for (int l = 0; l < mv->n_leapfrog; ++l) {
    for (int i = 0; i < 3 * N; ++i) p[i] = p[i] + h * F[i];
    for (int i = 0; i < 3 * N; ++i) nx[i] = nx[i] + s * p[i];
    or_energy_forces(N, nx, types, cell, c->pot, c->pressure, c->mu, &U,
                     &Hn, F);
    for (int i = 0; i < 3 * N; ++i) p[i] = p[i] + h * F[i];
  }
  double K1 = 0.5 * wave_sum_dot(p, p, N, scr);
  double dH = (Hn + K1) - (*E + K0);
  int accepted = (Hn < c->emax) && (dH < 1.0);
  if (accepted) {
    memcpy(x, nx, sizeof(double) * 3 * (size_t)N);
    *E = Hn;
  }
  free(p);
  free(nx);
  free(F);
  free(scr);
  return accepted;
}

/* sampling/moves/alchemical.py:40-79 (build_morph_kernel). */
static int morph_move(const chain_ctx* c, const double* x, int32_t* types,
                      const double* cell, double* E, const or_move* mv,
                      const double* u) {
  int N = c->N;
  int S = mv->n_species;
  int atom = (int)(u[0] * (double)N);
  if (atom > N - 1) atom = N - 1;
  int cand = (int)(u[1] * (double)(S - 1));
  if (cand > S - 2) cand = S - 2;
  int cur = types[atom];
  int nt = cand + (cand >= cur ? 1 : 0);
  types[atom] = nt;
  double U, Hn;
  or_energy_forces(N, x, types, cell, c->pot, c->pressure, c->mu, &U, &Hn,
                   NULL);
  int accepted = Hn < c->emax;
  if (accepted)
    *E = Hn;
  else
    types[atom] = cur;
  return accepted;
}

/* sampling/moves/swap.py:119-262 (species_swap).  Species pairs (a < b, both
 * in scope, ascending codes) weighted by n_a n_b: the first pair whose
 * cumulative weight exceeds u0 * total (swap.py:210-218); atom floor(u1 n_a)
 * of species a and floor(u2 n_b) of species b (the k-th atoms of their
 * species in atom order) exchange types; accepted iff an unlike pair exists
 * and H' < Emax. */
static int nth_of_species(const int32_t* types, int N, int code, int k) {
  int seen = 0;
  for (int a = 0; a < N; ++a)
    if (types[a] == code && seen++ == k) return a;
  return N - 1;
}

static int swap_move(const chain_ctx* c, const double* x, int32_t* types,
                     const double* cell, double* E, const or_move* mv,
                     const double* u) {
  int N = c->N, S = c->S;
  int scope = mv->species_mask ? mv->species_mask : (1 << S) - 1;
  int counts[16] = {0};
  for (int a = 0; a < N; ++a) counts[types[a]] += 1;
  double total = 0.0;
  for (int a = 0; a < S; ++a)
    for (int b = a + 1; b < S; ++b)
      if (((scope >> a) & 1) && ((scope >> b) & 1))
        total += (double)counts[a] * (double)counts[b];
  if (!(total > 0.0)) return 0;
  double target = u[0] * total, cum = 0.0;
  int lo = -1, hi = -1, last_lo = 0, last_hi = 1;
  for (int a = 0; a < S; ++a)
    for (int b = a + 1; b < S; ++b)
      if (((scope >> a) & 1) && ((scope >> b) & 1)) {
        cum += (double)counts[a] * (double)counts[b];
        last_lo = a;
        last_hi = b;
        if (lo < 0 && target < cum) {
          lo = a;
          hi = b;
        }
      }
  if (lo < 0) {
    lo = last_lo;
    hi = last_hi;
  }
  int nlo = counts[lo] > 1 ? counts[lo] : 1, nhi = counts[hi] > 1 ? counts[hi] : 1;
  int klo = (int)(u[1] * (double)nlo);
  if (klo > nlo - 1) klo = nlo - 1;
  int khi = (int)(u[2] * (double)nhi);
  if (khi > nhi - 1) khi = nhi - 1;
  int ia = nth_of_species(types, N, lo, klo), ib = nth_of_species(types, N, hi, khi);
  types[ia] = hi;
  types[ib] = lo;
  double U, Hn;
  or_energy_forces(N, x, types, cell, c->pot, c->pressure, c->mu, &U, &Hn, NULL);
  if (Hn < c->emax) {
    *E = Hn;
    return 1;
  }
  types[ia] = lo;
  types[ib] = hi;
  return 0;
}

static void count(int32_t* ctr, int reason, int n_eval, int n_grad) {
  ctr[0] += 1;
  if (reason < 4) ctr[1 + reason] += 1; /* 4: constraint, in no bucket */
  ctr[5] += n_eval;
  ctr[6] += n_grad;
}

int or_min_distance_ok(int N, const double* pos, const int32_t* types,
                       const double* cell_in, const double* floors, int S) {
  double det = or_det3(cell_in);
  double cell[9], inv[9];
  if (det == 0.0) { /* the stand-in cube of the energy, lj.py:138-139 */
    memset(cell, 0, sizeof(cell));
    cell[0] = cell[4] = cell[8] = 1.0e10;
    det = or_det3(cell);
  } else {
    memcpy(cell, cell_in, sizeof(cell));
  }
  inv3(cell, det, inv);
  for (int i = 0; i < N; ++i)
    for (int j = i + 1; j < N; ++j) {
      const double dx = pos[3 * j] - pos[3 * i];
      const double dy = pos[3 * j + 1] - pos[3 * i + 1];
      const double dz = pos[3 * j + 2] - pos[3 * i + 2];
      double f0 = fma(dz, inv[6], fma(dy, inv[3], dx * inv[0]));
      double f1 = fma(dz, inv[7], fma(dy, inv[4], dx * inv[1]));
      double f2 = fma(dz, inv[8], fma(dy, inv[5], dx * inv[2]));
      f0 -= rint(f0);
      f1 -= rint(f1);
      f2 -= rint(f2);
      const double ex = fma(f2, cell[6], fma(f1, cell[3], f0 * cell[0]));
      const double ey = fma(f2, cell[7], fma(f1, cell[4], f0 * cell[1]));
      const double ez = fma(f2, cell[8], fma(f1, cell[5], f0 * cell[2]));
      const double d2 = fma(ez, ez, fma(ey, ey, ex * ex));
      if (sqrt(d2) < floors[types[i] * S + types[j]]) return 0;
    }
  return 1;
}

/* sampling/mwg.py:108-150,227-257 + nested_sampling.py:533-609 (one chain). */
void or_walk_chain(int N, int Nmax, double* pos, int32_t* types, double* cell,
                   double* energy, double* dir, double* dirx, size_t xstride,
                   const or_pot* pot,
                   double pressure, const double* mu, double emax,
                   const double* step_sizes, const or_move* moves,
                   int n_moves, const int32_t* schedule, int n_steps,
                   const double* normals, const double* uniforms,
                   int32_t* counters) {
  chain_ctx c = {N, Nmax, pot->n_species, pot, pressure, mu, emax};
  (void)n_moves;
  double* pos0 = (double*)malloc(sizeof(double) * 3 * (size_t)(N > 0 ? N : 1));
  int32_t* types0 = (int32_t*)malloc(sizeof(int32_t) * (size_t)(N > 0 ? N : 1));
  for (int t = 0; t < n_steps; ++t) {
    int m = schedule[t];
    const or_move* mv = &moves[m];
    const double* xi = normals + (size_t)t * 3 * Nmax;
    const double* u = uniforms + (size_t)t * 4;
    double s = step_sizes[m];
    int32_t* ctr = counters + OR_NCOUNT * m;
    /* sampling/mwg.py:98-141: the constraint gate of this move, if any */
    const int species_move = mv->type == OR_MOVE_MORPH || mv->type == OR_MOVE_SWAP;
    const double* floors = species_move ? pot->dmin_types : pot->dmin_pos;
    double cell0[9], energy0 = *energy;
    if (floors) {
      memcpy(pos0, pos, sizeof(double) * 3 * (size_t)N);
      memcpy(types0, types, sizeof(int32_t) * (size_t)N);
      memcpy(cell0, cell, sizeof(cell0));
    }
    int reason = 0, ne = 1, ng = 0;
    switch (mv->type) {
      case OR_MOVE_GMC: {
        double* dd = (mv->dir_slot > 0 && dirx) ? dirx + (size_t)(mv->dir_slot - 1) * xstride : dir;
        reason = gmc_move(&c, pos, types, cell, energy, dd, s, mv, xi) ? 0 : 1;
        ne = ng = mv->n_reflect;
        break;
      }
      case OR_MOVE_VOLUME:
        reason = volume_move(&c, pos, types, cell, energy, s, mv, xi[0], u[0]);
        break;
      case OR_MOVE_HMC:
        reason = hmc_move(&c, pos, types, cell, energy, s, mv, xi) ? 0 : 1;
        ne = 2 * mv->n_leapfrog + 1;
        ng = 2 * mv->n_leapfrog;
        break;
      case OR_MOVE_MORPH:
        reason = morph_move(&c, pos, types, cell, energy, mv, u) ? 0 : 1;
        break;
      case OR_MOVE_SWAP:
        reason = swap_move(&c, pos, types, cell, energy, mv, u) ? 0 : 1;
        break;
      case OR_MOVE_SHEAR:
        reason = shear_move(&c, pos, types, cell, energy, s, mv, xi[0], xi[Nmax], u[0]);
        break;
      case OR_MOVE_STRETCH:
        reason = stretch_move(&c, pos, types, cell, energy, s, mv, xi[0], u[0]);
        break;
      default:
        continue;
    }
    /* an accepted proposal that breaks the constraint is undone (positions,
     * cell, species, energy; the GMC direction keeps the move's value) */
    if (floors && reason == 0 &&
        !or_min_distance_ok(N, pos, types, cell, floors, pot->n_species)) {
      memcpy(pos, pos0, sizeof(double) * 3 * (size_t)N);
      memcpy(types, types0, sizeof(int32_t) * (size_t)N);
      memcpy(cell, cell0, sizeof(cell0));
      *energy = energy0;
      reason = 4;
    }
    count(ctr, reason, ne, ng);
  }
  free(pos0);
  free(types0);
}

/* ---------------------------------------------------------------------- */
/* NS bookkeeping                                                         */
/* ---------------------------------------------------------------------- */

static void copy_walker(or_pop* pop, int dst, int src) {
  int Nm = pop->Nmax;
  memcpy(pop->pos + (size_t)dst * Nm * 3, pop->pos + (size_t)src * Nm * 3,
         sizeof(double) * 3 * (size_t)Nm);
  memcpy(pop->dir + (size_t)dst * Nm * 3, pop->dir + (size_t)src * Nm * 3,
         sizeof(double) * 3 * (size_t)Nm);
  const size_t xs = (size_t)pop->R * pop->K * Nm * 3;
  for (int e = 0; e < pop->n_xslots; ++e)
    memcpy(pop->dirx + e * xs + (size_t)dst * Nm * 3, pop->dirx + e * xs + (size_t)src * Nm * 3,
           sizeof(double) * 3 * (size_t)Nm);
  memcpy(pop->types + (size_t)dst * Nm, pop->types + (size_t)src * Nm,
         sizeof(int32_t) * (size_t)Nm);
  memcpy(pop->cell + (size_t)dst * 9, pop->cell + (size_t)src * 9,
         sizeof(double) * 9);
  pop->energy[dst] = pop->energy[src];
  pop->n_atoms[dst] = pop->n_atoms[src];
}

typedef struct {
  double key;
  int idx;
} keyed_t;

static int cmp_desc(const void* a, const void* b) {
  const keyed_t* x = (const keyed_t*)a;
  const keyed_t* y = (const keyed_t*)b;
  if (x->key > y->key) return -1;
  if (x->key < y->key) return 1;
  return (x->idx > y->idx) - (x->idx < y->idx);
}

/* nested_sampling.py:86-133,151-186 (n_cull = 1 path): worst with uniform
 * tie-break among exact maxima, clone shifted past worst, extras = top
 * n_extra of uniform noise excluding worst and clone. */
static void select_replica(const double* E, int K, int n_extra,
                           const double* sel_u, double clone_u,
                           const double* extra_u, int* worst, int* clone,
                           int* walk, double* emax) {
  double mx = E[0];
  for (int k = 1; k < K; ++k) mx = fmax(mx, E[k]);
  int w = -1;
  double best = -INFINITY;
  for (int k = 0; k < K; ++k) {
    if (E[k] == mx && (w < 0 || sel_u[k] > best)) {
      best = sel_u[k];
      w = k;
    }
  }
  int cl = (int)(clone_u * (double)(K - 1));
  if (cl > K - 2) cl = K - 2;
  if (cl >= w) cl += 1;
  walk[0] = w;
  if (n_extra > 0) {
    keyed_t* ks = (keyed_t*)malloc(sizeof(keyed_t) * (size_t)K);
    for (int k = 0; k < K; ++k) {
      ks[k].key = (k == w || k == cl) ? -INFINITY : extra_u[k];
      ks[k].idx = k;
    }
    qsort(ks, (size_t)K, sizeof(keyed_t), cmp_desc);
    for (int e = 0; e < n_extra; ++e) walk[1 + e] = ks[e].idx;
    free(ks);
  }
  *worst = w;
  *clone = cl;
  *emax = mx;
}

typedef struct {
  double e, u;
  int idx;
} lex_t;

/* potential descending, then noise descending, then index: the order of
 * jnp.lexsort((-noise, -potentials)), which is stable */
static int cmp_lex(const void* a, const void* b) {
  const lex_t* x = (const lex_t*)a;
  const lex_t* y = (const lex_t*)b;
  if (x->e != y->e) return x->e > y->e ? -1 : 1;
  if (x->u != y->u) return x->u > y->u ? -1 : 1;
  return (x->idx > y->idx) - (x->idx < y->idx);
}

static int cmp_asc(const void* a, const void* b) {
  const keyed_t* x = (const keyed_t*)a;
  const keyed_t* y = (const keyed_t*)b;
  if (x->key < y->key) return -1;
  if (x->key > y->key) return 1;
  return (x->idx > y->idx) - (x->idx < y->idx);
}

/* nested_sampling.py:188-212 (n_cull = C > 1): the first C walkers of the
 * lexicographic order are culled in that order and the contour is the
 * potential of the C-th; the survivors order[C:] are permuted
 * (jax.random.permutation) by sorting them on the keys perm_u[p] of their
 * positions p ascending, equal keys by walker index; the first C of the
 * permutation are the clone sources, the next n_extra the extra chains.
 * walk = culled slots, then extras. */
static double select_cull(const double* E, int K, int C, int n_extra,
                          const double* sel_u, const double* perm_u, int* worst,
                          int* clone, int* walk, double* dead_e) {
  lex_t* o = (lex_t*)malloc(sizeof(lex_t) * (size_t)K);
  for (int k = 0; k < K; ++k) {
    o[k].e = E[k];
    o[k].u = sel_u[k];
    o[k].idx = k;
  }
  qsort(o, (size_t)K, sizeof(lex_t), cmp_lex);
  for (int c = 0; c < C; ++c) {
    worst[c] = walk[c] = o[c].idx;
    dead_e[c] = o[c].e;
  }
  const double contour = o[C - 1].e;
  const int n_surv = K - C;
  keyed_t* sv = (keyed_t*)malloc(sizeof(keyed_t) * (size_t)n_surv);
  for (int p = 0; p < n_surv; ++p) {
    sv[p].key = perm_u[p];
    sv[p].idx = o[C + p].idx;
  }
  qsort(sv, (size_t)n_surv, sizeof(keyed_t), cmp_asc);
  for (int c = 0; c < C; ++c) clone[c] = sv[c].idx;
  for (int e = 0; e < n_extra; ++e) walk[C + e] = sv[C + e].idx;
  free(sv);
  free(o);
  return contour;
}

static void rebase_pair(double E, const double* cell, const int32_t* types,
                        int N, double Pf, const double* muf, double Pt,
                        const double* mut, int S, double* out) {
  double u = E - or_ensemble_energy(N, types, cell, Pf, muf, S);
  *out = u + or_ensemble_energy(N, types, cell, Pt, mut, S);
}

/* sampling/moves/replica_exchange.py:318-436 with EnsembleSwap (:120-164). */
void or_replica_exchange(or_pop* pop, const or_run* run, const double* emax,
                         int n_swap_cycles, const int32_t* pair_valid,
                         const double* re_u, int32_t* acc, int32_t* att) {
  int R = pop->R, K = pop->K, S = pop->S, Nm = pop->Nmax;
  int max_pairs = R / 2;
  if (R < 2) return;
  double* tmp_pos = (double*)malloc(sizeof(double) * 3 * (size_t)Nm);
  double* tmp_dir = (double*)malloc(sizeof(double) * 3 * (size_t)Nm);
  int32_t* tmp_t = (int32_t*)malloc(sizeof(int32_t) * (size_t)Nm);
  for (int cyc = 0; cyc < n_swap_cycles; ++cyc) {
    for (int ph = 0; ph < 2; ++ph) {
      int pi = 0;
      for (int i = ph; i + 1 < R; i += 2, ++pi) {
        int j = i + 1;
        const double* u = re_u + (((size_t)cyc * 2 + ph) * max_pairs + pi) * 2;
        if (!pair_valid[i]) continue;
        int wi = (int)(u[0] * (double)K);
        if (wi > K - 1) wi = K - 1;
        int wj = (int)(u[1] * (double)K);
        if (wj > K - 1) wj = K - 1;
        int si = i * K + wi, sj = j * K + wj;
        int Ni = pop->n_atoms[si], Nj = pop->n_atoms[sj];
        double into_i, into_j;
        rebase_pair(pop->energy[sj], pop->cell + 9 * (size_t)sj,
                    pop->types + (size_t)sj * Nm, Nj, run->pressure[j],
                    run->mu + (size_t)j * S, run->pressure[i],
                    run->mu + (size_t)i * S, S, &into_i);
        rebase_pair(pop->energy[si], pop->cell + 9 * (size_t)si,
                    pop->types + (size_t)si * Nm, Ni, run->pressure[i],
                    run->mu + (size_t)i * S, run->pressure[j],
                    run->mu + (size_t)j * S, S, &into_j);
        int ok = (into_i < emax[i]) && (into_j < emax[j]);
        att[i] += 1;
        if (ok) {
          acc[i] += 1;
          /* swap walker-bound fields (state/mc_state.py:117-137) */
          memcpy(tmp_pos, pop->pos + (size_t)si * Nm * 3,
                 sizeof(double) * 3 * (size_t)Nm);
          memcpy(pop->pos + (size_t)si * Nm * 3,
                 pop->pos + (size_t)sj * Nm * 3,
                 sizeof(double) * 3 * (size_t)Nm);
          memcpy(pop->pos + (size_t)sj * Nm * 3, tmp_pos,
                 sizeof(double) * 3 * (size_t)Nm);
          memcpy(tmp_dir, pop->dir + (size_t)si * Nm * 3,
                 sizeof(double) * 3 * (size_t)Nm);
          memcpy(pop->dir + (size_t)si * Nm * 3,
                 pop->dir + (size_t)sj * Nm * 3,
                 sizeof(double) * 3 * (size_t)Nm);
          memcpy(pop->dir + (size_t)sj * Nm * 3, tmp_dir,
                 sizeof(double) * 3 * (size_t)Nm);
          for (int e = 0; e < pop->n_xslots; ++e) {
            double* X = pop->dirx + (size_t)e * pop->R * pop->K * Nm * 3;
            memcpy(tmp_dir, X + (size_t)si * Nm * 3, sizeof(double) * 3 * (size_t)Nm);
            memcpy(X + (size_t)si * Nm * 3, X + (size_t)sj * Nm * 3,
                   sizeof(double) * 3 * (size_t)Nm);
            memcpy(X + (size_t)sj * Nm * 3, tmp_dir, sizeof(double) * 3 * (size_t)Nm);
          }
          memcpy(tmp_t, pop->types + (size_t)si * Nm,
                 sizeof(int32_t) * (size_t)Nm);
          memcpy(pop->types + (size_t)si * Nm, pop->types + (size_t)sj * Nm,
                 sizeof(int32_t) * (size_t)Nm);
          memcpy(pop->types + (size_t)sj * Nm, tmp_t,
                 sizeof(int32_t) * (size_t)Nm);
          double tc[9];
          memcpy(tc, pop->cell + 9 * (size_t)si, sizeof(tc));
          memcpy(pop->cell + 9 * (size_t)si, pop->cell + 9 * (size_t)sj,
                 sizeof(tc));
          memcpy(pop->cell + 9 * (size_t)sj, tc, sizeof(tc));
          pop->n_atoms[si] = Nj;
          pop->n_atoms[sj] = Ni;
          pop->energy[si] = into_i;
          pop->energy[sj] = into_j;
        }
      }
    }
  }
  free(tmp_pos);
  free(tmp_dir);
  free(tmp_t);
}

/* nested_sampling.py:391-686, all replicas; n_cull walkers culled each. */
int or_ns_iteration(or_pop* pop, const or_pot* pot, const or_run* run,
                    const or_move* moves, int n_moves, int n_steps,
                    int n_extra, int n_cull, int n_swap_cycles,
                    const int32_t* pair_valid, const or_draws* d,
                    or_iter_out* out, int n_threads) {
  int R = pop->R, K = pop->K, Nm = pop->Nmax, S = pop->S;
  const int C = n_cull;
  int n_walk = C + n_extra;
  if (K < 2 || C < 1 || n_extra < 0) return -1;
  if (C == 1 && n_extra > K - 2) return -1;
  if (C > 1 && (2 * C + n_extra > K || !d->extra_u)) return -1;
  for (int r = 0; r < R; ++r) {
    const size_t o = (size_t)r * C;
    if (C == 1) {
      int w, cl;
      double mx;
      select_replica(pop->energy + (size_t)r * K, K, n_extra,
                     d->sel_u + (size_t)r * K, d->clone_u[r],
                     d->extra_u ? d->extra_u + (size_t)r * K : NULL, &w, &cl,
                     out->walk + (size_t)r * n_walk, &mx);
      out->worst[o] = w;
      out->clone[o] = cl;
      out->emax[r] = mx;
      out->dead_energy[o] = pop->energy[(size_t)r * K + w];
    } else {
      out->emax[r] = select_cull(pop->energy + (size_t)r * K, K, C, n_extra,
                                 d->sel_u + (size_t)r * K, d->extra_u + (size_t)r * K,
                                 out->worst + o, out->clone + o,
                                 out->walk + (size_t)r * n_walk, out->dead_energy + o);
    }
    for (int c = 0; c < C; ++c)
      out->dead_volume[o + c] =
          fabs(or_det3(pop->cell + 9 * ((size_t)r * K + out->worst[o + c])));
    /* sources and culled slots are disjoint: copies in any order */
    for (int c = 0; c < C; ++c)
      copy_walker(pop, r * K + out->worst[o + c], r * K + out->clone[o + c]);
  }
  memset(out->counters, 0, sizeof(int32_t) * (size_t)R * n_moves * OR_NCOUNT);
  int total = R * n_walk;
#ifdef _OPENMP
  if (n_threads > 0) omp_set_num_threads(n_threads);
#else
  (void)n_threads;
#endif
#pragma omp parallel
  {
    int32_t* local =
        (int32_t*)calloc((size_t)R * n_moves * OR_NCOUNT, sizeof(int32_t));
#pragma omp for schedule(dynamic, 1)
    for (int q = 0; q < total; ++q) {
      int r = q / n_walk, ci = q % n_walk;
      int slot = r * K + out->walk[(size_t)r * n_walk + ci];
      const double* xi = d->normals + ((size_t)q * n_steps) * 3 * Nm;
      const double* u = d->uniforms + ((size_t)q * n_steps) * 4;
      or_walk_chain(pop->n_atoms[slot], Nm, pop->pos + (size_t)slot * Nm * 3,
                    pop->types + (size_t)slot * Nm, pop->cell + 9 * (size_t)slot,
                    pop->energy + slot, pop->dir + (size_t)slot * Nm * 3,
                    pop->n_xslots ? pop->dirx + (size_t)slot * Nm * 3 : NULL,
                    (size_t)R * K * Nm * 3, pot,
                    run->pressure[r], run->mu + (size_t)r * S, out->emax[r],
                    run->step_sizes + (size_t)r * n_moves, moves, n_moves,
                    d->schedule, n_steps, xi, u,
                    local + (size_t)r * n_moves * OR_NCOUNT);
    }
#pragma omp critical
    {
      for (int i = 0; i < R * n_moves * OR_NCOUNT; ++i)
        out->counters[i] += local[i];
    }
    free(local);
  }
  if (d->re_u != NULL && R > 1) {
    memset(out->re_acc, 0, sizeof(int32_t) * (size_t)(R - 1));
    memset(out->re_att, 0, sizeof(int32_t) * (size_t)(R - 1));
    or_replica_exchange(pop, run, out->emax, n_swap_cycles, pair_valid,
                        d->re_u, out->re_acc, out->re_att);
  }
  return 0;
}

/* sampling/adaptation/stepsize_handler.py:215-261: one trial move on each
 * sampled walker (a copy), reporting accepted and reject reason. */
/* init/burn_in.py:48-148 (_one_walk): every walker of replica r walks
 * walklength sub-steps at the fixed contour emax[r] with one schedule for
 * all; normals [R][K][walklength][3][Nmax], uniforms [R][K][walklength][4];
 * counters [R][n_moves][7] are overwritten with the walk's sums. */
void or_initial_walk(or_pop* pop, const or_pot* pot, const or_run* run,
                     const or_move* moves, int n_moves, const double* emax,
                     const int32_t* schedule, int walklength,
                     const double* normals, const double* uniforms,
                     int32_t* counters, int n_threads) {
  int R = pop->R, K = pop->K, Nm = pop->Nmax, S = pop->S;
  memset(counters, 0, sizeof(int32_t) * (size_t)R * n_moves * OR_NCOUNT);
#ifdef _OPENMP
  if (n_threads > 0) omp_set_num_threads(n_threads);
#else
  (void)n_threads;
#endif
#pragma omp parallel
  {
    int32_t* local =
        (int32_t*)calloc((size_t)R * n_moves * OR_NCOUNT, sizeof(int32_t));
#pragma omp for schedule(dynamic, 1)
    for (int q = 0; q < R * K; ++q) {
      int r = q / K;
      const double* xi = normals + (size_t)q * walklength * 3 * Nm;
      const double* u = uniforms + (size_t)q * walklength * 4;
      or_walk_chain(pop->n_atoms[q], Nm, pop->pos + (size_t)q * Nm * 3,
                    pop->types + (size_t)q * Nm, pop->cell + 9 * (size_t)q,
                    pop->energy + q, pop->dir + (size_t)q * Nm * 3,
                    pop->n_xslots ? pop->dirx + (size_t)q * Nm * 3 : NULL,
                    (size_t)R * K * Nm * 3, pot,
                    run->pressure[r], run->mu + (size_t)r * S, emax[r],
                    run->step_sizes + (size_t)r * n_moves, moves, n_moves,
                    schedule, walklength, xi, u,
                    local + (size_t)r * n_moves * OR_NCOUNT);
    }
#pragma omp critical
    {
      for (int i = 0; i < R * n_moves * OR_NCOUNT; ++i) counters[i] += local[i];
    }
    free(local);
  }
}

void or_trial_moves(const or_pop* pop, const or_pot* pot, const or_run* run,
                    const or_move* moves, int n_moves, int move_idx,
                    const double* trial_step, const double* emax,
                    int n_samples, const int32_t* sample_idx,
                    const double* normals, const double* uniforms,
                    int32_t* accepted, int32_t* reason, int n_threads) {
  int R = pop->R, K = pop->K, Nm = pop->Nmax, S = pop->S;
  int total = R * n_samples;
#ifdef _OPENMP
  if (n_threads > 0) omp_set_num_threads(n_threads);
#else
  (void)n_threads;
#endif
#pragma omp parallel
  {
    double* pos = (double*)malloc(sizeof(double) * 3 * (size_t)Nm);
    double* dir = (double*)malloc(sizeof(double) * 3 * (size_t)Nm);
    double* dx = (double*)malloc(sizeof(double) * 3 * (size_t)Nm * (size_t)(pop->n_xslots + 1));
    int32_t* ty = (int32_t*)malloc(sizeof(int32_t) * (size_t)Nm);
    double* ss = (double*)malloc(sizeof(double) * (size_t)n_moves);
#pragma omp for schedule(dynamic, 1)
    for (int q = 0; q < total; ++q) {
      int r = q / n_samples;
      int slot = r * K + sample_idx[q];
      double cell[9];
      memcpy(pos, pop->pos + (size_t)slot * Nm * 3, sizeof(double) * 3 * (size_t)Nm);
      memcpy(dir, pop->dir + (size_t)slot * Nm * 3, sizeof(double) * 3 * (size_t)Nm);
      for (int e = 0; e < pop->n_xslots; ++e)
        memcpy(dx + (size_t)e * Nm * 3,
               pop->dirx + ((size_t)e * R * K + slot) * Nm * 3, sizeof(double) * 3 * (size_t)Nm);
      memcpy(ty, pop->types + (size_t)slot * Nm, sizeof(int32_t) * (size_t)Nm);
      memcpy(cell, pop->cell + 9 * (size_t)slot, sizeof(cell));
      double E = pop->energy[slot];
      for (int m = 0; m < n_moves; ++m) ss[m] = trial_step[r];
      int32_t ctr[OR_NCOUNT * 16];
      memset(ctr, 0, sizeof(ctr));
      int32_t sched = move_idx;
      or_walk_chain(pop->n_atoms[slot], Nm, pos, ty, cell, &E, dir, dx, (size_t)Nm * 3, pot,
                    run->pressure[r], run->mu + (size_t)r * S, emax[r], ss,
                    moves, n_moves, &sched, 1,
                    normals + (size_t)q * 3 * Nm, uniforms + (size_t)q * 4,
                    ctr);
      int32_t* c = ctr + OR_NCOUNT * move_idx;
      /* no bucket: the constraint gate undid the move (reason 4) */
      int rs = c[1] ? 0 : (c[2] ? 1 : (c[3] ? 2 : (c[4] ? 3 : 4)));
      accepted[q] = rs == 0;
      reason[q] = rs;
    }
    free(pos);
    free(dir);
    free(dx);
    free(ty);
    free(ss);
  }
}

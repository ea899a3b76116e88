// Wave-level LJ energy/forces and the MWG move kernels (GMC, volume, HMC,
// alchemical morph) for one walker held by one wavefront.
#pragma once

#include "jr_device.h"
#include "../../include/jaxrens_amd.h"

namespace jr {

struct DevMove {
  int type, n_reflect, n_leapfrog, flat_v_prior, n_species, pad_;
  double perturb_angle, max_vpa, min_vpa, min_aspect;
};

struct PairTab {
  double rc2;
  const double* eps4;  // LDS or global, [S*S]
  const double* eps24;
  const double* sig2;
  int S;
};

// Per-wave LDS view: positions/types of all atoms plus a scratch row used to
// return travelling reaction forces home.
struct WaveLds {
  double* x;
  double* y;
  double* z;
  int* t;
  double* hx;
  double* hy;
  double* hz;
};

struct Geom {
  double c[9];
  double inv[9];
  bool diag;
};

__device__ __forceinline__ void make_geom(const double* cell, Geom& g) {
  double det = det3(cell);
  if (det == 0.0) {  // backends/lj.py:138-139: non-periodic stand-in cube
#pragma unroll
    for (int i = 0; i < 9; ++i) g.c[i] = 0.0;
    g.c[0] = g.c[4] = g.c[8] = 1.0e10;
    det = det3(g.c);
  } else {
#pragma unroll
    for (int i = 0; i < 9; ++i) g.c[i] = cell[i];
  }
  inv3(g.c, det, g.inv);
  g.diag = (g.c[1] == 0.0) && (g.c[2] == 0.0) && (g.c[3] == 0.0) &&
           (g.c[5] == 0.0) && (g.c[6] == 0.0) && (g.c[7] == 0.0);
}

// One pair, dr = MIC(x_j - x_i) (backends/lj.py:141-197).  inv2 is forced
// to zero outside the cutoff or for an inactive ring position, which zeroes
// both the energy and the force scalar.
template <bool FORCES, bool DIAG>
__device__ __forceinline__ void lj_pair(double xi, double yi, double zi,
                                        double xj, double yj, double zj,
                                        const Geom& g, double eps4,
                                        double eps24, double sig2, double rc2,
                                        bool active, double& e, double& fs,
                                        double& dx, double& dy, double& dz) {
  dx = xj - xi;
  dy = yj - yi;
  dz = zj - zi;
  if (DIAG) {
    double n0 = rint(dx * g.inv[0]);
    double n1 = rint(dy * g.inv[4]);
    double n2 = rint(dz * g.inv[8]);
    dx = fma(-n0, g.c[0], dx);
    dy = fma(-n1, g.c[4], dy);
    dz = fma(-n2, g.c[8], dz);
  } else {
    double f0 = fma(dz, g.inv[6], fma(dy, g.inv[3], dx * g.inv[0]));
    double f1 = fma(dz, g.inv[7], fma(dy, g.inv[4], dx * g.inv[1]));
    double f2 = fma(dz, g.inv[8], fma(dy, g.inv[5], dx * g.inv[2]));
    double n0 = rint(f0), n1 = rint(f1), n2 = rint(f2);
    double ex = fma(-n2, g.c[6], fma(-n1, g.c[3], fma(-n0, g.c[0], dx)));
    double ey = fma(-n2, g.c[7], fma(-n1, g.c[4], fma(-n0, g.c[1], dy)));
    double ez = fma(-n2, g.c[8], fma(-n1, g.c[5], fma(-n0, g.c[2], dz)));
    dx = ex;
    dy = ey;
    dz = ez;
  }
  double r2 = dot3(dx, dy, dz, dx, dy, dz);
  double inv2 = 1.0 / r2;
  inv2 = (active && (r2 < rc2)) ? inv2 : 0.0;
  double s2 = sig2 * inv2;
  double s6 = (s2 * s2) * s2;
  double s12 = s6 * s6;
  e = eps4 * (s12 - s6);
  if (FORCES) fs = (eps24 * fma(2.0, s12, -s6)) * inv2;
}

// Energy partial (per lane) and forces of the configuration held in (x,y,z,t)
// registers; writes the configuration into the wave's LDS first.
// Pair schedule: for k = 1 .. (N-1)/2 every ring position p pairs with
// (p + k) mod N, keeping its own force and adding the reaction to a
// travelling accumulator that rotates one position per k; for even N the
// k = N/2 pairs are evaluated from both ends with half the energy each.
template <int A, bool FORCES, bool MULTI, bool DIAG>
__device__ __forceinline__ void wave_eval_impl(
    const WaveLds& L, int N, const Geom& g, const PairTab& tab,
    const double (&x)[A], const double (&y)[A], const double (&z)[A],
    const int (&t)[A], double& e_lane, double (&fx)[A], double (&fy)[A],
    double (&fz)[A]) {
  const int lane = lane_id();
  double e[A], ox[A], oy[A], oz[A], tx[A], ty[A], tz[A];
#pragma unroll
  for (int s = 0; s < A; ++s) {
    e[s] = 0.0;
    ox[s] = oy[s] = oz[s] = 0.0;
    tx[s] = ty[s] = tz[s] = 0.0;
  }
  double eps4 = tab.eps4[0], eps24 = tab.eps24[0], sig2 = tab.sig2[0];
  const int kmax = (N - 1) >> 1;
  for (int k = 1; k <= kmax; ++k) {
    if (FORCES && k > 1) {
      ring_rot1<A>(tx, lane, N);
      ring_rot1<A>(ty, lane, N);
      ring_rot1<A>(tz, lane, N);
    }
#pragma unroll
    for (int s = 0; s < A; ++s) {
      const int a = lane + 64 * s;
      const bool active = a < N;
      int j = a + k;
      if (j >= N) j -= N;
      if (!active) j = 0;
      double xj = L.x[j], yj = L.y[j], zj = L.z[j];
      if (MULTI) {
        int ab = t[s] * tab.S + L.t[j];
        eps4 = tab.eps4[ab];
        eps24 = tab.eps24[ab];
        sig2 = tab.sig2[ab];
      }
      double ep, fs = 0.0, dx, dy, dz;
      lj_pair<FORCES, DIAG>(x[s], y[s], z[s], xj, yj, zj, g, eps4, eps24,
                            sig2, tab.rc2, active, ep, fs, dx, dy, dz);
      e[s] = e[s] + ep;
      if (FORCES) {
        ox[s] = fma(-fs, dx, ox[s]);
        oy[s] = fma(-fs, dy, oy[s]);
        oz[s] = fma(-fs, dz, oz[s]);
        tx[s] = fma(fs, dx, tx[s]);
        ty[s] = fma(fs, dy, ty[s]);
        tz[s] = fma(fs, dz, tz[s]);
      }
    }
  }
  if (N >= 2 && !(N & 1)) {
    const int h = N >> 1;
#pragma unroll
    for (int s = 0; s < A; ++s) {
      const int a = lane + 64 * s;
      const bool active = a < N;
      int j = a + h;
      if (j >= N) j -= N;
      if (!active) j = 0;
      double xj = L.x[j], yj = L.y[j], zj = L.z[j];
      if (MULTI) {
        int ab = t[s] * tab.S + L.t[j];
        eps4 = tab.eps4[ab];
        eps24 = tab.eps24[ab];
        sig2 = tab.sig2[ab];
      }
      double ep, fs = 0.0, dx, dy, dz;
      lj_pair<FORCES, DIAG>(x[s], y[s], z[s], xj, yj, zj, g, eps4, eps24,
                            sig2, tab.rc2, active, ep, fs, dx, dy, dz);
      e[s] = e[s] + 0.5 * ep;
      if (FORCES) {
        ox[s] = fma(-fs, dx, ox[s]);
        oy[s] = fma(-fs, dy, oy[s]);
        oz[s] = fma(-fs, dz, oz[s]);
      }
    }
  }
  double el = e[0];
#pragma unroll
  for (int s = 1; s < A; ++s) {
    const int a = lane + 64 * s;
    if (a < N) el = el + e[s];
  }
  e_lane = el;
  if (FORCES) {
    // position p holds the reaction of atom (p + kmax) mod N
#pragma unroll
    for (int s = 0; s < A; ++s) {
      const int a = lane + 64 * s;
      if (a < N) {
        int dst = a + kmax;
        if (dst >= N) dst -= N;
        L.hx[dst] = tx[s];
        L.hy[dst] = ty[s];
        L.hz[dst] = tz[s];
      }
    }
    wave_lds_sync();
#pragma unroll
    for (int s = 0; s < A; ++s) {
      const int a = lane + 64 * s;
      if (a < N) {
        fx[s] = ox[s] + L.hx[a];
        fy[s] = oy[s] + L.hy[a];
        fz[s] = oz[s] + L.hz[a];
      } else {
        fx[s] = fy[s] = fz[s] = 0.0;
      }
    }
    wave_lds_sync();
  }
}

template <int A, bool FORCES, bool MULTI>
__device__ __forceinline__ void wave_eval(const WaveLds& L, int N,
                                          const Geom& g, const PairTab& tab,
                                          const double (&x)[A],
                                          const double (&y)[A],
                                          const double (&z)[A],
                                          const int (&t)[A], double& e_lane,
                                          double (&fx)[A], double (&fy)[A],
                                          double (&fz)[A]) {
  const int lane = lane_id();
#pragma unroll
  for (int s = 0; s < A; ++s) {
    const int a = lane + 64 * s;
    if (a < N) {
      L.x[a] = x[s];
      L.y[a] = y[s];
      L.z[a] = z[s];
      if (MULTI) L.t[a] = t[s];
    }
  }
  wave_lds_sync();
  if (g.diag)
    wave_eval_impl<A, FORCES, MULTI, true>(L, N, g, tab, x, y, z, t, e_lane,
                                           fx, fy, fz);
  else
    wave_eval_impl<A, FORCES, MULTI, false>(L, N, g, tab, x, y, z, t, e_lane,
                                            fx, fy, fz);
  wave_lds_sync();
}

// Lane partial of dot3 over active slots, then the wave butterfly.
template <int A>
__device__ __forceinline__ double wave_dot(int N, const double (&ax)[A],
                                           const double (&ay)[A],
                                           const double (&az)[A],
                                           const double (&bx)[A],
                                           const double (&by)[A],
                                           const double (&bz)[A]) {
  const int lane = lane_id();
  double q = dot3(ax[0], ay[0], az[0], bx[0], by[0], bz[0]);
#pragma unroll
  for (int s = 1; s < A; ++s) {
    const int a = lane + 64 * s;
    double v = dot3(ax[s], ay[s], az[s], bx[s], by[s], bz[s]);
    q = (a < N) ? q + v : q;
  }
  if (lane >= N) q = 0.0;
  return wave_sum(q);
}

// Species counts via ballots (wave-uniform).
template <int A>
__device__ __forceinline__ void species_counts(int N, int S, const int (&t)[A],
                                               int* counts) {
  const int lane = lane_id();
#pragma unroll
  for (int sp = 0; sp < JR_MAX_SPECIES; ++sp) {
    int c = 0;
#pragma unroll
    for (int s = 0; s < A; ++s) {
      const int a = lane + 64 * s;
      unsigned long long b = __ballot((a < N) && (t[s] == sp));
      c += __popcll(b);
    }
    counts[sp] = (sp < S) ? c : 0;
  }
}

// P*V - mu.counts (backends/ensemble.py:34-59).
__device__ __forceinline__ double ensemble_term(const double* cell,
                                                double pressure,
                                                const double* mu, int S,
                                                const int* counts) {
  double e = pressure * fabs(det3(cell));
  double acc = 0.0;
#pragma unroll
  for (int sp = 0; sp < JR_MAX_SPECIES; ++sp)
    if (sp < S) acc = acc + mu[sp] * (double)counts[sp];
  return e - acc;
}

}  // namespace jr

// MWG move kernels for one walker per wavefront (register-resident state).
#pragma once

#include "jr_walk.h"

namespace jr {

template <int A>
struct WalkerRegs {
  double x[A], y[A], z[A];     // positions
  double dx[A], dy[A], dz[A];  // GMC persistent direction
  int t[A];                    // species
  double cell[9];
  double E;                    // full Hamiltonian U + P V - mu.N
  int counts[JR_MAX_SPECIES];
};

struct WalkCtx {
  WaveLds L;
  PairTab tab;
  int N;
  int xi_stride;  // row stride of the normals a move reads (xi[c * xi_stride + atom])
  int S;
  double pressure;
  lds_ptr<const double> mu;  // the replica's chemical potentials (LDS, JR_MAX_SPECIES)
  double emax;
};

__device__ __forceinline__ double tape_xi(const double* xi, int stride, int c,
                                          int a) {
  return xi[c * stride + a];
}

// constraints/min_distance.py:25-46: true iff no two atoms are closer than
// their floor floors[t_i * S + t_j] (a floor of 0 disables the pair).  The
// separation is the minimum image dr - rint(dr C^-1) C in Cartesian
// coordinates with its Euclidean norm; every unordered pair is tested once
// on the ring schedule (k = 1 .. N/2).  Writes the configuration into the
// wave's LDS rows, which every energy evaluation rewrites first.
template <int A>
__device__ bool min_distance_ok(const WalkCtx& C, const WalkerRegs<A>& W,
                                const double* floors) {
  const int lane = lane_id();
  const int N = C.N;
  LdsView<A> L{C.L.base};
  Geom g;
  make_geom(W.cell, g);
#pragma unroll
  for (int s = 0; s < A; ++s) {
    const int a = lane + 64 * s;
    if (a < N) {
      L.x()[a] = L.x()[a + N] = W.x[s];
      L.y()[a] = L.y()[a + N] = W.y[s];
      L.z()[a] = L.z()[a + N] = W.z[s];
      L.t()[a] = L.t()[a + N] = W.t[s];
    }
  }
  wave_lds_sync();
  bool bad = false;
  for (int k = 1; k <= (N >> 1); ++k) {
#pragma unroll
    for (int s = 0; s < A; ++s) {
      const int a = lane + 64 * s;
      if (a < N) {
        const int j = a + k;
        const double dx = L.x()[j] - W.x[s];
        const double dy = L.y()[j] - W.y[s];
        const double dz = L.z()[j] - W.z[s];
        double f0 = fma(dz, g.inv[6], fma(dy, g.inv[3], dx * g.inv[0]));
        double f1 = fma(dz, g.inv[7], fma(dy, g.inv[4], dx * g.inv[1]));
        double f2 = fma(dz, g.inv[8], fma(dy, g.inv[5], dx * g.inv[2]));
        f0 -= rint(f0);
        f1 -= rint(f1);
        f2 -= rint(f2);
        const double ex = fma(f2, g.c[6], fma(f1, g.c[3], f0 * g.c[0]));
        const double ey = fma(f2, g.c[7], fma(f1, g.c[4], f0 * g.c[1]));
        const double ez = fma(f2, g.c[8], fma(f1, g.c[5], f0 * g.c[2]));
        const double d2 = fma(ez, ez, fma(ey, ey, ex * ex));
        bad = bad || (sqrt(d2) < floors[W.t[s] * C.S + L.t()[j]]);
      }
    }
  }
  wave_lds_sync();
  return __ballot(bad) == 0;
}

// A walker's GMC direction to / from memory (active atoms only).
template <int A>
__device__ __forceinline__ void store_dir(const WalkerRegs<A>& W, double* D, int N) {
  const int lane = lane_id();
#pragma unroll
  for (int s = 0; s < A; ++s) {
    const int at = lane + 64 * s;
    if (at < N) {
      D[3 * at + 0] = W.dx[s];
      D[3 * at + 1] = W.dy[s];
      D[3 * at + 2] = W.dz[s];
    }
  }
}

template <int A>
__device__ __forceinline__ void load_dir(WalkerRegs<A>& W, const double* D, int N) {
  const int lane = lane_id();
#pragma unroll
  for (int s = 0; s < A; ++s) {
    const int at = lane + 64 * s;
    W.dx[s] = at < N ? D[3 * at + 0] : 0.0;
    W.dy[s] = at < N ? D[3 * at + 1] : 0.0;
    W.dz[s] = at < N ? D[3 * at + 2] : 0.0;
  }
}

// sampling/moves/galilean.py:142-300 (use_forces=True).  A species scope
// (mv.species_mask != 0) multiplies the direction before each normalisation
// and the reflection force by the moving-atom mask (1 for types in scope,
// else 0; galilean.py:58-72,171-180,233-240), so the other atoms stay fixed;
// refresh_direction draws a fresh direction on every call.
template <int A, int CFG>
__device__ bool gmc_move(const WalkCtx& C, WalkerRegs<A>& W, double step,
                         const DevMove& mv, const double* xi) {
  constexpr bool MULTI = (CFG & 1) != 0;
  constexpr bool F32 = (CFG & 2) != 0;
  (void)F32;
  const int lane = lane_id();
  const int N = C.N;
  const bool scoped = mv.species_mask != 0;
  // moving-atom mask m_a of a scoped move
  auto msk = [&](int s) -> double { return ((mv.species_mask >> W.t[s]) & 1) ? 1.0 : 0.0; };
  double dn = sqrt(wave_dot<A>(N, W.dx, W.dy, W.dz, W.dx, W.dy, W.dz));
  const bool fresh = dn < 1e-8 || mv.refresh_direction;
  double vx[A], vy[A], vz[A];
#pragma unroll
  for (int s = 0; s < A; ++s) {
    const int a = lane + 64 * s;
    double zx = 0.0, zy = 0.0, zz = 0.0;
    if (a < N) {
      zx = tape_xi(xi, C.xi_stride, 0, a);
      zy = tape_xi(xi, C.xi_stride, 1, a);
      zz = tape_xi(xi, C.xi_stride, 2, a);
    }
    if (fresh) {
      vx[s] = zx;
      vy[s] = zy;
      vz[s] = zz;
    } else {
      vx[s] = W.dx[s] + mv.perturb_angle * zx;
      vy[s] = W.dy[s] + mv.perturb_angle * zy;
      vz[s] = W.dz[s] + mv.perturb_angle * zz;
    }
    if (scoped) {
      const double m = msk(s);
      vx[s] = vx[s] * m;
      vy[s] = vy[s] * m;
      vz[s] = vz[s] * m;
    }
  }
  double vn = sqrt(wave_dot<A>(N, vx, vy, vz, vx, vy, vz));
  double den = fmax(vn, 1e-10);
  double x0[A], y0[A], z0[A];
#pragma unroll
  for (int s = 0; s < A; ++s) {
    vx[s] = vx[s] / den;
    vy[s] = vy[s] / den;
    vz[s] = vz[s] / den;
    W.dx[s] = vx[s];
    W.dy[s] = vy[s];
    W.dz[s] = vz[s];
    x0[s] = W.x[s];
    y0[s] = W.y[s];
    z0[s] = W.z[s];
  }
  Geom g;
  make_geom(W.cell, g);
  const double ens = ensemble_term(W.cell, C.pressure, C.mu, C.S, W.counts);
  double H = W.E;
  for (int r = 0; r < mv.n_reflect; ++r) {
#pragma unroll
    for (int s = 0; s < A; ++s) {
      W.x[s] = W.x[s] + step * W.dx[s];
      W.y[s] = W.y[s] + step * W.dy[s];
      W.z[s] = W.z[s] + step * W.dz[s];
    }
    double fx[A], fy[A], fz[A];
    double U = wave_eval<A, true, CFG>(C.L, N, g, C.tab, W.x, W.y, W.z, W.t, fx, fy, fz);
    double Hn = U + ens;
    if (__builtin_isnan(Hn)) Hn = __builtin_inf();
    if (Hn >= C.emax) {  // wave-uniform: the force sums only where a reflection needs them
      if (scoped) {
#pragma unroll
        for (int s = 0; s < A; ++s) {
          const double m = msk(s);
          fx[s] = fx[s] * m;
          fy[s] = fy[s] * m;
          fz[s] = fz[s] * m;
        }
      }
      double fsq = wave_dot<A>(N, fx, fy, fz, fx, fy, fz);
      double fdt = wave_dot<A>(N, fx, fy, fz, W.dx, W.dy, W.dz);
      double fn = fmax(sqrt(fsq), 1e-10);
      double tt = 2.0 * (fdt / fn);
#pragma unroll
      for (int s = 0; s < A; ++s) {
        W.dx[s] = fma(-tt, fx[s] / fn, W.dx[s]);
        W.dy[s] = fma(-tt, fy[s] / fn, W.dy[s]);
        W.dz[s] = fma(-tt, fz[s] / fn, W.dz[s]);
      }
    }
    H = Hn;
  }
  const bool accepted = H < C.emax;
  if (accepted) {
    W.E = H;
  } else {
#pragma unroll
    for (int s = 0; s < A; ++s) {
      W.x[s] = x0[s];
      W.y[s] = y0[s];
      W.z[s] = z0[s];
      W.dx[s] = -vx[s];
      W.dy[s] = -vy[s];
      W.dz[s] = -vz[s];
    }
  }
  return accepted;
}

// sampling/moves/volume.py:49-109 + _common.py:16-93 (kernels without shape
// moves); returns the reject reason.
template <int A, int CFG>
__device__ int volume_move(const WalkCtx& C, WalkerRegs<A>& W, double step,
                           const DevMove& mv, double xi0, double u0) {
  const int N = C.N;
  double oldV = fabs(det3(W.cell));
  double dV = (step * (double)N) * xi0;
  double newV = fabs(oldV + dV);
  double ratio = newV / oldV;
  double scale = cbrt_c(ratio);
  double nc[9];
#pragma unroll
  for (int i = 0; i < 9; ++i) nc[i] = W.cell[i] * scale;
  double nx[A], ny[A], nz[A];
#pragma unroll
  for (int s = 0; s < A; ++s) {
    nx[s] = W.x[s] * scale;
    ny[s] = W.y[s] * scale;
    nz[s] = W.z[s] * scale;
  }
  Geom g;
  make_geom(nc, g);
  double fx[A], fy[A], fz[A];
  double U = wave_eval<A, false, CFG>(C.L, N, g, C.tab, nx, ny, nz, W.t, fx, fy, fz);
  double Hn = U + ensemble_term(nc, C.pressure, C.mu, C.S, W.counts);
  bool cell_ok =
      check_cell_shape(nc, N, mv.max_vpa, mv.min_vpa, mv.min_aspect);
  double p_acc = mv.flat_v_prior ? 1.0 : fmin(1.0, ipow_c(ratio, N));
  bool energy_ok = Hn < C.emax;
  bool prior_ok = u0 < p_acc;
  if (energy_ok && cell_ok && prior_ok) {
#pragma unroll
    for (int s = 0; s < A; ++s) {
      W.x[s] = nx[s];
      W.y[s] = ny[s];
      W.z[s] = nz[s];
    }
#pragma unroll
    for (int i = 0; i < 9; ++i) W.cell[i] = nc[i];
    W.E = Hn;
    return 0;
  }
  return !energy_ok ? 1 : (!cell_ok ? 2 : 3);
}

// The cell moves of kernels whose move table holds shear or stretch (cfg
// bit 3): volume (sampling/moves/volume.py:49-109), shear (shear.py:56-120)
// and stretch (stretch.py:37-97), with the accept tail of _common.py:16-93.
// Draws: xi0 = normal [0][0], xi1 = normal [1][0], u0 = uniform 0.
//   volume   V' = |V + s N xi0|, cell' = cell r, x' = x r, r = (V'/V)^(1/3),
//            prior min(1, (V'/V)^N) unless flat;
//   shear    cell vector floor(3 u0) moved by (s xi0, s xi1) in the plane of
//            the other two, x' = x old^-1 new;
//   stretch  axis pair floor(3 u0) of (0,1), (0,2), (1,2) scaled by exp(rv)
//            and exp(-rv), rv = s xi0: cell' = cell D, x' = x D.
// All three proposals are formed (a few hundred flops) and the move's picked
// by wave-uniform selects, so a single straight-line energy evaluation serves
// the three.  Returns the reject reason: energy 1 > cell 2 > prior 3.
template <int A, int CFG>
__device__ int cell_move(const WalkCtx& C, WalkerRegs<A>& W, double step,
                         const DevMove& mv, double xi0, double xi1,
                         double u0) {
  const int N = C.N;
  const bool iso = mv.type == JR_MOVE_VOLUME;
  const bool shear = mv.type == JR_MOVE_SHEAR;
  const double oldV = fabs(det3(W.cell));
  const double newV = fabs(oldV + (step * (double)N) * xi0);
  const double ratio = newV / oldV;
  const double scale = cbrt_c(ratio);
  int sel = (int)(u0 * 3.0);
  if (sel > 2) sel = 2;
  double ncs[9], Ts[9];
  shear_cell(W.cell, sel, step * xi0, step * xi1, ncs);
  cell_transform(W.cell, ncs, Ts);
  const double up = exp_c(step * xi0), dn = exp_c(-(step * xi0));
  const double d[3] = {sel < 2 ? up : 1.0,
                       sel == 0 ? dn : (sel == 2 ? up : 1.0),
                       sel == 0 ? 1.0 : dn};
  double nc[9], T[9];
#pragma unroll
  for (int r = 0; r < 3; ++r)
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      const int i = 3 * r + c;
      const double cv = W.cell[i] * scale, ct = W.cell[i] * d[c];
      const double tt = r == c ? d[c] : 0.0;
      nc[i] = iso ? cv : (shear ? ncs[i] : ct);
      T[i] = shear ? Ts[i] : tt;
    }
  const double p_acc =
      (iso && !mv.flat_v_prior) ? fmin(1.0, ipow_c(ratio, N)) : 1.0;
  double nx[A], ny[A], nz[A];
#pragma unroll
  for (int s = 0; s < A; ++s) {
    const double tx = dot3(W.x[s], W.y[s], W.z[s], T[0], T[3], T[6]);
    const double ty = dot3(W.x[s], W.y[s], W.z[s], T[1], T[4], T[7]);
    const double tz = dot3(W.x[s], W.y[s], W.z[s], T[2], T[5], T[8]);
    nx[s] = iso ? W.x[s] * scale : tx;
    ny[s] = iso ? W.y[s] * scale : ty;
    nz[s] = iso ? W.z[s] * scale : tz;
  }
  const bool cell_ok =
      check_cell_shape(nc, N, mv.max_vpa, mv.min_vpa, mv.min_aspect);
  const bool prior_ok = u0 < p_acc;
  Geom g;
  make_geom(nc, g);
  double fx[A], fy[A], fz[A];
  double U = wave_eval<A, false, CFG>(C.L, N, g, C.tab, nx, ny, nz, W.t, fx, fy, fz);
  double Hn = U + ensemble_term(nc, C.pressure, C.mu, C.S, W.counts);
  const bool energy_ok = Hn < C.emax;
  if (energy_ok && cell_ok && prior_ok) {
#pragma unroll
    for (int s = 0; s < A; ++s) {
      W.x[s] = nx[s];
      W.y[s] = ny[s];
      W.z[s] = nz[s];
    }
#pragma unroll
    for (int i = 0; i < 9; ++i) W.cell[i] = nc[i];
    W.E = Hn;
    return 0;
  }
  return !energy_ok ? 1 : (!cell_ok ? 2 : 3);
}

// sampling/moves/hmc.py:51-152; forces at a position are evaluated once and
// reused by the following half kick (identical values to re-evaluating).
template <int A, int CFG>
__device__ bool hmc_move(const WalkCtx& C, WalkerRegs<A>& W, double step,
                         const DevMove& mv, const double* xi) {
  constexpr bool MULTI = (CFG & 1) != 0;
  constexpr bool F32 = (CFG & 2) != 0;
  (void)F32;
  const int lane = lane_id();
  const int N = C.N;
  double px[A], py[A], pz[A], qx[A], qy[A], qz[A];
#pragma unroll
  for (int s = 0; s < A; ++s) {
    const int a = lane + 64 * s;
    px[s] = py[s] = pz[s] = 0.0;
    if (a < N) {
      px[s] = tape_xi(xi, C.xi_stride, 0, a);
      py[s] = tape_xi(xi, C.xi_stride, 1, a);
      pz[s] = tape_xi(xi, C.xi_stride, 2, a);
    }
    qx[s] = W.x[s];
    qy[s] = W.y[s];
    qz[s] = W.z[s];
  }
  double K0 = 0.5 * wave_dot<A>(N, px, py, pz, px, py, pz);
  Geom g;
  make_geom(W.cell, g);
  const double ens = ensemble_term(W.cell, C.pressure, C.mu, C.S, W.counts);
  const double h = 0.5 * step;
  double fx[A], fy[A], fz[A];
  double U = wave_eval<A, true, CFG>(C.L, N, g, C.tab, qx, qy, qz, W.t, fx,
                                       fy, fz);
  for (int l = 0; l < mv.n_leapfrog; ++l) {
#pragma unroll
    for (int s = 0; s < A; ++s) {
      px[s] = px[s] + h * fx[s];
      py[s] = py[s] + h * fy[s];
      pz[s] = pz[s] + h * fz[s];
      qx[s] = qx[s] + step * px[s];
      qy[s] = qy[s] + step * py[s];
      qz[s] = qz[s] + step * pz[s];
    }
    U = wave_eval<A, true, CFG>(C.L, N, g, C.tab, qx, qy, qz, W.t, fx, fy,
                                  fz);
#pragma unroll
    for (int s = 0; s < A; ++s) {
      px[s] = px[s] + h * fx[s];
      py[s] = py[s] + h * fy[s];
      pz[s] = pz[s] + h * fz[s];
    }
  }
  double Hn = U + ens;
  double K1 = 0.5 * wave_dot<A>(N, px, py, pz, px, py, pz);
  double dH = (Hn + K1) - (W.E + K0);
  bool accepted = (Hn < C.emax) && (dH < 1.0);
  if (accepted) {
#pragma unroll
    for (int s = 0; s < A; ++s) {
      W.x[s] = qx[s];
      W.y[s] = qy[s];
      W.z[s] = qz[s];
    }
    W.E = Hn;
  }
  return accepted;
}

// sampling/moves/alchemical.py:40-79 (build_morph_kernel).
template <int A, int CFG>
__device__ bool morph_move(const WalkCtx& C, WalkerRegs<A>& W,
                           const DevMove& mv, double u0, double u1) {
  constexpr bool MULTI = (CFG & 1) != 0;
  constexpr bool F32 = (CFG & 2) != 0;
  (void)F32;
  const int lane = lane_id();
  const int N = C.N;
  const int S = mv.n_species;
  int atom = (int)(u0 * (double)N);
  if (atom > N - 1) atom = N - 1;
  int cand = (int)(u1 * (double)(S - 1));
  if (cand > S - 2) cand = S - 2;
  const int al = atom & 63, as = atom >> 6;
  int curv = 0;
#pragma unroll
  for (int s = 0; s < A; ++s)
    if (s == as) curv = W.t[s];
  const int cur = __builtin_amdgcn_readlane(curv, al);
  const int nt = cand + (cand >= cur ? 1 : 0);
#pragma unroll
  for (int s = 0; s < A; ++s)
    if (s == as && lane == al) W.t[s] = nt;
#pragma unroll
  for (int sp = 0; sp < JR_MAX_SPECIES; ++sp)
    W.counts[sp] += (sp == nt ? 1 : 0) - (sp == cur ? 1 : 0);
  Geom g;
  make_geom(W.cell, g);
  double fx[A], fy[A], fz[A];
  double U = wave_eval<A, false, CFG>(C.L, N, g, C.tab, W.x, W.y, W.z, W.t, fx, fy, fz);
  double Hn = U + ensemble_term(W.cell, C.pressure, C.mu, C.S, W.counts);
  bool accepted = Hn < C.emax;
  if (accepted) {
    W.E = Hn;
  } else {
#pragma unroll
    for (int s = 0; s < A; ++s)
      if (s == as && lane == al) W.t[s] = cur;
#pragma unroll
    for (int sp = 0; sp < JR_MAX_SPECIES; ++sp)
      W.counts[sp] += (sp == cur ? 1 : 0) - (sp == nt ? 1 : 0);
  }
  return accepted;
}

// Index of the k-th atom (0-based, atom order) whose type is code; N - 1
// when there are at most k of them (swap.py:88-108).
template <int A>
__device__ __forceinline__ int nth_of_species(const int (&t)[A], int N, int code,
                                              int k) {
  const int lane = lane_id();
  int idx = N - 1, seen = 0;
  bool found = false;
#pragma unroll
  for (int s = 0; s < A; ++s) {
    const bool m = (lane + 64 * s < N) && (t[s] == code);
    const unsigned long long bal = __ballot(m);
    const int c = __popcll(bal);
    if (!found && k < seen + c) {
      const int r = k - seen;
      const int pre = __popcll(bal & ((1ull << lane) - 1ull));
      const unsigned long long pick = __ballot(m && pre == r);
      idx = 64 * s + __ffsll((long long)pick) - 1;
      found = true;
    }
    seen += c;
  }
  return idx;
}

// Kernels with cfg bit 3: alchemical morph (alchemical.py:40-79) and species
// swap (swap.py:119-262) with one energy evaluation between them.  Morph as
// morph_move.  Swap: species pairs (a < b, both in scope, ascending codes)
// weighted by n_a n_b; the first pair whose cumulative weight exceeds
// u0 * total, then atom floor(u1 n_a) of species a and floor(u2 n_b) of
// species b exchange types; accepted iff an unlike pair exists and H < Emax.
template <int A, int CFG>
__device__ bool species_move(const WalkCtx& C, WalkerRegs<A>& W,
                             const DevMove& mv, double u0, double u1,
                             double u2) {
  const int lane = lane_id();
  const int N = C.N;
  const bool swap = mv.type == JR_MOVE_SWAP;
  int ia, ib = -1, ta, tb = 0;
  bool has_pair = true;
  if (swap) {
    const int S = C.S;
    const int scope = mv.species_mask ? mv.species_mask : ((1 << S) - 1);
    double total = 0.0;
    for (int a = 0; a < S; ++a)
      for (int b = a + 1; b < S; ++b)
        if (((scope >> a) & 1) && ((scope >> b) & 1))
          total += (double)W.counts[a] * (double)W.counts[b];
    has_pair = total > 0.0;
    const double target = u0 * (has_pair ? total : 1.0);
    int lo = -1, hi = -1, last_lo = 0, last_hi = 1;
    double cum = 0.0;
    for (int a = 0; a < S; ++a)
      for (int b = a + 1; b < S; ++b)
        if (((scope >> a) & 1) && ((scope >> b) & 1)) {
          cum += (double)W.counts[a] * (double)W.counts[b];
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
    const int nlo = max(W.counts[lo], 1), nhi = max(W.counts[hi], 1);
    int klo = (int)(u1 * (double)nlo);
    if (klo > nlo - 1) klo = nlo - 1;
    int khi = (int)(u2 * (double)nhi);
    if (khi > nhi - 1) khi = nhi - 1;
    ia = nth_of_species<A>(W.t, N, lo, klo);
    ib = nth_of_species<A>(W.t, N, hi, khi);
    ta = hi;
    tb = lo;
  } else {
    const int Sm = mv.n_species;
    ia = (int)(u0 * (double)N);
    if (ia > N - 1) ia = N - 1;
    int cand = (int)(u1 * (double)(Sm - 1));
    if (cand > Sm - 2) cand = Sm - 2;
    int curv = 0;
#pragma unroll
    for (int s = 0; s < A; ++s)
      if (s == (ia >> 6)) curv = W.t[s];
    const int cur = __builtin_amdgcn_readlane(curv, ia & 63);
    ta = cand + (cand >= cur ? 1 : 0);
  }
  if (!has_pair) return false;
  int old_a = 0, old_b = 0;
#pragma unroll
  for (int s = 0; s < A; ++s) {
    const int a = lane + 64 * s;
    if (a == ia) {
      old_a = W.t[s];
      W.t[s] = ta;
    }
    if (a == ib) {
      old_b = W.t[s];
      W.t[s] = tb;
    }
  }
  const int ca = __builtin_amdgcn_readlane(old_a, ia & 63);
  if (!swap) {
#pragma unroll
    for (int sp = 0; sp < JR_MAX_SPECIES; ++sp)
      W.counts[sp] += (sp == ta ? 1 : 0) - (sp == ca ? 1 : 0);
  }
  Geom g;
  make_geom(W.cell, g);
  double fx[A], fy[A], fz[A];
  double U = wave_eval<A, false, CFG>(C.L, N, g, C.tab, W.x, W.y, W.z, W.t, fx, fy, fz);
  double Hn = U + ensemble_term(W.cell, C.pressure, C.mu, C.S, W.counts);
  const bool accepted = Hn < C.emax;
  if (accepted) {
    W.E = Hn;
  } else {
#pragma unroll
    for (int s = 0; s < A; ++s) {
      const int a = lane + 64 * s;
      if (a == ia) W.t[s] = old_a;
      if (a == ib) W.t[s] = old_b;
    }
    if (!swap) {
#pragma unroll
      for (int sp = 0; sp < JR_MAX_SPECIES; ++sp)
        W.counts[sp] += (sp == ca ? 1 : 0) - (sp == ta ? 1 : 0);
    }
  }
  return accepted;
}

}  // namespace jr

// Wave-level LJ energy/forces and the MWG move kernels (GMC, volume, HMC,
// alchemical morph) for one walker held by one wavefront.
#pragma once

#include <type_traits>

#include "jr_device.h"
#include "jr_types.h"
#include "../../include/jaxrens_amd.h"

namespace jr {

struct PairTab {
  double rc2;
  const double* eps4;  // LDS or global, [S*S]
  const double* eps24;
  const double* sig2;
  int S;
  int img[3];  // image half-ranges (CFG bit 2 kernels only)
};

struct WaveLds {
  double* base;
  // Evaluation team (cfg bit 5 kernels): the block's `size` wavefronts walk
  // one chain together; of the evaluation's `chunks` chunks of ring steps
  // wave `rank` evaluates chunks/size consecutive ones.  xbuf holds the
  // block's exchange rows, [chunks][7][64 A] doubles (energy, own force,
  // travelling reaction per chunk).
  double* xbuf = nullptr;
  int rank = 0, size = 1, chunks = 1;
};

// bytes of a team's exchange rows
__host__ __device__ inline size_t team_xbuf_bytes(int A, int chunks) {
  return (size_t)chunks * 7 * 64 * A * sizeof(double);
}

// Per-wave LDS view, M = 64*A ring positions.  Positions (and types) are
// stored twice, entries j and j + N both holding atom j, so the partner of
// ring position p at step k is entry p + k with no wrap and every access is a
// base register plus an immediate offset.  hx/hy/hz return the travelling
// reaction forces home once per evaluation.
template <int A>
struct LdsView {
  static constexpr int M = 64 * A;
  double* b;
  __device__ __forceinline__ lds_ptr<double> x() const { return to_lds(b); }
  __device__ __forceinline__ lds_ptr<double> y() const { return to_lds(b + 2 * M); }
  __device__ __forceinline__ lds_ptr<double> z() const { return to_lds(b + 4 * M); }
  __device__ __forceinline__ lds_ptr<double> hx() const { return to_lds(b + 6 * M); }
  __device__ __forceinline__ lds_ptr<double> hy() const { return to_lds(b + 7 * M); }
  __device__ __forceinline__ lds_ptr<double> hz() const { return to_lds(b + 8 * M); }
  __device__ __forceinline__ lds_ptr<int> t() const {
    return to_lds(reinterpret_cast<int*>(b + 9 * M));
  }
};

// bytes of one wave's LDS region: arrays, JR_MAX_MOVES*JR_NCOUNT counters,
// then the walked chain's constants (JR_MAX_SPECIES chemical potentials and
// JR_MAX_MOVES step sizes, doubles)
__host__ __device__ inline size_t wave_lds_bytes_a(int A) {
  size_t b = (size_t)9 * 64 * A * sizeof(double) +
             (size_t)2 * 64 * A * sizeof(int) +
             (size_t)JR_MAX_MOVES * JR_NCOUNT * sizeof(int) +
             (size_t)(JR_MAX_SPECIES + JR_MAX_MOVES) * sizeof(double);
  return (b + 15) & ~(size_t)15;
}

__device__ __forceinline__ int* wave_counters(const WaveLds& L, int A) {
  return reinterpret_cast<int*>(L.base + 9 * 64 * A) + 2 * 64 * A;
}

// the walked chain's chemical potentials [JR_MAX_SPECIES], then its step
// sizes [JR_MAX_MOVES] (8-byte aligned: the counters end on one)
__device__ __forceinline__ double* wave_consts(const WaveLds& L, int A) {
  return reinterpret_cast<double*>(wave_counters(L, A) + JR_MAX_MOVES * JR_NCOUNT);
}

struct Geom {
  double c[9];
  double inv[9];
  bool diag;
  bool cubic;  // diag with c[0] == c[4] == c[8]
};

// Pair geometry kinds: minimum image in fractional coordinates for a cubic
// cell, Cartesian rint per axis for other diagonal cells, full triclinic.
enum { GEO_CUBIC = 0, GEO_DIAG = 1, GEO_TRI = 2 };

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
  // wave-uniform: keep in scalar registers
#pragma unroll
  for (int i = 0; i < 9; ++i) {
    g.c[i] = uniform_d(g.c[i]);
    g.inv[i] = uniform_d(g.inv[i]);
  }
  g.diag = (g.c[1] == 0.0) && (g.c[2] == 0.0) && (g.c[3] == 0.0) &&
           (g.c[5] == 0.0) && (g.c[6] == 0.0) && (g.c[7] == 0.0);
  g.cubic = g.diag && (g.c[0] == g.c[4]) && (g.c[0] == g.c[8]);
}

// dr = MIC(x_j - x_i) and r^2 (backends/lj.py:141-162).
// GEO_CUBIC works on fractional coordinates u = x * inv[0] (cell length L):
// du = u_j - u_i, du -= rint(du), r^2 / L^2 = du.du; lengths come back
// through L^2 folded into num and the cutoff, and L into the force scale.
template <int GEO>
__device__ __forceinline__ void pair_dr(double xi, double yi, double zi,
                                        double xj, double yj, double zj,
                                        const Geom& g, double& dx, double& dy,
                                        double& dz, double& r2) {
  dx = xj - xi;
  dy = yj - yi;
  dz = zj - zi;
  if (GEO == GEO_CUBIC) {
    dx = dx - rint(dx);
    dy = dy - rint(dy);
    dz = dz - rint(dz);
  } else if (GEO == GEO_DIAG) {
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
  r2 = dot3(dx, dy, dz, dx, dy, dz);
}

// Pair terms from q, where q = sigma^2/r^2 (one species: epsilon and
// 24 epsilon/sigma^2 are applied once per evaluation) or q = 1/r^2 (species
// table).  q == 0 outside the cutoff / for idle ring positions.
template <bool FORCES, bool MULTI>
__device__ __forceinline__ void pair_lj(double q, double eps4, double eps24,
                                        double sig2, double& e, double& fs) {
  double s2 = MULTI ? sig2 * q : q;
  double s6 = (s2 * s2) * s2;
  const double el = fma(s6, s6, -s6);  // s^12 - s^6
  if (MULTI) {
    e = eps4 * el;
    if (FORCES) fs = (eps24 * fma(s6, s6, el)) * q;  // 2 s^12 - s^6
  } else {
    e = el;
    if (FORCES) fs = fma(s6, s6, el) * q;
  }
}

// Same terms in float (the FP32-pair variant); accumulation stays FP64.
template <bool FORCES, bool MULTI>
__device__ __forceinline__ void pair_lj_f32(float q, float eps4, float eps24,
                                            float sig2, float& e, float& fs) {
  float s2 = MULTI ? sig2 * q : q;
  float s6 = (s2 * s2) * s2;
  const float el = fmaf(s6, s6, -s6);
  if (MULTI) {
    e = eps4 * el;
    if (FORCES) fs = (eps24 * fmaf(s6, s6, el)) * q;
  } else {
    e = el;
    if (FORCES) fs = fmaf(s6, s6, el) * q;
  }
}

// q for an idle or cut-off pair: only the high dword is cleared, leaving a
// denormal of at most 2^-1042 whose square underflows to +0, so s6, s12, the
// pair energy and fs all come out +0 exactly as for q = 0 (one v_cndmask
// instead of two).
__device__ __forceinline__ double mask_q(bool ok, double v) {
  return __hiloint2double(ok ? __double2hiint(v) : 0, __double2loint(v));
}

// Lattice translation i0 c0 + i1 c1 + i2 c2 (rows of the cell).
__device__ __forceinline__ void image_shift(const Geom& g, int i0, int i1,
                                            int i2, double& tx, double& ty,
                                            double& tz) {
  const double a = (double)i0, b = (double)i1, c = (double)i2;
  tx = fma(c, g.c[6], fma(b, g.c[3], a * g.c[0]));
  ty = fma(c, g.c[7], fma(b, g.c[4], a * g.c[1]));
  tz = fma(c, g.c[8], fma(b, g.c[5], a * g.c[2]));
}

struct PairIn {
  double dx, dy, dz, r2;
  bool ok;  // active and r2 < rc2
  double eps4, eps24, sig2;
};

// Per-evaluation constants: q numerator, cutoff (both divided by L^2 for
// GEO_CUBIC) and the force scale.
struct EvalConst {
  double num, rc2, cf;
};

// Pair of the atom in slot s (ring position a) with LDS entry j = a + k,
// whose species tj (species table kernels) the caller has read.
template <int A, int CFG, int GEO>
__device__ __forceinline__ PairIn load_pair_t(const LdsView<A>& L, int j,
                                              bool active, const Geom& g,
                                              const PairTab& tab,
                                              const EvalConst& K, double xs,
                                              double ys, double zs, int ts, int tj) {
  constexpr bool MULTI = (CFG & 1) != 0;
  PairIn P;
  pair_dr<GEO>(xs, ys, zs, L.x()[j], L.y()[j], L.z()[j], g, P.dx, P.dy, P.dz,
               P.r2);
  P.ok = active && (P.r2 < K.rc2);
  const lds_ptr<const double> e4 = to_lds(tab.eps4), e24 = to_lds(tab.eps24),
                              s2 = to_lds(tab.sig2);
  P.eps4 = e4[0];
  P.eps24 = e24[0];
  P.sig2 = s2[0];
  if (MULTI) {
    const int ab = ts * tab.S + tj;
    P.eps4 = e4[ab];
    P.eps24 = e24[ab];
    P.sig2 = s2[ab];
  }
  return P;
}

template <int A, int CFG, int GEO>
__device__ __forceinline__ PairIn load_pair(const LdsView<A>& L, int j,
                                            bool active, const Geom& g,
                                            const PairTab& tab,
                                            const EvalConst& K, double xs,
                                            double ys, double zs, int ts) {
  constexpr bool MULTI = (CFG & 1) != 0;
  return load_pair_t<A, CFG, GEO>(L, j, active, g, tab, K, xs, ys, zs, ts,
                                  MULTI ? L.t()[j] : 0);
}

// Energy partial (per lane) and forces of the configuration whose positions
// and types are in the wave's LDS (and in the x,y,z,t registers).
// Pair schedule: for k = 1 .. (N-1)/2 ring position p pairs with
// (p + k) mod N, keeping its own force (own = fma(-fs, dr, own)) and adding
// the reaction to a travelling accumulator (trav = fma(fs, dr, trav)) that
// rotates one ring position per step, so every atom's reaction is summed in
// step order; for even N the k = N/2 pairs are evaluated from both ends with
// half the energy each and no reaction.  Steps are taken in groups of G = 4
// while they last, then one group of 2, then a lone step.  A group shares one
// IEEE division:
//   G = 4: p01 = r2_0 r2_1, p23 = r2_2 r2_3, d = num / (p01 p23),
//          d01 = p23 d, d23 = p01 d, q0 = r2_1 d01, q1 = r2_0 d01,
//          q2 = r2_3 d23, q3 = r2_2 d23;
//   G = 2: d = num / (r2_0 r2_1), q0 = r2_1 d, q1 = r2_0 d;
//   G = 1 and k = N/2: q = num / r2,
// with num = sigma^2 (one species) or 1.
// FULL: N == 64*A, so no ring position is idle and the rotation is one DPP
// wave_rol per 32-bit half.
// Slot layout of the steps: with several atoms per lane (A > 1, no image
// sums) slot s of a lane walks ring position p = A lane + s, so the
// rotation moves slot s + 1 into slot s (a register renaming) and only slot 0
// crosses lanes, into slot A - 1 of the lane below; the reactions then rotate
// after every step.  Positions are read from the LDS copy, and the energy
// partials and forces go back to the p = lane + 64 s layout through LDS.
// Sums per ring position are the same in either layout.
template <int A, bool FORCES, int CFG, int GEO, bool FULL>
__device__ __forceinline__ void wave_eval_impl(
    const WaveLds& WL, int chunk, const LdsView<A>& L, int N, const Geom& g, const PairTab& tab,
    const EvalConst& K, const double (&x)[A], const double (&y)[A], const double (&z)[A],
    const int (&t)[A], double& e_lane, double (&fx)[A], double (&fy)[A],
    double (&fz)[A]) {
  constexpr bool MULTI = (CFG & 1) != 0;
  constexpr bool F32 = (CFG & 2) != 0;
  constexpr bool IMG = (CFG & 4) != 0;
  constexpr bool TEAM = (CFG & 32) != 0;
  static_assert(!(IMG && F32), "image sums are FP64 only");
  static_assert(!(TEAM && (IMG || F32)), "evaluation teams are FP64 minimum image only");
  constexpr bool ILV = A > 1 && !IMG;
  constexpr int LA = A == 1 ? 0 : A == 2 ? 1 : 2;  // log2 A
  const int lane = lane_id();
  auto pos_of = [&](int s) { return ILV ? A * lane + s : lane + 64 * s; };
  // own positions (and types) of the slots
  double xo[A], yo[A], zo[A];
  int to[A];
#pragma unroll
  for (int s = 0; s < A; ++s) {
    if constexpr (ILV) {
      const int a = pos_of(s);
      xo[s] = L.x()[a];
      yo[s] = L.y()[a];
      zo[s] = L.z()[a];
      to[s] = MULTI ? L.t()[a] : 0;
    } else {
      xo[s] = x[s];
      yo[s] = y[s];
      zo[s] = z[s];
      to[s] = t[s];
    }
  }
  double e[A], ox[A], oy[A], oz[A], tx[A], ty[A], tz[A];
#pragma unroll
  for (int s = 0; s < A; ++s) {
    e[s] = 0.0;
    ox[s] = oy[s] = oz[s] = 0.0;
    tx[s] = ty[s] = tz[s] = 0.0;
  }
  const double num = K.num;
  // Partial rings: ring position N - 1 (lane lw, slot sw of the layout)
  // takes position 0.  Slot 0 moves through the crossbar from lane + 1, or
  // from lane 0 into lane lw (A = 1, or interleaved slots).
  const int lw = (N - 1) >> (ILV ? LA : 0), sw = (N - 1) & (ILV ? A - 1 : 0);
  const int rot_addr = (lane == lw ? 0 : (lane + 1) & 63) << 2;
  bool wrap[A];
#pragma unroll
  for (int s = 0; s < A; ++s) wrap[s] = (s == sw) && (lane == lw);
  auto rotate = [&]() {
    if constexpr (ILV) {
      if (FULL) {
        ring_rot1_ilv<A, true>(tx, rot_addr, wrap);
        ring_rot1_ilv<A, true>(ty, rot_addr, wrap);
        ring_rot1_ilv<A, true>(tz, rot_addr, wrap);
      } else {
        ring_rot1_ilv<A, false>(tx, rot_addr, wrap);
        ring_rot1_ilv<A, false>(ty, rot_addr, wrap);
        ring_rot1_ilv<A, false>(tz, rot_addr, wrap);
      }
    } else if (FULL) {
      ring_rot1_full<A>(tx, lane);
      ring_rot1_full<A>(ty, lane);
      ring_rot1_full<A>(tz, lane);
    } else if constexpr (A == 1) {
      tx[0] = ring_perm(tx[0], rot_addr);
      ty[0] = ring_perm(ty[0], rot_addr);
      tz[0] = ring_perm(tz[0], rot_addr);
    } else {
      ring_rot1<A>(tx, lane, N);
      ring_rot1<A>(ty, lane, N);
      ring_rot1<A>(tz, lane, N);
    }
  };
  using QT = typename std::conditional<F32, float, double>::type;
  // q = num / r^2 for one pair (F32: hardware reciprocal, 1 ulp)
  auto q_single = [&](const PairIn& P) -> QT {
    if constexpr (F32)
      return P.ok ? (float)num * __builtin_amdgcn_rcpf((float)P.r2) : 0.0f;
    else
      return mask_q(P.ok, num / P.r2);
  };
  auto accumulate = [&](int s, const PairIn& P, QT q, bool reaction,
                        double escale) {
    double ep, fs = 0.0;
    if constexpr (F32) {
      float epf, fsf = 0.0f;
      pair_lj_f32<FORCES, MULTI>(q, (float)P.eps4, (float)P.eps24,
                                 (float)P.sig2, epf, fsf);
      ep = (double)epf;
      fs = (double)fsf;
    } else {
      pair_lj<FORCES, MULTI>(q, P.eps4, P.eps24, P.sig2, ep, fs);
    }
    e[s] = e[s] + (escale == 1.0 ? ep : escale * ep);
    if (FORCES) {
      ox[s] = fma(-fs, P.dx, ox[s]);
      oy[s] = fma(-fs, P.dy, oy[s]);
      oz[s] = fma(-fs, P.dz, oz[s]);
      if (reaction) {
        tx[s] = fma(fs, P.dx, tx[s]);
        ty[s] = fma(fs, P.dy, ty[s]);
        tz[s] = fma(fs, P.dz, tz[s]);
      }
    }
  };
  const int kmax = (N - 1) >> 1;
  // Steps of this evaluation: groups of four from kfirst to kend4, then (on
  // the last chunk) the group of 2, the lone step and the k = N/2 pairs.
  // Chunk c of C takes groups [ceil(c ng / C), ceil((c+1) ng / C)) of the
  // ng = kmax / 4 groups; one wave alone takes them all.
  int kfirst = 1, kend4 = 1 + 4 * (kmax >> 2);
  bool last = true;
  if constexpr (TEAM) {
    // C is a power of two: ceil(x / C) = (x + C - 1) >> log2(C)
    const int ng = kmax >> 2, C = WL.chunks, lc = __builtin_ctz(C);
    kfirst = 1 + 4 * ((chunk * ng + C - 1) >> lc);
    kend4 = 1 + 4 * (((chunk + 1) * ng + C - 1) >> lc);
    last = chunk == C - 1;
  }
  int k = kfirst;
  if constexpr (IMG) {
    // backends/lj.py:152-197 with supercell images: each ring pair is summed
    // over the translations T = i0 c0 + i1 c1 + i2 c2 (i_d ascending, i2
    // fastest) of its minimum-image displacement, one division per term;
    // then the self pairs at T != 0 with half weight and no force.
    const int h0 = tab.img[0], h1 = tab.img[1], h2 = tab.img[2];
    auto image_terms = [&](int s, const PairIn& P0, bool active, bool reaction,
                           double escale) {
      for (int i0 = -h0; i0 <= h0; ++i0)
        for (int i1 = -h1; i1 <= h1; ++i1)
          for (int i2 = -h2; i2 <= h2; ++i2) {
            double Tx, Ty, Tz;
            image_shift(g, i0, i1, i2, Tx, Ty, Tz);
            PairIn P = P0;
            P.dx = P0.dx + Tx;
            P.dy = P0.dy + Ty;
            P.dz = P0.dz + Tz;
            P.r2 = dot3(P.dx, P.dy, P.dz, P.dx, P.dy, P.dz);
            P.ok = active && (P.r2 < K.rc2);
            accumulate(s, P, mask_q(P.ok, num / P.r2), reaction, escale);
          }
    };
#pragma unroll 1
    for (; k <= kmax; ++k) {
      if (FORCES && k > 1) rotate();
#pragma unroll
      for (int s = 0; s < A; ++s) {
        const int a = pos_of(s);
        const bool active = FULL || a < N;
        image_terms(s, load_pair<A, CFG, GEO>(L, a + k, active, g, tab, K, xo[s],
                                             yo[s], zo[s], to[s]),
                    active, true, 1.0);
      }
    }
    if (N >= 2 && !(N & 1)) {
      const int hN = N >> 1;
#pragma unroll
      for (int s = 0; s < A; ++s) {
        const int a = pos_of(s);
        const bool active = FULL || a < N;
        image_terms(s, load_pair<A, CFG, GEO>(L, a + hN, active, g, tab, K,
                                             xo[s], yo[s], zo[s], to[s]),
                    active, false, 0.5);
      }
    }
#pragma unroll
    for (int s = 0; s < A; ++s) {
      const int ab = MULTI ? t[s] * tab.S + t[s] : 0;
      for (int i0 = -h0; i0 <= h0; ++i0)
        for (int i1 = -h1; i1 <= h1; ++i1)
          for (int i2 = -h2; i2 <= h2; ++i2) {
            if (i0 == 0 && i1 == 0 && i2 == 0) continue;
            double Tx, Ty, Tz;
            image_shift(g, i0, i1, i2, Tx, Ty, Tz);
            const double r2 = dot3(Tx, Ty, Tz, Tx, Ty, Tz);
            double ep, fs;
            pair_lj<false, MULTI>(mask_q(r2 < K.rc2, num / r2), to_lds(tab.eps4)[ab],
                                  to_lds(tab.eps24)[ab], to_lds(tab.sig2)[ab], ep, fs);
            e[s] = e[s] + 0.5 * ep;
          }
    }
    k = kmax + 1;
  }
  // Species table: the partner species of a group are read one group ahead
  // (entries up to 64 A + kmax + 3 < 2 M), so that the table reads of a
  // group do not wait on an LDS round trip of their own.
  int tn[4][A];
  if constexpr (MULTI && !IMG) {
#pragma unroll
    for (int s = 0; s < A; ++s)
#pragma unroll
      for (int i = 0; i < 4; ++i) tn[i][s] = L.t()[pos_of(s) + k + i];
  }
  // kept rolled: with N known (N == 64*A) a full unroll multiplies the
  // register pressure and the code size of every call site
#pragma unroll 1
  for (; !IMG && k < kend4; k += 4) {
    PairIn P[4][A];
    QT q[4][A];
#pragma unroll
    for (int s = 0; s < A; ++s) {
      const int a = pos_of(s);
      const bool active = FULL || a < N;
      // interleaved slots read their own positions again per group rather
      // than hold them through the loop (registers for the species table)
      const double xs = ILV ? L.x()[a] : xo[s], ys = ILV ? L.y()[a] : yo[s],
                   zs = ILV ? L.z()[a] : zo[s];
#pragma unroll
      for (int i = 0; i < 4; ++i)
        P[i][s] = load_pair_t<A, CFG, GEO>(L, a + k + i, active, g, tab, K, xs, ys,
                                            zs, to[s], MULTI ? tn[i][s] : 0);
      if constexpr (F32) {
#pragma unroll
        for (int i = 0; i < 4; ++i) q[i][s] = q_single(P[i][s]);
      } else {
        const double p01 = P[0][s].r2 * P[1][s].r2;
        const double p23 = P[2][s].r2 * P[3][s].r2;
        const double d = num / (p01 * p23);
        const double d01 = p23 * d, d23 = p01 * d;
        q[0][s] = mask_q(P[0][s].ok, P[1][s].r2 * d01);
        q[1][s] = mask_q(P[1][s].ok, P[0][s].r2 * d01);
        q[2][s] = mask_q(P[2][s].ok, P[3][s].r2 * d23);
        q[3][s] = mask_q(P[3][s].ok, P[2][s].r2 * d23);
      }
    }
    if constexpr (MULTI) {
#pragma unroll
      for (int s = 0; s < A; ++s)
#pragma unroll
        for (int i = 0; i < 4; ++i) tn[i][s] = L.t()[pos_of(s) + k + 4 + i];
    }
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      if (FORCES && !ILV && k + i > kfirst) rotate();
#pragma unroll
      for (int s = 0; s < A; ++s) accumulate(s, P[i][s], q[i][s], true, 1.0);
      if (FORCES && ILV) rotate();
    }
  }
  if (!IMG && last && k + 1 <= kmax) {
    PairIn P0[A], P1[A];
    QT q0[A], q1[A];
#pragma unroll
    for (int s = 0; s < A; ++s) {
      const int a = pos_of(s);
      const bool active = FULL || a < N;
      P0[s] = load_pair<A, CFG, GEO>(L, a + k, active, g, tab, K, xo[s], yo[s],
                                      zo[s], to[s]);
      P1[s] = load_pair<A, CFG, GEO>(L, a + k + 1, active, g, tab, K, xo[s],
                                      yo[s], zo[s], to[s]);
      if constexpr (F32) {
        q0[s] = q_single(P0[s]);
        q1[s] = q_single(P1[s]);
      } else {
        const double d = num / (P0[s].r2 * P1[s].r2);
        q0[s] = mask_q(P0[s].ok, P1[s].r2 * d);
        q1[s] = mask_q(P1[s].ok, P0[s].r2 * d);
      }
    }
    if (FORCES && !ILV && k > kfirst) rotate();
#pragma unroll
    for (int s = 0; s < A; ++s) accumulate(s, P0[s], q0[s], true, 1.0);
    if (FORCES) rotate();
#pragma unroll
    for (int s = 0; s < A; ++s) accumulate(s, P1[s], q1[s], true, 1.0);
    if (FORCES && ILV) rotate();
    k += 2;
  }
  if (!IMG && last && k <= kmax) {  // lone last step
    if (FORCES && !ILV && k > kfirst) rotate();
#pragma unroll
    for (int s = 0; s < A; ++s) {
      const int a = pos_of(s);
      const bool active = FULL || a < N;
      PairIn P = load_pair<A, CFG, GEO>(L, a + k, active, g, tab, K, xo[s], yo[s],
                                         zo[s], to[s]);
      accumulate(s, P, q_single(P), true, 1.0);
    }
    if (FORCES && ILV) rotate();
  }
  if (!IMG && last && N >= 2 && !(N & 1)) {
    const int h = N >> 1;
#pragma unroll
    for (int s = 0; s < A; ++s) {
      const int a = pos_of(s);
      const bool active = FULL || a < N;
      PairIn P = load_pair<A, CFG, GEO>(L, a + h, active, g, tab, K, xo[s], yo[s],
                                         zo[s], to[s]);
      accumulate(s, P, q_single(P), false, 0.5);
    }
  }
  if constexpr (TEAM) {
    // the chunk's sums into its exchange rows (the travelling reactions at
    // their atoms: position p holds atom (p + klast) mod N, p + klast + 1
    // when they rotate after every step); wave_eval adds the chunks once the
    // whole team has written them
    constexpr int M = 64 * A;
    const int klast = (last ? kmax : kend4 - 1) + (ILV ? 1 : 0);
    const lds_ptr<double> mine = to_lds(WL.xbuf + (size_t)chunk * 7 * M);
#pragma unroll
    for (int s = 0; s < A; ++s) {
      const int a = pos_of(s);
      if (FULL || a < N) {
        mine[a] = e[s];
        if (FORCES) {
          int dst = a + klast;
          if (dst >= N) dst -= N;
          mine[M + a] = ox[s];
          mine[2 * M + a] = oy[s];
          mine[3 * M + a] = oz[s];
          mine[4 * M + dst] = tx[s];
          mine[5 * M + dst] = ty[s];
          mine[6 * M + dst] = tz[s];
        }
      }
    }
    return;
  }
  const lds_ptr<double> HX = L.hx(), HY = L.hy(), HZ = L.hz();
  if constexpr (ILV) {
    // energy partials of the positions to the p = lane + 64 s layout (an
    // idle position's partial is +0)
    wave_lds_sync();
#pragma unroll
    for (int s = 0; s < A; ++s) HX[pos_of(s)] = e[s];
    wave_lds_sync();
    double el = HX[lane];
#pragma unroll
    for (int s = 1; s < A; ++s) {
      const int a = lane + 64 * s;
      if (FULL || a < N) el = el + HX[a];
    }
    e_lane = el;
    wave_lds_sync();
  } else {
    double el = e[0];
#pragma unroll
    for (int s = 1; s < A; ++s) {
      const int a = lane + 64 * s;
      if (FULL || a < N) el = el + e[s];
    }
    e_lane = el;
  }
  if (FORCES && ILV) {
    // position p holds the reaction of atom (p + kmax + 1) mod N; the sums
    // own + reaction go back through the same rows to p = lane + 64 s
#pragma unroll
    for (int s = 0; s < A; ++s) {
      const int a = pos_of(s);
      if (FULL || a < N) {
        int dst = a + kmax + 1;
        if (dst >= N) dst -= N;
        HX[dst] = tx[s];
        HY[dst] = ty[s];
        HZ[dst] = tz[s];
      }
    }
    wave_lds_sync();
#pragma unroll
    for (int s = 0; s < A; ++s) {
      const int a = pos_of(s);
      if (FULL || a < N) {
        const double sx = ox[s] + HX[a], sy = oy[s] + HY[a], sz = oz[s] + HZ[a];
        HX[a] = sx;
        HY[a] = sy;
        HZ[a] = sz;
      }
    }
    wave_lds_sync();
#pragma unroll
    for (int s = 0; s < A; ++s) {
      const int a = lane + 64 * s;
      if (FULL || a < N) {
        fx[s] = HX[a];
        fy[s] = HY[a];
        fz[s] = HZ[a];
        if (GEO == GEO_CUBIC || !MULTI) {
          fx[s] = K.cf * fx[s];
          fy[s] = K.cf * fy[s];
          fz[s] = K.cf * fz[s];
        }
      } else {
        fx[s] = fy[s] = fz[s] = 0.0;
      }
    }
  }
  if (FORCES && !ILV) {
    // position p holds the reaction of atom (p + kmax) mod N
#pragma unroll
    for (int s = 0; s < A; ++s) {
      const int a = pos_of(s);
      if (FULL || a < N) {
        int dst = a + kmax;
        if (dst >= N) dst -= N;
        HX[dst] = tx[s];
        HY[dst] = ty[s];
        HZ[dst] = tz[s];
      }
    }
    wave_lds_sync();
#pragma unroll
    for (int s = 0; s < A; ++s) {
      const int a = pos_of(s);
      if (FULL || a < N) {
        fx[s] = ox[s] + HX[a];
        fy[s] = oy[s] + HY[a];
        fz[s] = oz[s] + HZ[a];
        if (GEO == GEO_CUBIC || !MULTI) {
          fx[s] = K.cf * fx[s];
          fy[s] = K.cf * fy[s];
          fz[s] = K.cf * fz[s];
        }
      } else {
        fx[s] = fy[s] = fz[s] = 0.0;
      }
    }
  }
}

// Sum of a team's chunk rows in chunk order, ((x_0 + x_1) + x_2) + ..., for
// the energy partial of each ring position and the force on each atom (own
// plus travelling reaction, times cf where the evaluation scales).  C is a
// compile-time constant so every load is issued before the first add.
template <int A, bool FORCES, int C>
__device__ __forceinline__ void team_combine(lds_ptr<const double> xb, int N, bool scale,
                                             double cf, double (&e)[A], double (&fx)[A],
                                             double (&fy)[A], double (&fz)[A]) {
  constexpr int M = 64 * A;
  const int lane = lane_id();
#pragma unroll
  for (int s = 0; s < A; ++s) {
    const int a = lane + 64 * s;
    e[s] = 0.0;
    if (a < N) {
      double v[C][7];
#pragma unroll
      for (int c = 0; c < C; ++c)
#pragma unroll
        for (int r = 0; r < (FORCES ? 7 : 1); ++r) v[c][r] = xb[(c * 7 + r) * M + a];
      double acc[7];
#pragma unroll
      for (int r = 0; r < (FORCES ? 7 : 1); ++r) {
        acc[r] = v[0][r];
#pragma unroll
        for (int c = 1; c < C; ++c) acc[r] = acc[r] + v[c][r];
      }
      e[s] = acc[0];
      if (FORCES) {
        fx[s] = acc[1] + acc[4];
        fy[s] = acc[2] + acc[5];
        fz[s] = acc[3] + acc[6];
        if (scale) {
          fx[s] = cf * fx[s];
          fy[s] = cf * fy[s];
          fz[s] = cf * fz[s];
        }
      }
    } else if (FORCES) {
      fx[s] = fy[s] = fz[s] = 0.0;
    }
  }
}

// Model energy U (wave-uniform) and forces of the configuration in
// registers; writes the configuration (twice) into the wave's LDS first.
// A cubic cell is evaluated in fractional coordinates u = x * inv[0] with
//   num = num_0 / L^2, cutoff rc^2 / L^2, force scale 24 eps / sigma^2 * L
// (species table: L), where num_0 = sigma^2 (one species) or 1.
template <int A, bool FORCES, int CFG>
__device__ __forceinline__ double wave_eval(const WaveLds& W, int N,
                                            const Geom& g, const PairTab& tab,
                                            const double (&x)[A],
                                            const double (&y)[A],
                                            const double (&z)[A],
                                            const int (&t)[A],
                                            double (&fx)[A], double (&fy)[A],
                                            double (&fz)[A]) {
  constexpr bool MULTI = (CFG & 1) != 0;
  constexpr bool IMG = (CFG & 4) != 0;
  constexpr bool TEAM = (CFG & 32) != 0;
  const int lane = lane_id();
  const LdsView<A> L{W.base};
  EvalConst K;
  K.num = MULTI ? 1.0 : to_lds(tab.sig2)[0];
  K.rc2 = tab.rc2;
  K.cf = MULTI ? 1.0 : to_lds(tab.eps24)[0] / to_lds(tab.sig2)[0];
  double u[A], v[A], w[A];
  if (!IMG && g.cubic) {
    const double len = g.c[0], len2 = len * len;
    K.num = K.num / len2;
    K.rc2 = K.rc2 / len2;
    K.cf = K.cf * len;
#pragma unroll
    for (int s = 0; s < A; ++s) {
      u[s] = x[s] * g.inv[0];
      v[s] = y[s] * g.inv[0];
      w[s] = z[s] * g.inv[0];
    }
  } else {
#pragma unroll
    for (int s = 0; s < A; ++s) {
      u[s] = x[s];
      v[s] = y[s];
      w[s] = z[s];
    }
  }
  wave_lds_sync();  // earlier readers of this wave's LDS are done
#pragma unroll
  for (int s = 0; s < A; ++s) {
    const int a = lane + 64 * s;
    if (a < N) {
      L.x()[a] = u[s];
      L.y()[a] = v[s];
      L.z()[a] = w[s];
      L.x()[a + N] = u[s];
      L.y()[a + N] = v[s];
      L.z()[a + N] = w[s];
      if (MULTI) {
        L.t()[a] = t[s];
        L.t()[a + N] = t[s];
      }
    }
  }
  wave_lds_sync();
  double el = 0.0;
  const bool full = (N == 64 * A);
  auto eval_chunk = [&](int chunk) {
    if constexpr (IMG) {
      wave_eval_impl<A, FORCES, CFG, GEO_TRI, false>(W, chunk, L, N, g, tab, K, u, v, w,
                                                     t, el, fx, fy, fz);
    } else if (g.cubic) {
      if (full)
        wave_eval_impl<A, FORCES, CFG, GEO_CUBIC, true>(W, chunk, L, N, g, tab, K, u, v,
                                                        w, t, el, fx, fy, fz);
      else
        wave_eval_impl<A, FORCES, CFG, GEO_CUBIC, false>(W, chunk, L, N, g, tab, K, u, v,
                                                         w, t, el, fx, fy, fz);
    } else if (g.diag) {
      if (full)
        wave_eval_impl<A, FORCES, CFG, GEO_DIAG, true>(W, chunk, L, N, g, tab, K, u, v, w,
                                                       t, el, fx, fy, fz);
      else
        wave_eval_impl<A, FORCES, CFG, GEO_DIAG, false>(W, chunk, L, N, g, tab, K, u, v,
                                                        w, t, el, fx, fy, fz);
    } else {
      if (full)
        wave_eval_impl<A, FORCES, CFG, GEO_TRI, true>(W, chunk, L, N, g, tab, K, u, v, w,
                                                      t, el, fx, fy, fz);
      else
        wave_eval_impl<A, FORCES, CFG, GEO_TRI, false>(W, chunk, L, N, g, tab, K, u, v,
                                                       w, t, el, fx, fy, fz);
    }
  };
  if constexpr (TEAM) {
    // Each wave evaluates its chunks into the exchange rows (the first
    // barrier keeps them from overwriting rows the team still reads from the
    // previous evaluation), then every wave adds the chunks in chunk order,
    // so the whole team ends with the same energy partials and forces.
    constexpr int M = 64 * A;
    const int per = W.chunks / W.size;
    __syncthreads();
    for (int j = 0; j < per; ++j) eval_chunk(W.rank * per + j);
    __syncthreads();
    const lds_ptr<const double> xb = to_lds(static_cast<const double*>(W.xbuf));
    const bool scale = g.cubic || !MULTI;
    double e[A];
    switch (W.chunks) {
      case 2: team_combine<A, FORCES, 2>(xb, N, scale, K.cf, e, fx, fy, fz); break;
      case 4: team_combine<A, FORCES, 4>(xb, N, scale, K.cf, e, fx, fy, fz); break;
      default: team_combine<A, FORCES, 8>(xb, N, scale, K.cf, e, fx, fy, fz); break;
    }
    el = e[0];
#pragma unroll
    for (int s = 1; s < A; ++s) {
      const int a = lane + 64 * s;
      if (a < N) el = el + e[s];
    }
  } else {
    eval_chunk(0);
  }
  wave_lds_sync();
  double U = wave_sum((lane < N) ? el : 0.0);
  if (!MULTI) U = to_lds(tab.eps4)[0] * U;
  return U;
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

// The same with the chemical potentials in the wave's LDS.
__device__ __forceinline__ double ensemble_term(const double* cell,
                                                double pressure,
                                                lds_ptr<const double> mu, int S,
                                                const int* counts) {
  double e = pressure * fabs(det3(cell));
  double acc = 0.0;
#pragma unroll
  for (int sp = 0; sp < JR_MAX_SPECIES; ++sp)
    if (sp < S) acc = acc + mu[sp] * (double)counts[sp];
  return e - acc;
}

}  // namespace jr

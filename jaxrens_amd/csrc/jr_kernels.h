// The persistent walk kernel and the batched energy evaluator (templates,
// instantiated per kernel configuration in jr_walk_cfg*.hip and jr_eval_inst.hip).
#pragma once

#include "jr_moves.h"

namespace jr {

struct WalkArgs {
  PopPtrs pop;
  const double* pressure;
  const double* mu;
  const double* step_sizes;
  const double* emax;
  const int* walk;
  int n_walk, R, K, Nmax, S, n_steps, n_moves;
  const int* schedule;
  const double* normals;
  const double* uniforms;
  int* counters;  // [R][n_moves][JR_NCOUNT]
  int* queue;
  int total;
  const int* order;  // replica processing order (largest N first)
  double rc2;
  const double* eps4;
  const double* eps24;
  const double* sig2;
  const DevMove* moves;  // [n_moves], global (uniform scalar loads)
  // trial mode (step-size adaptation): one move per sampled walker
  int trial;
  int trial_move;
  int n_samples;
  const int* sample_idx;
  const double* trial_step;
  int* trial_reason;
  // image half-ranges (CFG bit 2); kept last: the register allocation of the
  // minimum-image walk kernels depends on the offsets of the arguments above
  int img[3];
  int lds_atoms;  // sizes each wave's LDS region (Nmax, or 64 A of a class launch)
  // minimum-distance floors [S][S] (cfg bit 4 kernels): for moves that move
  // atoms or the cell, and for moves that change species; NULL = not gated
  const double* dmin_pos;
  const double* dmin_types;
  // normals == NULL: the walk draws its normals and uniforms in place, the
  // values the fill kernels would write into the tape layout (Philox streams
  // 5 and 6 of iteration st->rng_iter under seed; see step_draws)
  const IterState* st;
  unsigned long long seed;
  // persistent directions of the species-scoped GMC moves: slot s > 0 of
  // walker w at dirx + ((s - 1) * n_pop + w) * Nmax * 3 (NULL: none)
  double* dirx;
  int n_pop;
  int chunks;  // evaluation chunks of a team kernel (cfg bit 5)
};

// Draws of sub-step t of chain tq in place of the tape (normals
// [tq][n_steps][3][Nmax] of stream 5 in the four-per-block layout,
// uniforms [tq][n_steps][4] of stream 6).  gmc and hmc: the walker's normals
// go to the wave's LDS rows hx..hz (free until the move's first energy
// evaluation) as xs[c * 64A + atom]; other moves: the normals [0][0] and
// [1][0] and the four uniforms, wave-uniform, from lanes 0..5 of one pass.
template <int A>
__device__ __forceinline__ void step_draws(const WalkArgs& a, long long it, int tq, int t,
                                           int N, bool full, lds_ptr<double> xs, double& xv0,
                                           double& xv1, double (&u)[4]) {
  constexpr int M = 64 * A;
  const int lane = lane_id();
  const long long blk = (long long)tq * a.n_steps + t;
  const long long E0 = blk * 3 * (long long)a.Nmax;
  if (full) {
    const long long q0 = E0 >> 2;
    const int nq = (int)(((E0 + 3LL * a.Nmax - 1) >> 2) - q0 + 1);
    const int n3 = 3 * a.Nmax;
    for (int i = lane; i < nq; i += 64) {
      double z[4];
      normal_quad(a.seed, it, 5u, q0 + i, z);
      const int e0 = (int)(4 * (q0 + i) - E0);  // in [-3, 3 Nmax + 3)
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        const int e = e0 + j;
        const int c = (e >= a.Nmax) + (e >= 2 * a.Nmax);
        const int at = e - c * a.Nmax;
        if (e >= 0 && e < n3 && at < N) xs[c * M + at] = z[j];
      }
    }
    wave_lds_sync();
  } else {
    const int l = lane & 7;
    const bool nrm = l < 2;
    const long long e = nrm ? E0 + (long long)l * a.Nmax : blk * 4 + (l - 2);
    uint32_t c[4];
    philox_block(a.seed, it, nrm ? 5u : 6u, nrm ? (e >> 2) : (e >> 1), c);
    const bool hi = nrm ? (e & 2) != 0 : (e & 1) != 0;
    const uint32_t w0 = hi ? c[2] : c[0], w1 = hi ? c[3] : c[1];
    double zc, zs;
    box_muller(w0, w1, zc, zs);
    const double v = nrm ? ((e & 1) ? zs : zc) : u53(w0, w1);
    xv0 = readlane_d(v, 0);
    xv1 = readlane_d(v, 1);
#pragma unroll
    for (int j = 0; j < 4; ++j) u[j] = readlane_d(v, 2 + j);
  }
}

__host__ __device__ inline size_t wave_lds_bytes(int Nmax, int n_moves) {
  (void)n_moves;
  return wave_lds_bytes_a(Nmax <= 64 ? 1 : (Nmax <= 128 ? 2 : 4));
}

// Dynamic LDS of a walk or evaluation block: per-wave regions, the pair
// table, and for an evaluation team (cfg bit 5, the block is one team of
// `wpb` wavefronts evaluating `chunks` chunks) its exchange rows and the
// work-item broadcast slot.
__host__ __device__ inline size_t team_xbuf_offset(int A, int wpb, int S) {
  return (wave_lds_bytes_a(A) * wpb + 3 * sizeof(double) * S * S + 15) & ~(size_t)15;
}

__host__ __device__ inline size_t block_lds_bytes(int A, int wpb, int S, int chunks) {
  if (chunks <= 1) return wave_lds_bytes_a(A) * wpb + 3 * sizeof(double) * S * S;
  return team_xbuf_offset(A, wpb, S) + team_xbuf_bytes(A, chunks) + 16;
}

// OCC > 0 asks the register allocator for at least OCC waves per SIMD.
// A team kernel (cfg bit 5) runs blocks of up to four waves, one per SIMD.
template <int A, int CFG, int OCC>
__global__ void __launch_bounds__(256)
    __attribute__((amdgpu_waves_per_eu(OCC > 0 ? OCC : 1, 8)))
    walk_kernel(WalkArgs a) {
  constexpr bool MULTI = (CFG & 1) != 0;
  constexpr bool F32 = (CFG & 2) != 0;
  constexpr bool TEAM = (CFG & 32) != 0;
  (void)F32;
  extern __shared__ double lds_raw[];
  const int lane = lane_id();
  // wave index within the block, wave-uniform (kept in a scalar register)
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int nwave_blk = blockDim.x >> 6;
  const size_t per_wave = wave_lds_bytes(a.lds_atoms, a.n_moves);
  char* base = reinterpret_cast<char*>(lds_raw) + per_wave * wave;
  double* tabs = reinterpret_cast<double*>(
      reinterpret_cast<char*>(lds_raw) + per_wave * nwave_blk);
  WaveLds L;
  L.base = reinterpret_cast<double*>(base);
  int* qslot = nullptr;
  if constexpr (TEAM) {
    // one team per block: every wave walks the block's chain, evaluating
    // its own chunks of the ring steps; wave 0 writes the results
    const size_t off = ((per_wave * nwave_blk + 3 * sizeof(double) * a.S * a.S) + 15) &
                       ~(size_t)15;
    L.xbuf = reinterpret_cast<double*>(reinterpret_cast<char*>(lds_raw) + off);
    L.rank = wave;
    L.size = nwave_blk;
    L.chunks = a.chunks;
    qslot = reinterpret_cast<int*>(reinterpret_cast<char*>(L.xbuf) +
                                   team_xbuf_bytes(A, a.chunks));
  }
  // the wave that writes results and counters (every wave without a team)
  const bool writer = !TEAM || wave == 0;
  const lds_ptr<int> lctr = to_lds(wave_counters(L, A));

  // pair table in LDS (block-shared)
  const int SS = a.S * a.S;
  for (int i = threadIdx.x; i < SS; i += blockDim.x) {
    tabs[i] = a.eps4[i];
    tabs[SS + i] = a.eps24[i];
    tabs[2 * SS + i] = a.sig2[i];
  }
  for (int i = lane; i < JR_MAX_MOVES * JR_NCOUNT; i += 64) lctr[i] = 0;
  __syncthreads();

  WalkCtx C;
  C.L = L;
  C.tab.rc2 = a.rc2;
  C.tab.eps4 = tabs;
  C.tab.eps24 = tabs + SS;
  C.tab.sig2 = tabs + 2 * SS;
  C.tab.S = a.S;
  C.tab.img[0] = a.img[0];
  C.tab.img[1] = a.img[1];
  C.tab.img[2] = a.img[2];
  const bool gen = a.normals == nullptr;
  const long long rng_it = gen ? a.st->rng_iter : 0;
  C.xi_stride = gen ? 64 * A : a.Nmax;
  C.S = a.S;

  for (;;) {
    // Every lane issues the queue atomic (lane 0 adds 1, the others 0) and
    // the first lane's result is the wave's work item.  No lane-dependent
    // branch may guard the fetch: with `if (lane == 0) q = atomicAdd(..)`
    // the compiler threads lanes 1..63 from an earlier `if (lane == 0)`
    // straight back to the loop head with q = 0 and no atomic, and those
    // lanes then repeat item 0 forever.  A team takes wave 0's item through
    // LDS, between two block barriers, so every wave of the block walks the
    // same chain and leaves the loop together.
    int q;
    if constexpr (TEAM) {
      if (wave == 0) {
        int v = atomicAdd(a.queue, lane == 0 ? 1 : 0);
        v = __builtin_amdgcn_readfirstlane(v);
        if (lane == 0) *qslot = v;
      }
      __syncthreads();
      q = __builtin_amdgcn_readfirstlane(*qslot);
      __syncthreads();
    } else {
      q = atomicAdd(a.queue, lane == 0 ? 1 : 0);
      q = __builtin_amdgcn_readfirstlane(q);
    }
    if (q >= a.total) break;
    int r, slot, tq;
    if (a.trial) {
      r = q / a.n_samples;
      slot = r * a.K + a.sample_idx[q];
      tq = q;
    } else {
      const int rr = q / a.n_walk, c = q - rr * a.n_walk;
      r = a.order ? a.order[rr] : rr;
      slot = r * a.K + a.walk[r * a.n_walk + c];
      tq = r * a.n_walk + c;
    }
    const int N = a.pop.n_atoms[slot];
    C.N = N;
    C.pressure = a.pressure[r];
    // the chain's chemical potentials and step sizes into the wave's LDS:
    // every lookup in the step loop would otherwise wait on a global load
    {
      const lds_ptr<double> cst = to_lds(wave_consts(L, A));
      wave_lds_sync();
      if (lane < JR_MAX_SPECIES)
        cst[lane] = lane < a.S ? a.mu[(size_t)r * a.S + lane] : 0.0;
      else if (lane < JR_MAX_SPECIES + JR_MAX_MOVES)
        cst[lane] = (!a.trial && lane - JR_MAX_SPECIES < a.n_moves)
                        ? a.step_sizes[(size_t)r * a.n_moves + lane - JR_MAX_SPECIES]
                        : 0.0;
      wave_lds_sync();
      C.mu = cst;
    }
    C.emax = a.emax[r];

    WalkerRegs<A> W;
    const double* P = a.pop.pos + (size_t)slot * a.Nmax * 3;
    const double* D = a.pop.dir + (size_t)slot * a.Nmax * 3;
    const int* T = a.pop.types + (size_t)slot * a.Nmax;
#pragma unroll
    for (int s = 0; s < A; ++s) {
      const int at = lane + 64 * s;
      if (at < N) {
        W.x[s] = P[3 * at + 0];
        W.y[s] = P[3 * at + 1];
        W.z[s] = P[3 * at + 2];
        W.dx[s] = D[3 * at + 0];
        W.dy[s] = D[3 * at + 1];
        W.dz[s] = D[3 * at + 2];
        W.t[s] = T[at];
      } else {
        W.x[s] = W.y[s] = W.z[s] = 0.0;
        W.dx[s] = W.dy[s] = W.dz[s] = 0.0;
        W.t[s] = 0;
      }
    }
#pragma unroll
    for (int i = 0; i < 9; ++i) W.cell[i] = a.pop.cell[(size_t)slot * 9 + i];
    W.E = a.pop.energy[slot];
    species_counts<A>(N, a.S, W.t, W.counts);

    const int n_steps = a.trial ? 1 : a.n_steps;
    int last_reason = 0;
    // the next sub-step's move index is fetched a step ahead
    int m_next = a.trial ? a.trial_move : a.schedule[0];
    for (int t = 0; t < n_steps; ++t) {
      const int m = m_next;
      if (t + 1 < n_steps) m_next = a.schedule[t + 1];
      const DevMove mv = a.moves[m];
      const double step = a.trial ? a.trial_step[r] : C.mu[JR_MAX_SPECIES + m];
      int reason = 0, n_eval = 1, n_grad = 0;
      int type = mv.type;
      if constexpr ((CFG & 8) != 0) {
        if (type == JR_MOVE_SHEAR || type == JR_MOVE_STRETCH) type = JR_MOVE_VOLUME;
        if (type == JR_MOVE_SWAP) type = JR_MOVE_MORPH;
      }
      const double* xi;
      const double* u;
      double xv0, xv1, ud[4];
      if (!gen) {
        xi = a.normals + ((size_t)tq * n_steps + t) * 3 * (size_t)a.Nmax;
        u = a.uniforms + ((size_t)tq * n_steps + t) * 4;
        xv0 = xi[0];
        xv1 = xi[a.Nmax];
      } else {
        const lds_ptr<double> xs = LdsView<A>{L.base}.hx();
        step_draws<A>(a, rng_it, tq, t, N, type == JR_MOVE_GMC || type == JR_MOVE_HMC, xs,
                      xv0, xv1, ud);
        xi = (const double*)xs;
        u = ud;
      }
      // configuration constraint gate (sampling/mwg.py:98-141)
      [[maybe_unused]] const double* floors = nullptr;
      [[maybe_unused]] WalkerRegs<A> W0;
      if constexpr ((CFG & 16) != 0) {
        floors = (type == JR_MOVE_MORPH) ? a.dmin_types : a.dmin_pos;
        if (floors) W0 = W;
      }
      switch (type) {
        case JR_MOVE_GMC: {
          // a scoped move's own persistent direction (galilean.py:107-131
          // direction_field); the population's stays aside in memory
          const bool own = mv.dir_slot > 0;
          double* Dh = a.pop.dir + (size_t)slot * a.Nmax * 3;
          double* Ds = own ? a.dirx + ((size_t)(mv.dir_slot - 1) * a.n_pop + slot) * a.Nmax * 3
                           : nullptr;
          if (own) {
            if (!a.trial) store_dir<A>(W, Dh, N);
            load_dir<A>(W, Ds, N);
          }
          reason = gmc_move<A, CFG>(C, W, step, mv, xi) ? 0 : 1;
          if (own && !a.trial) {
            store_dir<A>(W, Ds, N);
            load_dir<A>(W, Dh, N);
          }
          n_eval = n_grad = mv.n_reflect;
          break;
        }
        case JR_MOVE_VOLUME:  // with cfg bit 3: also shear and stretch
          if constexpr ((CFG & 8) != 0)
            reason = cell_move<A, CFG>(C, W, step, mv, xv0, xv1, u[0]);
          else
            reason = volume_move<A, CFG>(C, W, step, mv, xv0, u[0]);
          break;
        case JR_MOVE_HMC:
          reason = hmc_move<A, CFG>(C, W, step, mv, xi) ? 0 : 1;
          n_eval = 2 * mv.n_leapfrog + 1;
          n_grad = 2 * mv.n_leapfrog;
          break;
        case JR_MOVE_MORPH:  // with cfg bit 3: also species_swap
          if constexpr ((CFG & 8) != 0)
            reason = species_move<A, CFG>(C, W, mv, u[0], u[1], u[2]) ? 0 : 1;
          else
            reason = morph_move<A, CFG>(C, W, mv, u[0], u[1]) ? 0 : 1;
          break;

        default:
          break;
      }
      if constexpr ((CFG & 16) != 0) {
        // an accepted proposal that breaks the constraint is undone:
        // positions, cell, species and energy return, the GMC direction
        // keeps the move's value, and the proposal counts in no reject
        // bucket (its reason 4 falls outside the four the reference scatters)
        if (floors && reason == 0 && !min_distance_ok<A>(C, W, floors)) {
#pragma unroll
          for (int s = 0; s < A; ++s) {
            W.x[s] = W0.x[s];
            W.y[s] = W0.y[s];
            W.z[s] = W0.z[s];
            W.t[s] = W0.t[s];
          }
#pragma unroll
          for (int i = 0; i < 9; ++i) W.cell[i] = W0.cell[i];
#pragma unroll
          for (int i = 0; i < JR_MAX_SPECIES; ++i) W.counts[i] = W0.counts[i];
          W.E = W0.E;
          reason = 4;
        }
      }
      last_reason = reason;
      if (writer && lane == 0) {
        const lds_ptr<int> c = lctr + m * JR_NCOUNT;
        c[0] += 1;
        if ((CFG & 16) == 0 || reason < 4) c[1 + reason] += 1;
        c[5] += n_eval;
        c[6] += n_grad;
      }
    }
    if (a.trial) {
      if (writer && lane == 0) a.trial_reason[q] = last_reason;
      continue;
    }
    if (!writer) continue;
    double* Po = a.pop.pos + (size_t)slot * a.Nmax * 3;
    double* Do = a.pop.dir + (size_t)slot * a.Nmax * 3;
    int* To = a.pop.types + (size_t)slot * a.Nmax;
#pragma unroll
    for (int s = 0; s < A; ++s) {
      const int at = lane + 64 * s;
      if (at < N) {
        Po[3 * at + 0] = W.x[s];
        Po[3 * at + 1] = W.y[s];
        Po[3 * at + 2] = W.z[s];
        Do[3 * at + 0] = W.dx[s];
        Do[3 * at + 1] = W.dy[s];
        Do[3 * at + 2] = W.dz[s];
        To[at] = W.t[s];
      }
    }
    if (lane < 9) {
      double cv = W.cell[0];
#pragma unroll
      for (int i = 1; i < 9; ++i)
        if (lane == i) cv = W.cell[i];
      a.pop.cell[(size_t)slot * 9 + lane] = cv;
    }
    if (lane == 0) a.pop.energy[slot] = W.E;
    // counters are per replica: flush this chain's tallies
    wave_lds_sync();
    for (int i = lane; i < a.n_moves * JR_NCOUNT; i += 64) {
      int v = lctr[i];
      if (v) atomicAdd(a.counters + (size_t)r * a.n_moves * JR_NCOUNT + i, v);
      lctr[i] = 0;
    }
    wave_lds_sync();
  }
}

struct EvalArgs {
  const double* pos;
  const int* types;
  const int* n_atoms;
  const double* cell;
  const int* cfg_index;  // [B] configuration per item, or NULL (= item)
  const int* count;      // device item count (<= B), or NULL (= B)
  const double* pressure;
  const double* mu;
  double* U;
  double* H;
  double* F;
  int B, Nmax, S;
  double rc2;
  const double* eps4;
  const double* eps24;
  const double* sig2;
  int img[3];  // image half-ranges (CFG bit 2)
  int chunks;  // evaluation chunks of a team kernel (cfg bit 5)
  int team;    // waves per block of a team kernel
};

// Batched energy (+forces): one wave per configuration, or (cfg bit 5) one
// block per configuration whose waves evaluate one chunk each.
template <int A, int CFG>
__global__ void __launch_bounds__(256) eval_kernel(EvalArgs a) {
  constexpr bool MULTI = (CFG & 1) != 0;
  constexpr bool F32 = (CFG & 2) != 0;
  constexpr bool TEAM = (CFG & 32) != 0;
  (void)F32;
  extern __shared__ double lds_raw[];
  const int lane = lane_id();
  // wave index within the block, wave-uniform (kept in a scalar register)
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int nwave_blk = blockDim.x >> 6;
  const size_t per_wave = wave_lds_bytes(a.Nmax, 0);
  char* base = reinterpret_cast<char*>(lds_raw) + per_wave * wave;
  double* tabs = reinterpret_cast<double*>(
      reinterpret_cast<char*>(lds_raw) + per_wave * nwave_blk);
  WaveLds L;
  L.base = reinterpret_cast<double*>(base);
  if constexpr (TEAM) {
    const size_t off = ((per_wave * nwave_blk + 3 * sizeof(double) * a.S * a.S) + 15) &
                       ~(size_t)15;
    L.xbuf = reinterpret_cast<double*>(reinterpret_cast<char*>(lds_raw) + off);
    L.rank = wave;
    L.size = nwave_blk;
    L.chunks = a.chunks;
  }
  const int SS = a.S * a.S;
  for (int i = threadIdx.x; i < SS; i += blockDim.x) {
    tabs[i] = a.eps4[i];
    tabs[SS + i] = a.eps24[i];
    tabs[2 * SS + i] = a.sig2[i];
  }
  __syncthreads();
  PairTab tab;
  tab.rc2 = a.rc2;
  tab.eps4 = tabs;
  tab.eps24 = tabs + SS;
  tab.sig2 = tabs + 2 * SS;
  tab.S = a.S;
  tab.img[0] = a.img[0];
  tab.img[1] = a.img[1];
  tab.img[2] = a.img[2];

  const int item = TEAM ? (int)blockIdx.x : (int)blockIdx.x * nwave_blk + wave;
  const int n_items = a.count ? min(*a.count, a.B) : a.B;
  // wave-uniform (block-uniform for a team); no block barrier follows
  if (item >= n_items) return;
  const int b = a.cfg_index ? a.cfg_index[item] : item;
  const int N = a.n_atoms[b];
  double x[A], y[A], z[A];
  int t[A];
  const double* P = a.pos + (size_t)b * a.Nmax * 3;
#pragma unroll
  for (int s = 0; s < A; ++s) {
    const int at = lane + 64 * s;
    if (at < N) {
      x[s] = P[3 * at];
      y[s] = P[3 * at + 1];
      z[s] = P[3 * at + 2];
      t[s] = a.types[(size_t)b * a.Nmax + at];
    } else {
      x[s] = y[s] = z[s] = 0.0;
      t[s] = 0;
    }
  }
  double cell[9];
#pragma unroll
  for (int i = 0; i < 9; ++i) cell[i] = a.cell[(size_t)b * 9 + i];
  Geom g;
  make_geom(cell, g);
  double fx[A], fy[A], fz[A];
  double U;
  if (a.F)
    U = wave_eval<A, true, CFG>(L, N, g, tab, x, y, z, t, fx, fy, fz);
  else
    U = wave_eval<A, false, CFG>(L, N, g, tab, x, y, z, t, fx, fy, fz);
  int counts[JR_MAX_SPECIES];
  species_counts<A>(N, a.S, t, counts);
  const double pr = a.pressure ? a.pressure[b] : 0.0;
  double Hv;
  if (a.mu) {
    Hv = U + ensemble_term(cell, pr, a.mu + (size_t)b * a.S, a.S, counts);
  } else {
    Hv = U + pr * fabs(det3(cell));
  }
  if (TEAM && wave != 0) return;
  if (lane == 0) {
    if (a.U) a.U[b] = U;
    if (a.H) a.H[b] = Hv;
  }
  if (a.F) {
    double* Fo = a.F + (size_t)b * a.Nmax * 3;
#pragma unroll
    for (int s = 0; s < A; ++s) {
      const int at = lane + 64 * s;
      if (at < N) {
        Fo[3 * at] = fx[s];
        Fo[3 * at + 1] = fy[s];
        Fo[3 * at + 2] = fz[s];
      } else if (at < a.Nmax) {
        Fo[3 * at] = Fo[3 * at + 1] = Fo[3 * at + 2] = 0.0;
      }
    }
  }
}

}  // namespace jr

// Kernels of one NS iteration around the walk: select/clone, replica
// exchange, Philox draws and the iteration-state advance.  Included by
// jr_lib.hip only (non-template kernels: one definition in the library).
#pragma once

#include "jr_kernels.h"

namespace jr {

// ------------------------------------------------------------------ select
struct SelArgs {
  PopPtrs pop;
  int R, K, Nmax, n_extra, n_walk, n_moves, Kpad;
  const double* sel_u;
  const double* clone_u;
  const double* extra_u;
  double* emax;
  int* walk;
  int* counters;
  int* queue;
  IterState* st;
  // per-block outputs (indexed by st->block_pos)
  double* o_emax;
  int* o_worst;
  int* o_clone;
  double* o_dead_volume;
  double* o_dead_pos;
  int* o_dead_types;
  double* o_dead_cell;
  // n_cull culled walkers per replica; per-block outputs with a cull axis:
  // o_worst, o_clone, o_dead_* are [block][R][n_cull](...)
  int n_cull;
  double* o_dead_energy;
};

__device__ __forceinline__ bool key_before(double ka, int ia, double kb,
                                           int ib) {
  return (ka > kb) || (ka == kb && ia < ib);
}

// Block-wide bitonic sort of (key, index) pairs in LDS, Kpad a power of two,
// into the order `before` defines (a strict total order on the pairs).
template <class Before>
__device__ void bitonic_sort(double* skey, int* sidx, int Kpad, Before before) {
  const int tid = threadIdx.x, nt = blockDim.x;
  for (int size = 2; size <= Kpad; size <<= 1) {
    for (int stride = size >> 1; stride > 0; stride >>= 1) {
      for (int i = tid; i < Kpad; i += nt) {
        const int j = i ^ stride;
        if (j > i) {
          const bool up = ((i & size) == 0);
          const double ki = skey[i], kj = skey[j];
          const int ii = sidx[i], ij = sidx[j];
          const bool swap = up ? before(kj, ij, ki, ii) : before(ki, ii, kj, ij);
          if (swap) {
            skey[i] = kj;
            skey[j] = ki;
            sidx[i] = ij;
            sidx[j] = ii;
          }
        }
      }
      __syncthreads();
    }
  }
}

// nested_sampling.py:86-133,151-186 (n_cull = 1) and :486-518 (dead record,
// clone into the culled slot).  One block per replica.
__global__ void __launch_bounds__(1024) select_kernel(SelArgs a) {
  extern __shared__ double sh[];
  double* skey = sh;                                   // [Kpad]
  int* sidx = reinterpret_cast<int*>(sh + a.Kpad);     // [Kpad]
  __shared__ double red_d[32];
  __shared__ int red_i[32];
  __shared__ int s_worst, s_clone;
  __shared__ double s_mx;
  const int r = blockIdx.x;
  const int tid = threadIdx.x;
  const int nt = blockDim.x;
  const int lane = tid & 63, wv = tid >> 6, nwv = nt >> 6;
  const double* E = a.pop.energy + (size_t)r * a.K;

  // 1. contour = max energy
  double mx = -__builtin_inf();
  for (int k = tid; k < a.K; k += nt) mx = fmax(mx, E[k]);
#pragma unroll
  for (int m = 1; m < 64; m <<= 1) mx = fmax(mx, __shfl_xor(mx, m, 64));
  if (lane == 0) red_d[wv] = mx;
  __syncthreads();
  if (tid == 0) {
    double v = red_d[0];
    for (int w = 1; w < nwv; ++w) v = fmax(v, red_d[w]);
    s_mx = v;
  }
  __syncthreads();
  mx = s_mx;
  // 2. uniform tie-break among exact maxima (first index on equal noise)
  const double* su = a.sel_u + (size_t)r * a.K;
  double bk = -__builtin_inf();
  int bi = 0x7fffffff;
  for (int k = tid; k < a.K; k += nt) {
    if (E[k] == mx) {
      double kk = su[k];
      if (key_before(kk, k, bk, bi)) {
        bk = kk;
        bi = k;
      }
    }
  }
#pragma unroll
  for (int m = 1; m < 64; m <<= 1) {
    double ok = __shfl_xor(bk, m, 64);
    int oi = __shfl_xor(bi, m, 64);
    if (key_before(ok, oi, bk, bi)) {
      bk = ok;
      bi = oi;
    }
  }
  if (lane == 0) {
    red_d[wv] = bk;
    red_i[wv] = bi;
  }
  __syncthreads();
  if (tid == 0) {
    double k0 = red_d[0];
    int i0 = red_i[0];
    for (int w = 1; w < nwv; ++w)
      if (key_before(red_d[w], red_i[w], k0, i0)) {
        k0 = red_d[w];
        i0 = red_i[w];
      }
    int w = i0;
    int cl = (int)(a.clone_u[r] * (double)(a.K - 1));
    if (cl > a.K - 2) cl = a.K - 2;
    if (cl >= w) cl += 1;
    s_worst = w;
    s_clone = cl;
  }
  __syncthreads();
  const int w = s_worst, cl = s_clone;
  // 3. extras: stable descending sort of noise with worst/clone excluded
  if (a.n_extra > 0) {
    const double* eu = a.extra_u + (size_t)r * a.K;
    for (int k = tid; k < a.Kpad; k += nt) {
      if (k < a.K) {
        skey[k] = (k == w || k == cl) ? -__builtin_inf() : eu[k];
        sidx[k] = k;
      } else {
        skey[k] = -__builtin_inf();
        sidx[k] = k;
      }
    }
    __syncthreads();
    bitonic_sort(skey, sidx, a.Kpad, key_before);
    for (int e = tid; e < a.n_extra; e += nt)
      a.walk[(size_t)r * a.n_walk + 1 + e] = sidx[e];
  }
  const int bp = a.st->block_pos;
  const size_t ws = (size_t)r * a.K + w, cs = (size_t)r * a.K + cl;
  if (tid == 0) {
    a.walk[(size_t)r * a.n_walk] = w;
    a.emax[r] = mx;
    const size_t o = (size_t)bp * a.R + r;
    a.o_emax[o] = mx;
    a.o_worst[o] = w;
    a.o_clone[o] = cl;
    a.o_dead_volume[o] = fabs(det3(a.pop.cell + ws * 9));
    a.o_dead_energy[o] = mx;
  }
  for (int i = tid; i < a.n_moves * JR_NCOUNT; i += nt)
    a.counters[(size_t)r * a.n_moves * JR_NCOUNT + i] = 0;
  if (r == 0 && tid < JR_NQUEUE) a.queue[tid] = 0;
  if (a.o_dead_pos) {
    const size_t o = (size_t)bp * a.R + r;
    for (int i = tid; i < a.Nmax * 3; i += nt)
      a.o_dead_pos[o * a.Nmax * 3 + i] = a.pop.pos[ws * a.Nmax * 3 + i];
    for (int i = tid; i < a.Nmax; i += nt)
      a.o_dead_types[o * a.Nmax + i] = a.pop.types[ws * a.Nmax + i];
    for (int i = tid; i < 9; i += nt)
      a.o_dead_cell[o * 9 + i] = a.pop.cell[ws * 9 + i];
  }
  __syncthreads();
  // 4. clone the survivor into the culled slot (all walker-bound fields)
  for (int i = tid; i < a.Nmax * 3; i += nt) {
    a.pop.pos[ws * a.Nmax * 3 + i] = a.pop.pos[cs * a.Nmax * 3 + i];
    a.pop.dir[ws * a.Nmax * 3 + i] = a.pop.dir[cs * a.Nmax * 3 + i];
    for (int e = 0; e < a.pop.n_xslots; ++e) {
      double* X = a.pop.dirx + e * a.pop.xstride;
      X[ws * a.Nmax * 3 + i] = X[cs * a.Nmax * 3 + i];
    }
  }
  for (int i = tid; i < a.Nmax; i += nt)
    a.pop.types[ws * a.Nmax + i] = a.pop.types[cs * a.Nmax + i];
  if (tid < 9) a.pop.cell[ws * 9 + tid] = a.pop.cell[cs * 9 + tid];
  if (tid == 0) {
    a.pop.energy[ws] = a.pop.energy[cs];
    a.pop.n_atoms[ws] = a.pop.n_atoms[cs];
  }
}

// nested_sampling.py:188-212 (n_cull = C > 1) and :486-518.  The walkers in
// lexicographic order: potential descending, then the noise sel_u
// descending, then index (jnp.lexsort((-noise, -potentials)) is stable).
// The first C are culled in that order, and the contour is the potential
// of the C-th.  The survivors order[C:] are permuted by sorting on the keys
// extra_u[0 .. K-C) ascending (equal keys by walker index), where survivor
// position p takes key extra_u[p]; the first C of the permutation are the
// clone sources of the culled slots, the next n_extra the extra chains.
// One block per replica; LDS holds Kpad (key, index) pairs and the C culled
// indices.
__global__ void __launch_bounds__(1024) select_cull_kernel(SelArgs a) {
  extern __shared__ double sh[];
  double* skey = sh;                                      // [Kpad]
  int* sidx = reinterpret_cast<int*>(sh + a.Kpad);        // [Kpad]
  int* culled = sidx + a.Kpad;                            // [C]
  const int r = blockIdx.x;
  const int tid = threadIdx.x;
  const int nt = blockDim.x;
  const int K = a.K, C = a.n_cull;
  const double* E = a.pop.energy + (size_t)r * K;
  const double* su = a.sel_u + (size_t)r * K;
  const double* pu = a.extra_u + (size_t)r * K;
  const double inf = __builtin_inf();
  // 1. lexicographic order; padding sorts last (-inf potential and noise)
  for (int k = tid; k < a.Kpad; k += nt) {
    skey[k] = k < K ? E[k] : -inf;
    sidx[k] = k;
  }
  __syncthreads();
  bitonic_sort(skey, sidx, a.Kpad, [&](double ka, int ia, double kb, int ib) {
    if (ka != kb) return ka > kb;
    const double na = ia < K ? su[ia] : -inf, nb = ib < K ? su[ib] : -inf;
    if (na != nb) return na > nb;
    return ia < ib;
  });
  const int bp = a.st->block_pos;
  const size_t o = (size_t)bp * a.R + r;
  if (tid == 0) {
    a.emax[r] = skey[C - 1];
    a.o_emax[o] = skey[C - 1];
  }
  for (int c = tid; c < C; c += nt) {
    const int w = sidx[c];
    culled[c] = w;
    a.walk[(size_t)r * a.n_walk + c] = w;
    a.o_worst[o * C + c] = w;
    a.o_dead_energy[o * C + c] = skey[c];
    a.o_dead_volume[o * C + c] = fabs(det3(a.pop.cell + ((size_t)r * K + w) * 9));
  }
  __syncthreads();
  // 2. permutation of the survivors; culled and padding entries sort last
  for (int k = tid; k < a.Kpad; k += nt) skey[k] = (k >= C && k < K) ? pu[k - C] : inf;
  __syncthreads();
  bitonic_sort(skey, sidx, a.Kpad, [](double ka, int ia, double kb, int ib) {
    return (ka < kb) || (ka == kb && ia < ib);
  });
  for (int c = tid; c < C; c += nt) a.o_clone[o * C + c] = sidx[c];
  for (int e = tid; e < a.n_extra; e += nt)
    a.walk[(size_t)r * a.n_walk + C + e] = sidx[C + e];
  for (int i = tid; i < a.n_moves * JR_NCOUNT; i += nt)
    a.counters[(size_t)r * a.n_moves * JR_NCOUNT + i] = 0;
  if (r == 0 && tid < JR_NQUEUE) a.queue[tid] = 0;
  const int n3 = a.Nmax * 3;
  const size_t base = (size_t)r * K;
  if (a.o_dead_pos) {
    for (int i = tid; i < C * n3; i += nt) {
      const int c = i / n3, q = i - c * n3;
      a.o_dead_pos[(o * C + c) * n3 + q] = a.pop.pos[(base + culled[c]) * n3 + q];
    }
    for (int i = tid; i < C * a.Nmax; i += nt) {
      const int c = i / a.Nmax, q = i - c * a.Nmax;
      a.o_dead_types[(o * C + c) * a.Nmax + q] =
          a.pop.types[(base + culled[c]) * a.Nmax + q];
    }
    for (int i = tid; i < C * 9; i += nt) {
      const int c = i / 9, q = i - c * 9;
      a.o_dead_cell[(o * C + c) * 9 + q] = a.pop.cell[(base + culled[c]) * 9 + q];
    }
  }
  __syncthreads();
  // 3. clone each source into its culled slot (sources and slots are disjoint)
  for (int i = tid; i < C * n3; i += nt) {
    const int c = i / n3, q = i - c * n3;
    const size_t ws = base + culled[c], cs = base + sidx[c];
    a.pop.pos[ws * n3 + q] = a.pop.pos[cs * n3 + q];
    a.pop.dir[ws * n3 + q] = a.pop.dir[cs * n3 + q];
    for (int e = 0; e < a.pop.n_xslots; ++e) {
      double* X = a.pop.dirx + e * a.pop.xstride;
      X[ws * n3 + q] = X[cs * n3 + q];
    }
  }
  for (int i = tid; i < C * a.Nmax; i += nt) {
    const int c = i / a.Nmax, q = i - c * a.Nmax;
    a.pop.types[(base + culled[c]) * a.Nmax + q] = a.pop.types[(base + sidx[c]) * a.Nmax + q];
  }
  for (int i = tid; i < C * 9; i += nt) {
    const int c = i / 9, q = i - c * 9;
    a.pop.cell[(base + culled[c]) * 9 + q] = a.pop.cell[(base + sidx[c]) * 9 + q];
  }
  for (int c = tid; c < C; c += nt) {
    a.pop.energy[base + culled[c]] = a.pop.energy[base + sidx[c]];
    a.pop.n_atoms[base + culled[c]] = a.pop.n_atoms[base + sidx[c]];
  }
}

// ------------------------------------------------------ replica exchange
struct ReArgs {
  PopPtrs pop;
  int R, K, Nmax, S, n_cycles, max_pairs;
  const double* pressure;
  const double* mu;
  const double* emax;
  const int* pair_valid;
  const double* re_u;
  int* acc;  // [R-1] (per-block output row or standalone)
  int* att;
  IterState* st;  // NULL: always fire (standalone pass)
  int* o_acc;     // per-block outputs
  int* o_att;
};

// Species counts of one walker by wave ballots over its atoms.
__device__ __forceinline__ void wave_counts(const PopPtrs& p, size_t slot, int Nmax, int S,
                                            int* counts) {
  const int lane = lane_id();
  const int N = p.n_atoms[slot];
#pragma unroll
  for (int s = 0; s < JR_MAX_SPECIES; ++s) counts[s] = 0;
  for (int base = 0; base < N; base += 64) {
    const int at = base + lane;
    const int ty = at < N ? p.types[slot * Nmax + at] : -1;
#pragma unroll
    for (int s = 0; s < JR_MAX_SPECIES; ++s)
      if (s < S) counts[s] += __popcll(__ballot(ty == s));
  }
}

// sampling/moves/replica_exchange.py:318-436 with EnsembleSwap (:120-164).
// Per cycle the even pairs (0,1), (2,3), .. then the odd pairs; the pairs of
// one phase are disjoint, so each wave of the block takes pairs of the phase
// in parallel (pair pi of a phase is replicas (ph + 2 pi, ph + 2 pi + 1), its
// draws re_u[cycle][ph][pi][2]) and the block synchronises between phases.
__global__ void __launch_bounds__(1024) re_kernel(ReArgs a) {
  const int tid = threadIdx.x, lane = lane_id();
  const int wave = tid >> 6, n_waves = blockDim.x >> 6;
  int* acc = a.acc;
  int* att = a.att;
  if (a.st) {
    const long long i = a.st->i;
    const bool fires = a.st->re_interval > 0 && i > 0 && (i % a.st->re_interval) == 0;
    const int bp = a.st->block_pos;
    acc = a.o_acc + (size_t)bp * (a.R - 1);
    att = a.o_att + (size_t)bp * (a.R - 1);
    for (int k = tid; k < a.R - 1; k += blockDim.x) {
      acc[k] = 0;
      att[k] = 0;
    }
    if (!fires) return;  // block-uniform
  } else {
    for (int k = tid; k < a.R - 1; k += blockDim.x) {
      acc[k] = 0;
      att[k] = 0;
    }
  }
  __syncthreads();
  const size_t n3 = (size_t)a.Nmax * 3;
  for (int cyc = 0; cyc < a.n_cycles; ++cyc) {
    for (int ph = 0; ph < 2; ++ph) {
      for (int pi = wave; ph + 2 * pi + 1 < a.R; pi += n_waves) {
        const int i = ph + 2 * pi, j = i + 1;
        if (!a.pair_valid[i]) continue;  // wave-uniform
        const double* u = a.re_u + (((size_t)cyc * 2 + ph) * a.max_pairs + pi) * 2;
        const int wi = min((int)(u[0] * (double)a.K), a.K - 1);
        const int wj = min((int)(u[1] * (double)a.K), a.K - 1);
        const size_t si = (size_t)i * a.K + wi, sj = (size_t)j * a.K + wj;
        int ci[JR_MAX_SPECIES], cj[JR_MAX_SPECIES];
        wave_counts(a.pop, si, a.Nmax, a.S, ci);
        wave_counts(a.pop, sj, a.Nmax, a.S, cj);
        const double* mui = a.mu + (size_t)i * a.S;
        const double* muj = a.mu + (size_t)j * a.S;
        const double* celli = a.pop.cell + si * 9;
        const double* cellj = a.pop.cell + sj * 9;
        // rebase_ensemble_energy (backends/ensemble.py:62-87)
        const double ui = a.pop.energy[sj] - ensemble_term(cellj, a.pressure[j], muj, a.S, cj);
        const double into_i = ui + ensemble_term(cellj, a.pressure[i], mui, a.S, cj);
        const double uj = a.pop.energy[si] - ensemble_term(celli, a.pressure[i], mui, a.S, ci);
        const double into_j = uj + ensemble_term(celli, a.pressure[j], muj, a.S, ci);
        const bool ok = (into_i < a.emax[i]) && (into_j < a.emax[j]);
        if (lane == 0) {
          att[i] += 1;
          acc[i] += ok ? 1 : 0;
        }
        if (ok) {  // wave-uniform
          for (size_t k = lane; k < n3; k += 64) {
            const double p0 = a.pop.pos[si * n3 + k], p1 = a.pop.pos[sj * n3 + k];
            const double d0 = a.pop.dir[si * n3 + k], d1 = a.pop.dir[sj * n3 + k];
            a.pop.pos[si * n3 + k] = p1;
            a.pop.pos[sj * n3 + k] = p0;
            a.pop.dir[si * n3 + k] = d1;
            a.pop.dir[sj * n3 + k] = d0;
            for (int e = 0; e < a.pop.n_xslots; ++e) {
              double* X = a.pop.dirx + e * a.pop.xstride;
              const double x0 = X[si * n3 + k], x1 = X[sj * n3 + k];
              X[si * n3 + k] = x1;
              X[sj * n3 + k] = x0;
            }
          }
          for (int k = lane; k < a.Nmax; k += 64) {
            const int t0 = a.pop.types[si * a.Nmax + k], t1 = a.pop.types[sj * a.Nmax + k];
            a.pop.types[si * a.Nmax + k] = t1;
            a.pop.types[sj * a.Nmax + k] = t0;
          }
          if (lane < 9) {
            const double c0 = a.pop.cell[si * 9 + lane], c1 = a.pop.cell[sj * 9 + lane];
            a.pop.cell[si * 9 + lane] = c1;
            a.pop.cell[sj * 9 + lane] = c0;
          }
          if (lane == 0) {
            const int n0 = a.pop.n_atoms[si];
            a.pop.n_atoms[si] = a.pop.n_atoms[sj];
            a.pop.n_atoms[sj] = n0;
            a.pop.energy[si] = into_i;
            a.pop.energy[sj] = into_j;
          }
        }
      }
      __syncthreads();
    }
  }
}

// ------------------------------------------------------------ draws
// Philox counter = (element-pair index lo, hi, rng iteration, stream).
__global__ void fill_uniform_kernel(double* out, long long n,
                                    unsigned long long seed,
                                    const IterState* st, long long it_fixed,
                                    unsigned stream) {
  const long long it = st ? st->rng_iter : it_fixed;
  for (long long p = (long long)blockIdx.x * blockDim.x + threadIdx.x;
       2 * p < n; p += (long long)gridDim.x * blockDim.x) {
    uint32_t c[4] = {(uint32_t)p, (uint32_t)(p >> 32), (uint32_t)it, stream};
    philox4x32_10(c, (uint32_t)seed, (uint32_t)(seed >> 32));
    out[2 * p] = u53(c[0], c[1]);
    if (2 * p + 1 < n) out[2 * p + 1] = u53(c[2], c[3]);
  }
}

// Normals of a stream in the four-per-block layout of normal_quad (the
// same values the walk kernel draws in place); out[e], e < n.
__global__ void fill_normal_kernel(double* out, long long n,
                                   unsigned long long seed,
                                   const IterState* st, long long it_fixed,
                                   unsigned stream) {
  const long long it = st ? st->rng_iter : it_fixed;
  for (long long q = (long long)blockIdx.x * blockDim.x + threadIdx.x;
       4 * q < n; q += (long long)gridDim.x * blockDim.x) {
    double z[4];
    normal_quad(seed, it, stream, q, z);
#pragma unroll
    for (int j = 0; j < 4; ++j)
      if (4 * q + j < n) out[4 * q + j] = z[j];
  }
}

// Shared move schedule: m = first index with u < cum[m] (sampling/mwg.py:243).
__global__ void schedule_kernel(int* sched, int n_steps,
                                unsigned long long seed, const IterState* st,
                                long long it_fixed, unsigned stream,
                                const double* cum, int n_moves) {
  const long long it = st ? st->rng_iter : it_fixed;
  for (int t = threadIdx.x; t < n_steps; t += blockDim.x) {
    uint32_t c[4] = {(uint32_t)t, 0u, (uint32_t)it, stream};
    philox4x32_10(c, (uint32_t)seed, (uint32_t)(seed >> 32));
    double u = u53(c[0], c[1]);
    int m = 0;
    while (m < n_moves - 1 && !(u < cum[m])) ++m;
    sched[t] = m;
  }
}

// The iteration's selection, clone, extra, schedule and replica-exchange
// draws in one launch: the same Philox values as fill_uniform_kernel on
// streams 1, 2, 3, 7 and schedule_kernel on stream 4, over one flat index
// space of uniform pairs [sel | clone | extra | re] followed by the schedule.
struct DrawArgs {
  double* sel_u;
  double* clone_u;
  double* extra_u;  // NULL: no extras drawn
  double* re_u;     // NULL: one replica
  int* sched;
  const double* cum;
  const IterState* st;
  unsigned long long seed;
  long long n_sel, n_clone, n_extra, n_re;  // values per buffer
  int n_steps, n_moves;
};

__device__ __forceinline__ void draw_pair(double* out, long long n, long long p,
                                          unsigned long long seed, long long it,
                                          unsigned stream) {
  uint32_t c[4] = {(uint32_t)p, (uint32_t)(p >> 32), (uint32_t)it, stream};
  philox4x32_10(c, (uint32_t)seed, (uint32_t)(seed >> 32));
  out[2 * p] = u53(c[0], c[1]);
  if (2 * p + 1 < n) out[2 * p + 1] = u53(c[2], c[3]);
}

__global__ void iter_draws_kernel(DrawArgs a) {
  const long long it = a.st->rng_iter;
  const long long p_sel = (a.n_sel + 1) >> 1, p_clone = (a.n_clone + 1) >> 1;
  const long long p_extra = a.extra_u ? (a.n_extra + 1) >> 1 : 0;
  const long long p_re = a.re_u ? (a.n_re + 1) >> 1 : 0;
  const long long b1 = p_sel, b2 = b1 + p_clone, b3 = b2 + p_extra, b4 = b3 + p_re;
  const long long total = b4 + a.n_steps;
  for (long long g = (long long)blockIdx.x * blockDim.x + threadIdx.x; g < total;
       g += (long long)gridDim.x * blockDim.x) {
    if (g < b1) {
      draw_pair(a.sel_u, a.n_sel, g, a.seed, it, 1u);
    } else if (g < b2) {
      draw_pair(a.clone_u, a.n_clone, g - b1, a.seed, it, 2u);
    } else if (g < b3) {
      draw_pair(a.extra_u, a.n_extra, g - b2, a.seed, it, 3u);
    } else if (g < b4) {
      draw_pair(a.re_u, a.n_re, g - b3, a.seed, it, 7u);
    } else {
      const int t = (int)(g - b4);
      uint32_t c[4] = {(uint32_t)t, 0u, (uint32_t)it, 4u};
      philox4x32_10(c, (uint32_t)a.seed, (uint32_t)(a.seed >> 32));
      const double u = u53(c[0], c[1]);
      int m = 0;
      while (m < a.n_moves - 1 && !(u < a.cum[m])) ++m;
      a.sched[t] = m;
    }
  }
}

// Per-iteration output snapshot (move counters and walked indices) into the
// block's output slot: a compute-queue copy, so no copy engine has to be
// synchronised with the walk stream each iteration.
__global__ void snapshot_kernel(const int* counters, int nc, const int* walk, int nw,
                                int* o_counters, int* o_walk) {
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < nc + nw;
       i += gridDim.x * blockDim.x) {
    if (i < nc)
      o_counters[i] = counters[i];
    else
      o_walk[i - nc] = walk[i - nc];
  }
}

__global__ void advance_kernel(IterState* st) {
  if (threadIdx.x == 0) {
    st->i += 1;
    st->rng_iter += 1;
    st->block_pos += 1;
  }
}

}  // namespace jr

// Batched move arithmetic, delayed-acceptance decisions, training-ring and
// minibatch kernels of the learned-model walk.  The C walked chains of one
// NS iteration are gathered into chain-indexed working arrays (chain
// c = r * n_walk + j walks slot walk[r][j] of replica r); the judge and the
// surrogate evaluate them with separate launches.  FP64 move arithmetic.
//
// GMC (sampling/moves/galilean.py:142-300): d = normalize(d_prev + a xi)
// (fresh xi when |d_prev| < 1e-8); n_reflect times {x += s d; evaluate;
// if H >= threshold: d -= 2 (F.d / |F|) F / |F|}; accept on the final test,
// else x reverts and d = -(initial direction).
// Volume (sampling/moves/volume.py:49-109, moves/_common.py:16-93):
// V' = |V + s N xi|, cell and positions scaled by (V'/V)^(1/3); prior
// (V'/V)^N unless flat; volume-per-atom and aspect bounds.
// HMC (sampling/moves/hmc.py:51-152): p = xi; n_leapfrog x {p += s/2 F;
// x += s p; F = F(x); p += s/2 F}; accept iff H' < emax and
// (H' + K') - (H + K) < 1, else x reverts.
#pragma once

#include <hip/hip_runtime.h>

#include "jr_device.h"

// FP64 move arithmetic rounds as written (the CPU restatement does the same)
#pragma clang fp contract(off)
#include "jr_types.h"
#include "../../include/jaxrens_amd.h"

namespace jr {

struct LWChains {
  int C, n_walk, Nmax, n_moves;
  const int* n_atoms;  // [C]
};

// Device counters of the learned walk (integers: order-independent sums).
struct LWStats {
  unsigned long long proposals, stage1_pass, accepted, exact_fe, exact_e, sur_e;
  double loss;  // last minibatch loss
};

// Held-out sums per replica over fresh exact evaluations, in per-atom units
// e = (U_s - U) / N and v = V / N: count, sum e, sum e^2, sum v, sum v^2,
// sum e v.
constexpr int LW_HS = 6;

struct RingState {
  long long head, size, cap;
};

// Decision codes per chain: 1 = goes to the judge's test; 0 = stage-1
// reject; 2 = cell-shape reject; 3 = prior reject.
__device__ __forceinline__ int reason_of(int code, bool accepted) {
  if (accepted) return 0;
  return (code == 2 || code == 3) ? code : 1;
}

template <typename Fn>
__device__ __forceinline__ double lw_chain_sum(int N, Fn f) {
  double acc = 0.0;
  for (int a = (int)(threadIdx.x & 63); a < N; a += 64) acc += f(a);
  return wave_sum(acc);
}

// ------------------------------------------------------------ gather/scatter
struct LWGatherArgs {
  int C, n_walk, K, Nmax;
  const int* walk;  // [R][n_walk]
  PopPtrs pop;
  const double* emax;      // [R]
  const double* pressure;  // [R]
  double *cpos, *cdir, *ccell, *cH, *cem, *cpres;
  int *ctypes, *cn;
};

// one block of 256 threads per chain
__global__ void __launch_bounds__(256) lw_gather_kernel(LWGatherArgs a) {
  const int c = blockIdx.x;
  if (c >= a.C) return;
  const int r = c / a.n_walk;
  const size_t slot = (size_t)r * a.K + a.walk[c];
  const size_t n3 = (size_t)a.Nmax * 3;
  for (int i = threadIdx.x; i < (int)n3; i += blockDim.x) {
    a.cpos[c * n3 + i] = a.pop.pos[slot * n3 + i];
    a.cdir[c * n3 + i] = a.pop.dir[slot * n3 + i];
  }
  for (int i = threadIdx.x; i < a.Nmax; i += blockDim.x)
    a.ctypes[(size_t)c * a.Nmax + i] = a.pop.types[slot * a.Nmax + i];
  if (threadIdx.x < 9) a.ccell[(size_t)c * 9 + threadIdx.x] = a.pop.cell[slot * 9 + threadIdx.x];
  if (threadIdx.x == 0) {
    a.cn[c] = a.pop.n_atoms[slot];
    a.cH[c] = a.pop.energy[slot];
    a.cem[c] = a.emax[r];
    a.cpres[c] = a.pressure[r];
  }
}

__global__ void __launch_bounds__(256) lw_scatter_kernel(LWGatherArgs a) {
  const int c = blockIdx.x;
  if (c >= a.C) return;
  const int r = c / a.n_walk;
  const size_t slot = (size_t)r * a.K + a.walk[c];
  const size_t n3 = (size_t)a.Nmax * 3;
  for (int i = threadIdx.x; i < (int)n3; i += blockDim.x) {
    a.pop.pos[slot * n3 + i] = a.cpos[c * n3 + i];
    a.pop.dir[slot * n3 + i] = a.cdir[c * n3 + i];
  }
  if (threadIdx.x < 9) a.pop.cell[slot * 9 + threadIdx.x] = a.ccell[(size_t)c * 9 + threadIdx.x];
  if (threadIdx.x == 0) a.pop.energy[slot] = a.cH[c];
}

// ------------------------------------------------------------------- GMC
// xi: normals of this sub-step for chain 0, chain stride xi_stride,
// component-major [3][Nmax] per chain.  One wave per chain (4 per block).
__global__ void __launch_bounds__(256) lw_gmc_begin_kernel(LWChains ch, double* dir,
                                                           const double* xi, size_t xi_stride,
                                                           double perturb, const double* pos,
                                                           double* x0, double* d0) {
  const int c = blockIdx.x * 4 + (threadIdx.x >> 6);
  if (c >= ch.C) return;
  const int N = ch.n_atoms[c];
  const size_t o = (size_t)c * ch.Nmax * 3;
  double* d = dir + o;
  const double* z = xi + (size_t)c * xi_stride;
  const int Nm = ch.Nmax;
  const double dn = sqrt(lw_chain_sum(N, [&](int a) {
    return dot3(d[3 * a], d[3 * a + 1], d[3 * a + 2], d[3 * a], d[3 * a + 1], d[3 * a + 2]);
  }));
  const bool fresh = dn < 1e-8;
  auto v = [&](int a, int k) {
    const double zk = z[(size_t)k * Nm + a];
    return fresh ? zk : d[3 * a + k] + perturb * zk;
  };
  const double vn = sqrt(lw_chain_sum(N, [&](int a) {
    const double v0 = v(a, 0), v1 = v(a, 1), v2 = v(a, 2);
    return dot3(v0, v1, v2, v0, v1, v2);
  }));
  const double den = fmax(vn, 1e-10);
  for (int a = (int)(threadIdx.x & 63); a < N; a += 64) {
    const double w0 = v(a, 0) / den, w1 = v(a, 1) / den, w2 = v(a, 2) / den;
    d[3 * a] = w0;
    d[3 * a + 1] = w1;
    d[3 * a + 2] = w2;
    d0[o + 3 * a] = w0;
    d0[o + 3 * a + 1] = w1;
    d0[o + 3 * a + 2] = w2;
    x0[o + 3 * a] = pos[o + 3 * a];
    x0[o + 3 * a + 1] = pos[o + 3 * a + 1];
    x0[o + 3 * a + 2] = pos[o + 3 * a + 2];
  }
}

// pos += step * dir, step = step_sizes[r][m] of the chain's replica
__global__ void lw_advance_kernel(LWChains ch, double* pos, const double* dir,
                                  const double* step_sizes, int m) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  const size_t per = (size_t)ch.Nmax * 3;
  if (i >= (size_t)ch.C * per) return;
  const int c = (int)(i / per);
  if ((int)((i % per) / 3) >= ch.n_atoms[c]) return;
  const double s = step_sizes[(size_t)(c / ch.n_walk) * ch.n_moves + m];
  pos[i] = fma(s, dir[i], pos[i]);
}

// H = U + P |det cell| (NaN -> +inf) for all chains or the listed ones.
// With a surrogate energy U the corrected enthalpy
// H = U - (ca N + cv V) + P V removes the fitted per-atom error
// e = ca + cv V / N (ca = cv = 0 for the judge; n may then be NULL).
__global__ void lw_ham_kernel(int C, const int* idx, const int* count, const double* U,
                              const double* cell, const double* pres, const int* n,
                              double ca, double cv, double* H) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  const int cnt = count ? *count : C;
  if (k >= cnt) return;
  const int c = idx ? idx[k] : k;
  const double V = fabs(det3(cell + (size_t)c * 9));
  double h = U[c] + pres[c] * V;
  if (n) h -= ca * (double)n[c] + cv * V;
  if (__builtin_isnan(h)) h = __builtin_inf();
  H[c] = h;
}

// Reflect d on the plane normal to F where H >= thr (one wave per chain).
__global__ void __launch_bounds__(256) lw_reflect_kernel(LWChains ch, double* dir,
                                                         const double* F, const double* H,
                                                         const double* thr) {
  const int c = blockIdx.x * 4 + (threadIdx.x >> 6);
  if (c >= ch.C) return;
  if (!(H[c] >= thr[c])) return;  // wave-uniform
  const int N = ch.n_atoms[c];
  const size_t o = (size_t)c * ch.Nmax * 3;
  double* d = dir + o;
  const double* f = F + o;
  const double fsq = lw_chain_sum(N, [&](int a) {
    return dot3(f[3 * a], f[3 * a + 1], f[3 * a + 2], f[3 * a], f[3 * a + 1], f[3 * a + 2]);
  });
  const double fdt = lw_chain_sum(N, [&](int a) {
    return dot3(f[3 * a], f[3 * a + 1], f[3 * a + 2], d[3 * a], d[3 * a + 1], d[3 * a + 2]);
  });
  const double fn = fmax(sqrt(fsq), 1e-10);
  const double t = 2.0 * (fdt / fn);
  for (int a = (int)(threadIdx.x & 63); a < N; a += 64)
#pragma unroll
    for (int k = 0; k < 3; ++k) d[3 * a + k] -= t * (f[3 * a + k] / fn);
}

// Rejected chains: positions back to x0, direction -d0.
__global__ void lw_gmc_finish_kernel(LWChains ch, double* pos, double* dir, const double* x0,
                                     const double* d0, const int* accept) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  const size_t per = (size_t)ch.Nmax * 3;
  if (i >= (size_t)ch.C * per) return;
  const int c = (int)(i / per);
  if ((int)((i % per) / 3) >= ch.n_atoms[c] || accept[c]) return;
  pos[i] = x0[i];
  dir[i] = -d0[i];
}

// ------------------------------------------------------------------- HMC
// p = xi (component-major tape slot), x0 = x, K0 = p.p / 2 (one wave per chain).
__global__ void __launch_bounds__(256) lw_hmc_begin_kernel(LWChains ch, const double* xi,
                                                           size_t xi_stride, const double* pos,
                                                           double* x0, double* p, double* K0) {
  const int c = blockIdx.x * 4 + (threadIdx.x >> 6);
  if (c >= ch.C) return;
  const int N = ch.n_atoms[c];
  const size_t o = (size_t)c * ch.Nmax * 3;
  const double* z = xi + (size_t)c * xi_stride;
  for (int a = (int)(threadIdx.x & 63); a < N; a += 64)
#pragma unroll
    for (int k = 0; k < 3; ++k) {
      p[o + 3 * a + k] = z[(size_t)k * ch.Nmax + a];
      x0[o + 3 * a + k] = pos[o + 3 * a + k];
    }
  const double pp = lw_chain_sum(N, [&](int a) {
    const double* q = p + o + 3 * a;
    return dot3(q[0], q[1], q[2], q[0], q[1], q[2]);
  });
  if ((threadIdx.x & 63) == 0) K0[c] = 0.5 * pp;
}

// kicks (one, or two: the closing half kick of the previous leapfrog step and
// the opening one of this step, rounded separately), then the drift
__global__ void lw_hmc_kick_drift_kernel(LWChains ch, double* p, double* pos, const double* F,
                                         const double* step_sizes, int m, int kicks,
                                         int drift) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  const size_t per = (size_t)ch.Nmax * 3;
  if (i >= (size_t)ch.C * per) return;
  const int c = (int)(i / per);
  if ((int)((i % per) / 3) >= ch.n_atoms[c]) return;
  const double s = step_sizes[(size_t)(c / ch.n_walk) * ch.n_moves + m];
  const double h = 0.5 * s;
  double q = p[i];
  for (int k = 0; k < kicks; ++k) q = q + h * F[i];
  p[i] = q;
  if (drift) pos[i] = pos[i] + s * q;
}

// The HMC decision; rejected chains return to x0.  Plain (judge forces;
// code == NULL): the reference's rule H' < Emax and (H' + K') - (H + K) < 1
// (hmc.py:51-152, marked unvalidated there for this reason: a hard cut is
// no detailed-balance accept).  With unit masses, momenta N(0, 1) and a
// small step it accepts almost every move, and the exact flow keeps
// exp(-H - K), so the move leans toward exp(-H) inside the contour rather
// than the flat constrained prior.  Screened (code != NULL): the
// surrogate's forces drive the leapfrog, and the move is the exact
// Metropolis step of the constrained prior uniform(x) N(p) under the
// leapfrog-then-flip map, a volume-preserving involution for any force
// field: stage 1 agrees, H' < Emax and u < exp(K - K'), u the sub-step's
// uniform 0 (chain stride u_stride), exp_c so that every build decides
// alike.
__global__ void __launch_bounds__(256) lw_hmc_finish_kernel(LWChains ch, const int* code,
                                                            const double* Hy, double* Hcur,
                                                            const double* K0, const double* p,
                                                            const double* em, int* accept,
                                                            int* reason, double* pos,
                                                            const double* x0, double* Hs_cur,
                                                            const double* Hs_new,
                                                            const double* u, size_t u_stride) {
  const int c = blockIdx.x * 4 + (threadIdx.x >> 6);
  if (c >= ch.C) return;
  const int N = ch.n_atoms[c];
  const size_t o = (size_t)c * ch.Nmax * 3;
  const double pp = lw_chain_sum(N, [&](int a) {
    const double* q = p + o + 3 * a;
    return dot3(q[0], q[1], q[2], q[0], q[1], q[2]);
  });
  const int k = code ? code[c] : 1;
  bool acc = false;
  if (k == 1) {
    const double K1 = 0.5 * pp;
    if (code)
      acc = (Hy[c] < em[c]) && (u[(size_t)c * u_stride] < exp_c(K0[c] - K1));
    else
      acc = (Hy[c] < em[c]) && ((Hy[c] + K1) - (Hcur[c] + K0[c]) < 1.0);
  }
  if (!acc)
    for (int a = (int)(threadIdx.x & 63); a < N; a += 64)
#pragma unroll
      for (int q = 0; q < 3; ++q) pos[o + 3 * a + q] = x0[o + 3 * a + q];
  if ((threadIdx.x & 63) == 0) {
    accept[c] = acc;
    reason[c] = acc ? 0 : 1;
    if (acc) {
      Hcur[c] = Hy[c];
      if (Hs_cur) Hs_cur[c] = Hs_new[c];
    }
  }
}

// ---------------------------------------------------------------- volume
struct LWVolArgs {
  LWChains ch;
  const double* pos;
  const double* cell;
  const double* step_sizes;
  int m;
  const double* xi;  // normals of this sub-step, chain stride xi_stride (first value used)
  size_t xi_stride;
  const double* u;   // uniforms of this sub-step, chain stride u_stride (first value used)
  size_t u_stride;
  double* npos;
  double* ncell;
  int* code;         // 1 = to the energy test, 2 = cell reject, 3 = prior reject
  double max_vpa, min_vpa, min_aspect;
  int flat;
};

__global__ void __launch_bounds__(256) lw_volume_propose_kernel(LWVolArgs a) {
  const int c = blockIdx.x * 4 + (threadIdx.x >> 6);
  if (c >= a.ch.C) return;
  const int N = a.ch.n_atoms[c];
  const double* cl = a.cell + (size_t)c * 9;
  const double step = a.step_sizes[(size_t)(c / a.ch.n_walk) * a.ch.n_moves + a.m];
  const double oldV = fabs(det3(cl));
  const double dV = (step * (double)N) * a.xi[(size_t)c * a.xi_stride];
  const double newV = fabs(oldV + dV);
  const double ratio = newV / oldV;
  const double scale = cbrt_c(ratio);
  const size_t o = (size_t)c * a.ch.Nmax * 3;
  for (int i = (int)(threadIdx.x & 63); i < 3 * N; i += 64) a.npos[o + i] = a.pos[o + i] * scale;
  if ((threadIdx.x & 63) == 0) {
    double nc[9];
    for (int i = 0; i < 9; ++i) {
      nc[i] = cl[i] * scale;
      a.ncell[(size_t)c * 9 + i] = nc[i];
    }
    const bool cell_ok = check_cell_shape(nc, N, a.max_vpa, a.min_vpa, a.min_aspect);
    const double p_acc = a.flat ? 1.0 : fmin(1.0, ipow_c(ratio, N));
    a.code[c] = !cell_ok ? 2 : (a.u[(size_t)c * a.u_stride] < p_acc ? 1 : 3);
  }
}

// Accepted volume moves: the proposal replaces the chain's state.
__global__ void lw_volume_commit_kernel(LWChains ch, double* pos, double* cell,
                                        const double* npos, const double* ncell,
                                        const int* accept) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  const size_t per = (size_t)ch.Nmax * 3;
  if (i >= (size_t)ch.C * per) return;
  const int c = (int)(i / per);
  if (!accept[c]) return;
  const size_t k = i % per;
  if ((int)(k / 3) < ch.n_atoms[c]) pos[i] = npos[i];
  if (k < 9) cell[(size_t)c * 9 + k] = ncell[(size_t)c * 9 + k];
}

// ------------------------------------------------------------ decisions
// Stage 1: code_in (NULL = all 1) -> code_out; a chain whose code is 1 keeps
// it iff [Hs_x < thr] == [Hs_y < thr], else gets 0.
__global__ void lw_stage1_kernel(int C, const int* code_in, const double* Hs_x,
                                 const double* Hs_y, const double* thr, int* code_out) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= C) return;
  const int k = code_in ? code_in[c] : 1;
  if (k != 1) {
    code_out[c] = k;
    return;
  }
  code_out[c] = ((Hs_x[c] < thr[c]) == (Hs_y[c] < thr[c])) ? 1 : 0;
}

// Surrogate contours per chain: steering at the estimated contour
// emax + N bias_r, screening with slack emax + N (bias_r + k sigma_r).
__global__ void lw_thr_kernel(int C, int n_walk, const double* em, const int* n,
                              const double* bias, const double* sig, double k,
                              double* thr_steer, double* thr_screen) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= C) return;
  const int r = c / n_walk;
  const double N = (double)n[c];
  thr_steer[c] = em[c] + N * bias[r];
  thr_screen[c] = em[c] + N * (bias[r] + k * sig[r]);
}

// Upper bound on U / N over every state the chains' contours admit:
// U = H - P V < Emax - P V, with V fixed (no volume move, vol_lo = 0) or
// inside [vol_lo, vol_hi] N.  One block; the maximum over chains goes to
// *out.  The auto walk screens only while this bound stays inside the
// range the surrogate is trained on (U / N <= train_max_epa).
__global__ void __launch_bounds__(256) lw_contour_kernel(int C, const double* em,
                                                         const double* pres, const double* cell,
                                                         const int* n, double vol_lo,
                                                         double vol_hi, double* out) {
  __shared__ double part[256];
  double m = -INFINITY;
  for (int c = threadIdx.x; c < C; c += 256) {
    const double N = (double)n[c], P = pres[c];
    const double V = vol_lo > 0.0 ? (P >= 0.0 ? vol_lo : vol_hi) * N : fabs(det3(cell + 9 * c));
    m = fmax(m, (em[c] - P * V) / N);
  }
  part[threadIdx.x] = m;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) part[threadIdx.x] = fmax(part[threadIdx.x], part[threadIdx.x + w]);
    __syncthreads();
  }
  if (threadIdx.x == 0) *out = part[0];
}

// Final test: accept iff code == 1 (NULL = all) and H < emax.  Accepted
// chains take H (and the surrogate's Hs_new into Hs_cur when given).
__global__ void lw_accept_kernel(int C, const int* code, const double* H, const double* em,
                                 int* accept, int* reason, double* Hcur, double* Hs_cur,
                                 const double* Hs_new) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= C) return;
  const int k = code ? code[c] : 1;
  const bool acc = (k == 1) && (H[c] < em[c]);
  accept[c] = acc;
  reason[c] = reason_of(k, acc);
  if (acc) {
    Hcur[c] = H[c];
    if (Hs_cur) Hs_cur[c] = Hs_new[c];
  }
}

// Per-replica counters (sampling/nested_sampling.py:533-609 buckets) and the
// walk statistics.  n_eval/n_grad: per chain constant plus per_code1 for
// chains whose code is 1.
struct LWCountArgs {
  int C, n_walk, n_moves, m;
  const int* reason;
  const int* code;  // NULL = all 1
  int eval_const, eval_code1, grad_const;
  int* counters;    // [R][n_moves][JR_NCOUNT]
  LWStats* stats;
  int screened;     // count stage-1 passes
};

__global__ void lw_count_kernel(LWCountArgs a) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= a.C) return;
  const int r = c / a.n_walk;
  const int k = a.code ? a.code[c] : 1;
  const int rs = a.reason[c];
  int* ct = a.counters + ((size_t)r * a.n_moves + a.m) * JR_NCOUNT;
  atomicAdd(ct + 0, 1);
  atomicAdd(ct + 1 + rs, 1);
  const int ne = a.eval_const + (k == 1 ? a.eval_code1 : 0);
  if (ne) atomicAdd(ct + 5, ne);
  if (a.grad_const) atomicAdd(ct + 6, a.grad_const);
  if (!a.stats) return;
  atomicAdd(&a.stats->proposals, 1ull);
  if (rs == 0) atomicAdd(&a.stats->accepted, 1ull);
  if (a.screened && k == 1) atomicAdd(&a.stats->stage1_pass, 1ull);
}

// dst += (count ? *count : n) * mult  (single thread)
__global__ void lw_stat_add_kernel(unsigned long long* dst, const int* count, long long n,
                                   long long mult) {
  if (threadIdx.x == 0 && blockIdx.x == 0)
    *dst += (unsigned long long)((count ? (long long)*count : n) * mult);
}

// Deterministic stream compaction (one block of 1024 threads): idx lists,
// in increasing c, the chains with flag(c); count receives their number.
// mode 0: code[c] == want (all chains when code is NULL); mode 1: that and
// U[c] finite with U[c] / n[c] <= cap.
__global__ void __launch_bounds__(1024) lw_compact_kernel(int C, const int* code, int want,
                                                          int mode, const double* U,
                                                          const int* n, double cap, int* idx,
                                                          int* count) {
  __shared__ int part[1024];
  const int t = threadIdx.x;
  const int per = (C + 1023) / 1024;
  const int lo = t * per, hi = min(C, lo + per);
  auto flag = [&](int c) {
    if (code && code[c] != want) return false;
    if (mode == 0) return true;
    const double u = U[c];
    return __builtin_isfinite(u) && u <= cap * (double)n[c];
  };
  int local = 0;
  for (int c = lo; c < hi; ++c) local += flag(c) ? 1 : 0;
  part[t] = local;
  __syncthreads();
  for (int off = 1; off < 1024; off <<= 1) {
    const int v = t >= off ? part[t - off] : 0;
    __syncthreads();
    part[t] += v;
    __syncthreads();
  }
  int pos = part[t] - local;
  for (int c = lo; c < hi; ++c)
    if (flag(c)) idx[pos++] = c;
  if (t == 1023) *count = part[1023];
}

// ------------------------------------------------- held-out error + ring
// Held-out sums (LW_HS values per replica) over the listed chains, one
// block per replica, fixed order.  V from the evaluated configuration's cell.
__global__ void __launch_bounds__(256) lw_heldout_kernel(const int* idx, const int* count,
                                                         int n_walk, const double* U,
                                                         const double* Us, const int* n,
                                                         const double* cell, double* hs) {
  __shared__ double part[LW_HS][256];
  const int r = blockIdx.x;
  const int cnt = *count;
  double s[LW_HS] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
  for (int k = threadIdx.x; k < cnt; k += 256) {
    const int c = idx[k];
    if (c / n_walk != r) continue;
    const double N = (double)n[c];
    const double e = (Us[c] - U[c]) / N;
    const double v = fabs(det3(cell + (size_t)c * 9)) / N;
    s[0] += 1.0;
    s[1] += e;
    s[2] += e * e;
    s[3] += v;
    s[4] += v * v;
    s[5] += e * v;
  }
#pragma unroll
  for (int q = 0; q < LW_HS; ++q) part[q][threadIdx.x] = s[q];
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w)
#pragma unroll
      for (int q = 0; q < LW_HS; ++q) part[q][threadIdx.x] += part[q][threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x < LW_HS) hs[(size_t)r * LW_HS + threadIdx.x] += part[threadIdx.x][0];
}

// Listed chains' descriptor rows, training targets (judge energy minus the
// surrogate's pair baseline) and sizes into the ring at head + k (only the
// last cap entries of an oversized list are kept).
// One block per list entry.
__global__ void __launch_bounds__(256) lw_ring_append_kernel(const int* idx, const int* count,
                                                             int Nmax, const __bf16* xrows,
                                                             const double* U, const double* Ub,
                                                             const int* n,
                                                             const RingState* rs, __bf16* ring,
                                                             double* ringE, int* ringN) {
  const int k = blockIdx.x;
  const int cnt = *count;
  if (k >= cnt) return;
  const long long cap = rs->cap;
  if ((long long)cnt - k > cap) return;
  const long long slot = (rs->head + k) % cap;
  const int c = idx[k];
  const int nc = n[c];
  const size_t per = (size_t)Nmax * 80;
  const __bf16* src = xrows + (size_t)c * per;
  __bf16* dst = ring + (size_t)slot * per;
  for (size_t i = threadIdx.x; i < per; i += 256)
    dst[i] = (i / 80 < (size_t)nc) ? src[i] : (__bf16)0.0f;
  if (threadIdx.x == 0) {
    ringE[slot] = U[c] - Ub[c];  // the network's target: judge minus baseline
    ringN[slot] = nc;
  }
}

__global__ void lw_ring_commit_kernel(const int* count, RingState* rs) {
  if (threadIdx.x == 0 && blockIdx.x == 0) {
    const long long cnt = *count;
    rs->head = (rs->head + cnt) % rs->cap;
    rs->size = min(rs->size + cnt, rs->cap);
  }
}

// ------------------------------------------------------------- training
// Minibatch of Bt ring entries drawn uniformly (Philox, counter
// (b, step, stream 10)).
__global__ void lw_train_draw_kernel(int Bt, const RingState* rs, unsigned long long seed,
                                     long long step, int* idx) {
  const int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= Bt) return;
  uint32_t c[4] = {(uint32_t)b, 0u, (uint32_t)step, 10u};
  c[1] = (uint32_t)(step >> 32);
  philox4x32_10(c, (uint32_t)seed, (uint32_t)(seed >> 32));
  const long long size = rs->size > 0 ? rs->size : 1;
  long long k = (long long)(u53(c[0], c[1]) * (double)size);
  if (k >= size) k = size - 1;
  idx[b] = (int)k;
}

__global__ void lw_train_gather_kernel(int Bt, int Nmax, const int* idx, const __bf16* ring,
                                       const double* ringE, const int* ringN, __bf16* X,
                                       double* E, int* n) {
  const size_t per = (size_t)Nmax * 80;
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < (size_t)Bt * per) {
    const size_t b = i / per;
    X[i] = ring[(size_t)idx[b] * per + i % per];
  }
  if (i < (size_t)Bt) {
    E[i] = ringE[idx[i]];
    n[i] = ringN[idx[i]];
  }
}

// Loss 1/Bt sum_b ((E_hat_b - E_b) / n_b)^2 with E_hat_b the sum of the
// first n_b row energies; c_rows = dloss/de_row (0 for padding rows); sq[b].
__global__ void __launch_bounds__(256) lw_loss_weights_kernel(int Bt, int Nmax,
                                                              const float* e_rows,
                                                              const double* E, const int* n,
                                                              float* c_rows, double* sq) {
  const int b = blockIdx.x;
  if (b >= Bt) return;
  __shared__ double part[256];
  const int nb = n[b];
  double s = 0.0;
  for (int i = threadIdx.x; i < nb; i += 256) s += (double)e_rows[(size_t)b * Nmax + i];
  part[threadIdx.x] = s;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) part[threadIdx.x] += part[threadIdx.x + w];
    __syncthreads();
  }
  const double d = (part[0] - E[b]) / (double)nb;
  const float cb = (float)(2.0 * d / ((double)nb * Bt));
  for (int i = threadIdx.x; i < Nmax; i += 256) c_rows[(size_t)b * Nmax + i] = i < nb ? cb : 0.0f;
  if (threadIdx.x == 0) sq[b] = d * d;
}

__global__ void lw_loss_sum_kernel(int Bt, const double* sq, LWStats* st) {
  if (threadIdx.x == 0 && blockIdx.x == 0) {
    double s = 0.0;
    for (int b = 0; b < Bt; ++b) s += sq[b];
    st->loss = s / Bt;
  }
}

}  // namespace jr

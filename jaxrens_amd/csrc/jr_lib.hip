// Host side of libjrhip: handle, device buffers, launches, C ABI.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "jr_handle.h"
#include "jr_iter_kernels.h"
#include "jr_launch.h"
#include "jr_pot.h"

using namespace jr;

namespace jrh {

thread_local std::string g_err;

int set_global_err(int code, const char* fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof(buf), fmt, ap);
  va_end(ap);
  g_err = buf;
  return code;
}

}  // namespace jrh

using namespace jrh;



namespace {

long long pow2_at_least(long long n) {
  long long p = 1;
  while (p < n) p <<= 1;
  return p;
}

}  // namespace

extern "C" {

const char* jr_global_error(void) { return g_err.c_str(); }

const char* jr_last_error(const jr_handle* h) {
  return h ? h->err.c_str() : g_err.c_str();
}

int jr_eval_energy_forces(const jr_pot* pot, int32_t device, int32_t B,
                          int32_t Nmax, const double* pos,
                          const int32_t* types, const int32_t* n_atoms,
                          const double* cell, const double* pressure,
                          const double* mu, double* U, double* H, double* F) {
  int rc = check_eval_args(pot, B, Nmax, pos, types, n_atoms, cell);
  if (rc) return rc;
  if (B == 0) return JR_OK;
  HIPCHK_G(hipSetDevice(device));
  const int S = pot->n_species;
  DBuf<double> d_pos, d_cell, d_pr, d_mu, d_U, d_H, d_F, d_e4, d_e24, d_s2;
  DBuf<int> d_t, d_n;
  HIPCHK_G(d_pos.alloc((size_t)B * Nmax * 3));
  HIPCHK_G(d_cell.alloc((size_t)B * 9));
  HIPCHK_G(d_t.alloc((size_t)B * Nmax));
  HIPCHK_G(d_n.alloc((size_t)B));
  HIPCHK_G(d_U.alloc((size_t)B));
  HIPCHK_G(d_H.alloc((size_t)B));
  HIPCHK_G(d_e4.alloc((size_t)S * S));
  HIPCHK_G(d_e24.alloc((size_t)S * S));
  HIPCHK_G(d_s2.alloc((size_t)S * S));
  if (pressure) HIPCHK_G(d_pr.alloc((size_t)B));
  if (mu) HIPCHK_G(d_mu.alloc((size_t)B * S));
  if (F) HIPCHK_G(d_F.alloc((size_t)B * Nmax * 3));
  hipStream_t s;
  HIPCHK_G(hipStreamCreate(&s));
  auto h2d = [&](void* d, const void* hsrc, size_t bytes) {
    return hipMemcpyAsync(d, hsrc, bytes, hipMemcpyHostToDevice, s);
  };
  HIPCHK_G(h2d(d_pos.p, pos, sizeof(double) * B * Nmax * 3));
  HIPCHK_G(h2d(d_cell.p, cell, sizeof(double) * B * 9));
  HIPCHK_G(h2d(d_t.p, types, sizeof(int) * B * Nmax));
  HIPCHK_G(h2d(d_n.p, n_atoms, sizeof(int) * B));
  HIPCHK_G(h2d(d_e4.p, pot->eps4, sizeof(double) * S * S));
  HIPCHK_G(h2d(d_e24.p, pot->eps24, sizeof(double) * S * S));
  HIPCHK_G(h2d(d_s2.p, pot->sig2, sizeof(double) * S * S));
  if (pressure) HIPCHK_G(h2d(d_pr.p, pressure, sizeof(double) * B));
  if (mu) HIPCHK_G(h2d(d_mu.p, mu, sizeof(double) * B * S));
  EvalArgs a;
  std::memset(&a, 0, sizeof(a));
  a.pos = d_pos.p;
  a.types = d_t.p;
  a.n_atoms = d_n.p;
  a.cell = d_cell.p;
  a.pressure = pressure ? d_pr.p : nullptr;
  a.mu = mu ? d_mu.p : nullptr;
  a.U = d_U.p;
  a.H = d_H.p;
  a.F = F ? d_F.p : nullptr;
  a.B = B;
  a.Nmax = Nmax;
  a.S = S;
  a.rc2 = pot->rc2;
  a.eps4 = d_e4.p;
  a.eps24 = d_e24.p;
  a.sig2 = d_s2.p;
  set_images(a.img, *pot);
  a.chunks = pot_chunks(pot);
  a.team = team_waves(atoms_per_lane(Nmax), a.chunks, S, device);
  launch_eval(atoms_per_lane(Nmax), pot_cfg(pot), a, s);
  HIPCHK_G(hipGetLastError());
  if (U) HIPCHK_G(hipMemcpyAsync(U, d_U.p, sizeof(double) * B, hipMemcpyDeviceToHost, s));
  if (H) HIPCHK_G(hipMemcpyAsync(H, d_H.p, sizeof(double) * B, hipMemcpyDeviceToHost, s));
  if (F)
    HIPCHK_G(hipMemcpyAsync(F, d_F.p, sizeof(double) * B * Nmax * 3,
                            hipMemcpyDeviceToHost, s));
  HIPCHK_G(hipStreamSynchronize(s));
  HIPCHK_G(hipStreamDestroy(s));
  return JR_OK;
}

void jr_destroy(jr_handle* h) {
  if (!h) return;
  (void)hipSetDevice(h->cfg.device);
  if (h->stream) (void)hipStreamSynchronize(h->stream);
  jrl::learned_destroy(h);
  if (h->graph_exec) (void)hipGraphExecDestroy(h->graph_exec);
  if (h->graph) (void)hipGraphDestroy(h->graph);
  for (auto e : h->ev) (void)hipEventDestroy(e);
  if (h->stream2) (void)hipStreamSynchronize(h->stream2);
  if (h->ev_fork) (void)hipEventDestroy(h->ev_fork);
  if (h->ev_join) (void)hipEventDestroy(h->ev_join);
  if (h->stream2) (void)hipStreamDestroy(h->stream2);
  if (h->stream) (void)hipStreamDestroy(h->stream);
  delete h;  // DBuf members release in their owners' scope below
}

}  // extern "C"

namespace {

void release_all(jr_handle* h) {
  h->pos.release(); h->dir.release(); h->cell.release(); h->energy.release();
  h->types.release(); h->n_atoms.release();
  h->pressure.release(); h->mu.release(); h->step_sizes.release(); h->emax.release();
  h->eps4.release(); h->eps24.release(); h->sig2.release(); h->dmin.release(); h->moves.release(); h->cum.release();
  h->walk.release(); h->counters.release(); h->queue.release(); h->pair_valid.release();
  h->order.release(); h->sched.release();
  h->sel_u.release(); h->clone_u.release(); h->extra_u.release(); h->normals.release();
  h->uniforms.release(); h->re_u.release(); h->st.release();
  h->o_emax.release(); h->o_dead_volume.release(); h->o_dead_pos.release(); h->o_dead_cell.release();
  h->o_dead_energy.release();
  h->o_worst.release(); h->o_clone.release(); h->o_dead_types.release(); h->o_counters.release();
  h->o_re_acc.release(); h->o_re_att.release(); h->o_walk.release();
  h->t_idx.release(); h->t_reason.release(); h->t_step.release(); h->t_emax.release();
  h->t_normals.release(); h->t_uniforms.release();
}

int ensure_out(jr_handle* h, int n_iter) {
  if (n_iter <= h->out_cap) return JR_OK;
  const auto& c = h->cfg;
  const int R = c.n_replicas, Nm = c.n_atoms_max;
  // capacity in powers of two from 64 iterations, so that ordinary blocks
  // allocate nothing once the handle exists
  size_t n = 64;
  while (n < (size_t)n_iter) n *= 2;
  const size_t D = n * R * c.n_cull;  // dead points in the block
  HIPCHK(h, h->o_emax.alloc(n * R));
  HIPCHK(h, h->o_dead_energy.alloc(D));
  HIPCHK(h, h->o_dead_volume.alloc(D));
  HIPCHK(h, h->o_dead_pos.alloc(D * Nm * 3));
  HIPCHK(h, h->o_dead_cell.alloc(D * 9));
  HIPCHK(h, h->o_worst.alloc(D));
  HIPCHK(h, h->o_clone.alloc(D));
  HIPCHK(h, h->o_dead_types.alloc(D * Nm));
  HIPCHK(h, h->o_counters.alloc(n * R * c.n_moves * JR_NCOUNT));
  HIPCHK(h, h->o_re_acc.alloc(n * std::max(R - 1, 1)));
  HIPCHK(h, h->o_re_att.alloc(n * std::max(R - 1, 1)));
  HIPCHK(h, h->o_walk.alloc(n * R * h->n_walk));
  h->out_cap = (int)n;
  if (h->graph_exec) {
    (void)hipGraphExecDestroy(h->graph_exec);
    h->graph_exec = nullptr;
  }
  if (h->graph) {
    (void)hipGraphDestroy(h->graph);
    h->graph = nullptr;
  }
  return JR_OK;
}

// population energies on the device (needs run params)
int recompute_energies(jr_handle* h) {
  if (h->learned) return jrl::learned_recompute_energies(h);
  const auto& c = h->cfg;
  const int R = c.n_replicas, K = c.n_walkers;
  const int S = c.pot.n_species;
  // per-walker pressure/mu expanded on host
  std::vector<double> pr((size_t)R * K), mu((size_t)R * K * S);
  for (int r = 0; r < R; ++r)
    for (int k = 0; k < K; ++k) {
      pr[(size_t)r * K + k] = h->h_pressure[r];
      for (int s = 0; s < S; ++s) mu[((size_t)r * K + k) * S + s] = h->h_mu[(size_t)r * S + s];
    }
  DBuf<double> dpr, dmu;
  HIPCHK(h, dpr.alloc(pr.size()));
  HIPCHK(h, dmu.alloc(mu.size()));
  HIPCHK(h, hipMemcpyAsync(dpr.p, pr.data(), pr.size() * 8, hipMemcpyHostToDevice, h->stream));
  HIPCHK(h, hipMemcpyAsync(dmu.p, mu.data(), mu.size() * 8, hipMemcpyHostToDevice, h->stream));
  EvalArgs a;
  std::memset(&a, 0, sizeof(a));
  a.pos = h->pos.p;
  a.types = h->types.p;
  a.n_atoms = h->n_atoms.p;
  a.cell = h->cell.p;
  a.pressure = dpr.p;
  a.mu = dmu.p;
  a.U = nullptr;
  a.H = h->energy.p;
  a.F = nullptr;
  a.B = R * K;
  a.Nmax = c.n_atoms_max;
  a.S = S;
  a.rc2 = c.pot.rc2;
  a.eps4 = h->eps4.p;
  a.eps24 = h->eps24.p;
  a.sig2 = h->sig2.p;
  set_images(a.img, c.pot);
  a.chunks = pot_chunks(&c.pot);
  a.team = team_waves(h->A, a.chunks, S, c.device);
  launch_eval(h->A, h->cfgbits & 39, a, h->stream);
  HIPCHK(h, hipGetLastError());
  HIPCHK(h, hipStreamSynchronize(h->stream));
  return JR_OK;
}

PopPtrs pop_ptrs(jr_handle* h) {
  PopPtrs p{h->pos.p, h->types.p, h->n_atoms.p, h->cell.p, h->energy.p, h->dir.p};
  p.dirx = h->dirx.p;
  p.n_xslots = h->n_xslots;
  p.xstride = (size_t)h->cfg.n_replicas * h->cfg.n_walkers * h->cfg.n_atoms_max * 3;
  return p;
}

WalkArgs walk_args(jr_handle* h) {
  const auto& c = h->cfg;
  WalkArgs a;
  std::memset(&a, 0, sizeof(a));
  a.pop = pop_ptrs(h);
  a.pressure = h->pressure.p;
  a.mu = h->mu.p;
  a.step_sizes = h->step_sizes.p;
  a.emax = h->emax.p;
  a.walk = h->walk.p;
  a.n_walk = h->n_walk;
  a.R = c.n_replicas;
  a.K = c.n_walkers;
  a.Nmax = c.n_atoms_max;
  a.S = c.pot.n_species;
  a.n_steps = c.n_mcmc_steps;
  a.n_moves = c.n_moves;
  a.schedule = h->sched.p;
  a.normals = h->normals.p;
  a.uniforms = h->uniforms.p;
  a.counters = h->counters.p;
  a.queue = h->queue.p;
  a.total = c.n_replicas * h->n_walk;
  a.order = h->order.p;
  a.rc2 = c.pot.rc2;
  a.eps4 = h->eps4.p;
  a.eps24 = h->eps24.p;
  a.sig2 = h->sig2.p;
  set_images(a.img, c.pot);
  a.moves = h->moves.p;
  a.lds_atoms = c.n_atoms_max;
  // device draws: the walk draws its normals and uniforms in place
  a.st = h->st.p;
  a.seed = c.seed;
  a.dirx = h->dirx.p;
  a.n_pop = c.n_replicas * c.n_walkers;
  a.chunks = pot_chunks(&c.pot);
  if (!h->draws_buffered) {
    a.normals = nullptr;
    a.uniforms = nullptr;
  }
  if (h->cfgbits & 16) {
    const int SS = c.pot.n_species * c.pot.n_species;
    a.dmin_pos = h->dmin.p;
    a.dmin_types = h->dmin_types ? h->dmin.p + SS : nullptr;
  }
  return a;
}

SelArgs sel_args(jr_handle* h) {
  const auto& c = h->cfg;
  SelArgs a;
  std::memset(&a, 0, sizeof(a));
  a.pop = pop_ptrs(h);
  a.R = c.n_replicas;
  a.K = c.n_walkers;
  a.Nmax = c.n_atoms_max;
  a.n_extra = c.n_extra;
  a.n_walk = h->n_walk;
  a.n_moves = c.n_moves;
  a.Kpad = (c.n_extra > 0 || c.n_cull > 1) ? (int)pow2_at_least(c.n_walkers) : 0;
  a.n_cull = c.n_cull;
  a.o_dead_energy = h->o_dead_energy.p;
  a.sel_u = h->sel_u.p;
  a.clone_u = h->clone_u.p;
  a.extra_u = h->extra_u.p;
  a.emax = h->emax.p;
  a.walk = h->walk.p;
  a.counters = h->counters.p;
  a.queue = h->queue.p;
  a.st = h->st.p;
  a.o_emax = h->o_emax.p;
  a.o_worst = h->o_worst.p;
  a.o_clone = h->o_clone.p;
  a.o_dead_volume = h->o_dead_volume.p;
  a.o_dead_pos = h->o_dead_pos.p;
  a.o_dead_types = h->o_dead_types.p;
  a.o_dead_cell = h->o_dead_cell.p;
  return a;
}

// Selection and cloning of one iteration, one block per replica: the
// single-cull kernel, or with n_cull > 1 the lexicographic-order kernel
// (LDS: Kpad (key, index) pairs, plus the culled indices).
void launch_select(jr_handle* h, hipStream_t s) {
  const SelArgs sa = sel_args(h);
  const size_t pairs = sa.Kpad ? (size_t)sa.Kpad * (sizeof(double) + sizeof(int)) : 0;
  if (sa.n_cull > 1)
    hipLaunchKernelGGL(select_cull_kernel, dim3(sa.R), dim3(1024),
                       pairs + sizeof(int) * sa.n_cull, s, sa);
  else
    hipLaunchKernelGGL(select_kernel, dim3(sa.R), dim3(1024), pairs, s, sa);
}

ReArgs re_args(jr_handle* h, IterState* st) {
  const auto& c = h->cfg;
  ReArgs a;
  std::memset(&a, 0, sizeof(a));
  a.pop = pop_ptrs(h);
  a.R = c.n_replicas;
  a.K = c.n_walkers;
  a.Nmax = c.n_atoms_max;
  a.S = c.pot.n_species;
  a.n_cycles = c.n_swap_cycles;
  a.max_pairs = h->max_pairs;
  a.pressure = h->pressure.p;
  a.mu = h->mu.p;
  a.emax = h->emax.p;
  a.pair_valid = h->pair_valid.p;
  a.re_u = h->re_u.p;
  a.st = st;
  a.o_acc = h->o_re_acc.p;
  a.o_att = h->o_re_att.p;
  return a;
}

// RE block: one wave per pair of a phase, at most 16 waves
unsigned re_threads(const jr_handle* h) {
  return 64u * (unsigned)std::min(std::max(h->max_pairs, 1), 16);
}

// extra_u feeds the extras' noise (n_cull = 1) or the survivors' permutation
// keys (n_cull > 1)
bool uses_extra_u(const jr_config& c) { return c.n_extra > 0 || c.n_cull > 1; }

size_t n_normals(const jr_handle* h) {
  const auto& c = h->cfg;
  return (size_t)c.n_replicas * h->n_walk * c.n_mcmc_steps * 3 * c.n_atoms_max;
}
size_t n_uniforms(const jr_handle* h) {
  const auto& c = h->cfg;
  return (size_t)c.n_replicas * h->n_walk * c.n_mcmc_steps * 4;
}
size_t n_re_u(const jr_handle* h) {
  return (size_t)std::max(h->cfg.n_swap_cycles, 1) * 2 * std::max(h->max_pairs, 1) * 2;
}

unsigned grid_for(size_t n) {
  size_t pairs = (n + 1) / 2;
  size_t b = (pairs + 255) / 256;
  if (b < 1) b = 1;
  if (b > 8192) b = 8192;
  return (unsigned)b;
}

// Device draws for the current iteration (Philox keyed by seed, counter
// (index, rng_iter, stream)): selection, clone, extras, schedule and RE
// draws into their buffers; `it` is the iteration's draw counter, kept for
// jr_get_last_tape.
void enqueue_draws(jr_handle* h, long long it) {
  h->last_draw_iter = it;
  const auto& c = h->cfg;
  const int R = c.n_replicas, K = c.n_walkers;
  hipStream_t s = h->stream;
  IterState* st = h->st.p;
  unsigned long long seed = c.seed;
  DrawArgs da;
  da.sel_u = h->sel_u.p;
  da.clone_u = h->clone_u.p;
  da.extra_u = uses_extra_u(c) ? h->extra_u.p : nullptr;
  da.re_u = c.n_replicas > 1 ? h->re_u.p : nullptr;
  da.sched = h->sched.p;
  da.cum = h->cum.p;
  da.st = st;
  da.seed = seed;
  da.n_sel = (long long)R * K;
  da.n_clone = R;
  da.n_extra = (long long)R * K;
  da.n_re = c.n_replicas > 1 ? (long long)n_re_u(h) : 0;
  da.n_steps = c.n_mcmc_steps;
  da.n_moves = c.n_moves;
  {
    const size_t pairs = (size_t)(da.n_sel + 1) / 2 + (size_t)(R + 1) / 2 +
                         (da.extra_u ? (size_t)(da.n_extra + 1) / 2 : 0) +
                         (size_t)(da.n_re + 1) / 2 + (size_t)c.n_mcmc_steps;
    hipLaunchKernelGGL(iter_draws_kernel, dim3(grid_for(pairs)), dim3(256), 0, s, da);
  }
  if (h->learned) {
    // the learned walk's kernels read the draws from the tape buffers
    hipLaunchKernelGGL(fill_normal_kernel, dim3(grid_for(n_normals(h))), dim3(256), 0, s,
                       h->normals.p, (long long)n_normals(h), seed, st, 0LL, 5u);
    hipLaunchKernelGGL(fill_uniform_kernel, dim3(grid_for(n_uniforms(h))), dim3(256), 0, s,
                       h->uniforms.p, (long long)n_uniforms(h), seed, st, 0LL, 6u);
    h->draws_buffered = true;
  } else {
    // the pair walk draws its normals (stream 5) and uniforms (stream 6) in place
    h->draws_buffered = false;
  }
}

// The walk of one iteration: a single persistent launch, or with replicas of
// several atoms-per-lane classes one launch per class (largest A first), each
// with the kernel built for its A, its own queue counter and grid.  A chain's
// results do not depend on A, so the split changes no number.  The classes
// after the first run on stream2, forked from `s` and joined back into it,
// so they run concurrently with the first: a class whose kernel leaves
// registers free on a SIMD (the A = 2 species-table kernel holds 344 of 512,
// exactly enough for one 168-register A = 1 wave beside it) shares it with
// waves of the next class, and the classes share no data.
hipError_t enqueue_walk(jr_handle* h, const WalkArgs& base, hipStream_t s) {
  if (h->n_classes == 0) {
    launch_walk(h->A, h->cfgbits, h->occ, base, h->blocks, h->wpb, h->lds_walk, s);
    return hipSuccess;
  }
  const int R = h->cfg.n_replicas;
  hipError_t e;
  if ((e = hipEventRecord(h->ev_fork, s)) != hipSuccess) return e;
  if ((e = hipStreamWaitEvent(h->stream2, h->ev_fork, 0)) != hipSuccess) return e;
  for (int j = 0; j < h->n_classes; ++j) {
    const auto& k = h->cls[j];
    WalkArgs wa = base;
    wa.order = h->cls_order.p + (size_t)j * R;
    wa.total = k.n_rep * base.n_walk;
    wa.queue = h->queue.p + 1 + j;
    wa.lds_atoms = 64 * k.A;
    launch_walk(k.A, h->cfgbits, k.occ, wa, k.blocks, k.wpb, k.lds, j > 0 ? h->stream2 : s);
  }
  if ((e = hipEventRecord(h->ev_join, h->stream2)) != hipSuccess) return e;
  return hipStreamWaitEvent(s, h->ev_join, 0);
}

// select -> walk -> RE -> advance, all on the handle's stream.  The learned
// model's walk is sequenced from the host (one sync for the move schedule,
// one for its statistics); the pair-potential walk is one launch.
int enqueue_iteration_core(jr_handle* h, bool timed, int ev_idx) {
  const auto& c = h->cfg;
  launch_select(h, h->stream);
  if (timed) (void)hipEventRecord(h->ev[2 * ev_idx], h->stream);
  if (h->learned) {
    int rc = jrl::learned_walk(h);
    if (rc) return rc;
  } else {
    HIPCHK(h, enqueue_walk(h, walk_args(h), h->stream));
  }
  if (timed) (void)hipEventRecord(h->ev[2 * ev_idx + 1], h->stream);
  if (c.n_replicas > 1) {
    ReArgs ra = re_args(h, h->st.p);
    hipLaunchKernelGGL(re_kernel, dim3(1), dim3(re_threads(h)), 0, h->stream, ra);
  }
  hipLaunchKernelGGL(advance_kernel, dim3(1), dim3(64), 0, h->stream, h->st.p);
  return JR_OK;
}

}  // namespace

// counters/walk snapshots go into the block's output slot after each
// iteration (after every graph replay on the graph path), by a copy kernel
// on the walk stream.
static int snapshot_iter(jr_handle* h, int it) {
  const auto& c = h->cfg;
  const int R = c.n_replicas;
  const size_t nc = (size_t)R * c.n_moves * JR_NCOUNT;
  const size_t nw = (size_t)R * h->n_walk;
  const unsigned blocks = (unsigned)std::min<size_t>(256, (nc + nw + 255) / 256);
  hipLaunchKernelGGL(snapshot_kernel, dim3(blocks), dim3(256), 0, h->stream, h->counters.p,
                     (int)nc, h->walk.p, (int)nw, h->o_counters.p + (size_t)it * nc,
                     h->o_walk.p + (size_t)it * nw);
  HIPCHK(h, hipGetLastError());
  return JR_OK;
}

namespace {

// Persistent walk grid: waves per block, occupancy variant, and as many
// blocks as are resident (never more than the chains of one iteration).  An
// evaluation team (cfg bit 5) is one block of eval_chunks waves per chain.
int set_geometry(jr_handle* h, int wpb, int occ) {
  const auto& c = h->cfg;
  const bool team = (h->cfgbits & 32) != 0;
  const int S = c.pot.n_species;
  const int C = team ? pot_chunks(&c.pot) : 1;
  if (team) {
    wpb = team_waves(h->A, C, S, c.device);
    occ = 0;
  }
  // occupancy request 3: the build held to three waves per SIMD, which
  // exists for one atom per lane without cfg bits 2-5
  auto occ_for = [&](int A) { return (A == 1 && !(h->cfgbits & 60)) ? occ : 0; };
  h->occ_req = occ;
  h->wpb = wpb;
  h->ipb = team ? 1 : wpb;
  h->occ = occ_for(h->A);
  for (int j = 0; j < h->n_classes; ++j) {
    auto& k = h->cls[j];
    k.occ = occ_for(k.A);
    k.wpb = team ? team_waves(k.A, C, S, c.device) : wpb;
    k.lds = block_lds_bytes(k.A, k.wpb, S, C);
    int pc = 0;
    HIPCHK(h, hipOccupancyMaxActiveBlocksPerMultiprocessor(
                  &pc, walk_fn(k.A, h->cfgbits, k.occ), 64 * k.wpb, k.lds));
    if (pc < 1) pc = 1;
    const int need = (k.n_rep * h->n_walk + h->ipb - 1) / h->ipb;
    k.blocks = std::max(1, std::min(need, pc * h->n_cu));
    if (h->block_cap > 0) k.blocks = std::min(k.blocks, h->block_cap);
  }
  h->lds_walk = block_lds_bytes(h->A, wpb, S, C);
  int per_cu = 0;
  HIPCHK(h, hipOccupancyMaxActiveBlocksPerMultiprocessor(
                &per_cu, walk_fn(h->A, h->cfgbits, h->occ), 64 * wpb,
                h->lds_walk));
  if (per_cu < 1) per_cu = 1;
  const int total = c.n_replicas * h->n_walk;
  const int need = (total + h->ipb - 1) / h->ipb;
  h->max_blocks = per_cu * h->n_cu;
  if (h->block_cap > 0) h->max_blocks = std::min(h->max_blocks, h->block_cap);
  h->blocks = std::max(1, std::min(need, h->max_blocks));
  if (h->graph_exec) {
    (void)hipGraphExecDestroy(h->graph_exec);
    h->graph_exec = nullptr;
  }
  if (h->graph) {
    (void)hipGraphDestroy(h->graph);
    h->graph = nullptr;
  }
  return JR_OK;
}

}  // namespace

namespace jrl {

hipError_t lj_eval(jr_handle* h, const double* pos, const int* types, const int* n_atoms,
                   const double* cell, const int* idx, const int* count, int B, double* U,
                   double* F, hipStream_t s) {
  EvalArgs a;
  std::memset(&a, 0, sizeof(a));
  a.pos = pos;
  a.types = types;
  a.n_atoms = n_atoms;
  a.cell = cell;
  a.cfg_index = idx;
  a.count = count;
  a.U = U;
  a.F = F;
  a.B = B;
  a.Nmax = h->cfg.n_atoms_max;
  a.S = h->cfg.pot.n_species;
  a.rc2 = h->cfg.pot.rc2;
  a.eps4 = h->eps4.p;
  a.eps24 = h->eps24.p;
  a.sig2 = h->sig2.p;
  set_images(a.img, h->cfg.pot);
  a.chunks = pot_chunks(&h->cfg.pot);
  a.team = team_waves(h->A, a.chunks, a.S, h->cfg.device);
  launch_eval(h->A, h->cfgbits & 39, a, s);
  return hipGetLastError();
}

}  // namespace jrl

extern "C" {

int jr_create(const jr_config* cfg, jr_handle** out) {
  if (!cfg || !out) return set_global_err(JR_E_ARG, "cfg/out must not be NULL");
  *out = nullptr;
  jr_config c = *cfg;
  if (c.mpnn) {
    // the learned model replaces the pair potential
    if (c.mpnn->n_species < 1 || c.mpnn->n_species > JR_MAX_SPECIES)
      return set_global_err(JR_E_ARG, "mpnn n_species out of range");
    c.pot = jr_pot{c.mpnn->n_species, JR_PREC_FP64, INFINITY, nullptr, nullptr, nullptr};
  } else {
    int rc = validate_pot(&c.pot);
    if (rc) return rc;
    if (c.learned && pot_images(&c.pot))
      return set_global_err(JR_E_UNSUPPORTED,
                            "the surrogate-screened walk takes minimum-image pairs only");
  }
  if (c.n_replicas < 1) return set_global_err(JR_E_ARG, "n_replicas must be >= 1");
  if (c.n_walkers < 2 || c.n_walkers > 4096)
    return set_global_err(JR_E_ARG, "n_walkers must be in [2, 4096]");
  if (c.n_atoms_max < 1 || c.n_atoms_max > 256)
    return set_global_err(JR_E_ARG, "n_atoms_max must be in [1, 256]");
  if (c.n_moves < 1 || c.n_moves > JR_MAX_MOVES || !c.moves || !c.move_weights)
    return set_global_err(JR_E_ARG, "need 1..%d moves with weights", JR_MAX_MOVES);
  if (c.n_mcmc_steps < 1) return set_global_err(JR_E_ARG, "n_mcmc_steps must be >= 1");
  if (c.n_cull < 1) return set_global_err(JR_E_ARG, "n_cull must be >= 1");
  if (c.n_cull == 1 && (c.n_extra < 0 || c.n_extra > c.n_walkers - 2))
    return set_global_err(JR_E_ARG, "n_extra must be in [0, K-2]");
  if (c.n_cull > 1 && (c.n_extra < 0 || 2 * c.n_cull + c.n_extra > c.n_walkers))
    return set_global_err(JR_E_ARG,
                          "n_cull=%d, n_extra=%d: need 2*n_cull + n_extra <= K=%d (each culled "
                          "slot is refilled from a distinct survivor, extras from the rest)",
                          c.n_cull, c.n_extra, c.n_walkers);
  if (c.n_cull > 1 && (c.mpnn || c.learned))
    return set_global_err(JR_E_UNSUPPORTED, "the learned walk is built for n_cull = 1");
  if (c.min_distance_types && !c.min_distance)
    return set_global_err(JR_E_ARG, "min_distance_types needs min_distance");
  if (pot_chunks(&c.pot) > 1 && (c.learned || c.min_distance))
    return set_global_err(JR_E_UNSUPPORTED,
                          "eval_chunks > 1 is built for the pair-potential walk without the "
                          "minimum-distance constraint");
  if (c.min_distance) {
    if (c.mpnn || c.learned || c.pot.precision != JR_PREC_FP64 || pot_images(&c.pot))
      return set_global_err(JR_E_UNSUPPORTED,
                            "the minimum-distance constraint is built for the FP64 pair walk "
                            "without supercell images");
    const int S = c.pot.n_species;
    for (int i = 0; i < S * S; ++i)
      if (!(c.min_distance[i] >= 0.0) ||
          (c.min_distance_types && !(c.min_distance_types[i] >= 0.0)))
        return set_global_err(JR_E_ARG, "minimum distances must be >= 0");
  }
  if (c.n_swap_cycles < 0) return set_global_err(JR_E_ARG, "n_swap_cycles must be >= 0");
  double wsum = 0.0;
  for (int m = 0; m < c.n_moves; ++m) {
    const jr_move& mv = c.moves[m];
    if (mv.type < 0 || mv.type > JR_MOVE_SWAP)
      return set_global_err(JR_E_ARG, "move %d: unknown type %d", m, mv.type);
    if (mv.type == JR_MOVE_MORPH && (mv.n_species < 2 || mv.n_species > c.pot.n_species))
      return set_global_err(JR_E_ARG, "alchemical_morph needs 2 <= n_species <= S");
    if (mv.type == JR_MOVE_GMC && mv.n_reflect < 1)
      return set_global_err(JR_E_ARG, "gmc needs n_reflect >= 1");
    if (mv.type == JR_MOVE_GMC) {
      if (mv.species_mask & ~((1 << c.pot.n_species) - 1))
        return set_global_err(JR_E_ARG, "gmc species scope names codes >= S = %d",
                              c.pot.n_species);
      if (mv.dir_slot < 0 || mv.dir_slot >= JR_MAX_MOVES)
        return set_global_err(JR_E_ARG, "gmc dir_slot must be in [0, %d)", JR_MAX_MOVES);
      if ((c.mpnn || c.learned) && (mv.species_mask || mv.refresh_direction || mv.dir_slot))
        return set_global_err(JR_E_UNSUPPORTED,
                              "species-scoped gmc and refresh_direction are built for the "
                              "pair-potential walk");
    } else if (mv.dir_slot != 0 || mv.refresh_direction) {
      return set_global_err(JR_E_ARG, "move %d: dir_slot and refresh_direction are gmc-only", m);
    }
    if ((mv.type == JR_MOVE_SHEAR || mv.type == JR_MOVE_STRETCH || mv.type == JR_MOVE_SWAP) &&
        !c.mpnn && c.pot.precision != JR_PREC_FP64)
      return set_global_err(JR_E_UNSUPPORTED,
                            "shear, stretch and species_swap are built for FP64 pairs only");
    if (mv.type == JR_MOVE_SWAP) {
      const int S = c.pot.n_species;
      const int full = (1 << S) - 1;
      const int scope = mv.species_mask ? mv.species_mask : full;
      if (S < 2 || (scope & ~full) || __builtin_popcount(scope) < 2)
        return set_global_err(JR_E_ARG,
                              "species_swap needs >= 2 species in scope, codes below S = %d", S);
    }
    if (c.move_weights[m] < 0.0) return set_global_err(JR_E_ARG, "negative move weight");
    wsum += c.move_weights[m];
  }
  if (!(wsum > 0.0)) return set_global_err(JR_E_ARG, "move weights sum to zero");
  HIPCHK_G(hipSetDevice(c.device));

  jr_handle* h = new jr_handle();
  h->cfg = c;
  h->cfg.moves = nullptr;
  h->cfg.move_weights = nullptr;
  h->cfg.pot.eps4 = h->cfg.pot.eps24 = h->cfg.pot.sig2 = nullptr;
  h->n_walk = c.n_cull + c.n_extra;
  h->max_pairs = c.n_replicas / 2;
  h->A = atoms_per_lane(c.n_atoms_max);
  h->cfgbits = pot_cfg(&c.pot);
  for (int m = 0; m < c.n_moves; ++m)
    if (c.moves[m].type == JR_MOVE_SHEAR || c.moves[m].type == JR_MOVE_STRETCH ||
        c.moves[m].type == JR_MOVE_SWAP)
      h->cfgbits |= 8;
  if (c.min_distance) h->cfgbits |= 16;
  const int R = c.n_replicas, K = c.n_walkers, Nm = c.n_atoms_max, S = c.pot.n_species;
  auto fail = [&](int code) {
    g_err = h->err;
    release_all(h);
    jr_destroy(h);
    return code;
  };
#define TRY(call)                                                        \
  do {                                                                   \
    hipError_t e_ = (call);                                              \
    if (e_ != hipSuccess) {                                              \
      h->err = std::string(#call) + ": " + hipGetErrorString(e_);        \
      return fail(JR_E_HIP);                                             \
    }                                                                    \
  } while (0)
  TRY(hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking));
  TRY(hipStreamCreateWithFlags(&h->stream2, hipStreamNonBlocking));
  TRY(hipEventCreateWithFlags(&h->ev_fork, hipEventDisableTiming));
  TRY(hipEventCreateWithFlags(&h->ev_join, hipEventDisableTiming));
  const size_t W = (size_t)R * K;
  TRY(h->pos.alloc(W * Nm * 3));
  TRY(h->dir.alloc(W * Nm * 3));
  for (int m = 0; m < c.n_moves; ++m)
    if (c.moves[m].type == JR_MOVE_GMC) h->n_xslots = std::max(h->n_xslots, c.moves[m].dir_slot);
  if (h->n_xslots) {
    TRY(h->dirx.alloc((size_t)h->n_xslots * W * Nm * 3));
    TRY(hipMemset(h->dirx.p, 0, sizeof(double) * h->n_xslots * W * Nm * 3));
  }
  TRY(h->cell.alloc(W * 9));
  TRY(h->energy.alloc(W));
  TRY(h->types.alloc(W * Nm));
  TRY(h->n_atoms.alloc(W));
  TRY(h->pressure.alloc(R));
  TRY(h->mu.alloc((size_t)R * S));
  TRY(h->step_sizes.alloc((size_t)R * c.n_moves));
  TRY(h->emax.alloc(R));
  TRY(h->eps4.alloc((size_t)S * S));
  if (c.min_distance) {
    TRY(h->dmin.alloc((size_t)2 * S * S));
    TRY(hipMemcpy(h->dmin.p, c.min_distance, sizeof(double) * S * S, hipMemcpyHostToDevice));
    h->dmin_types = c.min_distance_types != nullptr;
    if (h->dmin_types)
      TRY(hipMemcpy(h->dmin.p + (size_t)S * S, c.min_distance_types, sizeof(double) * S * S,
                    hipMemcpyHostToDevice));
  }
  TRY(h->eps24.alloc((size_t)S * S));
  TRY(h->sig2.alloc((size_t)S * S));
  TRY(h->moves.alloc(c.n_moves));
  TRY(h->cum.alloc(c.n_moves));
  TRY(h->walk.alloc((size_t)R * h->n_walk));
  TRY(h->counters.alloc((size_t)R * c.n_moves * JR_NCOUNT));
  TRY(h->queue.alloc(JR_NQUEUE));
  TRY(h->cls_order.alloc((size_t)3 * R));
  TRY(h->pair_valid.alloc(std::max(R - 1, 1)));
  TRY(h->order.alloc(R));
  TRY(h->sched.alloc(c.n_mcmc_steps));
  TRY(h->sel_u.alloc(W));
  TRY(h->clone_u.alloc(R));
  TRY(h->extra_u.alloc(W));
  TRY(h->re_u.alloc(n_re_u(h)));
  TRY(h->st.alloc(1));
  if (c.mpnn || c.learned) {
    TRY(h->normals.alloc(n_normals(h)));
    TRY(h->uniforms.alloc(n_uniforms(h)));
  }

  if (!c.mpnn) {
    TRY(hipMemcpy(h->eps4.p, c.pot.eps4, sizeof(double) * S * S, hipMemcpyHostToDevice));
    TRY(hipMemcpy(h->eps24.p, c.pot.eps24, sizeof(double) * S * S, hipMemcpyHostToDevice));
    TRY(hipMemcpy(h->sig2.p, c.pot.sig2, sizeof(double) * S * S, hipMemcpyHostToDevice));
  }
  std::vector<DevMove> dm(c.n_moves);
  std::vector<double> cum(c.n_moves);
  double acc = 0.0;
  for (int m = 0; m < c.n_moves; ++m) {
    const jr_move& mv = c.moves[m];
    dm[m] = DevMove{mv.type, mv.n_reflect, mv.n_leapfrog, mv.flat_v_prior, mv.n_species,
                    (mv.type == JR_MOVE_SWAP || mv.type == JR_MOVE_GMC) ? mv.species_mask : 0,
                    mv.perturb_angle, mv.max_vpa, mv.min_vpa, mv.min_aspect,
                    mv.type == JR_MOVE_GMC ? mv.refresh_direction : 0,
                    mv.type == JR_MOVE_GMC ? mv.dir_slot : 0};
    acc += c.move_weights[m] / wsum;
    cum[m] = acc;
  }
  cum[c.n_moves - 1] = 2.0;  // last bucket absorbs rounding
  TRY(hipMemcpy(h->moves.p, dm.data(), sizeof(DevMove) * c.n_moves, hipMemcpyHostToDevice));
  TRY(hipMemcpy(h->cum.p, cum.data(), sizeof(double) * c.n_moves, hipMemcpyHostToDevice));
  std::vector<int> pv(std::max(R - 1, 1), 1), ord(R);
  for (int r = 0; r < R; ++r) ord[r] = r;
  TRY(hipMemcpy(h->pair_valid.p, pv.data(), sizeof(int) * pv.size(), hipMemcpyHostToDevice));
  TRY(hipMemcpy(h->order.p, ord.data(), sizeof(int) * R, hipMemcpyHostToDevice));
  h->h_pressure.assign(R, 0.0);
  h->h_mu.assign((size_t)R * S, 0.0);
  h->h_step.assign((size_t)R * c.n_moves, 0.1);
  TRY(hipMemcpy(h->pressure.p, h->h_pressure.data(), sizeof(double) * R, hipMemcpyHostToDevice));
  TRY(hipMemcpy(h->mu.p, h->h_mu.data(), sizeof(double) * R * S, hipMemcpyHostToDevice));
  TRY(hipMemcpy(h->step_sizes.p, h->h_step.data(), sizeof(double) * R * c.n_moves,
                hipMemcpyHostToDevice));
  TRY(hipMemset(h->energy.p, 0, sizeof(double) * W));
  TRY(hipMemset(h->emax.p, 0, sizeof(double) * R));
  std::vector<double> inf(R, INFINITY);
  TRY(hipMemcpy(h->emax.p, inf.data(), sizeof(double) * R, hipMemcpyHostToDevice));
  IterState st0{0, 0, 0, c.re_interval};
  TRY(hipMemcpy(h->st.p, &st0, sizeof(st0), hipMemcpyHostToDevice));

  // launch geometry: persistent grid sized by occupancy
  int dev = c.device;
  hipDeviceProp_t prop;
  TRY(hipGetDeviceProperties(&prop, dev));
  h->n_cu = prop.multiProcessorCount;
  {
    int g = set_geometry(h, c.waves_per_block > 0 ? std::min(c.waves_per_block, 4) : 4, 0);
    if (g) return fail(g);
  }
  int ret = ensure_out(h, 1);
  if (ret) return fail(ret);
  if (c.mpnn || c.learned) {
    h->cfg.mpnn = nullptr;
    h->cfg.learned = nullptr;
    ret = jrl::learned_create(h, &c);
    if (ret) return fail(ret);
  }
  *out = h;
  return JR_OK;
#undef TRY
}

int jr_set_population(jr_handle* h, const double* pos, const int32_t* types,
                      const int32_t* n_atoms, const double* cell,
                      const double* dir) {
  if (!h) return set_global_err(JR_E_ARG, "NULL handle");
  if (!pos || !types || !n_atoms || !cell)
    return h->fail(JR_E_ARG, "pos/types/n_atoms/cell must not be NULL");
  const auto& c = h->cfg;
  const size_t W = (size_t)c.n_replicas * c.n_walkers;
  const int Nm = c.n_atoms_max, S = c.pot.n_species;
  for (size_t w = 0; w < W; ++w) {
    if (n_atoms[w] < 1 || n_atoms[w] > Nm)
      return h->fail(JR_E_ARG, "n_atoms[%zu]=%d outside [1, %d]", w, n_atoms[w], Nm);
    for (int i = 0; i < n_atoms[w]; ++i) {
      int t = types[w * Nm + i];
      if (t < 0 || t >= S) return h->fail(JR_E_ARG, "species %d out of range", t);
    }
  }
  if (h->learned) {
    int rc = jrl::learned_check_population(h, types, n_atoms, cell);
    if (rc) return rc;
  }
  HIPCHK(h, hipSetDevice(c.device));
  HIPCHK(h, hipMemcpyAsync(h->pos.p, pos, sizeof(double) * W * Nm * 3, hipMemcpyHostToDevice, h->stream));
  HIPCHK(h, hipMemcpyAsync(h->types.p, types, sizeof(int) * W * Nm, hipMemcpyHostToDevice, h->stream));
  HIPCHK(h, hipMemcpyAsync(h->n_atoms.p, n_atoms, sizeof(int) * W, hipMemcpyHostToDevice, h->stream));
  HIPCHK(h, hipMemcpyAsync(h->cell.p, cell, sizeof(double) * W * 9, hipMemcpyHostToDevice, h->stream));
  if (dir)
    HIPCHK(h, hipMemcpyAsync(h->dir.p, dir, sizeof(double) * W * Nm * 3, hipMemcpyHostToDevice, h->stream));
  else
    HIPCHK(h, hipMemsetAsync(h->dir.p, 0, sizeof(double) * W * Nm * 3, h->stream));
  if (h->n_xslots)
    HIPCHK(h, hipMemsetAsync(h->dirx.p, 0, sizeof(double) * h->n_xslots * W * Nm * 3,
                             h->stream));
  // default RE validity: equal atom counts per replica (first walker)
  const int R = c.n_replicas, K = c.n_walkers;
  std::vector<int> pv(std::max(R - 1, 1), 1);
  std::vector<std::pair<int, int>> byn(R);
  for (int r = 0; r < R; ++r) byn[r] = {-n_atoms[(size_t)r * K], r};
  for (int r = 0; r + 1 < R; ++r) pv[r] = n_atoms[(size_t)r * K] == n_atoms[(size_t)(r + 1) * K];
  std::stable_sort(byn.begin(), byn.end());
  std::vector<int> ord(R);
  for (int r = 0; r < R; ++r) ord[r] = byn[r].second;
  HIPCHK(h, hipMemcpyAsync(h->pair_valid.p, pv.data(), sizeof(int) * pv.size(), hipMemcpyHostToDevice, h->stream));
  HIPCHK(h, hipMemcpyAsync(h->order.p, ord.data(), sizeof(int) * R, hipMemcpyHostToDevice, h->stream));
  // size classes by atoms per lane, each in the largest-N-first order
  h->n_classes = 0;
  if (!h->learned) {
    std::vector<int> corder((size_t)3 * R, 0);
    int used = 0;
    for (int A : {4, 2, 1}) {
      int n = 0;
      for (int r : ord)
        if (atoms_per_lane(n_atoms[(size_t)r * K]) == A) corder[(size_t)used * R + n++] = r;
      if (n > 0) {
        h->cls[used].A = A;
        h->cls[used].n_rep = n;
        ++used;
      }
    }
    if (used > 1) {
      h->n_classes = used;
      HIPCHK(h, hipMemcpyAsync(h->cls_order.p, corder.data(), sizeof(int) * 3 * R,
                               hipMemcpyHostToDevice, h->stream));
    }
  }
  HIPCHK(h, hipStreamSynchronize(h->stream));
  h->have_pop = true;
  // launch shapes follow the size classes; a captured graph is dropped
  int g = set_geometry(h, h->wpb, h->occ_req);
  if (g) return g;
  return recompute_energies(h);
}

int jr_get_population(jr_handle* h, double* pos, int32_t* types,
                      int32_t* n_atoms, double* cell, double* energy,
                      double* dir) {
  if (!h) return set_global_err(JR_E_ARG, "NULL handle");
  const auto& c = h->cfg;
  const size_t W = (size_t)c.n_replicas * c.n_walkers;
  const int Nm = c.n_atoms_max;
  HIPCHK(h, hipSetDevice(c.device));
  if (pos) HIPCHK(h, hipMemcpyAsync(pos, h->pos.p, sizeof(double) * W * Nm * 3, hipMemcpyDeviceToHost, h->stream));
  if (types) HIPCHK(h, hipMemcpyAsync(types, h->types.p, sizeof(int) * W * Nm, hipMemcpyDeviceToHost, h->stream));
  if (n_atoms) HIPCHK(h, hipMemcpyAsync(n_atoms, h->n_atoms.p, sizeof(int) * W, hipMemcpyDeviceToHost, h->stream));
  if (cell) HIPCHK(h, hipMemcpyAsync(cell, h->cell.p, sizeof(double) * W * 9, hipMemcpyDeviceToHost, h->stream));
  if (energy) HIPCHK(h, hipMemcpyAsync(energy, h->energy.p, sizeof(double) * W, hipMemcpyDeviceToHost, h->stream));
  if (dir) HIPCHK(h, hipMemcpyAsync(dir, h->dir.p, sizeof(double) * W * Nm * 3, hipMemcpyDeviceToHost, h->stream));
  HIPCHK(h, hipStreamSynchronize(h->stream));
  return JR_OK;
}

int jr_get_direction_slot(jr_handle* h, int32_t slot, double* dir) {
  if (!h || !dir) return set_global_err(JR_E_ARG, "NULL argument");
  if (slot < 1 || slot > h->n_xslots)
    return h->fail(JR_E_ARG, "direction slot %d: the moves use slots 1..%d", slot, h->n_xslots);
  const auto& c = h->cfg;
  const size_t n = (size_t)c.n_replicas * c.n_walkers * c.n_atoms_max * 3;
  HIPCHK(h, hipSetDevice(c.device));
  HIPCHK(h, hipMemcpyAsync(dir, h->dirx.p + (size_t)(slot - 1) * n, sizeof(double) * n,
                           hipMemcpyDeviceToHost, h->stream));
  HIPCHK(h, hipStreamSynchronize(h->stream));
  return JR_OK;
}

int jr_set_direction_slot(jr_handle* h, int32_t slot, const double* dir) {
  if (!h || !dir) return set_global_err(JR_E_ARG, "NULL argument");
  if (slot < 1 || slot > h->n_xslots)
    return h->fail(JR_E_ARG, "direction slot %d: the moves use slots 1..%d", slot, h->n_xslots);
  const auto& c = h->cfg;
  const size_t n = (size_t)c.n_replicas * c.n_walkers * c.n_atoms_max * 3;
  HIPCHK(h, hipSetDevice(c.device));
  HIPCHK(h, hipMemcpyAsync(h->dirx.p + (size_t)(slot - 1) * n, dir, sizeof(double) * n,
                           hipMemcpyHostToDevice, h->stream));
  HIPCHK(h, hipStreamSynchronize(h->stream));
  return JR_OK;
}

int jr_set_run_params(jr_handle* h, const double* pressure, const double* mu,
                      const double* step_sizes) {
  if (!h) return set_global_err(JR_E_ARG, "NULL handle");
  const auto& c = h->cfg;
  const int R = c.n_replicas, S = c.pot.n_species;
  HIPCHK(h, hipSetDevice(c.device));
  // stored energies carry P V - mu N: re-evaluate only when those change
  bool ens_changed = false;
  if (pressure && !std::equal(pressure, pressure + R, h->h_pressure.begin()))
    ens_changed = true;
  if (mu && !std::equal(mu, mu + (size_t)R * S, h->h_mu.begin()))
    ens_changed = true;
  if (pressure) h->h_pressure.assign(pressure, pressure + R);
  if (mu) h->h_mu.assign(mu, mu + (size_t)R * S);
  if (step_sizes) h->h_step.assign(step_sizes, step_sizes + (size_t)R * c.n_moves);
  HIPCHK(h, hipMemcpyAsync(h->pressure.p, h->h_pressure.data(), sizeof(double) * R, hipMemcpyHostToDevice, h->stream));
  HIPCHK(h, hipMemcpyAsync(h->mu.p, h->h_mu.data(), sizeof(double) * R * S, hipMemcpyHostToDevice, h->stream));
  HIPCHK(h, hipMemcpyAsync(h->step_sizes.p, h->h_step.data(), sizeof(double) * R * c.n_moves, hipMemcpyHostToDevice, h->stream));
  HIPCHK(h, hipStreamSynchronize(h->stream));
  if (h->have_pop && ens_changed) return recompute_energies(h);
  return JR_OK;
}

int jr_get_step_sizes(jr_handle* h, double* step_sizes) {
  if (!h || !step_sizes) return set_global_err(JR_E_ARG, "NULL argument");
  std::copy(h->h_step.begin(), h->h_step.end(), step_sizes);
  return JR_OK;
}

int jr_set_pair_valid(jr_handle* h, const int32_t* pair_valid) {
  if (!h || !pair_valid) return set_global_err(JR_E_ARG, "NULL argument");
  const int R = h->cfg.n_replicas;
  if (R < 2) return JR_OK;
  HIPCHK(h, hipSetDevice(h->cfg.device));
  HIPCHK(h, hipMemcpy(h->pair_valid.p, pair_valid, sizeof(int) * (R - 1), hipMemcpyHostToDevice));
  return JR_OK;
}

int jr_set_iteration(jr_handle* h, int64_t i) {
  if (!h) return set_global_err(JR_E_ARG, "NULL handle");
  h->iter_i = i;
  return JR_OK;
}

int jr_get_rng_state(jr_handle* h, int64_t* state) {
  if (!h || !state) return set_global_err(JR_E_ARG, "NULL argument");
  state[0] = h->rng_iter;
  state[1] = h->trial_rng;
  state[2] = h->burn_rng;
  return JR_OK;
}

int jr_set_rng_state(jr_handle* h, const int64_t* state) {
  if (!h || !state) return set_global_err(JR_E_ARG, "NULL argument");
  if (state[0] < 0 || state[1] < 0 || state[2] < 0)
    return h->fail(JR_E_ARG, "draw counters must be >= 0");
  h->rng_iter = state[0];
  h->trial_rng = state[1];
  h->burn_rng = state[2];
  return JR_OK;
}

int jr_set_emax(jr_handle* h, const double* emax) {
  if (!h || !emax) return set_global_err(JR_E_ARG, "NULL argument");
  HIPCHK(h, hipSetDevice(h->cfg.device));
  HIPCHK(h, hipMemcpy(h->emax.p, emax, sizeof(double) * h->cfg.n_replicas, hipMemcpyHostToDevice));
  return JR_OK;
}

int jr_get_emax(jr_handle* h, double* emax) {
  if (!h || !emax) return set_global_err(JR_E_ARG, "NULL argument");
  HIPCHK(h, hipSetDevice(h->cfg.device));
  HIPCHK(h, hipMemcpy(emax, h->emax.p, sizeof(double) * h->cfg.n_replicas, hipMemcpyDeviceToHost));
  return JR_OK;
}

}  // extern "C"

namespace {

int copy_outputs(jr_handle* h, int n_iter, jr_iter_out* out) {
  if (!out) return JR_OK;
  const auto& c = h->cfg;
  const size_t R = c.n_replicas, Nm = c.n_atoms_max;
  const size_t n = (size_t)n_iter;
  hipStream_t s = h->stream;
  auto d2h = [&](void* dst, const void* src, size_t bytes) {
    return hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, s);
  };
  const size_t D = n * R * c.n_cull;  // dead points in the block
  if (out->emax) HIPCHK(h, d2h(out->emax, h->o_emax.p, sizeof(double) * n * R));
  if (out->worst) HIPCHK(h, d2h(out->worst, h->o_worst.p, sizeof(int) * D));
  if (out->clone) HIPCHK(h, d2h(out->clone, h->o_clone.p, sizeof(int) * D));
  if (out->dead_volume) HIPCHK(h, d2h(out->dead_volume, h->o_dead_volume.p, sizeof(double) * D));
  if (out->dead_pos) HIPCHK(h, d2h(out->dead_pos, h->o_dead_pos.p, sizeof(double) * D * Nm * 3));
  if (out->dead_types) HIPCHK(h, d2h(out->dead_types, h->o_dead_types.p, sizeof(int) * D * Nm));
  if (out->dead_cell) HIPCHK(h, d2h(out->dead_cell, h->o_dead_cell.p, sizeof(double) * D * 9));
  if (out->dead_energy) HIPCHK(h, d2h(out->dead_energy, h->o_dead_energy.p, sizeof(double) * D));
  if (out->counters)
    HIPCHK(h, d2h(out->counters, h->o_counters.p, sizeof(int) * n * R * c.n_moves * JR_NCOUNT));
  if (R > 1) {
    if (out->re_acc) HIPCHK(h, d2h(out->re_acc, h->o_re_acc.p, sizeof(int) * n * (R - 1)));
    if (out->re_att) HIPCHK(h, d2h(out->re_att, h->o_re_att.p, sizeof(int) * n * (R - 1)));
  }
  if (out->walk) HIPCHK(h, d2h(out->walk, h->o_walk.p, sizeof(int) * n * R * h->n_walk));
  return JR_OK;
}

int begin_block(jr_handle* h, int n_iter) {
  int rc = ensure_out(h, n_iter);
  if (rc) return rc;
  IterState st{h->iter_i, h->rng_iter, 0, h->cfg.re_interval};
  HIPCHK(h, hipMemcpyAsync(h->st.p, &st, sizeof(st), hipMemcpyHostToDevice, h->stream));
  if ((int)h->ev.size() < 2 * n_iter) {
    size_t old = h->ev.size();
    h->ev.resize(std::max(2 * n_iter, 2 * h->out_cap));
    for (size_t i = old; i < h->ev.size(); ++i) HIPCHK(h, hipEventCreate(&h->ev[i]));
  }
  return JR_OK;
}

int upload_tape(jr_handle* h, const jr_tape* t, int it) {
  const auto& c = h->cfg;
  const size_t R = c.n_replicas, K = c.n_walkers;
  hipStream_t s = h->stream;
  auto h2d = [&](void* dst, const void* src, size_t bytes) {
    return hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, s);
  };
  if (!t->sel_u || !t->clone_u || !t->schedule || !t->normals || !t->uniforms)
    return h->fail(JR_E_ARG, "tape: sel_u/clone_u/schedule/normals/uniforms required");
  if (uses_extra_u(c) && !t->extra_u) return h->fail(JR_E_ARG, "tape: extra_u required");
  HIPCHK(h, h2d(h->sel_u.p, t->sel_u + (size_t)it * R * K, sizeof(double) * R * K));
  HIPCHK(h, h2d(h->clone_u.p, t->clone_u + (size_t)it * R, sizeof(double) * R));
  if (uses_extra_u(c))
    HIPCHK(h, h2d(h->extra_u.p, t->extra_u + (size_t)it * R * K, sizeof(double) * R * K));
  for (int s2 = 0; s2 < c.n_mcmc_steps; ++s2) {
    int m = t->schedule[(size_t)it * c.n_mcmc_steps + s2];
    if (m < 0 || m >= c.n_moves) return h->fail(JR_E_ARG, "tape: move index %d out of range", m);
  }
  HIPCHK(h, h2d(h->sched.p, t->schedule + (size_t)it * c.n_mcmc_steps, sizeof(int) * c.n_mcmc_steps));
  if (h->normals.n < n_normals(h)) HIPCHK(h, h->normals.alloc(n_normals(h)));
  if (h->uniforms.n < n_uniforms(h)) HIPCHK(h, h->uniforms.alloc(n_uniforms(h)));
  h->draws_buffered = true;
  HIPCHK(h, h2d(h->normals.p, t->normals + (size_t)it * n_normals(h), sizeof(double) * n_normals(h)));
  HIPCHK(h, h2d(h->uniforms.p, t->uniforms + (size_t)it * n_uniforms(h), sizeof(double) * n_uniforms(h)));
  if (t->re_u && R > 1)
    HIPCHK(h, h2d(h->re_u.p, t->re_u + (size_t)it * n_re_u(h), sizeof(double) * n_re_u(h)));
  return JR_OK;
}

int finish_block(jr_handle* h, int n_iter, bool timed, jr_iter_out* out) {
  int rc = copy_outputs(h, n_iter, out);
  if (rc) return rc;
  HIPCHK(h, hipStreamSynchronize(h->stream));
  if (timed) {
    for (int it = 0; it < n_iter; ++it) {
      float ms = 0.f;
      HIPCHK(h, hipEventElapsedTime(&ms, h->ev[2 * it], h->ev[2 * it + 1]));
      h->walk_ms += ms;
      h->walk_launches += 1;
    }
  }
  h->iter_i += n_iter;
  h->rng_iter += n_iter;
  return JR_OK;
}

}  // namespace

extern "C" {

int jr_ns_iterations(jr_handle* h, int32_t n_iter, const jr_tape* tape,
                     jr_iter_out* out) {
  if (!h) return set_global_err(JR_E_ARG, "NULL handle");
  if (!h->have_pop) return h->fail(JR_E_STATE, "population not set");
  if (n_iter < 0) return h->fail(JR_E_ARG, "n_iter < 0");
  if (n_iter == 0) return JR_OK;
  HIPCHK(h, hipSetDevice(h->cfg.device));
  int rc = begin_block(h, n_iter);
  if (rc) return rc;
  for (int it = 0; it < n_iter; ++it) {
    if (tape) {
      rc = upload_tape(h, tape, it);
      if (rc) return rc;
    } else {
      enqueue_draws(h, h->rng_iter + it);
    }
    rc = enqueue_iteration_core(h, true, it);
    if (rc) return rc;
    rc = snapshot_iter(h, it);
    if (rc) return rc;
  }
  HIPCHK(h, hipGetLastError());
  return finish_block(h, n_iter, true, out);
}

int jr_ns_iterations_graph(jr_handle* h, int32_t n_iter, jr_iter_out* out) {
  if (!h) return set_global_err(JR_E_ARG, "NULL handle");
  if (!h->have_pop) return h->fail(JR_E_STATE, "population not set");
  if (h->learned)
    return h->fail(JR_E_UNSUPPORTED,
                   "the learned-model walk is host-sequenced (data-dependent launches); "
                   "use jr_ns_iterations");
  if (n_iter <= 0) return n_iter == 0 ? JR_OK : h->fail(JR_E_ARG, "n_iter < 0");
  HIPCHK(h, hipSetDevice(h->cfg.device));
  int rc = begin_block(h, n_iter);
  if (rc) return rc;
  if (!h->graph_exec) {
    HIPCHK(h, hipStreamBeginCapture(h->stream, hipStreamCaptureModeThreadLocal));
    enqueue_draws(h, h->rng_iter);
    (void)enqueue_iteration_core(h, false, 0);
    hipError_t e = hipStreamEndCapture(h->stream, &h->graph);
    if (e != hipSuccess) return h->fail(JR_E_HIP, "graph capture: %s", hipGetErrorString(e));
    HIPCHK(h, hipGraphInstantiate(&h->graph_exec, h->graph, nullptr, nullptr, 0));
  }
  h->last_draw_iter = h->rng_iter + n_iter - 1;
  h->draws_buffered = false;
  for (int it = 0; it < n_iter; ++it) {
    HIPCHK(h, hipGraphLaunch(h->graph_exec, h->stream));
    rc = snapshot_iter(h, it);
    if (rc) return rc;
  }
  return finish_block(h, n_iter, false, out);
}

int jr_get_last_tape(jr_handle* h, double* sel_u, double* clone_u,
                     double* extra_u, int32_t* schedule, double* normals,
                     double* uniforms, double* re_u) {
  if (!h) return set_global_err(JR_E_ARG, "NULL handle");
  const auto& c = h->cfg;
  const size_t R = c.n_replicas, K = c.n_walkers;
  HIPCHK(h, hipSetDevice(c.device));
  hipStream_t s = h->stream;
  auto d2h = [&](void* dst, const void* src, size_t bytes) {
    return hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, s);
  };
  if (sel_u) HIPCHK(h, d2h(sel_u, h->sel_u.p, sizeof(double) * R * K));
  if (clone_u) HIPCHK(h, d2h(clone_u, h->clone_u.p, sizeof(double) * R));
  if (extra_u && uses_extra_u(c)) HIPCHK(h, d2h(extra_u, h->extra_u.p, sizeof(double) * R * K));
  if (schedule) HIPCHK(h, d2h(schedule, h->sched.p, sizeof(int) * c.n_mcmc_steps));
  if ((normals || uniforms) && !h->draws_buffered) {
    // the walk drew them in place: write the same values into the tape
    // buffers (allocated on first use)
    if (h->normals.n < n_normals(h)) HIPCHK(h, h->normals.alloc(n_normals(h)));
    if (h->uniforms.n < n_uniforms(h)) HIPCHK(h, h->uniforms.alloc(n_uniforms(h)));
    const unsigned long long seed = c.seed;
    hipLaunchKernelGGL(fill_normal_kernel, dim3(grid_for(n_normals(h))), dim3(256), 0, s,
                       h->normals.p, (long long)n_normals(h), seed, (const IterState*)nullptr,
                       h->last_draw_iter, 5u);
    hipLaunchKernelGGL(fill_uniform_kernel, dim3(grid_for(n_uniforms(h))), dim3(256), 0, s,
                       h->uniforms.p, (long long)n_uniforms(h), seed,
                       (const IterState*)nullptr, h->last_draw_iter, 6u);
    HIPCHK(h, hipGetLastError());
  }
  if (normals) HIPCHK(h, d2h(normals, h->normals.p, sizeof(double) * n_normals(h)));
  if (uniforms) HIPCHK(h, d2h(uniforms, h->uniforms.p, sizeof(double) * n_uniforms(h)));
  if (re_u && R > 1) HIPCHK(h, d2h(re_u, h->re_u.p, sizeof(double) * n_re_u(h)));
  HIPCHK(h, hipStreamSynchronize(s));
  return JR_OK;
}

int jr_replica_exchange(jr_handle* h, const double* re_u, int32_t* acc,
                        int32_t* att) {
  if (!h) return set_global_err(JR_E_ARG, "NULL handle");
  const auto& c = h->cfg;
  const int R = c.n_replicas;
  if (R < 2) return JR_OK;
  HIPCHK(h, hipSetDevice(c.device));
  DBuf<int> dacc, datt;
  HIPCHK(h, dacc.alloc(R - 1));
  HIPCHK(h, datt.alloc(R - 1));
  if (re_u) {
    HIPCHK(h, hipMemcpyAsync(h->re_u.p, re_u, sizeof(double) * n_re_u(h), hipMemcpyHostToDevice, h->stream));
  } else {
    // independent stream id so standalone passes do not reuse iteration draws
    hipLaunchKernelGGL(fill_uniform_kernel, dim3(grid_for(n_re_u(h))), dim3(256), 0, h->stream,
                       h->re_u.p, (long long)n_re_u(h), (unsigned long long)c.seed,
                       (const IterState*)nullptr, (long long)(h->trial_rng++), 9u);
  }
  ReArgs ra = re_args(h, nullptr);
  ra.acc = dacc.p;
  ra.att = datt.p;
  hipLaunchKernelGGL(re_kernel, dim3(1), dim3(re_threads(h)), 0, h->stream, ra);
  HIPCHK(h, hipGetLastError());
  if (acc) HIPCHK(h, hipMemcpyAsync(acc, dacc.p, sizeof(int) * (R - 1), hipMemcpyDeviceToHost, h->stream));
  if (att) HIPCHK(h, hipMemcpyAsync(att, datt.p, sizeof(int) * (R - 1), hipMemcpyDeviceToHost, h->stream));
  HIPCHK(h, hipStreamSynchronize(h->stream));
  return JR_OK;
}

int jr_trial_moves(jr_handle* h, int32_t move_idx, const double* trial_step,
                   const double* emax, int32_t n_samples,
                   const int32_t* sample_idx, const double* normals,
                   const double* uniforms, int32_t* accepted,
                   int32_t* reason) {
  if (!h) return set_global_err(JR_E_ARG, "NULL handle");
  const auto& c = h->cfg;
  if (!h->have_pop) return h->fail(JR_E_STATE, "population not set");
  if (move_idx < 0 || move_idx >= c.n_moves) return h->fail(JR_E_ARG, "bad move index");
  if (n_samples < 1 || !sample_idx || !trial_step)
    return h->fail(JR_E_ARG, "need n_samples >= 1, sample_idx, trial_step");
  if ((normals == nullptr) != (uniforms == nullptr))
    return h->fail(JR_E_ARG, "normals and uniforms must both be given or both NULL");
  const int R = c.n_replicas, Nm = c.n_atoms_max;
  const size_t T = (size_t)R * n_samples;
  for (size_t q = 0; q < T; ++q)
    if (sample_idx[q] < 0 || sample_idx[q] >= c.n_walkers)
      return h->fail(JR_E_ARG, "sample index out of range");
  HIPCHK(h, hipSetDevice(c.device));
  if (h->t_idx.n < T) {
    HIPCHK(h, h->t_idx.alloc(T));
    HIPCHK(h, h->t_reason.alloc(T));
    HIPCHK(h, h->t_normals.alloc(T * 3 * Nm));
    HIPCHK(h, h->t_uniforms.alloc(T * 4));
  }
  if (!h->t_step.p) {
    HIPCHK(h, h->t_step.alloc(R));
    HIPCHK(h, h->t_emax.alloc(R));
  }
  hipStream_t s = h->stream;
  HIPCHK(h, hipMemcpyAsync(h->t_idx.p, sample_idx, sizeof(int) * T, hipMemcpyHostToDevice, s));
  HIPCHK(h, hipMemcpyAsync(h->t_step.p, trial_step, sizeof(double) * R, hipMemcpyHostToDevice, s));
  if (emax) HIPCHK(h, hipMemcpyAsync(h->t_emax.p, emax, sizeof(double) * R, hipMemcpyHostToDevice, s));
  else HIPCHK(h, hipMemcpyAsync(h->t_emax.p, h->emax.p, sizeof(double) * R, hipMemcpyDeviceToDevice, s));
  if (normals) {
    HIPCHK(h, hipMemcpyAsync(h->t_normals.p, normals, sizeof(double) * T * 3 * Nm, hipMemcpyHostToDevice, s));
    HIPCHK(h, hipMemcpyAsync(h->t_uniforms.p, uniforms, sizeof(double) * T * 4, hipMemcpyHostToDevice, s));
  } else {
    long long it = h->trial_rng++;
    hipLaunchKernelGGL(fill_normal_kernel, dim3(grid_for(T * 3 * Nm)), dim3(256), 0, s,
                       h->t_normals.p, (long long)(T * 3 * Nm), (unsigned long long)c.seed,
                       (const IterState*)nullptr, it, 10u);
    hipLaunchKernelGGL(fill_uniform_kernel, dim3(grid_for(T * 4)), dim3(256), 0, s,
                       h->t_uniforms.p, (long long)(T * 4), (unsigned long long)c.seed,
                       (const IterState*)nullptr, it, 11u);
  }
  if (h->learned) {
    std::vector<int> rs(T);
    const int rc = jrl::learned_trial_moves(h, move_idx, n_samples, sample_idx, h->t_step.p,
                                            h->t_emax.p, h->t_normals.p, h->t_uniforms.p,
                                            rs.data());
    if (rc) return rc;
    for (size_t q = 0; q < T; ++q) {
      if (accepted) accepted[q] = rs[q] == 0;
      if (reason) reason[q] = rs[q];
    }
    return JR_OK;
  }
  HIPCHK(h, hipMemsetAsync(h->queue.p, 0, sizeof(int), s));
  WalkArgs a = walk_args(h);
  a.emax = h->t_emax.p;
  a.normals = h->t_normals.p;
  a.uniforms = h->t_uniforms.p;
  a.total = (int)T;
  a.trial = 1;
  a.trial_move = move_idx;
  a.n_samples = n_samples;
  a.sample_idx = h->t_idx.p;
  a.trial_step = h->t_step.p;
  a.trial_reason = h->t_reason.p;
  a.counters = nullptr;
  int blocks = std::max(1, std::min((int)((T + h->ipb - 1) / h->ipb), h->blocks));
  launch_walk(h->A, h->cfgbits, h->occ, a, blocks, h->wpb, h->lds_walk, s);
  HIPCHK(h, hipGetLastError());
  std::vector<int> rs(T);
  HIPCHK(h, hipMemcpyAsync(rs.data(), h->t_reason.p, sizeof(int) * T, hipMemcpyDeviceToHost, s));
  HIPCHK(h, hipStreamSynchronize(s));
  for (size_t q = 0; q < T; ++q) {
    if (accepted) accepted[q] = rs[q] == 0;
    if (reason) reason[q] = rs[q];
  }
  return JR_OK;
}

int jr_initial_walk(jr_handle* h, int32_t walklength, const double* emax,
                    const int32_t* schedule, const double* normals,
                    const double* uniforms, int32_t max_chunk,
                    int32_t* counters_out) {
  if (!h) return set_global_err(JR_E_ARG, "NULL handle");
  const auto& c = h->cfg;
  if (!h->have_pop) return h->fail(JR_E_STATE, "population not set");
  if (walklength < 1 || !emax) return h->fail(JR_E_ARG, "need walklength >= 1 and emax");
  const bool host = schedule != nullptr;
  if ((normals != nullptr) != host || (uniforms != nullptr) != host)
    return h->fail(JR_E_ARG, "schedule, normals and uniforms must all be given or all NULL");
  const int R = c.n_replicas, K = c.n_walkers, Nm = c.n_atoms_max, wl = walklength;
  if (host)
    for (int t = 0; t < wl; ++t)
      if (schedule[t] < 0 || schedule[t] >= c.n_moves)
        return h->fail(JR_E_ARG, "schedule entry %d out of range", t);
  HIPCHK(h, hipSetDevice(c.device));
  hipStream_t s = h->stream;
  // walkers per replica per launch: all K unless their normals pass 1 GiB
  const size_t per_walker = (size_t)wl * 3 * Nm;
  const size_t cap = ((size_t)1 << 30) / sizeof(double);
  int chunk = (int)std::min<size_t>(
      (size_t)K, std::max<size_t>(1, cap / ((size_t)R * per_walker)));
  if (max_chunk > 0) chunk = std::min(chunk, (int)max_chunk);
  // the learned walk takes passes of at most the chains its buffers hold
  if (h->learned) chunk = std::max(1, std::min(chunk, jrl::learned_chains(h) / R));
  if (h->b_chunk != chunk || h->b_steps != wl) {
    HIPCHK(h, h->b_normals.alloc((size_t)R * chunk * per_walker));
    HIPCHK(h, h->b_uniforms.alloc((size_t)R * chunk * wl * 4));
    HIPCHK(h, h->b_walk.alloc((size_t)R * chunk));
    HIPCHK(h, h->b_sched.alloc((size_t)wl));
    h->b_chunk = chunk;
    h->b_steps = wl;
  }
  if (!h->b_ctr.p) HIPCHK(h, h->b_ctr.alloc((size_t)R * c.n_moves * JR_NCOUNT));
  if (!h->t_emax.p) {
    HIPCHK(h, h->t_step.alloc(R));
    HIPCHK(h, h->t_emax.alloc(R));
  }
  HIPCHK(h, hipMemcpyAsync(h->t_emax.p, emax, sizeof(double) * R, hipMemcpyHostToDevice, s));
  HIPCHK(h, hipMemsetAsync(h->b_ctr.p, 0, sizeof(int) * R * c.n_moves * JR_NCOUNT, s));
  // one move schedule for every walker of every replica (burn_in.py:113-117)
  if (host) {
    HIPCHK(h, hipMemcpyAsync(h->b_sched.p, schedule, sizeof(int) * wl, hipMemcpyHostToDevice, s));
  } else {
    hipLaunchKernelGGL(schedule_kernel, dim3(1), dim3(64), 0, s, h->b_sched.p, wl,
                       (unsigned long long)c.seed, (const IterState*)nullptr,
                       (long long)h->burn_rng, 14u, h->cum.p, c.n_moves);
  }
  // the learned walk dispatches its launches by move type from the host
  std::vector<int> sched_host;
  if (h->learned) {
    sched_host.resize(wl);
    HIPCHK(h, hipMemcpyAsync(sched_host.data(), h->b_sched.p, sizeof(int) * wl,
                             hipMemcpyDeviceToHost, s));
    HIPCHK(h, hipStreamSynchronize(s));
  }
  std::vector<int> widx((size_t)R * chunk);
  for (int k0 = 0; k0 < K; k0 += chunk) {
    const int ck = std::min(chunk, K - k0);
    for (int r = 0; r < R; ++r)
      for (int j = 0; j < ck; ++j) widx[(size_t)r * ck + j] = k0 + j;
    HIPCHK(h, hipMemcpyAsync(h->b_walk.p, widx.data(), sizeof(int) * R * ck,
                             hipMemcpyHostToDevice, s));
    if (host) {
      // rows k0 .. k0+ck of [R][K][wl][3][Nmax] and [R][K][wl][4]
      HIPCHK(h, hipMemcpy2DAsync(h->b_normals.p, sizeof(double) * ck * per_walker,
                                 normals + (size_t)k0 * per_walker,
                                 sizeof(double) * K * per_walker,
                                 sizeof(double) * ck * per_walker, R,
                                 hipMemcpyHostToDevice, s));
      HIPCHK(h, hipMemcpy2DAsync(h->b_uniforms.p, sizeof(double) * ck * wl * 4,
                                 uniforms + (size_t)k0 * wl * 4,
                                 sizeof(double) * K * wl * 4,
                                 sizeof(double) * ck * wl * 4, R,
                                 hipMemcpyHostToDevice, s));
    } else {
      const long long it = h->burn_rng;
      const size_t nn = (size_t)R * ck * per_walker, nu = (size_t)R * ck * wl * 4;
      hipLaunchKernelGGL(fill_normal_kernel, dim3(grid_for(nn)), dim3(256), 0, s,
                         h->b_normals.p, (long long)nn, (unsigned long long)c.seed,
                         (const IterState*)nullptr, it * 4096 + k0 / chunk, 12u);
      hipLaunchKernelGGL(fill_uniform_kernel, dim3(grid_for(nu)), dim3(256), 0, s,
                         h->b_uniforms.p, (long long)nu, (unsigned long long)c.seed,
                         (const IterState*)nullptr, it * 4096 + k0 / chunk, 13u);
    }
    if (h->learned) {
      const int rc = jrl::learned_burn(h, wl, sched_host.data(), ck, h->b_walk.p, h->t_emax.p,
                                       h->b_normals.p, h->b_uniforms.p, h->b_ctr.p);
      if (rc) return rc;
      continue;
    }
    HIPCHK(h, hipMemsetAsync(h->queue.p, 0, sizeof(int), s));
    WalkArgs a = walk_args(h);
    a.emax = h->t_emax.p;
    a.walk = h->b_walk.p;
    a.n_walk = ck;
    a.n_steps = wl;
    a.schedule = h->b_sched.p;
    a.normals = h->b_normals.p;
    a.uniforms = h->b_uniforms.p;
    a.counters = h->b_ctr.p;
    a.total = R * ck;
    const int blocks = std::max(1, std::min((R * ck + h->ipb - 1) / h->ipb, h->max_blocks));
    launch_walk(h->A, h->cfgbits, h->occ, a, blocks, h->wpb, h->lds_walk, s);
    HIPCHK(h, hipGetLastError());
  }
  if (!host) h->burn_rng += 1;
  if (counters_out)
    HIPCHK(h, hipMemcpyAsync(counters_out, h->b_ctr.p, sizeof(int) * R * c.n_moves * JR_NCOUNT,
                             hipMemcpyDeviceToHost, s));
  HIPCHK(h, hipStreamSynchronize(s));
  return JR_OK;
}

int jr_autotune(jr_handle* h, int32_t reps, int32_t* best_wpb,
                int32_t* best_occ, double* best_ms) {
  if (!h) return set_global_err(JR_E_ARG, "NULL handle");
  if (h->learned)
    return h->fail(JR_E_UNSUPPORTED, "%s: pair-potential walk only", __func__);
  if (!h->have_pop) return h->fail(JR_E_STATE, "population not set");
  if (reps < 1) return h->fail(JR_E_ARG, "reps must be >= 1");
  const auto& c = h->cfg;
  HIPCHK(h, hipSetDevice(c.device));
  const size_t W = (size_t)c.n_replicas * c.n_walkers, Nm = c.n_atoms_max;
  hipStream_t s = h->stream;
  // save the population and contour; tuning leaves both unchanged
  DBuf<double> s_pos, s_dir, s_cell, s_e, s_emax, t_pos, t_dir, t_cell, t_e;
  DBuf<int> s_t, s_n, t_t, t_n;
  HIPCHK(h, s_pos.alloc(W * Nm * 3)); HIPCHK(h, s_dir.alloc(W * Nm * 3));
  HIPCHK(h, s_cell.alloc(W * 9)); HIPCHK(h, s_e.alloc(W)); HIPCHK(h, s_emax.alloc(c.n_replicas));
  HIPCHK(h, s_t.alloc(W * Nm)); HIPCHK(h, s_n.alloc(W));
  HIPCHK(h, t_pos.alloc(W * Nm * 3)); HIPCHK(h, t_dir.alloc(W * Nm * 3));
  HIPCHK(h, t_cell.alloc(W * 9)); HIPCHK(h, t_e.alloc(W));
  HIPCHK(h, t_t.alloc(W * Nm)); HIPCHK(h, t_n.alloc(W));
  auto copy_pop = [&](double* dp, double* dd, double* dc, double* de, int* dt, int* dn,
                      const double* sp, const double* sd, const double* sc,
                      const double* se, const int* st, const int* sn) -> hipError_t {
    hipError_t e;
    if ((e = hipMemcpyAsync(dp, sp, W * Nm * 3 * 8, hipMemcpyDeviceToDevice, s))) return e;
    if ((e = hipMemcpyAsync(dd, sd, W * Nm * 3 * 8, hipMemcpyDeviceToDevice, s))) return e;
    if ((e = hipMemcpyAsync(dc, sc, W * 9 * 8, hipMemcpyDeviceToDevice, s))) return e;
    if ((e = hipMemcpyAsync(de, se, W * 8, hipMemcpyDeviceToDevice, s))) return e;
    if ((e = hipMemcpyAsync(dt, st, W * Nm * 4, hipMemcpyDeviceToDevice, s))) return e;
    return hipMemcpyAsync(dn, sn, W * 4, hipMemcpyDeviceToDevice, s);
  };
  HIPCHK(h, copy_pop(s_pos.p, s_dir.p, s_cell.p, s_e.p, s_t.p, s_n.p, h->pos.p, h->dir.p,
                     h->cell.p, h->energy.p, h->types.p, h->n_atoms.p));
  HIPCHK(h, hipMemcpyAsync(s_emax.p, h->emax.p, c.n_replicas * 8, hipMemcpyDeviceToDevice, s));
  DBuf<double> s_dirx;
  const size_t nx = (size_t)h->n_xslots * W * Nm * 3;
  if (nx) {
    HIPCHK(h, s_dirx.alloc(nx));
    HIPCHK(h, hipMemcpyAsync(s_dirx.p, h->dirx.p, nx * 8, hipMemcpyDeviceToDevice, s));
  }
  // one iteration's draws and selection, then the post-selection state
  int rc = begin_block(h, 1);
  if (rc) return rc;
  enqueue_draws(h, h->rng_iter);
  launch_select(h, s);
  HIPCHK(h, copy_pop(t_pos.p, t_dir.p, t_cell.p, t_e.p, t_t.p, t_n.p, h->pos.p, h->dir.p,
                     h->cell.p, h->energy.p, h->types.p, h->n_atoms.p));
  hipEvent_t e0, e1;
  HIPCHK(h, hipEventCreate(&e0));
  HIPCHK(h, hipEventCreate(&e1));
  const int wpbs[3] = {4, 2, 1};
  const int occs[2] = {0, 3};
  int bw = h->wpb, bo = h->occ_req;
  double bms = 1e300;
  bool one_per_lane = h->A == 1;
  for (int j = 0; j < h->n_classes; ++j) one_per_lane = one_per_lane || h->cls[j].A == 1;
  // an evaluation team has one geometry (eval_chunks waves per block)
  const bool team = (h->cfgbits & 32) != 0;
  for (int oi = 0; oi < (one_per_lane && !team ? 2 : 1); ++oi) {
    for (int wi = 0; wi < (team ? 1 : 3); ++wi) {
      rc = set_geometry(h, wpbs[wi], occs[oi]);
      if (rc) return rc;
      double tot = 0.0;
      for (int r = 0; r <= reps; ++r) {  // r == 0: warm-up launch
        HIPCHK(h, copy_pop(h->pos.p, h->dir.p, h->cell.p, h->energy.p, h->types.p,
                           h->n_atoms.p, t_pos.p, t_dir.p, t_cell.p, t_e.p, t_t.p, t_n.p));
        HIPCHK(h, hipMemsetAsync(h->queue.p, 0, sizeof(int) * JR_NQUEUE, s));
        HIPCHK(h, hipEventRecord(e0, s));
        HIPCHK(h, enqueue_walk(h, walk_args(h), s));
        HIPCHK(h, hipEventRecord(e1, s));
        HIPCHK(h, hipEventSynchronize(e1));
        float ms = 0.f;
        HIPCHK(h, hipEventElapsedTime(&ms, e0, e1));
        if (r > 0) tot += ms;
      }
      const double avg = tot / reps;
      if (avg < bms) {
        bms = avg;
        bw = h->wpb;
        bo = h->occ_req;
      }
    }
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  rc = set_geometry(h, bw, bo);
  if (rc) return rc;
  HIPCHK(h, copy_pop(h->pos.p, h->dir.p, h->cell.p, h->energy.p, h->types.p, h->n_atoms.p,
                     s_pos.p, s_dir.p, s_cell.p, s_e.p, s_t.p, s_n.p));
  HIPCHK(h, hipMemcpyAsync(h->emax.p, s_emax.p, c.n_replicas * 8, hipMemcpyDeviceToDevice, s));
  if (nx) HIPCHK(h, hipMemcpyAsync(h->dirx.p, s_dirx.p, nx * 8, hipMemcpyDeviceToDevice, s));
  HIPCHK(h, hipStreamSynchronize(s));
  if (best_wpb) *best_wpb = bw;
  if (best_occ) *best_occ = bo;
  if (best_ms) *best_ms = bms;
  return JR_OK;
}

int jr_set_launch(jr_handle* h, int32_t waves_per_block, int32_t occupancy) {
  if (!h) return set_global_err(JR_E_ARG, "NULL handle");
  if (h->learned)
    return h->fail(JR_E_UNSUPPORTED, "%s: pair-potential walk only", __func__);
  if (waves_per_block < 1 || waves_per_block > 4)
    return h->fail(JR_E_ARG, "waves_per_block must be in [1, 4]");
  if (occupancy != 0 && occupancy != 3)
    return h->fail(JR_E_ARG, "occupancy variant must be 0 or 3");
  HIPCHK(h, hipSetDevice(h->cfg.device));
  return set_geometry(h, waves_per_block, occupancy);
}

int jr_set_walk_block_cap(jr_handle* h, int32_t max_blocks) {
  if (!h) return set_global_err(JR_E_ARG, "NULL handle");
  if (h->learned)
    return h->fail(JR_E_UNSUPPORTED, "%s: pair-potential walk only", __func__);
  if (max_blocks < 0) return h->fail(JR_E_ARG, "max_blocks must be >= 0");
  HIPCHK(h, hipSetDevice(h->cfg.device));
  h->block_cap = max_blocks;
  return set_geometry(h, h->wpb, h->occ_req);
}

int jr_walk_timing(jr_handle* h, double* total_ms, int64_t* launches,
                   int32_t reset) {
  if (!h) return set_global_err(JR_E_ARG, "NULL handle");
  if (total_ms) *total_ms = h->walk_ms;
  if (launches) *launches = h->walk_launches;
  if (reset) {
    h->walk_ms = 0.0;
    h->walk_launches = 0;
  }
  return JR_OK;
}

int jr_get_info(jr_handle* h, int32_t* n_cu, int32_t* blocks,
                int32_t* waves_per_block) {
  if (!h) return set_global_err(JR_E_ARG, "NULL handle");
  if (n_cu) *n_cu = h->n_cu;
  if (blocks) *blocks = h->blocks;
  if (waves_per_block) *waves_per_block = h->wpb;
  return JR_OK;
}

}  // extern "C"

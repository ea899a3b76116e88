// The library handle (one population, one stream) shared by the
// translation units of libjrhip: the pair-potential walk in jr_lib.hip and
// the learned-model walk in jr_learned.hip.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdarg>
#include <cstdio>
#include <string>
#include <vector>

#include "jr_host.h"
#include "jr_types.h"

namespace jrl {
struct LearnedState;  // defined in jr_learned.hip
}

struct jr_handle {
  jr_config cfg;
  jr_mpnn mpnn_cfg{};
  jr_learned learned_cfg{};
  std::string err;
  int n_walk = 1;
  int max_pairs = 0;
  int A = 1;
  int cfgbits = 0;
  int occ = 0;      // walk-kernel occupancy variant (0 or 3) of the single launch
  int occ_req = 0;  // variant asked for (a one-atom-per-lane class may use it)
  // Replicas of different atoms-per-lane A (N <= 64, <= 128, <= 256) walk in
  // one launch per class, each with its own kernel instantiation; n_classes
  // is 0 when every replica shares one A.
  struct WalkClass {
    int A = 1, occ = 0, blocks = 1, n_rep = 0;
    int wpb = 4;  // waves per block (an evaluation team's size depends on A)
    size_t lds = 0;
  };
  int n_classes = 0;
  WalkClass cls[3];
  jrh::DBuf<int> cls_order;  // [3][R] replica order of each class
  hipStream_t stream = nullptr;
  // size classes after the first walk on stream2, forked from and joined
  // back into `stream` by ev_fork / ev_join
  hipStream_t stream2 = nullptr;
  // minimum-distance floors [2][S][S] (atom/cell moves, species moves)
  jrh::DBuf<double> dmin;
  bool dmin_types = false;
  hipEvent_t ev_fork = nullptr, ev_join = nullptr;
  // population
  jrh::DBuf<double> pos, dir, cell, energy;
  jrh::DBuf<double> dirx;  // [n_xslots][R*K][Nmax][3] scoped-GMC directions
  int n_xslots = 0;
  jrh::DBuf<int> types, n_atoms;
  bool have_pop = false;
  // run params
  jrh::DBuf<double> pressure, mu, step_sizes, emax;
  std::vector<double> h_pressure, h_mu, h_step;
  // potential
  jrh::DBuf<double> eps4, eps24, sig2;
  jrh::DBuf<jr::DevMove> moves;
  jrh::DBuf<double> cum;  // cumulative move probabilities
  // iteration work
  jrh::DBuf<int> walk, counters, queue, pair_valid, order, sched;
  jrh::DBuf<double> sel_u, clone_u, extra_u, normals, uniforms, re_u;
  // draws_buffered: normals/uniforms hold the current iteration's draws (a
  // host tape, or the learned walk's device draws).  Otherwise the pair walk
  // draws them in place and the buffers are filled only when
  // jr_get_last_tape asks for the draws of iteration last_draw_iter.
  bool draws_buffered = false;
  long long last_draw_iter = 0;
  jrh::DBuf<jr::IterState> st;
  // per-block outputs
  int out_cap = 0;
  jrh::DBuf<double> o_emax, o_dead_volume, o_dead_pos, o_dead_cell, o_dead_energy;
  jrh::DBuf<int> o_worst, o_clone, o_dead_types, o_counters, o_re_acc, o_re_att,
      o_walk;
  // trial buffers
  jrh::DBuf<int> t_idx, t_reason;
  jrh::DBuf<double> t_step, t_emax, t_normals, t_uniforms;
  long long trial_rng = 0;
  // fixed-contour walk of the whole population (burn-in)
  jrh::DBuf<int> b_walk, b_sched, b_ctr;
  jrh::DBuf<double> b_normals, b_uniforms;
  int b_chunk = 0;  // walkers per replica the draw buffers hold
  int b_steps = 0;  // sub-steps they were sized for
  long long burn_rng = 0;
  // timing
  std::vector<hipEvent_t> ev;
  double walk_ms = 0.0;
  long long walk_launches = 0;
  // graph
  hipGraph_t graph = nullptr;
  hipGraphExec_t graph_exec = nullptr;
  // launch geometry
  int n_cu = 0;
  int wpb = 4;
  int ipb = 4;  // chains a block walks at once (wpb, or 1 for an evaluation team)
  int blocks = 0;
  int max_blocks = 0;  // resident blocks of the walk kernel at this geometry
  int block_cap = 0;   // jr_set_walk_block_cap (0: none)
  size_t lds_walk = 0;
  long long iter_i = 0;
  long long rng_iter = 0;
  // learned model (message-passing judge + surrogate); NULL for pair potentials
  jrl::LearnedState* learned = nullptr;

  int fail(int code, const char* fmt, ...) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    err = buf;
    return code;
  }
};

#define HIPCHK(h, call)                                                   \
  do {                                                                    \
    hipError_t e_ = (call);                                               \
    if (e_ != hipSuccess)                                                 \
      return (h)->fail(JR_E_HIP, "%s failed: %s (%s:%d)", #call,          \
                       hipGetErrorString(e_), __FILE__, __LINE__);        \
  } while (0)

// Learned-model hooks, implemented in jr_learned.hip.
namespace jrl {
// Validate cfg->mpnn / cfg->learned and build the model state.
int learned_create(jr_handle* h, const jr_config* cfg);
void learned_destroy(jr_handle* h);
// Checks a population for the judge (cubic cells, edge >= rc, species).
int learned_check_population(jr_handle* h, const int32_t* types, const int32_t* n_atoms,
                             const double* cell);
// Exact energies H = U + P V of the whole population into h->energy.
int learned_recompute_energies(jr_handle* h);
// The walk of one iteration (after select, before RE), on h->stream.
int learned_walk(jr_handle* h);
// Trial moves for step-size adaptation: move m from each of the n sampled
// walkers per replica (sample_idx [R][n], host) at the trial step sizes
// tstep [R] under the contours temax [R] with the device draws tnorm
// [R][n][3][Nmax] and tunif [R][n][4]; reject reasons [R][n] (0 = accepted)
// to the host; the population is not changed.
int learned_trial_moves(jr_handle* h, int m, int n, const int32_t* sample_idx,
                        const double* tstep, const double* temax, const double* tnorm,
                        const double* tunif, int32_t* reason);
// Chains the learned walk's buffers hold (R n_walk).
int learned_chains(jr_handle* h);
// Burn-in walk of the walkers walk [R][nw] (device, R nw <= learned_chains):
// wl sub-steps of the move schedule sched (host) at the fixed contours emax
// [R] (device) with the draws normals [R][nw][wl][3][Nmax] and uniforms
// [R][nw][wl][4] (device); counters added into ctr [R][n_moves][JR_NCOUNT];
// the walkers are written back.
int learned_burn(jr_handle* h, int wl, const int* sched, int nw, const int* walk,
                 const double* emax, const double* normals, const double* uniforms,
                 int* ctr);
// Pair-potential judge of the learned walk (jr_lib.hip): model energies U and
// forces F (NULL: energy only) of the configurations listed in idx (NULL:
// all B) with the device count (NULL: B); outputs indexed by configuration.
hipError_t lj_eval(jr_handle* h, const double* pos, const int* types, const int* n_atoms,
                   const double* cell, const int* idx, const int* count, int B, double* U,
                   double* F, hipStream_t s);
}  // namespace jrl

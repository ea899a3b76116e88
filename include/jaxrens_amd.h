/*
 * jaxrens_amd -- C ABI of the MI355X (gfx950) nested-sampling inner walk.
 *
 * Plain pointers and sizes only; every entry point returns 0 on success and a
 * negative JR_E_* code on failure, with a message from jr_last_error() (or
 * jr_global_error() when no handle exists yet).  No C++ exceptions cross the
 * ABI.  A handle owns all device buffers and one HIP stream; it is not
 * thread-safe.  Host buffers are owned by the caller.
 *
 * Reference interfaces replaced (nunglert/jaxrens, src/jaxrens/...):
 *   jr_eval_energy_forces  <- backends/lj.py:125-204 (LJBackend.__call__),
 *                             backends/base.py:74-113 (eval_energy_and_forces),
 *                             backends/hamiltonian.py:112-153,
 *                             backends/ensemble.py:34-133 (EnsembleTerm)
 *   jr_create / jr_set_population / jr_set_run_params
 *                          <- sampling/mwg.py:40-223 (build_mwg, init_fn),
 *                             sampling/nested_sampling.py:260-383 (init_ns)
 *   jr_ns_iterations       <- sampling/nested_sampling.py:391-686 (ns_step,
 *                             one call per iteration from run_loop.py:511-516)
 *   jr_replica_exchange    <- sampling/moves/replica_exchange.py:318-436
 *                             (replica_exchange_step + EnsembleSwap),
 *                             sampling/inter_re_manager.py:218-270
 *   jr_trial_moves         <- sampling/adaptation/stepsize_handler.py:215-261
 *                             (trial moves of adjust_step_size)
 *   jr_initial_walk        <- init/burn_in.py:48-148 (_one_walk, one walk of
 *                             initial_walk :154-407)
 */
#ifndef JAXRENS_AMD_H
#define JAXRENS_AMD_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define JR_OK 0
#define JR_E_ARG (-1)
#define JR_E_HIP (-2)
#define JR_E_STATE (-3)
#define JR_E_UNSUPPORTED (-4)

/* move types (cli/schema/moves.py: gmc|galilean, volume, hmc,
 * alchemical_morph, shear, stretch, species_swap) */
#define JR_MOVE_GMC 0
#define JR_MOVE_VOLUME 1
#define JR_MOVE_HMC 2
#define JR_MOVE_MORPH 3
#define JR_MOVE_SHEAR 4   /* sampling/moves/shear.py:19-120 */
#define JR_MOVE_STRETCH 5 /* sampling/moves/stretch.py:18-97 */
#define JR_MOVE_SWAP 6    /* sampling/moves/swap.py:119-262 (species_swap) */

#define JR_MAX_MOVES 8
#define JR_MAX_SPECIES 8
#define JR_NCOUNT 7 /* proposed, rr0..rr3, n_evaluations, n_grad_evaluations */

#define JR_PREC_FP64 0
#define JR_PREC_FP32_PAIRS 1

typedef struct jr_handle jr_handle;

/* Pair potential: per species pair (a, b) at index a*S+b.
 * images[d] = supercell_trafo[d] / 2 (integer division) of the reference's
 * LJBackend (backends/lj.py:106-113,152-166): every pair is summed over the
 * minimum-image displacement plus the lattice translations
 * (i0, i1, i2) . cell with -images[d] <= i_d <= images[d], self pairs only at
 * non-zero translations; all zero (supercell_trafo (1,1,1)) is plain minimum
 * image.  Each entry at most JR_MAX_IMAGES; not with JR_PREC_FP32_PAIRS. */
#define JR_MAX_IMAGES 4
typedef struct {
  int32_t n_species;   /* S, 1..JR_MAX_SPECIES */
  int32_t precision;   /* JR_PREC_* */
  double rc2;          /* squared cutoff, +inf for none */
  const double* eps4;  /* [S*S] 4*eps_ab  */
  const double* eps24; /* [S*S] 24*eps_ab */
  const double* sig2;  /* [S*S] sigma_ab^2 */
  int32_t images[3];   /* image half-ranges per lattice vector */
  int32_t eval_chunks; /* 0 or 1: each evaluation sums every atom's pair
                          terms in one running sum over the ring steps;
                          C = 2, 4 or 8: the steps are split into C
                          consecutive chunks summed on their own and added
                          in chunk order (a fixed evaluation order; FP64
                          without images).  With C > 1 a team of C
                          wavefronts walks each chain, one chunk each, for
                          walks with few chains per iteration (n_extra = 0). */
} jr_pot;

typedef struct {
  int32_t type;        /* JR_MOVE_* */
  int32_t n_reflect;   /* gmc */
  int32_t n_leapfrog;  /* hmc */
  int32_t flat_v_prior;/* volume */
  int32_t n_species;   /* alchemical_morph */
  int32_t species_mask;/* species_swap: bit c set = type code c may be
                          exchanged; gmc: bit c set = atoms of type code c
                          move, the rest stay fixed; 0 = every species */
  double perturb_angle;/* gmc */
  double max_vpa;      /* volume, shear, stretch: cell-geometry guard */
  double min_vpa;
  double min_aspect;
  int32_t refresh_direction; /* gmc: fresh random direction every call */
  int32_t dir_slot;    /* gmc: persistent-direction slot; 0 = the
                          population's direction, 1.. = extra slots (one per
                          species-scoped move), < JR_MAX_MOVES */
} jr_move;

/* Seeded random-init invariant message-passing potential, the expensive
 * exact judge of the surrogate-screened walk (FP32 weights and arithmetic):
 * 8 Bessel radial functions with a polynomial envelope out to rc, two
 * continuous-filter interaction layers of 64 channels with SiLU, per-atom
 * linear readout, plus a 12th-power core repulsion below core_rc.  Weight
 * layouts: emb [S][64], W1/W2 [8][64], U1/U2 [64][64] (z = m U + c),
 * c1/c2/wo [64]. */
typedef struct {
  int32_t n_species; /* S, 1..2 */
  int32_t channels;  /* must be 64 */
  int32_t n_basis;   /* must be 8 */
  int32_t pad_;
  double rc, core_eps, core_sig, core_rc;
  const float *emb, *W1, *U1, *c1, *W2, *U2, *c2, *wo;
} jr_mpnn;

/* Surrogate-screened walk of the learned model (used when jr_config.mpnn is
 * set).  Modes:
 *   JR_LEARNED_PLAIN     exact judge forces steer every GMC trajectory and
 *                        the judge decides every move (no surrogate);
 *   JR_LEARNED_AUTO      two-stage delayed acceptance in an iteration where
 *                        the surrogate's held-out per-atom RMSE is <=
 *                        rmse_bound and every chain's contour bounds U/N at
 *                        or below train_max_epa (U < Emax - P V over the
 *                        volumes the moves allow), the plain walk otherwise;
 *                        the surrogate trains online;
 *   JR_LEARNED_SCREENED  delayed acceptance in every iteration (the bound is
 *                        ignored); the surrogate trains online.
 * The surrogate's enthalpy carries an error model refitted after every
 * iteration from fresh judge evaluations (before they are trained on):
 * per-atom error e = a + c V/N by least squares over all replicas, then per
 * replica the residual mean b_r and spread s_r.  With H_s = U_s - (a N + c V)
 * + P V, GMC trajectories reflect where H_s >= Emax + N b_r, and stage 1
 * compares the verdicts [H_s < Emax + N (b_r + delta_k s_r)] at the start
 * and end of a proposal, rejecting when they differ; stage 2 accepts iff the
 * judge's H < Emax (GMC, volume) or is the Metropolis test of the judge's
 * H + K (HMC, whose reference rule samples exp(-H) inside the contour).
 * Both verdicts are symmetric in the two states and every constant is fixed
 * during an iteration, so each move keeps its plain counterpart's target.
 * rmse_bound applies to the residual per-atom RMSE. */
#define JR_LEARNED_PLAIN 0
#define JR_LEARNED_AUTO 1
#define JR_LEARNED_SCREENED 2

typedef struct {
  int32_t mode;            /* JR_LEARNED_* */
  int32_t ring_capacity;   /* exactly evaluated configurations kept */
  int32_t train_batch;     /* configurations per Adam step */
  int32_t train_steps;     /* Adam steps per NS iteration */
  int32_t min_train;       /* stored configurations before training starts */
  int32_t min_heldout;     /* fresh evaluations before the RMSE estimate counts */
  double lr;               /* Adam step size */
  double rmse_bound;       /* per-atom RMSE bound of JR_LEARNED_AUTO */
  double delta_k;          /* screening slack in residual spreads */
  double train_max_epa;    /* configurations with U/N above this are not stored */
  /* pair baseline of the surrogate, eps ((sig/r)^12 - (sig/rc)^12) for
   * r < rc <= 3 (eps = 0: none); the network learns the judge minus it */
  double base_eps, base_sig, base_rc;
  const float* sur_params; /* initial surrogate [JR_SUR_NPARAMS] */
} jr_learned;

typedef struct {
  int32_t n_replicas;    /* R */
  int32_t n_walkers;     /* K (n_live) */
  int32_t n_atoms_max;   /* Nmax, <= 256 */
  int32_t n_moves;
  int32_t n_mcmc_steps;
  int32_t n_extra;       /* n_cull = 1: 0 .. K-2; n_cull > 1: 2*n_cull + n_extra <= K */
  int32_t n_cull;        /* walkers culled per iteration, >= 1 (nested_sampling.py:151-212);
                            the learned walk takes n_cull = 1 */
  int32_t n_swap_cycles;
  int32_t re_interval;   /* RE fires when i > 0 and i % re_interval == 0; 0 = never */
  int32_t device;
  int32_t waves_per_block; /* 0 = library default / tuned */
  int32_t pad_;
  uint64_t seed;
  jr_pot pot;
  const jr_move* moves;      /* [n_moves] */
  const double* move_weights;/* [n_moves] relative weights */
  /* Learned model: when mpnn != NULL the message-passing judge replaces
   * pot (NPT only; gmc and volume moves; cubic cells with edge >= rc) and
   * learned configures the surrogate (NULL = JR_LEARNED_PLAIN). */
  const jr_mpnn* mpnn;
  const jr_learned* learned;
  /* Configuration constraint (constraints/min_distance.py:25-46, gated as in
   * sampling/mwg.py:98-141): an accepted proposal that brings two atoms i, j
   * closer than min_distance[t_i*S + t_j] (minimum image; 0 disables a pair)
   * is undone, positions, cell, species and energy, and counts in no reject
   * bucket.  min_distance gates the moves that move atoms or the cell;
   * min_distance_types (the same or another [S][S] table; NULL = not gated)
   * gates alchemical_morph and species_swap, which the reference gates only
   * when the floors depend on the species.  NULL min_distance: no constraint.
   * FP64 pair walk without supercell images only. */
  const double* min_distance;
  const double* min_distance_types;
} jr_config;

/* Host-side draw tape for n_iter iterations (parity / replay).  Layouts:
 *   sel_u     [n_iter][R][K]
 *   clone_u   [n_iter][R]
 *   extra_u   [n_iter][R][K]               n_cull = 1: the extras' noise (NULL
 *                                          allowed when n_extra == 0); n_cull > 1:
 *                                          [0, K-n_cull) the survivors' permutation
 *                                          keys (survivor position p takes key p)
 *   schedule  [n_iter][n_steps]            move index per MWG sub-step
 *   normals   [n_iter][R][n_walk][n_steps][3][Nmax]
 *   uniforms  [n_iter][R][n_walk][n_steps][4]
 *   re_u      [n_iter][n_cycles][2][R/2][2] (NULL: RE draws unused)
 * One sub-step reads, by move type: gmc and hmc the normals [3][Nmax];
 * volume normal [0][0] and uniform 0; morph uniforms 0 (atom) and 1
 * (species); shear uniform 0 (cell vector floor(3u)) and normals [0][0],
 * [1][0]; stretch uniform 0 (axis pair floor(3u)) and normal [0][0];
 * species_swap uniforms 0 (species pair), 1 and 2 (atom of each species). */
typedef struct {
  const double* sel_u;
  const double* clone_u;
  const double* extra_u;
  const int32_t* schedule;
  const double* normals;
  const double* uniforms;
  const double* re_u;
} jr_tape;

/* Per-iteration outputs for a block of n_iter iterations (any may be NULL).
 * C = n_cull; the dead points of an iteration are in descending energy. */
typedef struct {
  double* emax;         /* [n_iter][R] contour = energy of the C-th dead point */
  int32_t* worst;       /* [n_iter][R][C] culled slots */
  int32_t* clone;       /* [n_iter][R][C] clone source of each culled slot */
  double* dead_volume;  /* [n_iter][R][C] */
  double* dead_pos;     /* [n_iter][R][C][Nmax][3] */
  int32_t* dead_types;  /* [n_iter][R][C][Nmax] */
  double* dead_cell;    /* [n_iter][R][C][9] */
  int32_t* counters;    /* [n_iter][R][n_moves][JR_NCOUNT] */
  int32_t* re_acc;      /* [n_iter][R-1] */
  int32_t* re_att;      /* [n_iter][R-1] */
  int32_t* walk;        /* [n_iter][R][n_walk], n_walk = C + n_extra: culled slots first */
  double* dead_energy;  /* [n_iter][R][C] */
} jr_iter_out;

const char* jr_global_error(void);
const char* jr_last_error(const jr_handle* h);

/* Batched energy (+forces) of B configurations, host buffers.
 * pos [B][Nmax][3], types [B][Nmax], n_atoms [B], cell [B][9],
 * pressure [B] (NULL = 0), mu [B][S] (NULL = 0).
 * Outputs: U [B] model energy, H [B] = U + P V - mu.N, F [B][Nmax][3]
 * (NULL = energy only). */
int jr_eval_energy_forces(const jr_pot* pot, int32_t device, int32_t B,
                          int32_t Nmax, const double* pos,
                          const int32_t* types, const int32_t* n_atoms,
                          const double* cell, const double* pressure,
                          const double* mu, double* U, double* H, double* F);

/* Allocation-free variant of jr_eval_energy_forces (same arguments, same
 * results bit for bit).  No counterpart in the reference: it is for callers
 * that cannot allocate device memory (serving loops, harnesses that forbid
 * hipMalloc).  The device staging buffers are one fixed array per GPU
 * (about 3.9 MiB, held until the process ends): B * Nmax <= 65536 atom
 * slots and B <= 4096 configurations, larger calls fail with JR_E_ARG.
 * Runs on the device's default stream and returns after the copies back.
 * Calls that overlap on one GPU (from several threads) share that array and
 * race; serialise them. */
#define JR_NOALLOC_MAX_SLOTS 65536
#define JR_NOALLOC_MAX_CONFIGS 4096
int jr_eval_energy_forces_noalloc(const jr_pot* pot, int32_t device, int32_t B,
                                  int32_t Nmax, const double* pos,
                                  const int32_t* types, const int32_t* n_atoms,
                                  const double* cell, const double* pressure,
                                  const double* mu, double* U, double* H, double* F);

int jr_create(const jr_config* cfg, jr_handle** out);
void jr_destroy(jr_handle* h);

/* Population upload; energies are evaluated on the device.  dir may be NULL
 * (zero GMC direction).  Shapes as jr_tape with [R*K] leading. */
int jr_set_population(jr_handle* h, const double* pos, const int32_t* types,
                      const int32_t* n_atoms, const double* cell,
                      const double* dir);
int jr_get_population(jr_handle* h, double* pos, int32_t* types,
                      int32_t* n_atoms, double* cell, double* energy,
                      double* dir);
/* Extra persistent GMC directions (galilean.py:107-131 direction_field,
 * cli/schema/moves.py GMCMoveSpec._direction_field: each species-scoped move
 * keeps its own), slot 1 .. max dir_slot of the moves, [R*K][Nmax][3].
 * jr_set_population zeroes them; get/set carry them across checkpoints. */
int jr_get_direction_slot(jr_handle* h, int32_t slot, double* dir);
int jr_set_direction_slot(jr_handle* h, int32_t slot, const double* dir);
/* pressure [R], mu [R][S] (NULL = 0), step_sizes [R][n_moves] */
int jr_set_run_params(jr_handle* h, const double* pressure, const double* mu,
                      const double* step_sizes);
int jr_get_step_sizes(jr_handle* h, double* step_sizes);
/* pair_valid [R-1]: RE pair (k, k+1) allowed (default: equal atom counts). */
int jr_set_pair_valid(jr_handle* h, const int32_t* pair_valid);
/* Outer-loop iteration index i used for the RE firing rule. */
int jr_set_iteration(jr_handle* h, int64_t i);
/* Counters of the device draw streams (Philox4x32-10 under the handle's
 * seed): state[0] NS iterations drawn, state[1] trial moves and standalone RE
 * passes, state[2] burn-in walks.  With the population, run parameters and
 * iteration index they make a checkpoint resume draw exactly what an
 * uninterrupted run draws (io/checkpoint.py:187-197 persists the run's PRNG
 * key for the same purpose). */
int jr_get_rng_state(jr_handle* h, int64_t* state /* [3] */);
int jr_set_rng_state(jr_handle* h, const int64_t* state /* [3] */);
int jr_get_emax(jr_handle* h, double* emax);
/* Set the resident contour emax [R] (the last iteration's, as a checkpoint
 * holds it; io/checkpoint.py:347-365 restores NSState.emax).  A new handle
 * starts at +inf; the next NS iteration replaces it, but trial moves and RE
 * passes before that iteration read it. */
int jr_set_emax(jr_handle* h, const double* emax);

/* Run n_iter NS iterations (select -> clone -> walk -> RE).  tape == NULL
 * draws everything on the device (Philox4x32-10).  One host sync at the end. */
int jr_ns_iterations(jr_handle* h, int32_t n_iter, const jr_tape* tape,
                     jr_iter_out* out);
/* Same, replayed from a captured HIP graph of one iteration (device draws). */
int jr_ns_iterations_graph(jr_handle* h, int32_t n_iter, jr_iter_out* out);

/* Device draws of the most recent iteration, in jr_tape layout (n_iter=1). */
int jr_get_last_tape(jr_handle* h, double* sel_u, double* clone_u,
                     double* extra_u, int32_t* schedule, double* normals,
                     double* uniforms, double* re_u);

/* One replica-exchange pass on the resident population using the current
 * emax; re_u [n_cycles][2][R/2][2] or NULL for device draws. */
int jr_replica_exchange(jr_handle* h, const double* re_u, int32_t* acc,
                        int32_t* att);

/* One trial move per sample (population untouched).  sample_idx [R][n],
 * trial_step [R], emax [R] (NULL = resident emax), normals
 * [R][n][3][Nmax], uniforms [R][n][4] (both NULL = device draws).
 * accepted/reason [R][n].  On a learned-model handle the move runs as the
 * learned walk would run it (judge-steered, or surrogate-steered and
 * screened when the walk screens), without counters, statistics or
 * training records. */
int jr_trial_moves(jr_handle* h, int32_t move_idx, const double* trial_step,
                   const double* emax, int32_t n_samples,
                   const int32_t* sample_idx, const double* normals,
                   const double* uniforms, int32_t* accepted,
                   int32_t* reason);

/* Fixed-contour walk of the whole live population before nested sampling
 * (init/burn_in.py:48-148, one walk of initial_walk :154-407): every walker
 * of replica r takes walklength MWG sub-steps at Emax = emax[r], held fixed,
 * with the handle's move table and step sizes, and one move schedule shared
 * by all walkers and replicas.  emax [R] (host).  Draws: schedule
 * [walklength], normals [R][K][walklength][3][Nmax] and uniforms
 * [R][K][walklength][4] from the host, or all three NULL for device draws.
 * Walkers go in launches of max_chunk per replica (<= 0: as many as keep the
 * draws of one launch within 1 GiB).  counters_out [R][n_moves][JR_NCOUNT]
 * (may be NULL) sums this walk's counters.  The NS contour, iteration state
 * and per-iteration outputs are left unchanged.  On a learned-model handle
 * the walkers take the learned walk's moves (judge-steered, or surrogate-
 * steered and screened when the walk screens) in passes of at most the walk's
 * chain count per launch, without statistics or training records. */
int jr_initial_walk(jr_handle* h, int32_t walklength, const double* emax,
                    const int32_t* schedule, const double* normals,
                    const double* uniforms, int32_t max_chunk,
                    int32_t* counters_out);

/* Autotuner (no reference counterpart; the reference leaves launch shapes to
 * XLA): times `reps` walk launches per candidate (waves per block 4/2/1 x
 * occupancy variant 0/3 for N <= 64) from one post-selection state and keeps
 * the fastest.  Population, contour and draw counters are left unchanged. */
int jr_autotune(jr_handle* h, int32_t reps, int32_t* best_wpb,
                int32_t* best_occ, double* best_ms);
/* Set the walk launch shape directly (waves per block 1..4, occupancy 0|3). */
int jr_set_launch(jr_handle* h, int32_t waves_per_block, int32_t occupancy);

/* Cap the persistent walk grid (every walk launch: NS iterations, trial
 * moves, burn-in) at max_blocks blocks; 0 removes the cap (as many blocks
 * as are resident).  No reference counterpart (the reference leaves launch
 * shapes to XLA).  With a small grid each wave takes many chains from the
 * work queue in turn, so tests reach that path at sizes the oracle checks
 * quickly; results do not depend on the cap. */
int jr_set_walk_block_cap(jr_handle* h, int32_t max_blocks);

/* Walk-kernel timing accumulated since the last reset (HIP events on the
 * handle's stream): total milliseconds and launch count. */
int jr_walk_timing(jr_handle* h, double* total_ms, int64_t* launches,
                   int32_t reset);
/* Launch geometry in use: CU count, persistent blocks, waves per block. */
int jr_get_info(jr_handle* h, int32_t* n_cu, int32_t* blocks,
                int32_t* waves_per_block);

/* ------------------------------------------------------------------------
 * Learned part.  No reference counterpart: the reference has no surrogate
 * and its ML backends (backends/, out of scope here) are not this judge.
 * ---------------------------------------------------------------------- */


/* Batched energy (+forces) of the judge, host buffers, shapes as
 * jr_eval_energy_forces (mu unused: NPT).  Cells must be cubic with edge
 * >= rc.  H = U + P |det cell| when pressure != NULL. */
int jr_mpnn_eval(const jr_mpnn* net, int32_t device, int32_t B, int32_t Nmax,
                 const double* pos, const int32_t* types,
                 const int32_t* n_atoms, const double* cell,
                 const double* pressure, double* U, double* H, double* F);

/* Radial-descriptor MLP surrogate (BF16 matrix cores, FP32 accumulate):
 * per atom 32 Gaussians (centres 0.8..2.5, width 0.1, smooth cutoff 3.0) per
 * neighbour species scaled by 1/4, plus the species one-hot, padded to 80
 * inputs; 80 -> 128 -> 128 -> 1 with SiLU; E = sum of per-atom energies.
 * FP32 parameter vector of JR_SUR_NPARAMS values: W1 [80][128], b1 [128],
 * W2 [128][128], b2 [128], w3 [128], b3. */
#define JR_SUR_INPUTS 80
#define JR_SUR_NPARAMS 27009

/* Surrogate energy (+forces) of B cubic configurations (edge >= 3.0),
 * host buffers as jr_eval_energy_forces; x_out [B][Nmax][80] receives the
 * BF16 descriptor rows (bit patterns) when not NULL.  base = {eps, sig, rc}
 * adds the pair baseline eps ((sig/r)^12 - (sig/rc)^12), r < rc <= 3, to
 * the network's energy (NULL: none). */
int jr_sur_eval(const float* params, const double* base, int32_t device, int32_t B,
                int32_t Nmax, const double* pos, const int32_t* types,
                const int32_t* n_atoms, const double* cell, int32_t n_species,
                double* U, double* F, uint16_t* x_out);

/* Rows of descriptors x [n_rows][80] (BF16 bits): per-row energies e_out
 * (incl. b3) and, when grad_out != NULL, the gradient of sum_r c[r] e_r with
 * respect to the FP32 parameters (deterministic order). */
int jr_sur_grad(const float* params, int32_t device, int32_t n_rows,
                const uint16_t* x, const float* c, float* e_out, float* grad_out);

/* One Adam step (beta1 0.9, beta2 0.999, eps 1e-8, step t >= 1) in place on
 * host arrays of JR_SUR_NPARAMS values, computed on the device. */
int jr_sur_adam(int32_t device, float* params, float* m, float* v,
                const float* grad, double lr, int32_t t);

/* Counters of the learned-model walk, accumulated since the last reset. */
typedef struct {
  int64_t iterations_screened; /* iterations walked with delayed acceptance */
  int64_t iterations_plain;    /* iterations walked with exact forces */
  int64_t proposals;           /* chain moves proposed */
  int64_t stage1_pass;         /* screened proposals passing stage 1 */
  int64_t accepted;            /* accepted moves (both walks) */
  int64_t exact_force_evals;   /* judge energy+force evaluations */
  int64_t exact_energy_evals;  /* judge energy-only evaluations */
  int64_t surrogate_evals;     /* surrogate evaluations */
  int64_t train_steps;         /* Adam steps */
  int64_t stored;              /* configurations in the training ring */
  double heldout_rmse;         /* latest held-out per-atom RMSE after the
                                  error model (inf: none) */
  double train_loss;           /* latest minibatch loss (per-atom MSE) */
  double walk_ms;              /* device time of the walks (HIP events) */
  double train_ms;             /* device time of training */
} jr_learned_stats;

int jr_learned_get_stats(jr_handle* h, jr_learned_stats* out, int32_t reset);
/* Current FP32 surrogate parameters [JR_SUR_NPARAMS]. */
int jr_learned_get_params(jr_handle* h, float* params);
/* Replace the surrogate (reset_adam != 0 also clears the Adam moments). */
int jr_learned_set_params(jr_handle* h, const float* params, int32_t reset_adam);
/* Switch JR_LEARNED_* mode between iterations. */
int jr_learned_set_mode(jr_handle* h, int32_t mode);

#ifdef __cplusplus
}
#endif
#endif

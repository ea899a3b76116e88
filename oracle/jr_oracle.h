/*
 * CPU oracle for the nested-sampling inner walk: an FP64 C restatement of the
 * reference algorithm (nunglert/jaxrens), used ONLY as test infrastructure
 * (parity checks and the CPU baseline in bench.py).  Nothing in the product
 * path links or calls this library.
 *
 * Every function fixes one canonical floating-point evaluation order (listed
 * in jr_oracle.c next to each step) so that a device implementation following
 * the same order can be compared bit-for-bit.  The reference itself leaves
 * that order to XLA; the per-function comments cite the reference lines whose
 * semantics are restated.
 */
#ifndef JR_ORACLE_H
#define JR_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum { OR_MOVE_GMC = 0, OR_MOVE_VOLUME = 1, OR_MOVE_HMC = 2, OR_MOVE_MORPH = 3,
       OR_MOVE_SHEAR = 4, OR_MOVE_STRETCH = 5, OR_MOVE_SWAP = 6 };

/* counters per (replica, move): proposed, rr0 (=accepted), rr1 energy,
 * rr2 cell, rr3 prior, n_evaluations, n_grad_evaluations */
enum { OR_NCOUNT = 7 };

typedef struct {
  int n_species;        /* S */
  double rc2;           /* squared cutoff; +inf for no cutoff */
  const double* eps4;   /* [S*S] 4*eps_ab */
  const double* eps24;  /* [S*S] 24*eps_ab */
  const double* sig2;   /* [S*S] sigma_ab^2 */
  int img[3];           /* image half-ranges supercell_trafo[d] / 2; all 0:
                           minimum image only (backends/lj.py:106-113) */
  const double* dmin_pos;   /* [S*S] minimum-distance floors gating moves
                               that move atoms or the cell; NULL = none */
  const double* dmin_types; /* [S*S] floors gating morph and swap; NULL = none */
  int chunks;           /* evaluation chunks C (0 or 1: one running sum per
                           atom over all ring steps; see or_energy_forces) */
} or_pot;

typedef struct {
  int type;
  int n_reflect;
  int n_leapfrog;
  int flat_v_prior;
  int n_species;
  int species_mask; /* species_swap / gmc scope, 0 = all */
  double perturb_angle;
  double max_vpa, min_vpa, min_aspect;
  int refresh_direction; /* gmc: fresh direction every call */
  int dir_slot;          /* gmc: 0 = dir, s > 0 = extra slot s */
} or_move;

typedef struct {
  int R, K, Nmax, S;
  double* pos;     /* [R*K][Nmax][3] */
  int32_t* types;  /* [R*K][Nmax] */
  int32_t* n_atoms;/* [R*K] */
  double* cell;    /* [R*K][9] rows = lattice vectors */
  double* energy;  /* [R*K] full Hamiltonian U + P V - mu.N */
  double* dir;     /* [R*K][Nmax][3] GMC persistent direction */
  double* dirx;    /* [n_xslots][R*K][Nmax][3] directions of the
                      species-scoped GMC moves (slots 1..n_xslots) */
  int n_xslots;
} or_pop;

typedef struct {
  const double* pressure;   /* [R] */
  const double* mu;         /* [R*S] */
  const double* step_sizes; /* [R*n_moves] */
} or_run;

typedef struct {
  const double* sel_u;     /* [R*K] worst-walker tie-break noise */
  const double* clone_u;   /* [R] */
  const double* extra_u;   /* [R*K] (unused when n_extra == 0) */
  const int32_t* schedule; /* [n_steps] shared move indices */
  const double* normals;   /* [R][n_walk][n_steps][3][Nmax] */
  const double* uniforms;  /* [R][n_walk][n_steps][4] */
  const double* re_u;      /* [n_cycles][2][max_pairs][2] (NULL: no RE) */
} or_draws;

typedef struct {
  double* emax;         /* [R] contour */
  int32_t* worst;       /* [R][n_cull] */
  int32_t* clone;       /* [R][n_cull] */
  int32_t* walk;        /* [R][n_walk] */
  double* dead_energy;  /* [R][n_cull] */
  double* dead_volume;  /* [R][n_cull] */
  int32_t* counters;    /* [R*n_moves*7] */
  int32_t* re_acc;      /* [R-1] */
  int32_t* re_att;      /* [R-1] */
} or_iter_out;

double or_det3(const double* c);
/* constraints/min_distance.py:25-46: 1 iff no pair i != j is closer than
 * floors[t_i*S + t_j] (minimum image dr - rint(dr C^-1) C, Euclidean norm). */
int or_min_distance_ok(int N, const double* pos, const int32_t* types,
                       const double* cell, const double* floors, int S);
double or_cbrt(double x);
double or_exp(double x);
double or_ipow(double x, int n);
int or_check_cell_shape(const double* cell, int n_atoms, double max_vpa,
                        double min_vpa, double min_aspect);
double or_ensemble_energy(int N, const int32_t* types, const double* cell,
                          double pressure, const double* mu, int S);

/* U (model), H = U + ensemble term, F = -dU/dx (forces of the ensemble term
 * are exactly zero).  F may be NULL for an energy-only evaluation. */
void or_energy_forces(int N, const double* pos, const int32_t* types,
                      const double* cell, const or_pot* pot, double pressure,
                      const double* mu, double* U, double* H, double* F);

/* Fixed-contour walk of every walker (burn-in), see jr_oracle.c. */
void or_initial_walk(or_pop* pop, const or_pot* pot, const or_run* run,
                     const or_move* moves, int n_moves, const double* emax,
                     const int32_t* schedule, int walklength,
                     const double* normals, const double* uniforms,
                     int32_t* counters, int n_threads);

/* Walk one chain for n_steps MWG sub-steps (state updated in place). */
/* dirx: the chain's slot-1 direction (NULL without scoped moves), slot s at
 * dirx + (s - 1) * xstride. */
void or_walk_chain(int N, int Nmax, double* pos, int32_t* types, double* cell,
                   double* energy, double* dir, double* dirx, size_t xstride,
                   const or_pot* pot,
                   double pressure, const double* mu, double emax,
                   const double* step_sizes, const or_move* moves,
                   int n_moves, const int32_t* schedule, int n_steps,
                   const double* normals, const double* uniforms,
                   int32_t* counters);

/* One NS iteration over all R replicas: select (n_cull walkers each),
 * clone, walk, optional replica exchange.  worst, clone, dead_energy and
 * dead_volume of `out` are [R][n_cull]; walk is [R][n_cull + n_extra]. */
int or_ns_iteration(or_pop* pop, const or_pot* pot, const or_run* run,
                    const or_move* moves, int n_moves, int n_steps,
                    int n_extra, int n_cull, int n_swap_cycles,
                    const int32_t* pair_valid, const or_draws* d,
                    or_iter_out* out, int n_threads);

/* Replica exchange pass (EnsembleSwap). */
void or_replica_exchange(or_pop* pop, const or_run* run, const double* emax,
                         int n_swap_cycles, const int32_t* pair_valid,
                         const double* re_u, int32_t* acc, int32_t* att);

/* One trial move per sampled walker, population untouched (step-size
 * adaptation).  accepted/reason: [R*n_samples]. */
void or_trial_moves(const or_pop* pop, const or_pot* pot, const or_run* run,
                    const or_move* moves, int n_moves, int move_idx,
                    const double* trial_step, const double* emax,
                    int n_samples, const int32_t* sample_idx,
                    const double* normals, const double* uniforms,
                    int32_t* accepted, int32_t* reason, int n_threads);

#ifdef __cplusplus
}
#endif
#endif

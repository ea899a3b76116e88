"""log Z of the surrogate-screened walk against the plain walk at the
benchmark's system size for the message-passing judge (matched log Z).  The
screened GMC and volume moves keep the constrained prior exact (delayed
acceptance with symmetric stage-1 verdicts), so the two must agree within
the run-to-run spread.  Runs are engines with their own seeds (the move
schedule is shared inside an engine), compared by Welch's t over the engine
means, |t| < 3 (about 1 % two-sided at these counts).

The 256-atom configuration (its judge, density and cell bounds): 8 engines
per walk of 8 replicas (P = 0.5, no exchange), K = 8, 8 sub-steps
gmc(n_reflect 4):volume 3:1, 100 iterations, each engine from its own
jittered lattice; plain against screened with the surrogate training online
from a seeded network, on a generic r^-12 wall baseline of its own (not the
judge's core term).  Prints a ``LOGZ_RESULT`` JSON line with delta log Z,
sigma and t."""

import numpy as np
import pytest

from jaxrens_amd.backends import EnsembleTerm, Hamiltonian
from jaxrens_amd.engine import NSEngine, RunConfig
from jaxrens_amd.mpnn import MPNNBackend, init_mpnn
from jaxrens_amd.surrogate import LearnedConfig, init_surrogate
from logz_stats import report

pytestmark = pytest.mark.gpu

M, R, K, N_ITER = 8, 8, 8, 100
T_MAX = 3.0
# the surrogate's own wall: eps ((sig/r)^12 - (sig/rc)^12) below rc
GENERIC_WALL = (1.0, 0.8, 1.0)


def _stats_sum(acc, st):
    for k in ("iterations_screened", "iterations_plain", "proposals", "stage1_pass",
              "accepted", "exact_energy_evals"):
        acc[k] = acc.get(k, 0) + int(st[k])
    acc.setdefault("heldout_rmse", []).append(float(st["heldout_rmse"]))


def test_mpnn256_screened_matches_plain_log_z():
    import bench
    N, P = 256, 0.5
    jp = init_mpnn(seed=11)
    moves = [{"type": "galilean", "n_reflect": 4, "step_size": 0.1, "weight": 3},
             {"type": "volume", "step_size": 0.3, "weight": 1}]
    means = {"plain": [], "screened": []}
    acc = {}
    for mode, base in (("plain", 610), ("screened", 620)):
        for m in range(M):
            lc = None
            if mode != "plain":
                lc = LearnedConfig(mode=mode, ring_capacity=8192, train_batch=64,
                                   train_steps=50, min_train=128, min_heldout=16, lr=1e-3,
                                   rmse_bound=0.05, delta_k=2.0, train_max_epa=50.0,
                                   baseline=GENERIC_WALL, params=init_surrogate(1))
            run = RunConfig(n_replicas=R, n_walkers=K, n_atoms=N, n_mcmc_steps=8, n_extra=0,
                            seed=base + m, pressures=[P] * R, re_interval=0)
            eng = NSEngine(Hamiltonian(MPNNBackend(jp), [EnsembleTerm(P)]), moves, run,
                           cell=bench.SUR_CELL, learned=lc)
            eng.set_population(*bench.sur_population(R, K, N, seed=70 + m))
            eng.ns_iterations(N_ITER)
            lz = eng.final_log_evidence()
            means[mode].append(float(lz.mean()))
            if mode != "plain":
                _stats_sum(acc, eng.learned_stats())
            print(f"  N=256 {mode} engine {m}: mean log Z {lz.mean():.3f}, contour median "
                  f"{np.median(eng.emax()):.3f}", flush=True)
            eng.close()
    t = report("mpnn256", means["plain"], means["screened"],
               {"mode": "screened, online training", "baseline": list(GENERIC_WALL),
                "replicas_per_engine": R, "iterations": N_ITER, "n_walkers": K,
                "accepted_per_judge_eval": acc["accepted"] / max(acc["exact_energy_evals"], 1),
                "stage1_pass": acc["stage1_pass"] / max(acc["proposals"], 1)})
    assert acc["iterations_screened"] == M * N_ITER
    assert abs(t) < T_MAX

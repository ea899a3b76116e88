"""Statistical agreement of the device walk with the CPU oracle on
independent draws, on the LJ-8 NPT setting: 8 LJ atoms, P = 0.1, K = 128,
20 sub-steps of galilean (n_reflect 10, step 0.1, weight 4) and volume (step
0.3, weight 1), flat volume prior on [1, 512], 1200 iterations.

16 independent runs on the GPU (device Philox draws, one replica each) and 16
on the oracle (host draws from another generator); log Z and the heat
capacity C_v(T) of ``postprocess/thermodynamics.py`` from each run's dead
points and final live set must agree: |mean_gpu - mean_cpu| below 2 sigma of
the difference of means.  The replicas of one engine (and of one oracle
call) share the move schedule; over 1200 x 20 sub-steps its common share of
volume moves fluctuates by a few tenths of a per cent, well below what these
summaries resolve.  The tape parity tests show the two compute the same
function of the draws; this shows the device draws sample like the host's.
Needs an MI355X."""

import numpy as np
import pytest

from jaxrens_amd import thermodynamics as T
from jaxrens_amd.backends import EnsembleTerm, Hamiltonian, LJBackend
from jaxrens_amd.draws import make_tape
from jaxrens_amd.engine import NSEngine, RunConfig
from jaxrens_amd.run_loop import run_ns
from jaxrens_amd.termination import IterationTermination

pytestmark = pytest.mark.gpu

R, K, N, STEPS, ITERS, P = 16, 128, 8, 20, 1200, 0.1
MOVES = [{"type": "galilean", "n_reflect": 10, "step_size": 0.1, "weight": 4},
         {"type": "volume", "step_size": 0.3, "weight": 1}]
CELL = {"min_volume_per_atom": 1.0 / 8, "max_volume_per_atom": 64.0,
        "flat_V_prior": True}
TEMPS = [0.5, 1.0, 2.0]


def _summaries(emax, live):
    """(log Z, C_v(T)...) per run from dead energies (iters, R) and live (R, K)."""
    out = []
    for r in range(emax.shape[1]):
        out.append([T.log_evidence(emax[:, r], live[r], K)] +
                   list(T.heat_capacity_curve(TEMPS, emax[:, r], live[r], K)))
    return np.array(out)


def _gpu_runs():
    lj = Hamiltonian(LJBackend(1.0, 1.0, 2.5), [EnsembleTerm(P)])
    run = RunConfig(n_replicas=R, n_walkers=K, n_atoms=N, n_mcmc_steps=STEPS,
                    n_extra=0, seed=2024, pressures=[P] * R, re_interval=0)
    eng = NSEngine(lj, MOVES, run, cell=CELL)
    eng.init_walkers(np.random.default_rng(2024), 1.0 / 8, 64.0, True)
    res = run_ns(eng, [IterationTermination(ITERS)], block=50)
    return res.emax, eng.get_population()["energy"]


def _oracle_runs(oracle):
    from oracle.oracle import Population
    from jaxrens_amd.moves import MWG
    rng = np.random.default_rng(777)
    lj = LJBackend(1.0, 1.0, 2.5)
    pot = oracle.Pot.from_table(lj.table)
    vol = rng.uniform(1.0, 512.0, (R, K))
    L = vol ** (1 / 3)
    pos = rng.random((R, K, N, 3)) * L[..., None, None]
    cell = L[..., None, None] * np.eye(3)
    energy = np.empty((R, K))
    for r in range(R):
        for k in range(K):
            energy[r, k] = oracle.energy_forces(pot, pos[r, k], np.zeros(N), cell[r, k],
                                                pressure=P, forces=False)[1]
    pop = Population(pos, np.zeros((R, K, N), np.int32), np.full((R, K), N, np.int32),
                     cell, energy, np.zeros((R, K, N, 3)))
    mw = MWG(MOVES, CELL)
    ss = np.tile(mw.default_step_sizes, (R, 1))
    emax = np.empty((ITERS, R))
    for it in range(ITERS):
        tape = make_tape(rng, 1, R, K, N, STEPS, 0, mw.move_probs)
        o = oracle.ns_iteration(pop, pot, np.full(R, P), np.zeros((R, 1)), ss,
                                mw.as_oracle_moves(), STEPS, 0, tape, do_re=False)
        emax[it] = o["emax"]
    return emax, pop.energy


def test_device_draws_sample_like_the_oracle(oracle):
    print("statistics: LJ-8 NPT GPU vs oracle, independent draws", flush=True)
    g = _summaries(*_gpu_runs())
    c = _summaries(*_oracle_runs(oracle))
    names = ["log Z"] + [f"C_v(T={t})" for t in TEMPS]
    for j, name in enumerate(names):
        mg, mc = g[:, j].mean(), c[:, j].mean()
        se = np.sqrt(g[:, j].var(ddof=1) / R + c[:, j].var(ddof=1) / R)
        z = abs(mg - mc) / se
        print(f"{name}: gpu {mg:.4f} +- {g[:, j].std(ddof=1) / np.sqrt(R):.4f}, "
              f"oracle {mc:.4f} +- {c[:, j].std(ddof=1) / np.sqrt(R):.4f}, z = {z:.2f}",
              flush=True)
        assert z < 2.0, (name, mg, mc, se)

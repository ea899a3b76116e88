"""Parity at the sizes the benchmark measures (bit-for-bit, ``==`` on every
float64 output), on device draws read back with ``last_tape`` and replayed on
the CPU oracle.

* LJ-64 NPT replica stack, P x K = 16 x 1024, n_extra = K - 2 (the whole
  population walked), 20 sub-steps galilean:volume 4:1, RE every iteration:
  16 368 chains per iteration on a resident grid of a few thousand waves, so
  every wave walks several chains from the work queue in turn.
* Binary LJ semi-grand muPT batch, P x K = 20 x 512, N in {32, 48, 64, 96,
  128}, galilean:volume:morph 4:1:1, whole population walked: two size
  classes in concurrent launches, each through its own work queue.

The oracle replays each iteration on all host threads (a few seconds)."""

import numpy as np
import pytest

from jaxrens_amd.backends import LJBackend
from jaxrens_amd.engine import NSEngine, RunConfig

pytestmark = pytest.mark.gpu

GMC_VOL = [{"type": "galilean", "n_reflect": 10, "step_size": 0.1, "weight": 4},
           {"type": "volume", "step_size": 0.3, "weight": 1}]


def _replay(oracle, eng, launches, pair_valid=None):
    """One NS iteration per launch shape (waves per block, occupancy variant)
    on device draws, each replayed on the oracle from the same state."""
    from oracle.oracle import Population
    opop = Population.from_engine_dict(eng.get_population())
    pot = oracle.Pot.from_table(eng.model.table)
    moves = eng.mwg.as_oracle_moves()
    for wpb, occ in launches:
        eng.set_launch(wpb, occ)
        info = eng.launch_info()
        chains = eng.n_replicas * eng.n_walk
        print(f"  iteration {eng.iteration}: launch wpb={wpb} occ={occ} "
              f"blocks={info['blocks']} chains={chains}", flush=True)
        out = eng.ns_iterations(1)
        tape = eng.last_tape()
        i = eng.iteration - 1
        fires = eng.run.re_interval > 0 and i > 0 and i % eng.run.re_interval == 0
        o = oracle.ns_iteration(opop, pot, eng.pressure, eng.mu, eng.step_sizes, moves,
                                eng.run.n_mcmc_steps, eng.run.n_extra, tape, it=0,
                                n_swap_cycles=eng.run.n_swap_cycles, pair_valid=pair_valid,
                                do_re=fires)
        for k in ("emax", "worst", "clone", "walk", "dead_volume", "counters"):
            assert np.array_equal(out[k][0], o[k]), (k, i)
        if eng.n_replicas > 1:
            assert np.array_equal(out["re_acc"][0], o["re_acc"]), i
            assert np.array_equal(out["re_att"][0], o["re_att"]), i
        d = eng.get_population()
        assert np.array_equal(d["energy"], opop.energy), i
        assert np.array_equal(d["positions"], opop.pos), i
        assert np.array_equal(d["direction"], opop.dir), i
        assert np.array_equal(d["types"], opop.types), i
        assert np.array_equal(d["cell"].reshape(opop.cell.shape), opop.cell), i
        assert np.array_equal(d["n_atoms"], opop.n_atoms), i
        acc = out["counters"][0][:, :, 1].sum()
        print(f"  iteration {i}: {acc} of {out['counters'][0][:, :, 0].sum()} moves "
              f"accepted, RE {out['re_acc'][0].sum()}/{out['re_att'][0].sum()}", flush=True)
        if fires:
            assert out["re_att"][0].sum() > 0


def test_lj64_16x1024_whole_population(oracle):
    """The headline workload (as bench.py builds it), three iterations: the
    tuned launch shape (two waves per block, three waves per SIMD) twice,
    then the default one; the second and third fire RE."""
    print("full-size parity: LJ-64 NPT 16x1024, n_extra=1022", flush=True)
    import bench
    eng = bench.build_engine(n_extra=1022)
    assert eng.n_replicas * eng.n_walk == 16368
    _replay(oracle, eng, [(2, 3), (2, 3), (4, 0)])
    eng.close()


def test_semigrand_20x512_whole_population(oracle):
    """The binary semi-grand batch (as bench.py builds it), two iterations,
    the second with RE between neighbours of equal size."""
    print("full-size parity: binary semi-grand 20x512, N in {32..128}", flush=True)
    R, K = 20, 512
    sizes = [n for n in (32, 48, 64, 96, 128) for _ in range(4)]
    eps = np.array([[1.0, 1.5], [1.5, 0.5]])
    sig = np.array([[1.0, 0.8], [0.8, 0.88]])
    mu = np.stack([np.zeros(R), np.linspace(-2.0, 2.0, R)], axis=1)
    run = RunConfig(n_replicas=R, n_walkers=K, n_atoms=sizes, n_mcmc_steps=20,
                    n_extra=K - 2, seed=42, pressures=[1.0] * R, chemical_potentials=mu,
                    re_interval=1)
    moves = GMC_VOL + [{"type": "alchemical_morph", "n_species": 2, "weight": 1}]
    eng = NSEngine(LJBackend(pair_epsilon=eps, pair_sigma=sig, cutoff=2.5), moves, run,
                   cell={"min_volume_per_atom": 0.5, "max_volume_per_atom": 8.0,
                         "flat_V_prior": True})
    eng.init_walkers(np.random.default_rng(42), 1.0, 8.0, True, species_p=[0.5, 0.5])
    pv = np.array([sizes[i] == sizes[i + 1] for i in range(R - 1)], np.int32)
    _replay(oracle, eng, [(2, 3), (4, 0)], pair_valid=pv)
    eng.close()

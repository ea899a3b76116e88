"""Minimum-distance constraint on the device (``constraints/min_distance.py``,
gated as in ``sampling/mwg.py:98-141``): whole NS iterations bit for bit
against the CPU oracle on host tapes, with floors high enough that the gate
fires, for atom/cell moves and for species moves.  Needs an MI355X."""

import numpy as np
import pytest

from jaxrens_amd.backends import LJBackend
from jaxrens_amd.draws import make_tape
from jaxrens_amd.engine import NSEngine, RunConfig

pytestmark = pytest.mark.gpu


def _parity(oracle, eng, n_iter, seed):
    from oracle.oracle import Population
    R, K, Nm = eng.n_replicas, eng.n_walkers, eng.Nmax
    pot = oracle.Pot.from_table(eng.model.table).with_min_distance(
        eng.min_distance, eng.min_distance_types)
    moves = eng.mwg.as_oracle_moves()
    opop = Population.from_engine_dict(eng.get_population())
    tape = make_tape(np.random.default_rng(seed), n_iter, R, K, Nm, eng.run.n_mcmc_steps,
                     eng.run.n_extra, eng.mwg.move_probs)
    gated = 0
    for it in range(n_iter):
        out = eng.ns_iterations(1, tape=tape.iteration(it))
        fires = (eng.iteration - 1) > 0 and eng.run.re_interval > 0
        o = oracle.ns_iteration(opop, pot, eng.pressure, eng.mu, eng.step_sizes, moves,
                                eng.run.n_mcmc_steps, eng.run.n_extra, tape, it=it,
                                do_re=fires and R > 1)
        assert np.array_equal(out["emax"][0], o["emax"]), it
        assert np.array_equal(out["counters"][0], o["counters"]), it
        d = eng.get_population()
        for k, v in (("energy", opop.energy), ("positions", opop.pos), ("types", opop.types),
                     ("direction", opop.dir)):
            assert np.array_equal(d[k], v), (it, k)
        assert np.array_equal(d["cell"].reshape(opop.cell.shape), opop.cell), it
        c = out["counters"][0]
        gated += int((c[..., 0] - c[..., 1:5].sum(-1)).sum())  # proposals in no bucket
    return gated


def test_min_distance_gate_atom_moves(oracle):
    print("min-distance gate: gmc + volume + hmc", flush=True)
    moves = [{"type": "galilean", "n_reflect": 4, "step_size": 0.15, "weight": 3},
             {"type": "volume", "step_size": 0.3, "weight": 1},
             {"type": "hmc", "n_leapfrog": 3, "step_size": 0.01, "weight": 1}]
    run = RunConfig(n_replicas=2, n_walkers=12, n_atoms=20, n_mcmc_steps=8, n_extra=6,
                    seed=4, pressures=[0.5, 2.0], re_interval=1)
    eng = NSEngine(LJBackend(1.0, 1.0, 2.5), moves, run,
                   cell={"min_volume_per_atom": 0.5, "max_volume_per_atom": 20.0,
                         "flat_V_prior": True},
                   constraints=[{"type": "minimum_distance", "d_min": 0.95}])
    eng.init_walkers(np.random.default_rng(4), 1.5, 6.0, True)
    assert eng.min_distance_types is None
    assert _parity(oracle, eng, 4, 21) > 0


@pytest.mark.parametrize("N", [40, 96])
def test_min_distance_gate_species_moves(oracle, N):
    print(f"min-distance gate: semi-grand N={N}", flush=True)
    eps = np.array([[1.0, 1.5], [1.5, 0.5]])
    sig = np.array([[1.0, 0.8], [0.8, 0.88]])
    moves = [{"type": "galilean", "n_reflect": 3, "step_size": 0.1, "weight": 2},
             {"type": "alchemical_morph", "n_species": 2, "weight": 2},
             {"type": "species_swap", "weight": 1},
             {"type": "shear", "step_size": 0.05, "weight": 1}]
    R = 2
    mu = np.stack([np.zeros(R), np.linspace(-1, 1, R)], axis=1)
    run = RunConfig(n_replicas=R, n_walkers=8, n_atoms=N, n_mcmc_steps=8, n_extra=4,
                    seed=6, pressures=[1.0] * R, chemical_potentials=mu, re_interval=1)
    eng = NSEngine(LJBackend(pair_epsilon=eps, pair_sigma=sig, cutoff=2.5), moves, run,
                   cell={"min_volume_per_atom": 0.5, "max_volume_per_atom": 8.0,
                         "min_aspect_ratio": 0.6, "flat_V_prior": True},
                   constraints=[{"type": "minimum_distance",
                                 "d_min": {"default": 0.7, "A-A": 0.95, "A-B": 0.8}}],
                   symbol_map={0: "A", 1: "B"})
    eng.init_walkers(np.random.default_rng(6), 1.2, 6.0, True, species_p=[0.5, 0.5])
    assert eng.min_distance_types is not None
    assert _parity(oracle, eng, 4, 33) > 0


def _gated_engine():
    moves = [{"type": "galilean", "n_reflect": 4, "step_size": 0.15, "weight": 3},
             {"type": "volume", "step_size": 0.3, "weight": 1}]
    run = RunConfig(n_replicas=2, n_walkers=16, n_atoms=24, n_mcmc_steps=4, n_extra=4,
                    seed=9, pressures=[0.5, 2.0], re_interval=1)
    eng = NSEngine(LJBackend(1.0, 1.0, 2.5), moves, run,
                   cell={"min_volume_per_atom": 0.5, "max_volume_per_atom": 20.0,
                         "flat_V_prior": True},
                   constraints=[{"type": "minimum_distance", "d_min": 1.0}])
    eng.init_walkers(np.random.default_rng(9), 1.5, 6.0, True)
    return eng


def test_min_distance_gate_trial_moves(oracle):
    """The adaptation's trial moves report reason 4 where the gate undoes
    a move, as the oracle does."""
    print("min-distance gate: trial moves", flush=True)
    from oracle.oracle import Population
    eng = _gated_engine()
    eng.ns_iterations(2)
    rng = np.random.default_rng(2)
    idx = rng.integers(0, 16, size=(2, 40)).astype(np.int32)
    nrm = rng.standard_normal((2, 40, 3, 24))
    uni = rng.random((2, 40, 4))
    pop = Population.from_engine_dict(eng.get_population())
    pot = oracle.Pot.from_table(eng.model.table).with_min_distance(eng.min_distance)
    seen = set()
    for m, ss in [(0, 0.3), (1, 0.8)]:
        acc, rs = eng.trial_moves(m, [ss, ss * 2], idx, normals=nrm, uniforms=uni)
        oacc, ors = oracle.trial_moves(pop, pot, eng.pressure, eng.mu, eng.step_sizes,
                                       eng.mwg.as_oracle_moves(), m, [ss, ss * 2], eng.emax(),
                                       idx, nrm, uni)
        assert np.array_equal(acc, oacc) and np.array_equal(rs, ors), m
        seen |= set(np.unique(rs).tolist())
    assert 4 in seen


def test_min_distance_gate_initial_walk(oracle):
    print("min-distance gate: burn-in walk", flush=True)
    from oracle.oracle import Population
    from jaxrens_amd.burn_in import burn_in_emax
    eng = _gated_engine()
    pot = oracle.Pot.from_table(eng.model.table).with_min_distance(eng.min_distance)
    opop = Population.from_engine_dict(eng.get_population())
    emax = burn_in_emax(eng, 1.0)
    rng = np.random.default_rng(5)
    R, K, Nm, wl = 2, 16, 24, 6
    sched = rng.choice(2, size=wl, p=eng.mwg.move_probs).astype(np.int32)
    nn = rng.standard_normal((R, K, wl, 3, Nm))
    uu = rng.random((R, K, wl, 4))
    ctr = eng.initial_walk(wl, emax, sched, nn, uu)
    octr = oracle.initial_walk(opop, pot, eng.pressure, eng.mu, eng.step_sizes,
                               eng.mwg.as_oracle_moves(), emax, sched, nn, uu)
    assert np.array_equal(ctr, octr)
    d = eng.get_population()
    assert np.array_equal(d["positions"], opop.pos) and np.array_equal(d["energy"], opop.energy)
    assert (ctr[..., 0] - ctr[..., 1:5].sum(-1)).sum() > 0  # the gate fired


def test_grid_initial_walkers_satisfy_the_floor(oracle):
    """Walkers placed on the grid spaced by the floor start valid, then walk
    gated (the invariant the reference's gate relies on)."""
    print("min-distance gate: grid initial walkers", flush=True)
    import ctypes as C
    from oracle.oracle import _d, _i, f64, i32
    moves = [{"type": "galilean", "n_reflect": 4, "step_size": 0.15, "weight": 3},
             {"type": "volume", "step_size": 0.3, "weight": 1}]
    run = RunConfig(n_replicas=2, n_walkers=16, n_atoms=27, n_mcmc_steps=4, n_extra=4,
                    seed=12, pressures=[0.5, 2.0], re_interval=1)
    eng = NSEngine(LJBackend(1.0, 1.0, 2.5), moves, run,
                   cell={"min_volume_per_atom": 1.0, "max_volume_per_atom": 20.0,
                         "flat_V_prior": True},
                   constraints=[{"type": "minimum_distance", "d_min": 0.9}])
    eng.init_walkers(np.random.default_rng(12), 1.0, 6.0, True, grid_distance=0.9)
    L = oracle.lib()
    L.or_min_distance_ok.restype = C.c_int
    d = eng.get_population()
    fl = f64(eng.min_distance)
    for r in range(2):
        for k in range(16):
            p, t = f64(d["positions"][r, k]), i32(d["types"][r, k])
            c = f64(np.asarray(d["cell"][r, k]).reshape(9))
            assert L.or_min_distance_ok(27, _d(p), _i(t), _d(c), _d(fl), 1) == 1, (r, k)
    out = eng.ns_iterations(3)
    assert np.all(np.isfinite(out["emax"]))

"""HIP path vs the CPU oracle, bit for bit, for the supercell image sums
(``backends/lj.py:106-113,152-197``) and the volume-preserving cell moves
shear and stretch (``sampling/moves/shear.py``, ``stretch.py``).  Needs an
MI355X."""

import numpy as np
import pytest

from jaxrens_amd.backends import EnsembleTerm, Hamiltonian, LJBackend
from jaxrens_amd.draws import make_tape
from jaxrens_amd.engine import NSEngine, RunConfig

pytestmark = pytest.mark.gpu

TUTORIAL_MOVES = [{"type": "gmc", "n_reflect": 6, "step_size": 0.1, "weight": 4.0},
                  {"type": "volume", "step_size": 0.2, "weight": 1.0},
                  {"type": "shear", "step_size": 0.08, "weight": 1.0},
                  {"type": "stretch", "step_size": 0.08, "weight": 1.0}]
TUTORIAL_CELL = {"max_volume_per_atom": 30.0, "min_volume_per_atom": 1.2,
                 "min_aspect_ratio": 0.7, "flat_V_prior": True}


def _cfgs(rng, B, N, L, skew):
    pos = rng.random((B, N, 3)) * L + rng.integers(-2, 3, size=(B, N, 3)) * L
    cell = np.tile(np.eye(3) * L, (B, 1, 1)) + rng.uniform(-skew, skew, (B, 3, 3)) * L
    return pos, cell


@pytest.mark.parametrize("N", [1, 2, 5, 8, 33, 64, 65, 128])
@pytest.mark.parametrize("sc", [(2, 2, 2), (4, 4, 4), (3, 1, 5)])
def test_eval_with_images(oracle, N, sc):
    print(f"eval images N={N} sc={sc}", flush=True)
    rng = np.random.default_rng(N * 31 + sc[0])
    lj = LJBackend(1.0, 1.0, 2.5, supercell_trafo=sc)
    L = max(2.0, (N / 0.9) ** (1 / 3))
    for skew in (0.0, 0.2):
        pos, cell = _cfgs(rng, 4, N, L, skew)
        ty = np.zeros((4, N), np.int32)
        U, H, F = lj._eval(pos, ty, cell, pressure=0.1, forces=True)
        pot = oracle.Pot.from_table(lj.table)
        for b in range(4):
            u, h, f = oracle.energy_forces(pot, pos[b], ty[b], cell[b], pressure=0.1)
            assert U[b] == u and H[b] == h, (skew, b, U[b], u)
            assert np.array_equal(F[b], f), (skew, b, np.abs(F[b] - f).max())


@pytest.mark.parametrize("N", [7, 64, 96])
def test_eval_with_images_pair_table(oracle, N):
    print(f"eval images pair table N={N}", flush=True)
    rng = np.random.default_rng(N)
    eps = np.array([[1.0, 1.5], [1.5, 0.5]])
    sig = np.array([[1.0, 0.8], [0.8, 0.88]])
    lj = LJBackend(pair_epsilon=eps, pair_sigma=sig, cutoff=2.5, supercell_trafo=(2, 4, 2))
    pos, cell = _cfgs(rng, 3, N, max(2.2, (N / 0.9) ** (1 / 3)), 0.15)
    ty = rng.integers(0, 2, (3, N))
    mu = rng.normal(size=(3, 2))
    U, H, F = lj._eval(pos, ty, cell, pressure=1.0, mu=mu, forces=True)
    pot = oracle.Pot.from_table(lj.table)
    for b in range(3):
        u, h, f = oracle.energy_forces(pot, pos[b], ty[b], cell[b], pressure=1.0, mu=mu[b])
        assert U[b] == u and H[b] == h
        assert np.array_equal(F[b], f)


def _engine(lj, moves, R, K, N, n_extra, n_steps, cell, v_range, pressures, seed):
    run = RunConfig(n_replicas=R, n_walkers=K, n_atoms=N, n_mcmc_steps=n_steps,
                    n_extra=n_extra, seed=seed, pressures=pressures, re_interval=1)
    eng = NSEngine(lj, moves, run, cell=cell)
    eng.init_walkers(np.random.default_rng(seed), v_range[0], v_range[1], True)
    return eng


def _run_parity(oracle, eng, n_iter, tape):
    from oracle.oracle import Population
    opop = Population.from_engine_dict(eng.get_population())
    pot = oracle.Pot.from_table(eng.model.table)
    moves = eng.mwg.as_oracle_moves()
    sheared, total = 0, 0
    for it in range(n_iter):
        out = eng.ns_iterations(1, tape=tape.iteration(it))
        fires = (eng.iteration - 1) > 0 and eng.run.re_interval > 0
        o = oracle.ns_iteration(opop, pot, eng.pressure, eng.mu, eng.step_sizes, moves,
                                eng.run.n_mcmc_steps, eng.run.n_extra, tape, it=it,
                                n_swap_cycles=eng.run.n_swap_cycles, do_re=fires)
        assert np.array_equal(out["emax"][0], o["emax"]), it
        assert np.array_equal(out["counters"][0], o["counters"]), it
        total = total + o["counters"]
        d = eng.get_population()
        assert np.array_equal(d["energy"], opop.energy), it
        assert np.array_equal(d["positions"], opop.pos), it
        assert np.array_equal(d["cell"].reshape(opop.cell.shape), opop.cell), it
        assert np.array_equal(d["direction"], opop.dir), it
        c = d["cell"].reshape(-1, 9)
        sheared = int(np.sum(np.any(c[:, [1, 2, 3, 5, 6, 7]] != 0.0, axis=1)))
    return sheared, total


def test_ns_iterations_tutorial_npt_with_images(oracle):
    """The LJ-8 NPT tutorial setting: gmc + volume + shear + stretch, cell
    prior [1.2, 30] per atom, supercell (4, 4, 4) images."""
    print("ns tutorial lj8 images", flush=True)
    lj = Hamiltonian(LJBackend(1.0, 1.0, 2.5, supercell_trafo=(4, 4, 4)), [EnsembleTerm(0.1)])
    eng = _engine(lj, TUTORIAL_MOVES, 2, 16, 8, 8, 20, TUTORIAL_CELL, (1.2, 30.0),
                  [0.1, 1.0], 21)
    tape = make_tape(np.random.default_rng(22), 5, 2, 16, 8, 20, 8, eng.mwg.move_probs)
    sheared, ctr = _run_parity(oracle, eng, 5, tape)
    assert sheared > 0
    # shear and stretch proposals were accepted
    assert ctr[:, 2, 1].sum() > 0 and ctr[:, 3, 1].sum() > 0


@pytest.mark.parametrize("N,sc", [(64, (1, 1, 1)), (96, (1, 1, 1)), (40, (2, 2, 2))])
def test_ns_iterations_cell_moves(oracle, N, sc):
    print(f"ns cell moves N={N} sc={sc}", flush=True)
    lj = LJBackend(1.0, 1.0, 2.5, supercell_trafo=sc)
    moves = [{"type": "gmc", "n_reflect": 4, "step_size": 0.1, "weight": 2.0},
             {"type": "shear", "step_size": 0.1, "weight": 1.0},
             {"type": "stretch", "step_size": 0.1, "weight": 1.0}]
    cell = {"max_volume_per_atom": 64.0, "min_volume_per_atom": 0.5,
            "min_aspect_ratio": 0.6, "flat_V_prior": True}
    eng = _engine(lj, moves, 2, 6, N, 3, 8, cell, (1.0, 8.0), [0.05, 0.5], N)
    tape = make_tape(np.random.default_rng(N + 1), 3, 2, 6, N, 8, 3, eng.mwg.move_probs)
    sheared, _ = _run_parity(oracle, eng, 3, tape)
    assert sheared > 0


def test_trial_moves_shear_stretch(oracle):
    print("trial shear/stretch with images", flush=True)
    from oracle.oracle import Population
    lj = LJBackend(1.0, 1.0, 2.5, supercell_trafo=(4, 4, 4))
    eng = _engine(lj, TUTORIAL_MOVES, 2, 16, 8, 4, 6, TUTORIAL_CELL, (1.2, 30.0),
                  [0.1, 1.0], 5)
    eng.ns_iterations(2)
    rng = np.random.default_rng(3)
    idx = rng.integers(0, 16, size=(2, 30)).astype(np.int32)
    nrm = rng.standard_normal((2, 30, 3, 8))
    uni = rng.random((2, 30, 4))
    emax = eng.emax()
    pop = Population.from_engine_dict(eng.get_population())
    pot = oracle.Pot.from_table(lj.table)
    for m, ss in [(2, 0.2), (3, 0.2)]:
        acc, rs = eng.trial_moves(m, [ss, ss * 2], idx, normals=nrm, uniforms=uni)
        oacc, ors = oracle.trial_moves(pop, pot, eng.pressure, eng.mu, eng.step_sizes,
                                       eng.mwg.as_oracle_moves(), m, [ss, ss * 2], emax,
                                       idx, nrm, uni)
        assert np.array_equal(acc, oacc) and np.array_equal(rs, ors), m
        assert acc.any()


def test_ns_iterations_species_swap_and_morph(oracle):
    """species_swap next to alchemical_morph in a binary pair-table system
    with mixed sizes (N = 32 and 96, two atoms per lane), bit-exact."""
    print("ns species swap + morph", flush=True)
    from oracle.oracle import Population
    eps = np.array([[1.0, 1.5], [1.5, 0.5]])
    sig = np.array([[1.0, 0.8], [0.8, 0.88]])
    lj = LJBackend(pair_epsilon=eps, pair_sigma=sig, cutoff=2.5)
    moves = [{"type": "gmc", "n_reflect": 4, "step_size": 0.1, "weight": 3},
             {"type": "volume", "step_size": 0.3, "weight": 1},
             {"type": "species_swap", "weight": 2},
             {"type": "alchemical_morph", "n_species": 2, "weight": 1}]
    R, K = 4, 6
    mu = np.stack([np.zeros(R), np.linspace(-1, 1, R)], axis=1)
    run = RunConfig(n_replicas=R, n_walkers=K, n_atoms=[32, 32, 96, 96], n_mcmc_steps=8,
                    n_extra=3, seed=9, pressures=[1.0] * R, chemical_potentials=mu,
                    re_interval=0)
    eng = NSEngine(lj, moves, run, cell={"min_volume_per_atom": 0.5,
                                         "max_volume_per_atom": 8.0, "flat_V_prior": True})
    eng.init_walkers(np.random.default_rng(9), 0.9, 3.0, True, species_p=[0.5, 0.5])
    tape = make_tape(np.random.default_rng(19), 4, R, K, 96, 8, 3, eng.mwg.move_probs)
    opop = Population.from_engine_dict(eng.get_population())
    pot = oracle.Pot.from_table(lj.table)
    total = 0
    for it in range(4):
        out = eng.ns_iterations(1, tape=tape.iteration(it))
        o = oracle.ns_iteration(opop, pot, eng.pressure, eng.mu, eng.step_sizes,
                                eng.mwg.as_oracle_moves(), 8, 3, tape, it=it, do_re=False)
        assert np.array_equal(out["emax"][0], o["emax"]), it
        assert np.array_equal(out["counters"][0], o["counters"]), it
        total = total + o["counters"]
        d = eng.get_population()
        assert np.array_equal(d["types"], opop.types), it
        assert np.array_equal(d["energy"], opop.energy), it
        assert np.array_equal(d["positions"], opop.pos), it
    assert total[:, 2, 1].sum() > 0 and total[:, 3, 1].sum() > 0  # swaps, morphs accepted

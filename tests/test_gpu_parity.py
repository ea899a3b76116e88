"""HIP path vs the CPU oracle on identical inputs (bit-for-bit: compared
with ``==`` on every float64 output; the two follow one canonical
floating-point order).  Needs an MI355X.

``cap = 1`` runs a case on one persistent block (``set_walk_block_cap``), so
that each wave takes several chains from the work queue in turn and reuses
its LDS, counters and registers between them: the path the full-size
workloads take."""

import math

import numpy as np
import pytest

from jaxrens_amd import _lib
from jaxrens_amd.backends import EnsembleTerm, Hamiltonian, LJBackend
from jaxrens_amd.draws import make_tape
from jaxrens_amd.engine import NSEngine, RunConfig

pytestmark = pytest.mark.gpu


def _random_cfgs(rng, B, N, triclinic, n_species=1, dense=0.6):
    L = (N / dense) ** (1 / 3)
    pos = rng.random((B, N, 3)) * L
    pos += rng.integers(-2, 3, size=(B, N, 3)) * L
    cell = np.tile(np.eye(3) * L, (B, 1, 1))
    if triclinic:
        cell = cell + rng.uniform(-0.15, 0.15, size=(B, 3, 3)) * L
    ty = rng.integers(0, n_species, size=(B, N))
    return pos, ty, cell


@pytest.mark.parametrize("N", [1, 2, 3, 7, 8, 13, 32, 33, 63, 64, 65, 96, 127, 128, 129, 200, 256])
@pytest.mark.parametrize("triclinic", [False, True])
def test_eval_single_species(oracle, N, triclinic):
    print(f"eval N={N} triclinic={triclinic}", flush=True)
    rng = np.random.default_rng(N * 7 + triclinic)
    lj = LJBackend(1.0, 1.0, 2.5)
    pos, ty, cell = _random_cfgs(rng, 6, N, triclinic)
    U, H, F = lj._eval(pos, ty, cell, pressure=np.linspace(0.01, 1.0, 6), forces=True)
    pot = oracle.Pot.from_table(lj.table)
    for b in range(6):
        u, h, f = oracle.energy_forces(pot, pos[b], ty[b], cell[b],
                                       pressure=np.linspace(0.01, 1.0, 6)[b])
        assert U[b] == u and H[b] == h, (b, U[b], u)
        assert np.array_equal(F[b], f), (b, np.abs(F[b] - f).max())


@pytest.mark.parametrize("N", [8, 48, 64, 96, 128])
def test_eval_pair_table_and_mu(oracle, N):
    print(f"eval pair-table N={N}", flush=True)
    rng = np.random.default_rng(N)
    eps = np.array([[1.0, 1.5], [1.5, 0.5]])
    sig = np.array([[1.0, 0.8], [0.8, 0.88]])
    lj = LJBackend(pair_epsilon=eps, pair_sigma=sig, cutoff=2.5)
    pos, ty, cell = _random_cfgs(rng, 5, N, True, n_species=2, dense=0.8)
    mu = rng.normal(size=(5, 2))
    U, H, F = lj._eval(pos, ty, cell, pressure=1.0, mu=mu, forces=True)
    pot = oracle.Pot.from_table(lj.table)
    for b in range(5):
        u, h, f = oracle.energy_forces(pot, pos[b], ty[b], cell[b], pressure=1.0, mu=mu[b])
        assert U[b] == u and H[b] == h
        assert np.array_equal(F[b], f)


def test_hamiltonian_api_matches_terms():
    rng = np.random.default_rng(1)
    lj = LJBackend(1.0, 1.0, 2.5)
    H = Hamiltonian(lj, [EnsembleTerm(pressure=0.3)])
    pos, ty, cell = _random_cfgs(rng, 1, 20, False)
    r = H.energy_and_forces(pos[0], ty[0], cell[0])
    u = lj.energy_and_forces(pos[0], ty[0], cell[0])
    V = abs(np.linalg.det(cell[0]))
    assert r.energy == pytest.approx(u.energy + 0.3 * V, rel=1e-14)
    assert np.array_equal(r.forces, u.forces)


def _engine(lj, moves, R, K, N, n_extra, n_steps, seed=3, cell_spec=None,
            pressures=None, mu=None, re_interval=1, species_p=None,
            v_range=(0.125, 64.0)):
    run = RunConfig(n_replicas=R, n_walkers=K, n_atoms=N, n_mcmc_steps=n_steps,
                    n_extra=n_extra, seed=seed, pressures=pressures,
                    chemical_potentials=mu, re_interval=re_interval)
    eng = NSEngine(lj, moves, run, cell=cell_spec)
    rng = np.random.default_rng(seed)
    eng.init_walkers(rng, v_range[0], v_range[1], True, species_p=species_p)
    return eng


GMC_VOL = [{"type": "galilean", "n_reflect": 10, "step_size": 0.1, "weight": 4},
           {"type": "volume", "step_size": 0.3, "weight": 1}]
CELL = {"min_volume_per_atom": 0.125, "max_volume_per_atom": 64.0, "flat_V_prior": True}


def _compare_population(eng, opop):
    d = eng.get_population()
    assert np.array_equal(d["energy"], opop.energy)
    assert np.array_equal(d["positions"], opop.pos)
    assert np.array_equal(d["direction"], opop.dir)
    assert np.array_equal(d["types"], opop.types)
    assert np.array_equal(d["cell"].reshape(opop.cell.shape), opop.cell)
    assert np.array_equal(d["n_atoms"], opop.n_atoms)
    if "direction_slots" in d:
        assert np.array_equal(d["direction_slots"], opop.dirx)


def _run_parity(oracle, eng, n_iter, tape, pair_valid=None):
    from oracle.oracle import Population
    opop = Population.from_engine_dict(eng.get_population())
    pot = oracle.Pot.from_table(eng.model.table)
    moves = eng.mwg.as_oracle_moves()
    for it in range(n_iter):
        eng_out = eng.ns_iterations(1, tape=tape.iteration(it))
        fires = eng.run.re_interval > 0 and (eng.iteration - 1) > 0 and \
            (eng.iteration - 1) % eng.run.re_interval == 0
        o = oracle.ns_iteration(opop, pot, eng.pressure, eng.mu, eng.step_sizes, moves,
                                eng.run.n_mcmc_steps, eng.run.n_extra, tape, it=it,
                                n_swap_cycles=eng.run.n_swap_cycles,
                                pair_valid=pair_valid, do_re=fires)
        assert np.array_equal(eng_out["emax"][0], o["emax"]), it
        assert np.array_equal(eng_out["worst"][0], o["worst"])
        assert np.array_equal(eng_out["clone"][0], o["clone"])
        assert np.array_equal(eng_out["walk"][0], o["walk"])
        assert np.array_equal(eng_out["dead_volume"][0], o["dead_volume"])
        assert np.array_equal(eng_out["counters"][0], o["counters"]), it
        if eng.n_replicas > 1:
            assert np.array_equal(eng_out["re_acc"][0], o["re_acc"])
            assert np.array_equal(eng_out["re_att"][0], o["re_att"])
        _compare_population(eng, opop)


@pytest.mark.parametrize("cap", [0, 1])
@pytest.mark.parametrize("n_extra", [0, 14])
def test_ns_iterations_lj8(oracle, n_extra, cap):
    print(f"ns lj8 n_extra={n_extra} cap={cap}", flush=True)
    lj = Hamiltonian(LJBackend(1.0, 1.0, 2.5), [EnsembleTerm(0.1)])
    eng = _engine(lj, GMC_VOL, 1, 16, 8, n_extra, 20, cell_spec=CELL)
    eng.set_walk_block_cap(cap)
    tape = make_tape(np.random.default_rng(9), 6, 1, 16, 8, 20, n_extra,
                     eng.mwg.move_probs, ties=True)
    _run_parity(oracle, eng, 6, tape)


@pytest.mark.parametrize("cap", [0, 1])
def test_ns_iterations_lj64_replicas_with_re(oracle, cap):
    print(f"ns lj64 R=4 with RE cap={cap}", flush=True)
    lj = LJBackend(1.0, 1.0, 2.5)
    P = list(10 ** (-2 + 3 * np.arange(4) / 3))
    eng = _engine(lj, GMC_VOL, 4, 8, 64, 6, 6, cell_spec=CELL, pressures=P)
    eng.set_walk_block_cap(cap)
    tape = make_tape(np.random.default_rng(10), 4, 4, 8, 64, 6, 6, eng.mwg.move_probs)
    _run_parity(oracle, eng, 4, tape)


@pytest.mark.parametrize("cap", [0, 1])
def test_ns_iterations_hmc_gmc_n128(oracle, cap):
    print(f"ns hmc+gmc N=128 cap={cap}", flush=True)
    lj = LJBackend(1.0, 1.0, 2.5)
    moves = [{"type": "hmc", "n_leapfrog": 4, "step_size": 0.005, "weight": 1},
             {"type": "gmc", "n_reflect": 3, "step_size": 0.05, "weight": 1}]
    eng = _engine(lj, moves, 2, 4, 128, 2, 4, pressures=[0.05, 0.5],
                  cell_spec={"min_volume_per_atom": 0.5, "max_volume_per_atom": 4.0},
                  v_range=(1.2, 3.0))
    eng.set_walk_block_cap(cap)
    tape = make_tape(np.random.default_rng(11), 3, 2, 4, 128, 4, 2, eng.mwg.move_probs)
    _run_parity(oracle, eng, 3, tape)


@pytest.mark.parametrize("occ,cap", [(0, 0), (3, 0), (3, 1)])
def test_ns_iterations_semigrand_mixed_n(oracle, occ, cap):
    """Two size classes (N <= 64 and 65..128), each in its own launch, with
    the N <= 64 class on its default and its three-waves-per-SIMD build."""
    print(f"ns semi-grand mixed N occupancy={occ} cap={cap}", flush=True)
    eps = np.array([[1.0, 1.5], [1.5, 0.5]])
    sig = np.array([[1.0, 0.8], [0.8, 0.88]])
    lj = LJBackend(pair_epsilon=eps, pair_sigma=sig, cutoff=2.5)
    moves = [{"type": "gmc", "n_reflect": 4, "step_size": 0.1, "weight": 4},
             {"type": "volume", "step_size": 0.3, "weight": 1},
             {"type": "alchemical_morph", "n_species": 2, "weight": 1}]
    R = 6
    n_atoms = [32, 32, 48, 48, 96, 96]
    mu = np.stack([np.zeros(R), np.linspace(-2, 2, R)], axis=1)
    run = RunConfig(n_replicas=R, n_walkers=6, n_atoms=n_atoms, n_mcmc_steps=6,
                    n_extra=3, seed=5, pressures=[1.0] * R, chemical_potentials=mu,
                    re_interval=1)
    eng = NSEngine(lj, moves, run, cell=CELL)
    eng.init_walkers(np.random.default_rng(5), 0.5, 8.0, True, species_p=[0.5, 0.5])
    eng.set_launch(2, occ)
    eng.set_walk_block_cap(cap)
    tape = make_tape(np.random.default_rng(12), 4, R, 6, 96, 6, 3, eng.mwg.move_probs)
    pv = np.array([n_atoms[i] == n_atoms[i + 1] for i in range(R - 1)], np.int32)
    _run_parity(oracle, eng, 4, tape, pair_valid=pv)


@pytest.mark.parametrize("cap", [0, 1])
def test_ns_iterations_odd_and_four_per_lane(oracle, cap):
    """Interleaved slots on odd rings (N = 97 under two atoms per lane, the
    wrap in slot 0) and under four atoms per lane (N = 150 partial, N = 256
    full), species table with mu, gmc + volume + morph."""
    print(f"ns N = 97, 150, 256 cap={cap}", flush=True)
    lj = LJBackend(**BINARY)
    moves = [{"type": "gmc", "n_reflect": 3, "step_size": 0.05, "weight": 3},
             {"type": "volume", "step_size": 0.2, "weight": 1},
             {"type": "alchemical_morph", "n_species": 2, "weight": 1}]
    R = 4
    n_atoms = [97, 97, 150, 256]
    mu = np.stack([np.zeros(R), np.linspace(-1, 1, R)], axis=1)
    run = RunConfig(n_replicas=R, n_walkers=4, n_atoms=n_atoms, n_mcmc_steps=4,
                    n_extra=2, seed=13, pressures=[1.0] * R, chemical_potentials=mu,
                    re_interval=1)
    eng = NSEngine(lj, moves, run, cell=CELL)
    eng.init_walkers(np.random.default_rng(13), 0.8, 4.0, True, species_p=[0.5, 0.5])
    eng.set_walk_block_cap(cap)
    tape = make_tape(np.random.default_rng(14), 3, R, 4, 256, 4, 2, eng.mwg.move_probs)
    pv = np.array([n_atoms[i] == n_atoms[i + 1] for i in range(R - 1)], np.int32)
    _run_parity(oracle, eng, 3, tape, pair_valid=pv)


@pytest.mark.parametrize("cap", [0, 1])
def test_device_draws_replay(oracle, cap):
    """Device Philox draws (normals and uniforms drawn in place by the walk),
    read back and replayed on the oracle."""
    print(f"device draws replay cap={cap}", flush=True)
    from oracle.oracle import Population
    lj = LJBackend(1.0, 1.0, 2.5)
    eng = _engine(lj, GMC_VOL, 3, 10, 20, 8, 5, cell_spec=CELL, pressures=[0.1, 0.5, 1.0])
    eng.set_walk_block_cap(cap)
    opop = Population.from_engine_dict(eng.get_population())
    pot = oracle.Pot.from_table(lj.table)
    for it in range(3):
        out = eng.ns_iterations(1)
        tape = eng.last_tape()
        fires = (eng.iteration - 1) > 0
        o = oracle.ns_iteration(opop, pot, eng.pressure, eng.mu, eng.step_sizes,
                                eng.mwg.as_oracle_moves(), 5, 8, tape, it=0,
                                do_re=fires)
        assert np.array_equal(out["emax"][0], o["emax"])
        assert np.array_equal(out["counters"][0], o["counters"])
        _compare_population(eng, opop)
    # draws look like uniforms / standard normals
    assert 0.0 <= tape.uniforms.min() and tape.uniforms.max() < 1.0
    z = tape.normals.ravel()
    assert abs(z.mean()) < 0.05 and abs(z.std() - 1.0) < 0.05


# species-scoped GMC (each with its own persistent direction), a refreshed
# scoped one, a refreshed unscoped one, and a volume move
SCOPED_MOVES = [
    {"type": "gmc", "species": 0, "n_reflect": 4, "step_size": 0.1, "weight": 2},
    {"type": "gmc", "species": [1], "n_reflect": 4, "step_size": 0.1, "weight": 2,
     "refresh_direction": True},
    {"type": "gmc", "n_reflect": 3, "step_size": 0.08, "weight": 1, "refresh_direction": True},
    {"type": "volume", "step_size": 0.3, "weight": 1}]
BINARY = dict(pair_epsilon=np.array([[1.0, 1.5], [1.5, 0.5]]),
              pair_sigma=np.array([[1.0, 0.8], [0.8, 0.88]]), cutoff=2.5)


def _scoped_engine(cap, seed=21):
    lj = LJBackend(**BINARY)
    eng = _engine(lj, SCOPED_MOVES, 4, 8, 32, 4, 8, seed=seed, cell_spec=CELL,
                  pressures=[0.1, 0.3, 1.0, 3.0], species_p=[0.5, 0.5], v_range=(0.5, 8.0))
    eng.set_walk_block_cap(cap)
    return eng, lj


@pytest.mark.parametrize("cap", [0, 1])
def test_ns_iterations_species_scoped_gmc(oracle, cap):
    """Tape-driven iterations with species-scoped GMC moves, bit for bit
    against the oracle, including each scoped move's persistent direction
    through cloning and replica exchange."""
    print(f"ns species-scoped gmc cap={cap}", flush=True)
    eng, _ = _scoped_engine(cap)
    assert eng.mwg.n_dir_slots == 3
    tape = make_tape(np.random.default_rng(13), 6, 4, 8, 32, 8, 4, eng.mwg.move_probs)
    _run_parity(oracle, eng, 6, tape)
    slots = eng.direction_slots()
    assert slots[0].any() and slots[1].any()
    pop = eng.get_population()
    for k, code in ((0, 0), (1, 1)):  # a scoped direction is zero off its species
        assert not slots[k][pop["types"] != code].any()


def test_species_scoped_gmc_device_draws_and_trials(oracle):
    """The in-place device draws replayed on the oracle, then trial moves of
    a scoped move (step-size adaptation) leave the population and the
    scoped directions untouched."""
    print("species-scoped gmc: device draws and trial moves", flush=True)
    from oracle.oracle import Population
    eng, lj = _scoped_engine(0, seed=22)
    opop = Population.from_engine_dict(eng.get_population())
    pot = oracle.Pot.from_table(lj.table)
    for it in range(3):
        out = eng.ns_iterations(1)
        tape = eng.last_tape()
        fires = (eng.iteration - 1) > 0
        o = oracle.ns_iteration(opop, pot, eng.pressure, eng.mu, eng.step_sizes,
                                eng.mwg.as_oracle_moves(), 8, 4, tape, it=0, do_re=fires)
        assert np.array_equal(out["counters"][0], o["counters"]), it
        _compare_population(eng, opop)
    rng = np.random.default_rng(23)
    idx = rng.integers(0, 8, size=(4, 20)).astype(np.int32)
    nrm = rng.standard_normal((4, 20, 3, 32))
    uni = rng.random((4, 20, 4))
    before = eng.get_population()
    pop = Population.from_engine_dict(before)
    for m in (0, 1):
        acc, rs = eng.trial_moves(m, [0.1] * 4, idx, normals=nrm, uniforms=uni)
        oacc, ors = oracle.trial_moves(pop, pot, eng.pressure, eng.mu, eng.step_sizes,
                                       eng.mwg.as_oracle_moves(), m, [0.1] * 4, eng.emax(),
                                       idx, nrm, uni)
        assert np.array_equal(acc, oacc) and np.array_equal(rs, ors), m
    after = eng.get_population()
    for k in ("positions", "direction", "direction_slots"):
        assert np.array_equal(after[k], before[k]), k


def test_graph_replay_equals_stream_path():
    print("graph replay", flush=True)
    lj = LJBackend(1.0, 1.0, 2.5)
    a = _engine(lj, GMC_VOL, 2, 12, 16, 10, 5, cell_spec=CELL, pressures=[0.1, 1.0], seed=21)
    b = _engine(lj, GMC_VOL, 2, 12, 16, 10, 5, cell_spec=CELL, pressures=[0.1, 1.0], seed=21)
    oa = a.ns_iterations(4)
    ob = b.ns_iterations(4, graph=True)
    for k in ("emax", "worst", "clone", "counters", "re_acc", "re_att", "walk"):
        assert np.array_equal(oa[k], ob[k]), k
    pa, pb = a.get_population(), b.get_population()
    for k in pa:
        assert np.array_equal(pa[k], pb[k]), k


def test_trial_moves_match_oracle(oracle):
    print("trial moves", flush=True)
    from oracle.oracle import Population
    lj = LJBackend(1.0, 1.0, 2.5)
    eng = _engine(lj, GMC_VOL, 2, 16, 24, 4, 4, cell_spec=CELL, pressures=[0.1, 1.0])
    eng.ns_iterations(2)
    rng = np.random.default_rng(2)
    idx = rng.integers(0, 16, size=(2, 30)).astype(np.int32)
    nrm = rng.standard_normal((2, 30, 3, 24))
    uni = rng.random((2, 30, 4))
    emax = eng.emax()
    pop = Population.from_engine_dict(eng.get_population())
    pot = oracle.Pot.from_table(lj.table)
    for m, ss in [(0, 0.07), (1, 0.5)]:
        acc, rs = eng.trial_moves(m, [ss, ss * 2], idx, normals=nrm, uniforms=uni)
        oacc, ors = oracle.trial_moves(pop, pot, eng.pressure, eng.mu, eng.step_sizes,
                                       eng.mwg.as_oracle_moves(), m, [ss, ss * 2], emax,
                                       idx, nrm, uni)
        assert np.array_equal(acc, oacc) and np.array_equal(rs, ors), m
    after = eng.get_population()
    assert np.array_equal(after["positions"], pop.pos)  # population untouched


def test_replica_exchange_standalone(oracle):
    print("replica exchange", flush=True)
    from oracle.oracle import Population
    lj = LJBackend(1.0, 1.0, 2.5)
    eng = _engine(lj, GMC_VOL, 5, 8, 12, 6, 3, cell_spec=CELL,
                  pressures=[0.01, 0.05, 0.1, 0.5, 1.0])
    eng.ns_iterations(3)
    pop = Population.from_engine_dict(eng.get_population())
    emax = eng.emax()
    re_u = np.random.default_rng(4).random((1, 2, 2, 2))
    st = eng.replica_exchange(re_u)
    acc, att = oracle.replica_exchange(pop, eng.pressure, eng.mu, emax, re_u, 1)
    assert np.array_equal(st["n_accepted_per_pair"], acc[:4])
    assert np.array_equal(st["n_attempted_per_pair"], att[:4])
    _compare_population(eng, pop)


def test_autotune_keeps_state_and_results():
    """Launch shape does not change results; tuning leaves the state alone."""
    print("autotune", flush=True)
    lj = LJBackend(1.0, 1.0, 2.5)
    a = _engine(lj, GMC_VOL, 3, 32, 64, 30, 4, cell_spec=CELL, pressures=[0.1, 1.0, 5.0], seed=8)
    b = _engine(lj, GMC_VOL, 3, 32, 64, 30, 4, cell_spec=CELL, pressures=[0.1, 1.0, 5.0], seed=8)
    before = a.get_population()
    t = a.autotune(reps=1)
    assert t["waves_per_block"] in (1, 2, 4) and t["occupancy"] in (0, 3)
    after = a.get_population()
    for k in before:
        assert np.array_equal(before[k], after[k]), k
    assert np.array_equal(a.emax(), b.emax())
    b.set_launch(1, 3)
    oa = a.ns_iterations(3)
    ob = b.ns_iterations(3)
    for k in ("emax", "counters", "walk"):
        assert np.array_equal(oa[k], ob[k]), k
    pa, pb = a.get_population(), b.get_population()
    for k in pa:
        assert np.array_equal(pa[k], pb[k]), k


def test_graph_replay_mixed_n_classes():
    """Two size classes (N = 32 and 96) walked in a captured graph equal the
    stream path."""
    print("graph replay mixed N", flush=True)
    eps = np.array([[1.0, 1.5], [1.5, 0.5]])
    sig = np.array([[1.0, 0.8], [0.8, 0.88]])
    moves = [{"type": "gmc", "n_reflect": 4, "step_size": 0.1, "weight": 4},
             {"type": "volume", "step_size": 0.3, "weight": 1},
             {"type": "alchemical_morph", "n_species": 2, "weight": 1}]
    mu = np.stack([np.zeros(4), np.linspace(-1, 1, 4)], axis=1)

    def make():
        run = RunConfig(n_replicas=4, n_walkers=16, n_atoms=[32, 32, 96, 96], n_mcmc_steps=6,
                        n_extra=8, seed=17, pressures=[1.0] * 4, chemical_potentials=mu,
                        re_interval=1)
        e = NSEngine(LJBackend(pair_epsilon=eps, pair_sigma=sig, cutoff=2.5), moves, run,
                     cell=CELL)
        e.init_walkers(np.random.default_rng(17), 0.5, 8.0, True, species_p=[0.5, 0.5])
        return e
    a, b = make(), make()
    oa = a.ns_iterations(4)
    ob = b.ns_iterations(4, graph=True)
    for k in ("emax", "worst", "clone", "counters", "re_acc", "re_att", "walk"):
        assert np.array_equal(oa[k], ob[k]), k
    pa, pb = a.get_population(), b.get_population()
    for k in pa:
        assert np.array_equal(pa[k], pb[k]), k


def test_eval_sizes_below_the_kernel_ring(oracle):
    """Batched evaluation where the kernel's ring (64 atoms per lane slot,
    sized by Nmax) is longer than a configuration's N, N a multiple of 64
    among them: N = 64 under two atoms per lane, N = 64 and 128 under four."""
    print("eval N below the kernel ring", flush=True)
    rng = np.random.default_rng(91)
    lj = LJBackend(1.0, 1.0, 2.5)
    pot = oracle.Pot.from_table(lj.table)
    st, keep = lj.table.as_struct(lj._prec_code)
    lib = _lib.load_library()
    for Nm, sizes in ((128, [64, 63, 65, 1, 100, 128]), (256, [64, 128, 192, 129, 200, 256])):
        B = len(sizes)
        pos = np.zeros((B, Nm, 3))
        ty = np.zeros((B, Nm), np.int32)
        cell = np.zeros((B, 3, 3))
        for b, n in enumerate(sizes):
            p, t, c = _random_cfgs(rng, 1, n, b % 2 == 1)
            pos[b, :n], ty[b, :n], cell[b] = p[0], t[0], c[0]
        na = np.array(sizes, np.int32)
        U, H, F = np.zeros(B), np.zeros(B), np.full((B, Nm, 3), np.nan)
        rc = lib.jr_eval_energy_forces(st, 0, B, Nm, _lib.dptr(pos), _lib.iptr(ty),
                                       _lib.iptr(na), _lib.dptr(cell), None, None,
                                       _lib.dptr(U), _lib.dptr(H), _lib.dptr(F))
        assert rc == 0, rc
        for b, n in enumerate(sizes):
            u, h, f = oracle.energy_forces(pot, pos[b, :n], ty[b, :n], cell[b])
            assert U[b] == u and H[b] == h, (Nm, n)
            assert np.array_equal(F[b, :n], f), (Nm, n, np.abs(F[b, :n] - f).max())


def test_trial_moves_mixed_sizes(oracle):
    """Trial moves of a mixed-size population run every chain in the kernel
    built for the largest N (here two atoms per lane), N = 64 and 32 among
    them."""
    print("trial moves, mixed sizes", flush=True)
    from oracle.oracle import Population
    lj = LJBackend(**BINARY)
    R, K = 4, 12
    sizes = [32, 64, 64, 128]
    run = RunConfig(n_replicas=R, n_walkers=K, n_atoms=sizes, n_mcmc_steps=4, n_extra=4,
                    seed=31, pressures=[1.0] * R, re_interval=0)
    eng = NSEngine(lj, GMC_VOL, run, cell=CELL)
    eng.init_walkers(np.random.default_rng(31), 0.5, 8.0, True, species_p=[0.5, 0.5])
    eng.ns_iterations(1)
    rng = np.random.default_rng(32)
    idx = rng.integers(0, K, size=(R, 16)).astype(np.int32)
    nrm = rng.standard_normal((R, 16, 3, 128))
    uni = rng.random((R, 16, 4))
    emax = eng.emax()
    pop = Population.from_engine_dict(eng.get_population())
    pot = oracle.Pot.from_table(lj.table)
    for m, ss in [(0, 0.05), (1, 0.3)]:
        acc, rs = eng.trial_moves(m, [ss] * R, idx, normals=nrm, uniforms=uni)
        oacc, ors = oracle.trial_moves(pop, pot, eng.pressure, eng.mu, eng.step_sizes,
                                       eng.mwg.as_oracle_moves(), m, [ss] * R, emax,
                                       idx, nrm, uni)
        assert np.array_equal(acc, oacc) and np.array_equal(rs, ors), m

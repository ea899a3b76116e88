"""Evaluation teams (``eval_chunks`` C > 1): a block of C wavefronts walks
each chain, every wave evaluating one chunk of the ring steps, and the
chunk sums are added in chunk order.  The HIP path is compared bit for bit
(``==`` on every float64 output) with the CPU oracle in the same chunked
order, for the batched evaluator, NS iterations on host tapes and on device
draws (the reference-default n_extra = 0 setting among them), trial moves,
the burn-in walk and graph replay.  Needs an MI355X."""

import ctypes as C

import numpy as np
import pytest

from jaxrens_amd.backends import EnsembleTerm, Hamiltonian, LJBackend
from jaxrens_amd.draws import make_tape
from jaxrens_amd.engine import NSEngine, RunConfig
from test_gpu_parity import (CELL, GMC_VOL, SCOPED_MOVES, _compare_population, _engine,
                             _random_cfgs, _run_parity)

pytestmark = pytest.mark.gpu

BINARY = dict(pair_epsilon=np.array([[1.0, 1.5], [1.5, 0.5]]),
              pair_sigma=np.array([[1.0, 0.8], [0.8, 0.88]]), cutoff=2.5)


@pytest.mark.parametrize("chunks", [2, 4, 8])
@pytest.mark.parametrize("N", [1, 2, 7, 8, 9, 13, 33, 64, 65, 96, 128, 129, 200, 256])
def test_eval_team_single_species(oracle, N, chunks):
    print(f"team eval N={N} chunks={chunks}", flush=True)
    rng = np.random.default_rng(N * 11 + chunks)
    lj = LJBackend(1.0, 1.0, 2.5, eval_chunks=chunks)
    pos, ty, cell = _random_cfgs(rng, 5, N, N % 2 == 1)
    pr = np.linspace(0.01, 1.0, 5)
    U, H, F = lj._eval(pos, ty, cell, pressure=pr, forces=True)
    pot = oracle.Pot.from_table(lj.table)
    assert pot.s.chunks == chunks
    for b in range(5):
        u, h, f = oracle.energy_forces(pot, pos[b], ty[b], cell[b], pressure=pr[b])
        assert U[b] == u and H[b] == h, (b, U[b], u)
        assert np.array_equal(F[b], f), (b, np.abs(F[b] - f).max())


@pytest.mark.parametrize("N", [8, 48, 96, 128])
def test_eval_team_pair_table_and_mu(oracle, N):
    print(f"team eval pair table N={N}", flush=True)
    rng = np.random.default_rng(N + 3)
    lj = LJBackend(**BINARY, eval_chunks=8)
    pos, ty, cell = _random_cfgs(rng, 4, N, True, n_species=2, dense=0.8)
    mu = rng.normal(size=(4, 2))
    U, H, F = lj._eval(pos, ty, cell, pressure=1.0, mu=mu, forces=True)
    pot = oracle.Pot.from_table(lj.table)
    for b in range(4):
        u, h, f = oracle.energy_forces(pot, pos[b], ty[b], cell[b], pressure=1.0, mu=mu[b])
        assert U[b] == u and H[b] == h
        assert np.array_equal(F[b], f)


@pytest.mark.parametrize("chunks,cap", [(8, 0), (8, 1), (2, 0)])
@pytest.mark.parametrize("n_extra", [0, 14])
def test_team_ns_iterations_lj8(oracle, n_extra, chunks, cap):
    print(f"team ns lj8 n_extra={n_extra} chunks={chunks} cap={cap}", flush=True)
    lj = Hamiltonian(LJBackend(1.0, 1.0, 2.5, eval_chunks=chunks), [EnsembleTerm(0.1)])
    eng = _engine(lj, GMC_VOL, 1, 16, 8, n_extra, 20, cell_spec=CELL)
    eng.set_walk_block_cap(cap)
    assert eng.launch_info()["waves_per_block"] == min(chunks, 4)
    tape = make_tape(np.random.default_rng(9), 5, 1, 16, 8, 20, n_extra,
                     eng.mwg.move_probs, ties=True)
    _run_parity(oracle, eng, 5, tape)


@pytest.mark.parametrize("chunks,cap", [(8, 0), (4, 1)])
def test_team_ns_iterations_lj64_with_re(oracle, chunks, cap):
    print(f"team ns lj64 R=4 RE chunks={chunks} cap={cap}", flush=True)
    lj = LJBackend(1.0, 1.0, 2.5, eval_chunks=chunks)
    P = list(10 ** (-2 + 3 * np.arange(4) / 3))
    eng = _engine(lj, GMC_VOL, 4, 8, 64, 6, 6, cell_spec=CELL, pressures=P)
    eng.set_walk_block_cap(cap)
    tape = make_tape(np.random.default_rng(10), 3, 4, 8, 64, 6, 6, eng.mwg.move_probs)
    _run_parity(oracle, eng, 3, tape)


@pytest.mark.parametrize("chunks", [4, 8])
def test_team_ns_iterations_hmc_gmc_n128(oracle, chunks):
    print(f"team ns hmc+gmc N=128 chunks={chunks}", flush=True)
    lj = LJBackend(1.0, 1.0, 2.5, eval_chunks=chunks)
    moves = [{"type": "hmc", "n_leapfrog": 4, "step_size": 0.005, "weight": 1},
             {"type": "gmc", "n_reflect": 3, "step_size": 0.05, "weight": 1}]
    eng = _engine(lj, moves, 2, 4, 128, 2, 4, pressures=[0.05, 0.5],
                  cell_spec={"min_volume_per_atom": 0.5, "max_volume_per_atom": 4.0},
                  v_range=(1.2, 3.0))
    tape = make_tape(np.random.default_rng(11), 2, 2, 4, 128, 4, 2, eng.mwg.move_probs)
    _run_parity(oracle, eng, 2, tape)


def test_team_ns_iterations_n256(oracle):
    print("team ns gmc+volume N=256 chunks=8", flush=True)
    lj = LJBackend(1.0, 1.0, 2.5, eval_chunks=8)
    eng = _engine(lj, GMC_VOL, 2, 4, 256, 1, 3, pressures=[0.5, 1.0],
                  cell_spec={"min_volume_per_atom": 0.5, "max_volume_per_atom": 4.0},
                  v_range=(1.2, 3.0))
    tape = make_tape(np.random.default_rng(14), 2, 2, 4, 256, 3, 1, eng.mwg.move_probs)
    _run_parity(oracle, eng, 2, tape)


def test_team_semigrand_mixed_n_with_swap(oracle):
    """Two size classes (N = 32, 48 and 96) as concurrent team launches,
    gmc, volume, species_swap and alchemical_morph."""
    print("team ns semi-grand mixed N with species_swap", flush=True)
    lj = LJBackend(**BINARY, eval_chunks=8)
    moves = [{"type": "gmc", "n_reflect": 4, "step_size": 0.1, "weight": 3},
             {"type": "volume", "step_size": 0.3, "weight": 1},
             {"type": "species_swap", "weight": 2},
             {"type": "alchemical_morph", "n_species": 2, "weight": 1}]
    R = 6
    n_atoms = [32, 32, 48, 48, 96, 96]
    mu = np.stack([np.zeros(R), np.linspace(-2, 2, R)], axis=1)
    run = RunConfig(n_replicas=R, n_walkers=6, n_atoms=n_atoms, n_mcmc_steps=6,
                    n_extra=3, seed=5, pressures=[1.0] * R, chemical_potentials=mu,
                    re_interval=1)
    eng = NSEngine(lj, moves, run, cell=CELL)
    eng.init_walkers(np.random.default_rng(5), 0.5, 8.0, True, species_p=[0.5, 0.5])
    tape = make_tape(np.random.default_rng(12), 3, R, 6, 96, 6, 3, eng.mwg.move_probs)
    pv = np.array([n_atoms[i] == n_atoms[i + 1] for i in range(R - 1)], np.int32)
    _run_parity(oracle, eng, 3, tape, pair_valid=pv)


def test_team_species_scoped_gmc(oracle):
    """Scoped moves keep their directions in global slots that every wave of
    the team reads and writes."""
    print("team ns species-scoped gmc", flush=True)
    lj = LJBackend(**BINARY, eval_chunks=4)
    eng = _engine(lj, SCOPED_MOVES, 4, 8, 32, 4, 8, seed=21, cell_spec=CELL,
                  pressures=[0.1, 0.3, 1.0, 3.0], species_p=[0.5, 0.5], v_range=(0.5, 8.0))
    tape = make_tape(np.random.default_rng(13), 4, 4, 8, 32, 8, 4, eng.mwg.move_probs)
    _run_parity(oracle, eng, 4, tape)


def _device_draws(oracle, eng, n_iter):
    from oracle.oracle import Population
    opop = Population.from_engine_dict(eng.get_population())
    pot = oracle.Pot.from_table(eng.model.table)
    for it in range(n_iter):
        out = eng.ns_iterations(1)
        tape = eng.last_tape()
        i = eng.iteration - 1
        fires = eng.run.re_interval > 0 and i > 0 and i % eng.run.re_interval == 0
        o = oracle.ns_iteration(opop, pot, eng.pressure, eng.mu, eng.step_sizes,
                                eng.mwg.as_oracle_moves(), eng.run.n_mcmc_steps,
                                eng.run.n_extra, tape, it=0, do_re=fires)
        for k in ("emax", "worst", "clone", "walk", "counters"):
            assert np.array_equal(out[k][0], o[k]), (k, it)
        if eng.n_replicas > 1:
            assert np.array_equal(out["re_acc"][0], o["re_acc"]), it
        _compare_population(eng, opop)


def test_team_lj64_16x1024_reference_default_device_draws(oracle):
    """The LJ-64 NPT stack at the reference's default n_extra = 0 (one chain
    per replica per iteration) on device draws: 8 iterations with RE."""
    print("team LJ-64 16x1024 n_extra=0 chunks=8, device draws", flush=True)
    import bench
    eng = bench.build_engine(n_extra=0, eval_chunks=8)
    _device_draws(oracle, eng, 8)
    eng.close()


def test_team_trial_moves(oracle):
    print("team trial moves", flush=True)
    from oracle.oracle import Population
    lj = LJBackend(1.0, 1.0, 2.5, eval_chunks=8)
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
    assert np.array_equal(eng.get_population()["positions"], pop.pos)


def test_team_initial_walk(oracle):
    print("team initial walk", flush=True)
    from oracle.oracle import Population
    from jaxrens_amd.burn_in import burn_in_emax
    moves = [{"type": "gmc", "n_reflect": 6, "step_size": 0.1, "weight": 4.0},
             {"type": "volume", "step_size": 0.2, "weight": 1.0},
             {"type": "shear", "step_size": 0.08, "weight": 1.0},
             {"type": "stretch", "step_size": 0.08, "weight": 1.0}]
    cell = {"max_volume_per_atom": 30.0, "min_volume_per_atom": 1.2,
            "min_aspect_ratio": 0.7, "flat_V_prior": True}
    lj = Hamiltonian(LJBackend(1.0, 1.0, 2.5, eval_chunks=4), [EnsembleTerm(0.1)])
    run = RunConfig(n_replicas=2, n_walkers=12, n_atoms=20, n_mcmc_steps=10, n_extra=4,
                    seed=7, pressures=[0.1, 1.0], re_interval=1)
    eng = NSEngine(lj, moves, run, cell=cell)
    eng.init_walkers(np.random.default_rng(7), 1.2, 30.0, True)
    pot = oracle.Pot.from_table(eng.model.table)
    opop = Population.from_engine_dict(eng.get_population())
    emax = burn_in_emax(eng, 1.0)
    rng = np.random.default_rng(3)
    Nm = eng.Nmax
    for w in range(2):
        sc = rng.choice(4, size=7, p=eng.mwg.move_probs).astype(np.int32)
        nn = rng.standard_normal((2, 12, 7, 3, Nm))
        uu = rng.random((2, 12, 7, 4))
        ctr = eng.initial_walk(7, emax, sc, nn, uu, max_chunk=5)
        octr = oracle.initial_walk(opop, pot, eng.pressure, eng.mu, eng.step_sizes,
                                   eng.mwg.as_oracle_moves(), emax, sc, nn, uu)
        assert np.array_equal(ctr, octr), w
        _compare_population(eng, opop)


def test_team_graph_replay_equals_stream_path():
    print("team graph replay", flush=True)
    lj = LJBackend(1.0, 1.0, 2.5, eval_chunks=8)
    a = _engine(lj, GMC_VOL, 2, 12, 16, 0, 5, cell_spec=CELL, pressures=[0.1, 1.0], seed=21)
    b = _engine(lj, GMC_VOL, 2, 12, 16, 0, 5, cell_spec=CELL, pressures=[0.1, 1.0], seed=21)
    oa = a.ns_iterations(4)
    ob = b.ns_iterations(4, graph=True)
    for k in ("emax", "worst", "clone", "counters", "re_acc", "re_att", "walk"):
        assert np.array_equal(oa[k], ob[k]), k
    pa, pb = a.get_population(), b.get_population()
    for k in pa:
        assert np.array_equal(pa[k], pb[k]), k


def test_team_eval_mixed_sizes_in_one_call(oracle):
    """One batched call over configurations of different atom counts (one
    team per configuration, the kernel built for the largest)."""
    print("team eval mixed sizes", flush=True)
    rng = np.random.default_rng(77)
    lj = LJBackend(**BINARY, eval_chunks=4)
    pot = oracle.Pot.from_table(lj.table)
    sizes = [1, 7, 40, 64, 65, 100, 128]
    Nm = max(sizes)
    B = len(sizes)
    pos = np.zeros((B, Nm, 3))
    ty = np.zeros((B, Nm), np.int32)
    cell = np.zeros((B, 3, 3))
    for b, n in enumerate(sizes):
        p, t, c = _random_cfgs(rng, 1, n, b % 2 == 0, n_species=2, dense=0.7)
        pos[b, :n], ty[b, :n], cell[b] = p[0], t[0], c[0]
    from jaxrens_amd import _lib
    lib = _lib.load_library()
    st, keep = lj.table.as_struct(lj._prec_code)
    U, H, F = np.zeros(B), np.zeros(B), np.zeros((B, Nm, 3))
    na = np.array(sizes, np.int32)
    rc = lib.jr_eval_energy_forces(C.byref(st), 0, B, Nm, _lib.dptr(_lib.f64(pos)),
                                   _lib.iptr(_lib.i32(ty)), _lib.iptr(na),
                                   _lib.dptr(_lib.f64(cell)), None, None, _lib.dptr(U),
                                   _lib.dptr(H), _lib.dptr(F))
    assert rc == 0, lib.jr_global_error()
    for b, n in enumerate(sizes):
        u, h, f = oracle.energy_forces(pot, pos[b, :n], ty[b, :n], cell[b])
        assert U[b] == u and H[b] == h, (b, n)
        assert np.array_equal(F[b, :n], f), (b, n)
        assert not F[b, n:].any()


def test_team_autotune_keeps_state_and_results():
    """The autotuner on a team handle times its one geometry and leaves the
    population and contour alone; results equal an untuned engine's."""
    print("team autotune", flush=True)
    lj = LJBackend(1.0, 1.0, 2.5, eval_chunks=4)
    a = _engine(lj, GMC_VOL, 3, 32, 64, 0, 4, cell_spec=CELL, pressures=[0.1, 1.0, 5.0], seed=8)
    b = _engine(lj, GMC_VOL, 3, 32, 64, 0, 4, cell_spec=CELL, pressures=[0.1, 1.0, 5.0], seed=8)
    before = a.get_population()
    t = a.autotune(reps=1)
    assert t["waves_per_block"] == 4
    after = a.get_population()
    for k in before:
        assert np.array_equal(before[k], after[k]), k
    oa = a.ns_iterations(3)
    ob = b.ns_iterations(3)
    for k in ("emax", "counters", "walk"):
        assert np.array_equal(oa[k], ob[k]), k

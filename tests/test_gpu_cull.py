"""n_cull > 1 on the device (``sampling/nested_sampling.py:188-212``): the
lexicographic-order select kernel, its draws and outputs, bit for bit
against the CPU oracle on host tapes and on the device's own draws, the
hipGraph path, and the run loop's one ``.energies`` row and frame per dead
point with the evidence in the pymatnest weights.  Needs an MI355X."""

import numpy as np
import pytest

from jaxrens_amd import prior_mass as pm
from jaxrens_amd.backends import LJBackend
from jaxrens_amd.draws import make_tape
from jaxrens_amd.engine import NSEngine, RunConfig
from jaxrens_amd.run_loop import EnergyLogger, run_ns
from jaxrens_amd.termination import IterationTermination
from jaxrens_amd.trajectory import read_extxyz

pytestmark = pytest.mark.gpu

GMC_VOL = [{"type": "galilean", "n_reflect": 6, "step_size": 0.1, "weight": 4},
           {"type": "volume", "step_size": 0.3, "weight": 1}]
CELL = {"min_volume_per_atom": 0.125, "max_volume_per_atom": 64.0, "flat_V_prior": True}


def _engine(R, K, N, C, n_extra, n_steps=5, seed=3):
    run = RunConfig(n_replicas=R, n_walkers=K, n_atoms=N, n_mcmc_steps=n_steps,
                    n_extra=n_extra, n_cull=C, seed=seed,
                    pressures=list(np.linspace(0.1, 1.0, R)), re_interval=1)
    eng = NSEngine(LJBackend(1.0, 1.0, 2.5), GMC_VOL, run, cell=CELL)
    eng.init_walkers(np.random.default_rng(seed), 0.5, 8.0, True)
    return eng


def _compare_population(eng, opop):
    d = eng.get_population()
    assert np.array_equal(d["energy"], opop.energy)
    assert np.array_equal(d["positions"], opop.pos)
    assert np.array_equal(d["direction"], opop.dir)
    assert np.array_equal(d["types"], opop.types)
    assert np.array_equal(d["cell"].reshape(opop.cell.shape), opop.cell)


def _check_iteration(eng_out, o, C):
    assert np.array_equal(eng_out["emax"][0], o["emax"])
    for k in ("worst", "clone", "dead_energy", "dead_volume"):
        assert np.array_equal(np.reshape(eng_out[k][0], (-1, C)),
                              np.reshape(o[k], (-1, C))), k
    assert np.array_equal(eng_out["walk"][0], o["walk"])
    assert np.array_equal(eng_out["counters"][0], o["counters"])
    if eng_out["re_acc"].shape[1] == o["re_acc"].shape[0]:
        assert np.array_equal(eng_out["re_att"][0], o["re_att"])
        assert np.array_equal(eng_out["re_acc"][0], o["re_acc"])


@pytest.mark.parametrize("R,K,N,C,n_extra", [(3, 24, 8, 3, 2), (2, 20, 33, 4, 5),
                                            (2, 64, 64, 8, 0), (1, 12, 96, 2, 3)])
def test_cull_tape_parity(oracle, R, K, N, C, n_extra):
    print(f"n_cull parity R={R} K={K} N={N} C={C} n_extra={n_extra}", flush=True)
    from oracle.oracle import Population
    eng = _engine(R, K, N, C, n_extra)
    pot = oracle.Pot.from_table(eng.model.table)
    moves = eng.mwg.as_oracle_moves()
    opop = Population.from_engine_dict(eng.get_population())
    tape = make_tape(np.random.default_rng(13), 4, R, K, N, 5, n_extra,
                     eng.mwg.move_probs, ties=True, n_cull=C)
    for it in range(4):
        before = opop.copy()
        out = eng.ns_iterations(1, tape=tape.iteration(it), dead_records=True)
        fires = (eng.iteration - 1) > 0
        o = oracle.ns_iteration(opop, pot, eng.pressure, eng.mu, eng.step_sizes, moves,
                                5, n_extra, tape, it=it, do_re=fires and R > 1, n_cull=C)
        _check_iteration(out, o, C)
        _compare_population(eng, opop)
        # dead records are the culled walkers as they were before cloning
        worst = np.reshape(out["worst"][0], (R, C))
        dpos = np.reshape(out["dead_pos"][0], (R, C, N, 3))
        dcell = np.reshape(out["dead_cell"][0], (R, C, 3, 3))
        for r in range(R):
            for c in range(C):
                assert np.array_equal(dpos[r, c], before.pos[r, worst[r, c]])
                assert np.array_equal(dcell[r, c].reshape(9), before.cell[r, worst[r, c]])


def test_cull_device_draws_replay(oracle):
    """Device Philox draws (including the permutation keys) read back and
    replayed on the oracle."""
    print("n_cull device draws replay", flush=True)
    from oracle.oracle import Population
    eng = _engine(2, 16, 12, 3, 2, seed=8)
    pot = oracle.Pot.from_table(eng.model.table)
    opop = Population.from_engine_dict(eng.get_population())
    for it in range(3):
        out = eng.ns_iterations(1)
        tape = eng.last_tape()
        fires = (eng.iteration - 1) > 0
        o = oracle.ns_iteration(opop, pot, eng.pressure, eng.mu, eng.step_sizes,
                                eng.mwg.as_oracle_moves(), 5, 2, tape, it=0, do_re=fires,
                                n_cull=3)
        _check_iteration(out, o, 3)
        _compare_population(eng, opop)
    keys = tape.extra_u[0, :, :13]
    assert 0.0 <= keys.min() and keys.max() < 1.0 and len(np.unique(keys)) == keys.size


def test_cull_graph_replay_equals_stream_path():
    print("n_cull graph replay", flush=True)
    a = _engine(2, 24, 16, 4, 3, seed=21)
    b = _engine(2, 24, 16, 4, 3, seed=21)
    oa = a.ns_iterations(4)
    ob = b.ns_iterations(4, graph=True)
    for k in ("emax", "worst", "clone", "dead_energy", "counters", "walk"):
        assert np.array_equal(oa[k], ob[k]), k
    pa, pb = a.get_population(), b.get_population()
    for k in pa:
        assert np.array_equal(pa[k], pb[k]), k


def test_cull_run_writes_every_dead_point(tmp_path):
    """n_cull = 3: three ``.energies`` rows and frames per iteration in
    descending energy, the last equal to the contour; log Z accumulated on
    the host equals the pymatnest weights applied to the written rows."""
    print("n_cull run loop", flush=True)
    C, K, n_iter = 3, 32, 12
    eng = _engine(2, K, 8, C, 4, seed=5)
    res = run_ns(eng, [IterationTermination(n_iter)], energies_prefix=str(tmp_path / "c"),
                 traj_prefix=str(tmp_path / "c"), traj_interval=4, symbol_map={0: "Ar"})
    assert res.iterations == n_iter
    for r in range(2):
        log = EnergyLogger.read(tmp_path / f"c.{r}.energies")
        assert log["n_cull"] == C and log["n_walkers"] == K
        assert np.array_equal(log["iterations"], np.repeat(np.arange(n_iter), C))
        e = log["energies"].reshape(n_iter, C)
        assert np.all(np.diff(e, axis=1) <= 0)
        np.testing.assert_allclose(e[:, -1], res.emax[:, r], rtol=1e-10)
        assert np.all(np.diff(e[:, -1]) <= 0)  # the contour only falls
        lw = pm.log_dead_weights(n_iter * C, K, C)
        terms = lw - log["energies"]
        m = terms.max()
        log_z = m + np.log(np.exp(terms - m).sum())
        np.testing.assert_allclose(eng.log_evidence[r], log_z, rtol=1e-9, atol=1e-9)
        frames = read_extxyz(tmp_path / f"c.{r}.traj.extxyz")
        assert [f["iter"] for f in frames] == [0, 0, 0, 4, 4, 4, 8, 8, 8]
        assert [f["cull_rank"] for f in frames] == [0, 1, 2] * 3
        for f in frames:
            row = (f["iter"]) * C + f["cull_rank"]
            assert abs(f["ns_energy"] - log["energies"][row]) <= 1e-9 * abs(f["ns_energy"])


def test_cull_config_needs_distinct_sources():
    from jaxrens_amd._lib import JrError
    run = RunConfig(n_replicas=1, n_walkers=8, n_atoms=4, n_mcmc_steps=2, n_extra=1,
                    n_cull=4, seed=1, pressures=[0.1], re_interval=0)
    with pytest.raises(JrError, match="2\\*n_cull"):
        NSEngine(LJBackend(1.0, 1.0, 2.5), GMC_VOL, run, cell=CELL)


def test_cull_log_z_agrees_with_single_cull():
    """LJ-8 NPT (P = 0.1, K = 128, 16 independent replicas): log Z with
    n_cull = 4 over 300 iterations and with n_cull = 1 over 1200 (the same
    number of dead points), live share included, agree within 2 sigma of
    the difference of the replica means."""
    print("n_cull log Z vs single cull", flush=True)
    moves = [{"type": "galilean", "n_reflect": 10, "step_size": 0.1, "weight": 4},
             {"type": "volume", "step_size": 0.3, "weight": 1}]
    cell = {"min_volume_per_atom": 1.0 / 8, "max_volume_per_atom": 64.0, "flat_V_prior": True}
    R = 16

    def log_z(C, iters, seed):
        run = RunConfig(n_replicas=R, n_walkers=128, n_atoms=8, n_mcmc_steps=20, n_extra=0,
                        n_cull=C, seed=seed, pressures=[0.1] * R, re_interval=0)
        eng = NSEngine(LJBackend(1.0, 1.0, 2.5), moves, run, cell=cell)
        eng.init_walkers(np.random.default_rng(seed), 1.0 / 8, 64.0, True)
        run_ns(eng, [IterationTermination(iters)], block=50)
        return eng.final_log_evidence()

    a, b = log_z(1, 1200, 101), log_z(4, 300, 202)
    se = np.sqrt(a.var(ddof=1) / R + b.var(ddof=1) / R)
    z = abs(a.mean() - b.mean()) / se
    print(f"log Z: n_cull=1 {a.mean():.4f} +- {a.std(ddof=1) / 4:.4f}, "
          f"n_cull=4 {b.mean():.4f} +- {b.std(ddof=1) / 4:.4f}, z = {z:.2f}", flush=True)
    assert z < 2.0

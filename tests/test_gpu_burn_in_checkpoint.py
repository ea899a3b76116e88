"""Burn-in (``init/burn_in.py``) and checkpoint / resume
(``io/checkpoint.py``, ``io/restart_truncate.py``) on the device engine.

The fixed-contour walk is compared bit for bit with the CPU oracle on host
draws, in whole-replica launches and in ragged chunks; a run split at a
checkpoint must reproduce the uninterrupted run bit for bit.  Needs an
MI355X."""

import numpy as np
import pytest

from jaxrens_amd.adaptation import AdaptationConfig
from jaxrens_amd.backends import EnsembleTerm, Hamiltonian, LJBackend
from jaxrens_amd.burn_in import burn_in_emax, initial_walk
from jaxrens_amd.checkpoint import load_checkpoint
from jaxrens_amd.engine import NSEngine, RunConfig
from jaxrens_amd.run_loop import EnergyLogger, run_ns
from jaxrens_amd.termination import IterationTermination
from jaxrens_amd.trajectory import read_extxyz

pytestmark = pytest.mark.gpu

MOVES = [{"type": "gmc", "n_reflect": 6, "step_size": 0.1, "weight": 4.0},
         {"type": "volume", "step_size": 0.2, "weight": 1.0},
         {"type": "shear", "step_size": 0.08, "weight": 1.0},
         {"type": "stretch", "step_size": 0.08, "weight": 1.0}]
CELL = {"max_volume_per_atom": 30.0, "min_volume_per_atom": 1.2,
        "min_aspect_ratio": 0.7, "flat_V_prior": True}


def _engine(R=2, K=12, N=8, seed=7, sc=(4, 4, 4), n_extra=4, moves=MOVES):
    lj = Hamiltonian(LJBackend(1.0, 1.0, 2.5, supercell_trafo=sc), [EnsembleTerm(0.1)])
    run = RunConfig(n_replicas=R, n_walkers=K, n_atoms=N, n_mcmc_steps=10,
                    n_extra=n_extra, seed=seed, pressures=list(np.linspace(0.1, 1.0, R)),
                    re_interval=1)
    eng = NSEngine(lj, moves, run, cell=CELL)
    eng.init_walkers(np.random.default_rng(seed), 1.2, 30.0, True)
    return eng


def _tape(rng, eng, wl):
    R, K, Nm = eng.n_replicas, eng.n_walkers, eng.Nmax
    probs = eng.mwg.move_probs
    sched = rng.choice(len(probs), size=wl, p=probs).astype(np.int32)
    return (sched, rng.standard_normal((R, K, wl, 3, Nm)), rng.random((R, K, wl, 4)))


@pytest.mark.parametrize("max_chunk", [0, 5])
def test_initial_walk_matches_oracle(oracle, max_chunk):
    print(f"initial walk parity max_chunk={max_chunk}", flush=True)
    from oracle.oracle import Population
    eng = _engine()
    pot = oracle.Pot.from_table(eng.model.table)
    moves = eng.mwg.as_oracle_moves()
    opop = Population.from_engine_dict(eng.get_population())
    emax = burn_in_emax(eng, 1.0)
    assert np.array_equal(emax, opop.energy.max(axis=1) + 8.0)
    contour = eng.emax()
    rng = np.random.default_rng(3)
    for w in range(2):
        sc, nn, uu = _tape(rng, eng, 7)
        ctr = eng.initial_walk(7, emax, sc, nn, uu, max_chunk=max_chunk)
        octr = oracle.initial_walk(opop, pot, eng.pressure, eng.mu, eng.step_sizes, moves,
                                   emax, sc, nn, uu)
        assert np.array_equal(ctr, octr), w
        d = eng.get_population()
        assert np.array_equal(d["energy"], opop.energy), w
        assert np.array_equal(d["positions"], opop.pos), w
        assert np.array_equal(d["cell"].reshape(opop.cell.shape), opop.cell), w
        assert np.array_equal(d["direction"], opop.dir), w
    assert ctr[:, :, 0].sum() == 2 * 12 * 7 and ctr[:, :, 1].sum() > 0
    assert np.all(opop.energy < emax[:, None])
    assert np.array_equal(eng.emax(), contour)  # the NS contour is untouched


def test_burn_in_device_draws_and_adaptation():
    print("burn-in device draws + adaptation", flush=True)
    eng = _engine(R=3, K=24, seed=11)
    e0 = eng.get_population()["energy"].copy()
    ss0 = eng.step_sizes.copy()
    res = initial_walk(eng, n_walks=4, walklength=10, adjust_interval=2,
                       emax_offset_per_atom=1.0,
                       adaptation=AdaptationConfig(adjust_n_samples=20, step_size_max=0.3),
                       rng=np.random.default_rng(1))
    e1 = eng.get_population()["energy"]
    assert np.all(e1 < res["emax"][:, None])
    assert not np.array_equal(e0, e1)
    assert res["counters"][:, :, 0].sum() == 4 * 3 * 24 * 10
    assert len(res["adaptation"]) == 2
    ss = eng.step_sizes
    assert not np.array_equal(ss, ss0)
    assert np.all(ss > 0) and np.all(ss <= 0.3 + 1e-12)
    # the walked population starts a normal NS run
    out = eng.ns_iterations(2)
    assert np.all(np.isfinite(out["emax"]))


@pytest.mark.parametrize("ck_at", [7, 10])
def test_checkpoint_resume_is_exact(tmp_path, ck_at):
    """Run 14 iterations with adaptation (every 5) and device draws,
    checkpointing at ``ck_at``; a fresh engine resumed from the checkpoint
    must end bit-identical, and its .energies files must equal the
    uninterrupted run's.  At 10 the first iteration after resume adapts, so
    the trial moves read the restored contour."""
    print(f"checkpoint resume at {ck_at}", flush=True)
    ad = AdaptationConfig(adjust_interval=5, adjust_n_samples=20)
    a = _engine(seed=5)
    ra = run_ns(a, [IterationTermination(14)], adaptation=ad, seed=9,
                energies_prefix=str(tmp_path / "a"),
                checkpoint_path=tmp_path / "ck.npz", checkpoint_interval=ck_at,
                traj_prefix=str(tmp_path / "a"), traj_interval=3, symbol_map={0: "Ar"})
    assert ra.iterations == 14
    ck = load_checkpoint(tmp_path / "ck.npz")
    assert ck["iteration"] == (14 if ck_at == 7 else 10)  # the last write
    # rerun A's first ck_at iterations to get the mid-run checkpoint
    b = _engine(seed=5)
    run_ns(b, [IterationTermination(ck_at)], adaptation=ad, seed=9,
           energies_prefix=str(tmp_path / "b"),
           checkpoint_path=tmp_path / "mid.npz", checkpoint_interval=ck_at,
           traj_prefix=str(tmp_path / "b"), traj_interval=3, symbol_map={0: "Ar"})
    mid = load_checkpoint(tmp_path / "mid.npz")
    assert mid["iteration"] == ck_at and mid["n_dead"] == ck_at
    assert np.array_equal(mid["emax"], b.emax()) and np.isfinite(mid["emax"]).all()
    # a third, fresh engine resumes at ck_at and appends to b's energies
    c = _engine(seed=5)
    # stale rows past the checkpoint are cut on resume
    with open(tmp_path / "b.0.energies", "a") as f:
        f.write(f"{ck_at} 1.0 1.0\n{ck_at + 1} 1.0 1.0\n")
    rc = run_ns(c, [IterationTermination(14)], adaptation=ad,
                energies_prefix=str(tmp_path / "b"), resume=tmp_path / "mid.npz",
                traj_prefix=str(tmp_path / "b"), traj_interval=3, symbol_map={0: "Ar"})
    assert rc.iterations == 14
    pa, pc = a.get_population(), c.get_population()
    for k in ("energy", "positions", "cell", "direction", "types"):
        assert np.array_equal(pa[k], pc[k]), k
    assert np.array_equal(a.log_evidence, c.log_evidence)
    assert np.array_equal(a.step_sizes, c.step_sizes)
    for r in range(2):
        ea = EnergyLogger.read(tmp_path / f"a.{r}.energies")
        eb = EnergyLogger.read(tmp_path / f"b.{r}.energies")
        assert np.array_equal(ea["iterations"], eb["iterations"])
        assert np.array_equal(ea["energies"], eb["energies"])
        assert np.array_equal(ea["volumes"], eb["volumes"])
        ta = (tmp_path / f"a.{r}.traj.extxyz").read_text()
        assert ta == (tmp_path / f"b.{r}.traj.extxyz").read_text()
        assert [f["iter"] for f in read_extxyz(tmp_path / f"a.{r}.traj.extxyz")] == \
            [0, 3, 6, 9, 12]


def test_run_writes_trajectory(tmp_path):
    """Dead points of every second iteration: the frame energy is the
    .energies row, and the LJ energy of the written (wrapped, 8-decimal)
    geometry plus P V reproduces it."""
    print("trajectory", flush=True)
    eng = _engine(R=2, K=12, seed=3)
    run_ns(eng, [IterationTermination(9)], energies_prefix=str(tmp_path / "t"),
           traj_prefix=str(tmp_path / "t"), traj_interval=2, symbol_map={0: "Ar"})
    lj = LJBackend(1.0, 1.0, 2.5, supercell_trafo=(4, 4, 4))
    for r in range(2):
        fr = read_extxyz(tmp_path / f"t.{r}.traj.extxyz")
        en = EnergyLogger.read(tmp_path / f"t.{r}.energies")
        assert [f["iter"] for f in fr] == [0, 2, 4, 6, 8]
        for f in fr:
            assert f["ns_energy"] == en["energies"][f["iter"]] or \
                abs(f["ns_energy"] - en["energies"][f["iter"]]) <= 1e-9 * abs(f["ns_energy"])
            frac = np.linalg.solve(f["cell"].T, f["positions"].T).T
            assert frac.min() > -1e-6 and frac.max() < 1 + 1e-6
            u = lj(f["positions"], np.zeros(8, np.int32), f["cell"]).energy
            h = u + eng.pressure[r] * abs(np.linalg.det(f["cell"]))
            assert h == pytest.approx(f["ns_energy"], rel=1e-5, abs=1e-5)

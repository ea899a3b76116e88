"""The learned-model walk on the GPU.

* Parity: one tape-driven NS iteration of the plain walk (judge-steered)
  and of the screened walk (surrogate-steered, two-stage test) against the
  CPU restatement in oracle/learned_walk.py.  The judge runs in FP32 on the
  device and FP64 in the oracle (stated accuracy |dE| <= 1e-4 (|E| + N)),
  so accept decisions must agree wherever the margin |H - emax| exceeds
  that error, positions to 1e-6 relative to the box.
* Exactness: delayed acceptance samples the same constrained prior, so
  log Z over independent runs agrees with the plain walk within 2 sigma,
  for an untrained surrogate and for one trained online.
* Gain: what the trained network buys over the walk without it.
"""

import numpy as np
import pytest

from jaxrens_amd.backends import EnsembleTerm, Hamiltonian, LJBackend
from jaxrens_amd.draws import make_tape
from jaxrens_amd.engine import NSEngine, RunConfig
from jaxrens_amd.mpnn import MPNNBackend, init_mpnn
from jaxrens_amd.surrogate import LearnedConfig, init_surrogate
from oracle import learned_walk as OW
from oracle.oracle import Pot, Population

pytestmark = pytest.mark.gpu

MOVES = [{"type": "galilean", "n_reflect": 4, "step_size": 0.08, "weight": 3},
         {"type": "volume", "step_size": 0.6, "weight": 1}]
# the surrogate's pair baseline: the judge's core form (eps, sig, rc)
BASE = (1.0, 0.85, 1.2)


def _moves_dicts(eng):
    out = []
    for m in eng.mwg.device_table():
        out.append({"type": m.type, "n_reflect": m.n_reflect, "n_leapfrog": m.n_leapfrog,
                    "perturb_angle": m.perturb_angle, "flat_v_prior": m.flat_v_prior,
                    "max_vpa": m.max_vpa, "min_vpa": m.min_vpa,
                    "min_aspect": m.min_aspect})
    return out


def _engine(mode, R, K, N, n_extra, n_steps, seed, pressures, vpa=(8.0, 20.0),
            moves=MOVES, judge_seed=3, learned_kw=None):
    judge = MPNNBackend(init_mpnn(seed=judge_seed))
    cell = {"min_volume_per_atom": vpa[0], "max_volume_per_atom": vpa[1],
            "flat_V_prior": True}
    run = RunConfig(n_replicas=R, n_walkers=K, n_atoms=N, n_mcmc_steps=n_steps,
                    n_extra=n_extra, seed=seed, pressures=pressures, re_interval=0)
    learned = None
    if mode != "plain":
        kw = dict(mode=mode, ring_capacity=4096, train_batch=16, train_steps=0,
                  min_train=32, min_heldout=8, lr=2e-3, rmse_bound=0.05, delta_k=2.0,
                  baseline=BASE, params=init_surrogate(5))
        kw.update(learned_kw or {})
        learned = LearnedConfig(**kw)
    eng = NSEngine(Hamiltonian(judge, [EnsembleTerm(pressures[0])]), moves, run,
                   cell=cell, learned=learned)
    rng = np.random.default_rng(seed)
    eng.init_walkers(rng, vpa[0] * 1.001, vpa[1], True, species_p=[0.5, 0.5])
    return eng, judge


@pytest.mark.parametrize("mode", ["plain", "auto", "screened"])
def test_learned_walk_matches_oracle(mode):
    """One iteration from a host tape: select/clone, every sub-step's
    decision and counters, the walk statistics, the stored training records
    and the final population against the oracle ("auto" starts without a
    held-out record, so it walks plain and predicts with the surrogate)."""
    R, K, N, n_extra, T = 2, 4, 16, 2, 6
    print(f"learned walk parity mode={mode} R={R} K={K} N={N} steps={T}", flush=True)
    eng, judge = _engine(mode, R, K, N, n_extra, T, seed=11, pressures=[0.5, 1.0])
    pop0 = Population.from_engine_dict(eng.get_population())
    tape = make_tape(np.random.default_rng(7), 1, R, K, N, T, n_extra, [3, 1])
    out = eng.ns_iterations(1, tape=tape)
    st = eng.learned_stats()
    pop_gpu = eng.get_population()
    models = OW.Models(judge.params, init_surrogate(5), base=BASE)
    omode = {"plain": "plain", "auto": "plain+sur", "screened": "screened"}[mode]
    oout, ost, recs = OW.ns_iteration(pop0, models, eng.pressure, eng.step_sizes,
                                      _moves_dicts(eng), tape, 0, n_extra, omode,
                                      train_max_epa=50.0)
    assert np.array_equal(out["worst"][0], oout["worst"])
    assert np.array_equal(out["clone"][0], oout["clone"])
    assert np.array_equal(out["walk"][0], oout["walk"])
    assert np.array_equal(out["emax"][0], oout["emax"])
    print("  counters gpu", out["counters"][0].tolist(), flush=True)
    print("  counters ora", oout["counters"].tolist(), flush=True)
    assert np.array_equal(out["counters"][0], oout["counters"])
    for k in ("proposals", "stage1_pass", "accepted"):
        assert st[k] == ost[k], k
    assert st["exact_force_evals"] == ost["exact_fe"]
    assert st["exact_energy_evals"] == ost["exact_e"]
    assert st["surrogate_evals"] == ost["sur_e"]
    assert st["stored"] == (len(recs) if mode != "plain" else 0)
    if mode != "plain" and len(recs) >= 8:
        # error model: per-atom error e = a + c V/N by least squares, residual RMSE
        e = np.array([(us - u) / n for u, us, n, v in recs])
        v = np.array([v / n for u, us, n, v in recs])
        A = np.stack([np.ones_like(v), v], 1)
        res = e - A @ np.linalg.lstsq(A, e, rcond=None)[0]
        rmse = np.sqrt(np.mean(res * res))
        print(f"  held-out rmse gpu {st['heldout_rmse']:.6g} oracle {rmse:.6g}", flush=True)
        assert st["heldout_rmse"] == pytest.approx(rmse, rel=2e-2, abs=1e-6)
    L = pop_gpu["cell"][..., 0, 0].max()
    np.testing.assert_allclose(pop_gpu["positions"], pop0.pos, atol=1e-6 * L, rtol=0)
    np.testing.assert_allclose(pop_gpu["cell"].reshape(R, K, 9), pop0.cell, rtol=1e-12)
    np.testing.assert_allclose(pop_gpu["energy"], pop0.energy,
                               atol=1e-4 * (np.abs(pop0.energy).max() + N))


@pytest.mark.parametrize("mode", ["plain", "screened"])
def test_learned_walk_256_matches_oracle(mode):
    """The 256-atom message-passing workload at its size and density (the
    benchmark's judge, initial jittered lattice at rho = 0.9 and cell bounds;
    ~1.2e5 edges per configuration), one tape-driven iteration of a short
    walk against the oracle, with the tolerances of the small case."""
    import bench
    R, K, N, n_extra, T = 1, 3, 256, 1, 3
    print(f"learned walk parity at N={N}: mode={mode} K={K} steps={T}", flush=True)
    jp = init_mpnn(seed=11)
    judge = MPNNBackend(jp)
    base = (jp.core_eps, jp.core_sig, jp.core_rc)
    moves = [{"type": "galilean", "n_reflect": 3, "step_size": 0.1, "weight": 4},
             {"type": "volume", "step_size": 0.3, "weight": 1}]
    learned = None
    if mode != "plain":
        learned = LearnedConfig(mode=mode, ring_capacity=256, train_batch=16, train_steps=0,
                                min_train=32, min_heldout=8, lr=1e-3, rmse_bound=0.05,
                                delta_k=2.0, train_max_epa=50.0, baseline=base,
                                params=init_surrogate(1))
    run = RunConfig(n_replicas=R, n_walkers=K, n_atoms=N, n_mcmc_steps=T, n_extra=n_extra,
                    seed=42, pressures=[0.5], re_interval=0)
    eng = NSEngine(Hamiltonian(judge, [EnsembleTerm(0.5)]), moves, run, cell=bench.SUR_CELL,
                   learned=learned)
    eng.set_population(*bench.sur_population(R, K, N, seed=7))
    pop0 = Population.from_engine_dict(eng.get_population())
    tape = make_tape(np.random.default_rng(17), 1, R, K, N, T, n_extra, [4, 1])
    out = eng.ns_iterations(1, tape=tape)
    st = eng.learned_stats()
    pop_gpu = eng.get_population()
    models = OW.Models(judge.params, init_surrogate(1), base=base)
    oout, ost, recs = OW.ns_iteration(pop0, models, eng.pressure, eng.step_sizes,
                                      _moves_dicts(eng), tape, 0, n_extra, mode,
                                      train_max_epa=50.0)
    for k in ("worst", "clone", "walk", "emax"):
        assert np.array_equal(out[k][0], oout[k]), k
    print("  counters gpu", out["counters"][0].tolist(), flush=True)
    print("  counters ora", oout["counters"].tolist(), flush=True)
    assert np.array_equal(out["counters"][0], oout["counters"])
    for k in ("proposals", "stage1_pass", "accepted"):
        assert st[k] == ost[k], k
    assert st["exact_force_evals"] == ost["exact_fe"]
    assert st["exact_energy_evals"] == ost["exact_e"]
    assert st["surrogate_evals"] == ost["sur_e"]
    L = pop_gpu["cell"][..., 0, 0].max()
    np.testing.assert_allclose(pop_gpu["positions"], pop0.pos, atol=1e-6 * L, rtol=0)
    np.testing.assert_allclose(pop_gpu["cell"].reshape(R, K, 9), pop0.cell, rtol=1e-12)
    np.testing.assert_allclose(pop_gpu["energy"], pop0.energy,
                               atol=1e-4 * (np.abs(pop0.energy).max() + N))


# LJ judge with HMC + GMC (+ volume): the surrogate's pair baseline is the
# r^-12 part of LJ (eps = sigma = 1, rc = 2.5), the network learns the rest
LJ_BASE = (4.0, 1.0, 2.5)
LJ_MOVES = [{"type": "hmc", "n_leapfrog": 5, "step_size": 0.005, "weight": 2},
            {"type": "galilean", "n_reflect": 4, "step_size": 0.05, "weight": 2},
            {"type": "volume", "step_size": 0.3, "weight": 1}]


def _lj_engine(mode, R, K, N, n_extra, n_steps, seed, pressures, vpa=(1.2, 3.0),
               moves=LJ_MOVES, learned_kw=None):
    lj = LJBackend(1.0, 1.0, 2.5)
    cell = {"min_volume_per_atom": vpa[0], "max_volume_per_atom": vpa[1],
            "flat_V_prior": True}
    run = RunConfig(n_replicas=R, n_walkers=K, n_atoms=N, n_mcmc_steps=n_steps,
                    n_extra=n_extra, seed=seed, pressures=pressures, re_interval=0)
    kw = dict(mode=mode, ring_capacity=4096, train_batch=16, train_steps=0, min_train=32,
              min_heldout=8, lr=2e-3, rmse_bound=0.05, delta_k=2.0, baseline=LJ_BASE,
              params=init_surrogate(6, S=1))
    kw.update(learned_kw or {})
    eng = NSEngine(Hamiltonian(lj, [EnsembleTerm(pressures[0])]), moves, run, cell=cell,
                   learned=LearnedConfig(**kw))
    eng.init_walkers(np.random.default_rng(seed), vpa[0] * 1.001, vpa[1], True)
    return eng, lj


@pytest.mark.parametrize("N", [32, 128])
@pytest.mark.parametrize("mode", ["plain", "auto", "screened"])
def test_learned_walk_lj_hmc_matches_oracle(mode, N):
    """The learned walk with the FP64 pair potential as the judge and HMC,
    GMC and volume moves: one tape-driven iteration against the oracle (the
    judge is bit-exact with the oracle's, so decisions and counters must
    match exactly and positions to rounding); N = 128 is the LJ-128
    configuration's system size."""
    R, K, n_extra, T = 2, 4, 2, 8
    print(f"learned walk, LJ judge, mode={mode} R={R} K={K} N={N} steps={T}", flush=True)
    moves = LJ_MOVES
    if N > 64:
        # shorter leapfrog steps keep Delta H below the hard cut at 128 atoms
        moves = [dict(m, step_size=0.002) if m["type"] == "hmc" else m for m in LJ_MOVES]
    eng, lj = _lj_engine(mode, R, K, N, n_extra, T, seed=12, pressures=[0.5, 1.0],
                         moves=moves)
    if N > 64:
        # a jittered lattice (density 0.75) instead of uniform positions
        import bench
        pos, _, cell0 = bench.sur_population(R, K, N, seed=12, rho=0.75)
        eng.set_population(pos, np.zeros((R, K, N), np.int32), cell0)
    # relax the initial walkers with the pair-potential walk first, so that
    # HMC proposals are accepted and trajectories are well conditioned
    cell = {"min_volume_per_atom": 1.2, "max_volume_per_atom": 3.0, "flat_V_prior": True}
    pre = NSEngine(Hamiltonian(lj, [EnsembleTerm(0.5)]), moves,
                   RunConfig(n_replicas=R, n_walkers=K, n_atoms=N, n_mcmc_steps=T,
                             n_extra=n_extra, seed=13, pressures=[0.5, 1.0]), cell=cell)
    pre.set_population(*(eng.get_population()[k] for k in ("positions", "types", "cell")))
    pre.ns_iterations(40)
    rel = pre.get_population()
    eng.set_population(rel["positions"], rel["types"], rel["cell"], direction=rel["direction"])
    pop0 = Population.from_engine_dict(eng.get_population())
    tape = make_tape(np.random.default_rng(8), 1, R, K, N, T, n_extra, [2, 2, 1])
    out = eng.ns_iterations(1, tape=tape)
    st = eng.learned_stats()
    pop_gpu = eng.get_population()
    models = OW.Models(Pot.from_table(lj.table), init_surrogate(6, S=1), S=1, base=LJ_BASE)
    omode = {"plain": "plain", "auto": "plain+sur", "screened": "screened"}[mode]
    oout, ost, recs = OW.ns_iteration(pop0, models, eng.pressure, eng.step_sizes,
                                      _moves_dicts(eng), tape, 0, n_extra, omode,
                                      train_max_epa=50.0)
    print("  counters gpu", out["counters"][0].tolist(), flush=True)
    print("  counters ora", oout["counters"].tolist(), flush=True)
    assert np.array_equal(out["walk"][0], oout["walk"])
    assert np.array_equal(out["counters"][0], oout["counters"])
    hmc_acc = out["counters"][0][:, 0, 1].sum()
    assert hmc_acc > 0, "no HMC proposal accepted: the accept path is untested"
    for k in ("proposals", "stage1_pass", "accepted"):
        assert st[k] == ost[k], k
    assert st["exact_force_evals"] == ost["exact_fe"]
    assert st["exact_energy_evals"] == ost["exact_e"]
    L = pop_gpu["cell"][..., 0, 0].max()
    # screened: the surrogate's BF16 matrix-core forces (stated accuracy 2e-2 |F|
    # against the BF16-emulating oracle) steer HMC and GMC, so trajectories
    # agree to that order; plain/auto: FP64 judge forces, rounding only
    tol = 1e-9 if mode != "screened" else 1e-3
    np.testing.assert_allclose(pop_gpu["positions"], pop0.pos, atol=tol * L, rtol=0)
    np.testing.assert_allclose(pop_gpu["cell"].reshape(R, K, 9), pop0.cell, rtol=1e-12)
    if mode != "screened":
        np.testing.assert_allclose(pop_gpu["energy"], pop0.energy, rtol=1e-10, atol=1e-10)
    else:
        # positions differ at the surrogate-force level: every stored energy must
        # be the FP64 judge's enthalpy of the stored configuration itself
        from oracle.oracle import energy_forces as or_ef
        pot = Pot.from_table(lj.table)
        for r in range(R):
            for k in range(K):
                c = pop_gpu["cell"][r, k]
                _, Hr, _ = or_ef(pot, pop_gpu["positions"][r, k], pop_gpu["types"][r, k], c,
                                 pressure=eng.pressure[r], forces=False)
                assert pop_gpu["energy"][r, k] == pytest.approx(Hr, rel=1e-12, abs=1e-10)


def test_lj_hmc_screened_keeps_log_z():
    """LJ judge, HMC + GMC: log Z of 16 independent runs per walk, the
    screened walk with online training against the plain walk (2 sigma)."""
    R, K, N, n_iter = 16, 32, 8, 320
    moves = [{"type": "hmc", "n_leapfrog": 5, "step_size": 0.01, "weight": 1},
             {"type": "galilean", "n_reflect": 4, "step_size": 0.15, "weight": 2},
             {"type": "volume", "step_size": 1.0, "weight": 1}]
    res = {}
    for mode, seed in (("plain", 41), ("auto", 42)):
        print(f"log Z LJ: {mode}", flush=True)
        eng, _ = _lj_engine(mode, R, K, N, 0, 12, seed, [1.0] * R, vpa=(4.0, 40.0),
                            moves=moves,
                            learned_kw=dict(train_steps=4, min_train=64, rmse_bound=0.2))
        eng.ns_iterations(n_iter)
        res[mode] = (eng.final_log_evidence(), eng.learned_stats())
    se = lambda a: np.std(a, ddof=1) / np.sqrt(len(a))  # noqa: E731
    lz_p, lz_a = res["plain"][0], res["auto"][0]
    z = (lz_a.mean() - lz_p.mean()) / np.hypot(se(lz_a), se(lz_p))
    st = res["auto"][1]
    print(f"  plain log Z {lz_p.mean():.4f} +- {se(lz_p):.4f}; auto {lz_a.mean():.4f} +- "
          f"{se(lz_a):.4f}; z = {z:+.2f}; screened iters {st['iterations_screened']} "
          f"stage-1 pass {st['stage1_pass']} of {st['proposals']} accepted {st['accepted']} "
          f"rmse {st['heldout_rmse']:.4f}", flush=True)
    assert abs(z) < 2.0
    assert st["iterations_screened"] > 0


def _log_z_runs(mode, n_iter, seed, learned_kw=None):
    """R independent NS runs (one per replica, same pressure, no exchange)."""
    R, K, N = 16, 32, 8
    eng, _ = _engine(mode, R, K, N, 0, 12, seed, [1.0] * R, vpa=(16.0, 40.0),
                     moves=[{"type": "galilean", "n_reflect": 4, "step_size": 0.15,
                             "weight": 3},
                            {"type": "volume", "step_size": 1.0, "weight": 1}],
                     learned_kw=learned_kw)
    eng.ns_iterations(n_iter)
    lz = eng.final_log_evidence()
    st = eng.learned_stats() if mode != "plain" else None
    return lz, st


def test_screened_walk_keeps_log_z():
    """log Z of 16 independent runs per walk: the screened walk with an
    untrained surrogate, and the auto walk training online, each agree with
    the plain walk within 2 sigma (sigma from the run-to-run spread)."""
    n_iter = 320
    print("log Z: plain", flush=True)
    lz_p, _ = _log_z_runs("plain", n_iter, 21)
    res = {}
    print("log Z: screened (untrained surrogate)", flush=True)
    res["screened"] = _log_z_runs("screened", n_iter, 22)
    print("log Z: auto (online training)", flush=True)
    res["auto"] = _log_z_runs("auto", n_iter, 23,
                              dict(train_steps=4, min_train=64, rmse_bound=0.2))
    se = lambda a: np.std(a, ddof=1) / np.sqrt(len(a))  # noqa: E731
    print(f"  plain    log Z {lz_p.mean():.4f} +- {se(lz_p):.4f}", flush=True)
    for k, (lz, st) in res.items():
        z = (lz.mean() - lz_p.mean()) / np.hypot(se(lz), se(lz_p))
        print(f"  {k:8s} log Z {lz.mean():.4f} +- {se(lz):.4f}  z = {z:+.2f}  "
              f"screened iters {st['iterations_screened']} plain iters "
              f"{st['iterations_plain']} stage-1 pass {st['stage1_pass']} of "
              f"{st['proposals']} accepted {st['accepted']} rmse {st['heldout_rmse']:.4f}",
              flush=True)
        assert abs(z) < 2.0, k
    assert res["screened"][1]["iterations_screened"] == n_iter
    assert res["auto"][1]["iterations_screened"] > 0


def test_trained_surrogate_gain():
    """What the trained network buys.  From one population compressed by
    plain NS (so the contour binds), the same iterations are walked (a)
    without the network (plain: n_reflect judge forces per GMC step),
    (b) screened by the pair baseline with an untrained network and (c)
    screened with the network trained online on this workload.  Measured:
    accepted moves per judge evaluation and the held-out error.  The trained
    walk needs less than half the judge work per accepted move of the plain
    walk; training cuts the held-out error; at this size (N = 32) the
    baseline alone already screens well, so (c) only has to match (b)
    there (the 256-atom comparison is part of the benchmark)."""
    R, K, N = 4, 32, 32
    kw = dict(R=R, K=K, N=N, n_extra=K - 2, n_steps=10, pressures=[1.0] * R,
              vpa=(4.0, 10.0))
    print("gain: compressing with plain NS", flush=True)
    base, _ = _engine("plain", seed=31, **kw)
    base.ns_iterations(200)
    pop = base.get_population()
    print(f"  contour after 200 iterations {base.emax()}", flush=True)

    def from_pop(mode, learned_kw=None):
        eng, _ = _engine(mode, seed=32, learned_kw=learned_kw, **kw)
        eng.set_population(pop["positions"], pop["types"], pop["cell"],
                           direction=pop["direction"])
        return eng

    print("gain: training online (auto walk)", flush=True)
    eng_t = from_pop("auto", dict(train_steps=200, min_train=64, train_batch=32,
                                  rmse_bound=0.0))
    eng_t.ns_iterations(8)
    trained = eng_t.surrogate_params()
    print(f"  trained held-out rmse {eng_t.learned_stats()['heldout_rmse']:.4f}", flush=True)
    rows = {}
    eng = from_pop("plain")
    eng.ns_iterations(6)
    rows["plain"] = eng.learned_stats()
    for name, params in (("untrained", init_surrogate(5)), ("trained", trained)):
        print(f"gain: screened {name}", flush=True)
        eng = from_pop("screened", dict(params=params))
        eng.ns_iterations(6)
        rows[name] = eng.learned_stats()
    eff = {}
    for name, st in rows.items():
        judge = st["exact_force_evals"] + st["exact_energy_evals"]
        eff[name] = st["accepted"] / max(judge, 1)
        print(f"  {name:9s} proposals {st['proposals']} accepted {st['accepted']} "
              f"judge evals {judge} stage-1 pass {st['stage1_pass']} screened iters "
              f"{st['iterations_screened']} accepted per judge eval {eff[name]:.4f} "
              f"rmse {st['heldout_rmse']:.4f}", flush=True)
    assert eff["trained"] > 2.0 * eff["plain"]
    assert rows["trained"]["heldout_rmse"] < 0.7 * rows["untrained"]["heldout_rmse"]
    assert eff["trained"] >= 0.98 * eff["untrained"]

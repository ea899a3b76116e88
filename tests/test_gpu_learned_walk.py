"""The learned-model walk on the GPU.

* Parity: one tape-driven NS iteration of the plain walk (judge-steered)
  and of the screened walk (surrogate-steered, two-stage test) against the
  CPU restatement in oracle/learned_walk.py.  The judge runs in FP32 on the
  device and FP64 in the oracle (stated accuracy |dE| <= 1e-4 (|E| + N)),
  so accept decisions must agree wherever the margin |H - emax| exceeds
  that error, positions to 1e-6 relative to the box.
* Exactness: each screened move keeps its plain counterpart's target, so
  log Z over independent runs (engines) agrees with the plain walk's,
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
from logz_stats import report

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


M_ENG, R_ENG = 6, 8  # engines per walk, replicas per engine


def test_lj_hmc_screened_tracks_the_random_walk():
    """LJ judge: the screened HMC move is the exact constrained-prior move,
    so a walk of HMC + GMC + volume screened by the surrogate (auto, online
    training) must give the log Z of the plain random-walk moves (GMC +
    volume) on the same system: 6 engines per walk of 8 replicas (engines
    are the independent runs, the move schedule being shared inside one),
    24 sub-steps, Welch's t over engine means, |t| < 3.  This is the
    reference's own test for clearing its HMC move (hmc.py:21-35: accept
    rate and energy trace tracking the random-walk move).  The plain walk
    with the reference's HMC rule is printed beside them: its hard cut on
    the energy error leans toward low H and lifts log Z."""
    K, N, n_iter, T = 32, 8, 320, 24
    hmc = {"type": "hmc", "n_leapfrog": 5, "step_size": 0.01, "weight": 1}
    gal = {"type": "galilean", "n_reflect": 4, "step_size": 0.15, "weight": 2}
    vol = {"type": "volume", "step_size": 1.0, "weight": 1}
    arms = (("random_walk", "plain", [gal, vol], 410),
            ("screened_hmc", "auto", [hmc, gal, vol], 420),
            ("reference_hmc", "plain", [hmc, gal, vol], 430))
    means, st = {}, None
    for name, mode, moves, base in arms:
        means[name] = []
        for m in range(M_ENG):
            eng, _ = _lj_engine(mode, R_ENG, K, N, 0, T, base + m, [1.0] * R_ENG,
                                vpa=(4.0, 40.0), moves=moves,
                                learned_kw=dict(train_steps=4, min_train=64, rmse_bound=0.2))
            eng.ns_iterations(n_iter)
            means[name].append(float(eng.final_log_evidence().mean()))
            if mode == "auto":
                s_ = eng.learned_stats()
                keys = ("iterations_screened", "stage1_pass", "proposals", "accepted")
                st = {k: s_[k] for k in keys} if st is None else {k: st[k] + s_[k] for k in keys}
            eng.close()
    t = report("lj8_hmc", means["random_walk"], means["screened_hmc"],
               {"iterations": n_iter, "replicas_per_engine": R_ENG, "n_mcmc_steps": T, **st},
               labels=("random_walk", "screened_hmc"))
    report("lj8_reference_hmc", means["random_walk"], means["reference_hmc"],
           {"iterations": n_iter, "replicas_per_engine": R_ENG, "n_mcmc_steps": T},
           labels=("random_walk", "reference_hmc"))
    assert abs(t) < 3.0
    assert st["iterations_screened"] > 0


def _log_z_runs(mode, n_iter, seed, learned_kw=None):
    """Mean log Z of each of M_ENG engines of R_ENG replicas (one NS run
    each, same pressure, no exchange), and the summed learned counters."""
    K, N = 32, 8
    means, st = [], None
    for m in range(M_ENG):
        eng, _ = _engine(mode, R_ENG, K, N, 0, 12, seed + m, [1.0] * R_ENG,
                         vpa=(16.0, 40.0),
                         moves=[{"type": "galilean", "n_reflect": 4, "step_size": 0.15,
                                 "weight": 3},
                                {"type": "volume", "step_size": 1.0, "weight": 1}],
                         learned_kw=learned_kw)
        eng.ns_iterations(n_iter)
        means.append(float(eng.final_log_evidence().mean()))
        if mode != "plain":
            s_ = eng.learned_stats()
            keys = ("iterations_screened", "iterations_plain", "stage1_pass", "proposals",
                    "accepted")
            st = {k: s_[k] for k in keys} if st is None else {k: st[k] + s_[k] for k in keys}
            st["heldout_rmse"] = s_["heldout_rmse"]
        eng.close()
    return means, st


def test_screened_walk_keeps_log_z():
    """log Z per engine (6 engines of 8 replicas per walk): the screened
    walk with an untrained surrogate, and the auto walk training online,
    each agree with the plain walk (Welch's t over engine means, |t| < 3)."""
    n_iter = 320
    print("log Z: plain", flush=True)
    lz_p, _ = _log_z_runs("plain", n_iter, 210)
    res = {}
    print("log Z: screened (untrained surrogate)", flush=True)
    res["screened"] = _log_z_runs("screened", n_iter, 220)
    print("log Z: auto (online training)", flush=True)
    res["auto"] = _log_z_runs("auto", n_iter, 230,
                              dict(train_steps=4, min_train=64, rmse_bound=0.2))
    for k, (lz, st) in res.items():
        t = report(f"mpnn8_{k}", lz_p, lz, {"iterations": n_iter,
                                             "replicas_per_engine": R_ENG, **st})
        assert abs(t) < 3.0, k
    assert res["screened"][1]["iterations_screened"] == M_ENG * n_iter
    assert res["auto"][1]["iterations_screened"] > 0


def test_auto_falls_back_when_the_error_rises():
    """The auto walk screens while the surrogate's held-out error stays
    within the bound and walks plain once it exceeds it.  After screening
    has started, the network is replaced mid-run by one whose readout
    weights are random and 50 times the trained ones' scale.  A replaced
    network has no held-out record, so the next iteration walks plain and
    measures its error (still predicting every judge evaluation); that error
    exceeds the bound, so every further iteration walks plain too.  Putting
    the trained network back, the plain iteration measures an error within
    the bound and screening resumes: the bound, not the replacement,
    decides."""
    print("auto walk fallback after a bad surrogate is set mid-run", flush=True)
    R, K, N = 4, 32, 8
    bound = 0.2
    eng, _ = _engine("auto", R, K, N, 0, 12, 51, [1.0] * R, vpa=(16.0, 40.0),
                     moves=[{"type": "galilean", "n_reflect": 4, "step_size": 0.15,
                             "weight": 3},
                            {"type": "volume", "step_size": 1.0, "weight": 1}],
                     learned_kw=dict(train_steps=4, min_train=64, rmse_bound=bound))
    for it in range(200):
        eng.ns_iterations(1)
        if eng.learned_stats()["iterations_screened"] >= 3:
            break
    st0 = eng.learned_stats()
    print(f"  screening after {it + 1} iterations: held-out rmse {st0['heldout_rmse']:.4f}, "
          f"screened {st0['iterations_screened']}, plain {st0['iterations_plain']}", flush=True)
    assert st0["iterations_screened"] >= 3 and st0["heldout_rmse"] <= bound
    from jaxrens_amd.surrogate import N_PARAMS
    good = eng.surrogate_params().copy()
    bad = good.copy()
    # w3 [128], the per-atom readout, randomised at 50 times its largest
    # weight: errors that vary with the environment, which neither the error
    # model's per-atom offset nor a few online training steps take out
    w3 = slice(N_PARAMS - 129, N_PARAMS - 1)
    bad[w3] = np.random.default_rng(0).normal(scale=50.0 * np.abs(good[w3]).max(), size=128)
    eng.set_surrogate_params(bad, reset_adam=True)
    plain0 = st0["iterations_plain"]
    for k in range(4):
        eng.ns_iterations(1)
        st = eng.learned_stats()
        print(f"  bad network, iteration {k}: plain {st['iterations_plain']}, screened "
              f"{st['iterations_screened']}, held-out rmse {st['heldout_rmse']:.4f}", flush=True)
        assert st["iterations_plain"] == plain0 + k + 1, k
        assert st["iterations_screened"] == st0["iterations_screened"]
        assert st["heldout_rmse"] > bound
    eng.set_surrogate_params(good, reset_adam=True)
    eng.ns_iterations(1)
    st1 = eng.learned_stats()
    assert st1["iterations_plain"] == plain0 + 5 and st1["heldout_rmse"] <= bound
    eng.ns_iterations(1)
    st2 = eng.learned_stats()
    print(f"  trained network back: held-out rmse {st1['heldout_rmse']:.4f}, then screened "
          f"{st2['iterations_screened']}", flush=True)
    assert st2["iterations_screened"] == st0["iterations_screened"] + 1
    pop = eng.get_population()
    assert np.all(pop["energy"] < eng.emax()[:, None]), "live walkers inside the contour"


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


def _relaxed_lj_engine(mode, R, K, N, n_extra, T, seed):
    """The LJ-judge learned engine of the parity test, its walkers relaxed by
    the pair-potential walk first."""
    eng, lj = _lj_engine(mode, R, K, N, n_extra, T, seed=seed, pressures=[0.5, 1.0])
    cell = {"min_volume_per_atom": 1.2, "max_volume_per_atom": 3.0, "flat_V_prior": True}
    pre = NSEngine(Hamiltonian(lj, [EnsembleTerm(0.5)]), LJ_MOVES,
                   RunConfig(n_replicas=R, n_walkers=K, n_atoms=N, n_mcmc_steps=T,
                             n_extra=n_extra, seed=seed + 1, pressures=[0.5, 1.0]), cell=cell)
    pre.set_population(*(eng.get_population()[k] for k in ("positions", "types", "cell")))
    pre.ns_iterations(40)
    rel = pre.get_population()
    eng.set_population(rel["positions"], rel["types"], rel["cell"], direction=rel["direction"])
    return eng, lj


@pytest.mark.parametrize("mode", ["plain", "screened"])
def test_learned_trial_moves_match_oracle(mode):
    """Trial moves of the learned walk (the step-size adaptation's moves):
    HMC, GMC and volume from sampled walkers at trial step sizes, from host
    draws, against the oracle's moves; seven samples per replica take three
    passes of the walk's six chains.  The FP64 pair-potential judge is
    bit-exact with the oracle's, so every reason must match.  The population,
    the learned statistics and the training ring stay as they were."""
    R, K, N, n_extra, T = 2, 4, 32, 2, 4
    print(f"learned trial moves mode={mode} R={R} K={K} N={N}", flush=True)
    eng, lj = _relaxed_lj_engine(mode, R, K, N, n_extra, T, seed=21)
    before = eng.get_population()
    st0 = eng.learned_stats()
    pop = Population.from_engine_dict(before)
    models = OW.Models(Pot.from_table(lj.table), init_surrogate(6, S=1), S=1, base=LJ_BASE)
    omode = {"plain": "plain", "screened": "screened"}[mode]
    rng = np.random.default_rng(22)
    n = 7
    emax = pop.energy.max(axis=1)
    seen = set()
    for m in range(len(LJ_MOVES)):
        idx = rng.integers(0, K, size=(R, n)).astype(np.int32)
        nrm = rng.standard_normal((R, n, 3, N))
        uni = rng.random((R, n, 4))
        ts = eng.step_sizes[:, m] * np.array([1.0, 1.5])
        acc, rs = eng.trial_moves(m, ts, idx, emax=emax, normals=nrm, uniforms=uni)
        ors = OW.trial_moves(pop, models, eng.pressure, _moves_dicts(eng), m, ts, emax, idx,
                             nrm, uni, omode)
        print(f"  move {m}: gpu {rs.tolist()} oracle {ors.tolist()}", flush=True)
        assert np.array_equal(rs, ors), m
        assert np.array_equal(acc, ors == 0), m
        seen.update(np.unique(rs).tolist())
    assert 0 in seen and len(seen) > 1, seen  # accepted and rejected proposals
    after = eng.get_population()
    for k in before:
        assert np.array_equal(before[k], after[k]), k
    st1 = eng.learned_stats()
    for k in ("proposals", "accepted", "exact_force_evals", "exact_energy_evals",
              "surrogate_evals", "stored"):
        assert st1[k] == st0[k], k


def test_learned_run_with_adaptation():
    """run_ns with step-size adaptation on a learned handle (device draws):
    the trial moves run, the step sizes move, the population stays finite."""
    from jaxrens_amd.adaptation import AdaptationConfig
    from jaxrens_amd.run_loop import run_ns
    from jaxrens_amd.termination import IterationTermination
    print("learned run with adaptation", flush=True)
    eng, _ = _relaxed_lj_engine("auto", 2, 8, 32, 2, 4, seed=31)
    s0 = eng.step_sizes.copy()
    res = run_ns(eng, [IterationTermination(6)], adaptation=AdaptationConfig(
        adjust_interval=3, adjust_n_samples=10, adjust_max_rounds=4), seed=3)
    assert res.iterations == 6
    assert len(res.adaptation) >= 2
    assert not np.array_equal(eng.step_sizes, s0)
    assert np.all(np.isfinite(eng.get_population()["energy"]))


@pytest.mark.parametrize("mode", ["plain", "screened"])
def test_learned_burn_in_matches_oracle(mode):
    """The burn-in walk on a learned handle (init/burn_in.py:48-148): every
    walker takes three sub-steps of one schedule at fixed contours, in passes
    of the walk's six chains (eight walkers per replica: 3 + 3 + 2), against
    the oracle's walk with the FP64 pair-potential judge: counters exactly,
    the walked population to rounding; the learned statistics untouched.
    The surrogate (BF16 rows) matches its oracle only to rounding, and an
    untrained one drives HMC's leapfrog far enough that the rounding decides
    its stage-1 verdicts, so the screened walk takes gmc and volume moves
    only, well inside the contour."""
    R, K, N, n_extra, T, wl = 2, 8, 32, 2, 4, 3
    print(f"learned burn-in mode={mode} R={R} K={K} N={N} walklength={wl}", flush=True)
    eng, lj = _relaxed_lj_engine(mode, R, K, N, n_extra, T, seed=41)
    pop0 = Population.from_engine_dict(eng.get_population())
    st0 = eng.learned_stats()
    rng = np.random.default_rng(42)
    sched = rng.integers(0 if mode == "plain" else 1, len(LJ_MOVES), size=wl).astype(np.int32)
    nrm = rng.standard_normal((R, K, wl, 3, N))
    uni = rng.random((R, K, wl, 4))
    emax = pop0.energy.max(axis=1) + (0.5 if mode == "plain" else 50.0)
    ctr = eng.initial_walk(wl, emax, schedule=sched, normals=nrm, uniforms=uni)
    pop_gpu = eng.get_population()
    models = OW.Models(Pot.from_table(lj.table), init_surrogate(6, S=1), S=1, base=LJ_BASE)
    octr = OW.initial_walk(pop0, models, eng.pressure, eng.step_sizes, _moves_dicts(eng), emax,
                           sched, nrm, uni, mode)
    print("  counters gpu", ctr.tolist(), flush=True)
    print("  counters ora", octr.tolist(), flush=True)
    assert np.array_equal(ctr, octr)
    assert ctr[:, :, 1].sum() > 0  # some moves accepted
    L = pop_gpu["cell"][..., 0, 0].max()
    np.testing.assert_allclose(pop_gpu["positions"], pop0.pos, atol=1e-6 * L, rtol=0)
    np.testing.assert_allclose(pop_gpu["cell"].reshape(R, K, 9), pop0.cell, rtol=1e-12)
    np.testing.assert_allclose(pop_gpu["energy"], pop0.energy,
                               atol=1e-9 * (np.abs(pop0.energy).max() + N))
    st1 = eng.learned_stats()
    for k in ("proposals", "accepted", "exact_force_evals", "exact_energy_evals", "stored"):
        assert st1[k] == st0[k], k

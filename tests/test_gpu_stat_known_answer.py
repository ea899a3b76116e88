"""Known-answer checks of the whole engine + host outer loop on the GPU.

One atom in a periodic cubic box under pressure: U = 0 (minimum image
excludes self images), H = P V, and with a flat volume prior on
[V_lo, V_hi] the evidence is closed form,
  Z = (exp(-P V_lo) - exp(-P V_hi)) / (P (V_hi - V_lo)).
This exercises selection, cloning, the volume move with its cell guard,
adaptation by trial moves, evidence accumulation and the live-walker term
(the reference's known-answer test of this kind is the harmonic one,
tests/sampling/test_nested_sampling.py:124-166).
"""

import math

import numpy as np
import pytest

from jaxrens_amd.adaptation import AdaptationConfig
from jaxrens_amd.backends import EnsembleTerm, Hamiltonian, LJBackend
from jaxrens_amd.engine import NSEngine, RunConfig
from jaxrens_amd.run_loop import EnergyLogger, run_ns
from jaxrens_amd.termination import IterationTermination, PriorMassTermination

pytestmark = pytest.mark.gpu


def test_single_atom_npt_log_evidence(tmp_path):
    print("known answer: single atom NPT", flush=True)
    R, K = 4, 128
    P = np.array([0.5, 1.0, 2.0, 4.0])
    v_lo, v_hi = 1.0, 50.0
    lj = Hamiltonian(LJBackend(1.0, 1.0, 2.5), [EnsembleTerm(0.0)])
    run = RunConfig(n_replicas=R, n_walkers=K, n_atoms=1, n_mcmc_steps=12,
                    n_extra=0, seed=11, pressures=list(P), re_interval=0)
    cell = {"min_volume_per_atom": v_lo, "max_volume_per_atom": v_hi,
            "min_aspect_ratio": 0.5, "flat_V_prior": True}
    eng = NSEngine(lj, [{"type": "volume", "step_size": 2.0}], run, cell=cell)
    eng.init_walkers(np.random.default_rng(11), v_lo, v_hi, True)
    def progress(i, engine, out):
        if (i + 1) % 500 == 0:
            print(f"  iteration {i + 1}: emax {out['emax'][-1]}", flush=True)

    res = run_ns(eng, [PriorMassTermination(K, 1e-5), IterationTermination(24 * K)],
                 adaptation=AdaptationConfig(adjust_interval=25, adjust_max_rounds=8,
                                             step_size_max=50.0),
                 energies_prefix=str(tmp_path / "one"), callbacks=[progress], seed=3)
    exact = np.log((np.exp(-P * v_lo) - np.exp(-P * v_hi)) / (P * (v_hi - v_lo)))
    info = np.log(P * (v_hi - v_lo))  # ~ information, sets the NS error
    sigma = np.sqrt(info / K)
    err = res.log_evidence - exact
    print("log Z", res.log_evidence, "exact", exact, "sigma", sigma, "iters", res.iterations)
    assert np.all(np.abs(err) < 4 * sigma + 0.02), err
    # contour energies decrease and the dead volumes follow H = P V
    e = res.emax
    assert np.all(np.diff(e, axis=0) <= 0)
    np.testing.assert_allclose(e, P[None, :] * res.dead_volume, rtol=1e-12)
    # .energies files in the reference format
    log = EnergyLogger.read(tmp_path / "one.2.energies")
    assert log["n_walkers"] == K and log["n_atoms"] == 1
    assert np.array_equal(log["iterations"], np.arange(res.iterations))
    np.testing.assert_allclose(log["energies"], e[:, 2], rtol=1e-9)


def test_lj_contour_decreases_and_counters_add_up():
    print("LJ-8 run: monotone contour, counter identities", flush=True)
    lj = Hamiltonian(LJBackend(1.0, 1.0, 2.5), [EnsembleTerm(0.1)])
    moves = [{"type": "galilean", "n_reflect": 10, "step_size": 0.1, "weight": 4},
             {"type": "volume", "step_size": 0.3, "weight": 1}]
    cell = {"min_volume_per_atom": 0.125, "max_volume_per_atom": 64.0,
            "flat_V_prior": True}
    run = RunConfig(n_replicas=2, n_walkers=128, n_atoms=8, n_mcmc_steps=20,
                    n_extra=0, seed=42, pressures=[0.1, 0.1], re_interval=1)
    eng = NSEngine(lj, moves, run, cell=cell)
    eng.init_walkers(np.random.default_rng(42), 0.125, 64.0, True)
    res = run_ns(eng, [IterationTermination(400)],
                 adaptation=AdaptationConfig(adjust_interval=50, adjust_max_rounds=8),
                 block=1)
    assert res.iterations == 400
    assert np.all(np.diff(res.emax, axis=0) <= 0)
    c = res.counters
    # proposed = sum of the four reason buckets; evaluations per move
    assert np.array_equal(c[..., 0], c[..., 1:5].sum(-1))
    assert np.array_equal(c[:, 0, 5], 10 * c[:, 0, 0])
    assert np.array_equal(c[:, 1, 5], c[:, 1, 0])
    assert c[..., 0].sum() == 400 * 2 * 20
    # equal pressures: every attempted swap is accepted only below emax
    assert res.re_att.sum() == 399
    pop = eng.get_population()
    assert np.all(pop["energy"] <= res.emax[-1][:, None])

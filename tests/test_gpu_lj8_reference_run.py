"""The reference's own LJ-8 NPT run (tests/_assets/data/lj8_npt/output of the
reference; its contour log is the fixture tests/golden/lj8_npt.energies)
against an ensemble of independent runs of the same setup on the GPU.

The setup is the run's config snapshot (lj8_npt.config.snapshot.yaml): 8 LJ
atoms (eps = sigma = 1, cutoff 2.5, periodic), pressure 0.1, K = 32 live
walkers, 10 sub-steps per iteration, n_extra = 0, galilean (n_reflect 4, step
0.1, weight 2) and volume (step 0.2, weight 1) moves, cell volume per atom in
[1.5, 30] with the V^N prior and minimum aspect ratio 0.6, initial cells from
that prior with positions on a grid of spacing 1.0, one initial walk of 5
sub-steps at max(E) + 1 per atom after one step-size adaptation, then 200 NS
iterations adapting every 20 (50 samples, 15 rounds, window 0.3-0.5, factor
1.5, step cap 0.5).  The draws differ (the JAX PRNG stream cannot be
reproduced offline), so the check is distributional: at iterations 20, 40,
.., 200 the reference's contour lies inside the ensemble's 2.5-97.5 %
band, and the evidence of its 200 dead points lies within 2 sigma of the
ensemble's (sigma the run-to-run spread)."""

from pathlib import Path

import numpy as np
import pytest

from jaxrens_amd import prior_mass as pm
from jaxrens_amd.adaptation import AdaptationConfig
from jaxrens_amd.backends import EnsembleTerm, Hamiltonian, LJBackend
from jaxrens_amd.burn_in import initial_walk
from jaxrens_amd.engine import NSEngine, RunConfig
from jaxrens_amd.run_loop import EnergyLogger, run_ns
from jaxrens_amd.termination import IterationTermination

pytestmark = pytest.mark.gpu

FIXTURE = Path(__file__).parent / "golden" / "lj8_npt.energies"
N_RUNS, K, N, N_ITER = 128, 32, 8, 200
MOVES = [{"type": "galilean", "n_reflect": 4, "step_size": 0.1, "weight": 2.0},
         {"type": "volume", "step_size": 0.2, "weight": 1.0}]
CELL = {"min_volume_per_atom": 1.5, "max_volume_per_atom": 30.0, "min_aspect_ratio": 0.6,
        "flat_V_prior": False}
ADAPT = AdaptationConfig(adjust_interval=20, adjust_n_samples=50, adjust_max_rounds=15,
                         min_rate=0.3, max_rate=0.5, adjust_factor=1.5, step_size_max=0.5)


def _log_z(h):
    """Evidence (beta = 1) of the dead points h[0..n) of one run (K live, one
    cull per iteration), without the live share."""
    return np.logaddexp.reduce(pm.log_dead_weights(len(h), K, 1) - h)


def test_reference_lj8_run_lies_in_the_ensemble(tmp_path):
    print(f"LJ-8 NPT reference run vs {N_RUNS} independent GPU runs", flush=True)
    ref = EnergyLogger.read(FIXTURE)
    ref_h = np.asarray(ref["energies"], dtype=np.float64)
    assert len(ref_h) == N_ITER and np.array_equal(ref["iterations"], np.arange(N_ITER))
    run = RunConfig(n_replicas=N_RUNS, n_walkers=K, n_atoms=N, n_mcmc_steps=10, n_extra=0,
                    seed=20260, pressures=[0.1] * N_RUNS, re_interval=0)
    eng = NSEngine(Hamiltonian(LJBackend(1.0, 1.0, 2.5), [EnsembleTerm(0.1)]), MOVES, run,
                   cell=CELL)
    rng = np.random.default_rng(20261)
    eng.init_walkers(rng, 1.5, 30.0, flat_V_prior=False, grid_distance=1.0)
    initial_walk(eng, n_walks=1, walklength=5, adjust_interval=1, emax_offset_per_atom=1.0,
                 adaptation=ADAPT, rng=rng)
    res = run_ns(eng, [IterationTermination(N_ITER)], adaptation=ADAPT, seed=20262)
    assert res.iterations == N_ITER
    h = np.asarray(res.emax)  # (iterations, runs)
    assert h.shape == (N_ITER, N_RUNS)
    bad = []
    for t in range(19, N_ITER, 20):
        lo, hi = np.percentile(h[t], [2.5, 97.5])
        rank = 100.0 * np.mean(h[t] < ref_h[t])
        print(f"  iteration {t + 1:3d}: reference H {ref_h[t]:9.4f}  ensemble 2.5-97.5% "
              f"[{lo:9.4f}, {hi:9.4f}]  median {np.median(h[t]):9.4f}  rank {rank:5.1f}%",
              flush=True)
        if not lo <= ref_h[t] <= hi:
            bad.append(t + 1)
    lz = np.array([_log_z(h[:, r]) for r in range(N_RUNS)])
    lz_ref = _log_z(ref_h)
    z = (lz_ref - lz.mean()) / lz.std(ddof=1)
    print(f"  log Z of 200 dead points: reference {lz_ref:.4f}, ensemble {lz.mean():.4f} +- "
          f"{lz.std(ddof=1):.4f} (run-to-run), z = {z:+.2f}", flush=True)
    assert not bad, f"reference contour outside the 95 % band at iterations {bad}"
    assert abs(z) < 2.0

"""``NSEngine.ns_step``: the reference-shaped per-replica ``info`` dict
(``sampling/nested_sampling.py:653-685``) that callers and callbacks read by
key, for n_cull = 1 and n_cull > 1, against the block outputs of
``ns_iterations`` on an identical engine and the CPU oracle."""

import numpy as np
import pytest

from jaxrens_amd.backends import EnsembleTerm, Hamiltonian, LJBackend
from jaxrens_amd.draws import make_tape
from jaxrens_amd.engine import NSEngine, RunConfig

pytestmark = pytest.mark.gpu

# the keys of the reference's info dict
REF_KEYS = {"emax", "hmax", "worst_idx", "clone_idx", "worst_indices", "clone_indices",
            "acceptance_rate", "dead_walker", "dead_walkers", "n_accepted_per_move",
            "n_proposed_per_move", "reject_reason_counts_per_move",
            "n_evaluations_per_move", "n_grad_evaluations_per_move"}
MOVES = [{"type": "galilean", "n_reflect": 4, "step_size": 0.1, "weight": 4},
         {"type": "volume", "step_size": 0.3, "weight": 1}]
CELL = {"min_volume_per_atom": 0.125, "max_volume_per_atom": 64.0, "flat_V_prior": True}


def _engine(n_cull, n_extra):
    run = RunConfig(n_replicas=2, n_walkers=16, n_atoms=12, n_mcmc_steps=6, n_extra=n_extra,
                    n_cull=n_cull, seed=4, pressures=[0.2, 1.0], re_interval=1)
    eng = NSEngine(Hamiltonian(LJBackend(1.0, 1.0, 2.5), [EnsembleTerm(0.2)]), MOVES, run,
                   cell=CELL)
    eng.init_walkers(np.random.default_rng(4), 0.125, 64.0, True)
    return eng


@pytest.mark.parametrize("n_cull,n_extra", [(1, 5), (3, 4)])
def test_ns_step_info_keys_and_values(oracle, n_cull, n_extra):
    print(f"ns_step info n_cull={n_cull} n_extra={n_extra}", flush=True)
    from oracle.oracle import Population
    a, b = _engine(n_cull, n_extra), _engine(n_cull, n_extra)
    n_walk = n_cull + n_extra
    tape = make_tape(np.random.default_rng(6), 2, 2, 16, 12, 6, n_extra, a.mwg.move_probs,
                     n_cull=n_cull)
    opop = Population.from_engine_dict(a.get_population())
    for it in range(2):
        infos = a.ns_step(tape=tape.iteration(it))
        out = b.ns_iterations(1, tape=tape.iteration(it), dead_records=True)
        o = oracle.ns_iteration(opop, oracle.Pot.from_table(a.model.table), a.pressure, a.mu,
                                a.step_sizes, a.mwg.as_oracle_moves(), 6, n_extra, tape,
                                it=it, do_re=it > 0, n_cull=n_cull)
        assert len(infos) == 2
        for r, info in enumerate(infos):
            assert REF_KEYS <= set(info), REF_KEYS - set(info)
            w = np.reshape(out["worst"][0, r], (n_cull,))
            c = np.reshape(out["clone"][0, r], (n_cull,))
            assert info["emax"] == info["hmax"] == out["emax"][0, r] == o["emax"][r]
            assert info["worst_indices"].shape == (n_cull,)
            assert np.array_equal(info["worst_indices"], w)
            assert np.array_equal(info["clone_indices"], c)
            assert info["worst_idx"] == w[-1] and info["clone_idx"] == c[-1]
            rr = info["reject_reason_counts_per_move"]
            assert rr.shape == (a.n_moves, 4)
            assert np.array_equal(rr.sum(1), info["n_proposed_per_move"])
            assert np.array_equal(rr[:, 0], info["n_accepted_per_move"])
            assert np.array_equal(out["counters"][0, r], o["counters"][r])
            assert np.array_equal(info["n_evaluations_per_move"], o["counters"][r][:, 5])
            assert np.array_equal(info["n_grad_evaluations_per_move"], o["counters"][r][:, 6])
            assert info["n_proposed_per_move"].sum() == n_walk * 6
            assert info["acceptance_rate"] == pytest.approx(
                info["n_accepted_per_move"].sum() / (n_walk * 6))
            dw, dws = info["dead_walker"], info["dead_walkers"]
            assert dws["energy"].shape == (n_cull,)
            assert np.all(np.diff(dws["energy"]) <= 0), "dead points in descending energy"
            assert dw["energy"] == dws["energy"][-1] == info["emax"]
            np.testing.assert_array_equal(dw["positions"], dws["positions"][-1])
            assert dws["positions"].shape == (n_cull, 12, 3)
            assert dws["cell"].shape == (n_cull, 3, 3)
        pa, pb = a.get_population(), b.get_population()
        for k in pa:
            assert np.array_equal(pa[k], pb[k]), k

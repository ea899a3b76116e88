"""Host logic of the outer loop against known answers: the prior-mass ladder
(exact rational arithmetic of pymatnest's Eq. 3 and order-statistics closed
forms), the stopping rules and the step-size bisection rule (the
reference's hand-calculated cases, tests/golden/host_logic_cases.json written
by tests/golden/make_host_vectors.py)."""

from __future__ import annotations

import json
import math
from pathlib import Path

import numpy as np
import pytest

from jaxrens_amd import prior_mass as pm
from jaxrens_amd.adaptation import FLOOR_THRESHOLD, process_rate
from jaxrens_amd.termination import (EnergyTermination, IterationTermination,
                                     PriorMassTermination, TempTermination, check_any)

CASES = json.loads((Path(__file__).parent / "golden" / "host_logic_cases.json").read_text())
PM_IDS = [f"K{c['n_live']}-c{c['n_cull']}" for c in CASES["prior_mass"]]


# ------------------------------------------------------------- prior mass
@pytest.mark.parametrize("case", CASES["prior_mass"], ids=PM_IDS)
def test_ladder_matches_exact_rationals(case):
    K, c, n = case["n_live"], case["n_cull"], case["n_dead"]
    np.testing.assert_allclose(pm.log_dead_weights(n, K, c), case["log_dead_weights"],
                               rtol=0, atol=1e-12)
    np.testing.assert_allclose(pm.log_prior_mass(n, K, c), case["log_prior_mass"],
                               rtol=0, atol=1e-12)
    assert pm.log_alpha(K, c) == pytest.approx(case["log_alpha"], abs=1e-15)


@pytest.mark.parametrize("K,c", [(10, 1), (600, 1), (50, 3), (600, 10), (7, 3)])
def test_kth_largest_order_statistic(K, c):
    """Within an iteration the k-th cull leaves the expected mass of the k-th
    largest of K uniforms, (K + 1 - k) / (K + 1); iterations multiply it by
    alpha."""
    lx = pm.log_prior_mass(3 * c, K, c)
    for it in range(3):
        for k in range(1, c + 1):
            want = it * math.log((K + 1 - c) / (K + 1)) + math.log((K + 1 - k) / (K + 1))
            assert lx[it * c + k - 1] == pytest.approx(want, abs=1e-12)


def test_single_cull_is_geometric():
    K = 600
    n = np.arange(50)
    np.testing.assert_allclose(pm.log_prior_mass(50, K, 1), (n + 1) * math.log(K / (K + 1)),
                               rtol=0, atol=1e-12)


@pytest.mark.parametrize("K,c", [(10, 1), (600, 1), (50, 3), (600, 10), (7, 3)])
@pytest.mark.parametrize("n_iter", [0, 1, 13])
def test_dead_plus_live_conserves_mass(K, c, n_iter):
    """Dead weights plus the K live shares sum to X_0 = K / (K + 1)."""
    d = n_iter * c
    dead = np.logaddexp.reduce(pm.log_dead_weights(d, K, c)) if d else -np.inf
    total = np.logaddexp(dead, float(pm.log_live_weight(d, K, c)) + math.log(K))
    assert total == pytest.approx(math.log(K / (K + 1)), abs=1e-10)


@pytest.mark.parametrize("K,c", [(10, 0), (10, 10), (10, 11)])
def test_invalid_n_cull_rejected(K, c):
    with pytest.raises(ValueError):
        pm.log_alpha(K, c)


@pytest.mark.parametrize("K,c", [(16, 1), (50, 3)])
def test_per_iteration_weights_follow_the_ladder(K, c):
    w = pm.dead_log_weight(np.arange(20), K, c)
    np.testing.assert_allclose(w.ravel(), pm.log_dead_weights(20 * c, K, c), atol=1e-12)


def test_final_evidence_of_a_flat_population():
    """With every live walker at enthalpy h the live share is X_D exp(-h)."""
    K, c, it, h = 32, 2, 7, 1.5
    live = np.full((1, K), h)
    got = pm.final_log_evidence(np.array([-3.0]), it, live, K, c)
    x_d = pm.log_prior_mass(it * c + 1, K, c)[-1]
    assert got[0] == pytest.approx(np.logaddexp(-3.0, x_d - h), abs=1e-12)


# ------------------------------------------------------------ termination
def _rule(name, args):
    return {"iteration": IterationTermination, "prior_mass": PriorMassTermination,
            "energy": EnergyTermination}[name](**args)


@pytest.mark.parametrize("case", CASES["termination"])
def test_stopping_rules_hand_cases(case):
    rule = _rule(case["rule"], case["args"])
    if "log_evidence" in case:
        rule.update_evidence(case["log_evidence"])
    for i, h, want in case["checks"]:
        assert rule.check(i, h) is want, (case, i, h)


@pytest.mark.parametrize("case", CASES["first_trigger"])
def test_prior_mass_first_trigger(case):
    rule = _rule(case["rule"], case["args"])
    rule.update_evidence(case["log_evidence"])
    first = next(i for i in range(100_000) if rule.check(i, 0.0))
    assert first == case["first"]


def test_more_culls_contract_faster():
    first = {c["args"]["n_cull"]: c["first"] for c in CASES["first_trigger"]}
    assert first[10] * 8 < first[1]


def test_messages_name_their_parameters():
    assert "10" in IterationTermination(10).message()
    assert "-2.0" in EnergyTermination(-2.0).message()


def test_temp_weight_is_the_contour_dead_point():
    K, c, i = 40, 4, 17
    t = TempTermination(n_walkers=K, target_temp=1000.0, n_cull=c)
    t.check(i, 0.0)
    assert t.log_Z_term_last == pytest.approx(pm.log_dead_weights((i + 1) * c, K, c)[-1])


def _first(rule, energies):
    return next((i for i, e in enumerate(energies) if rule.check(i, e)), None)


def test_temp_converges_on_decaying_and_flat_contours():
    decay = [10.0 * math.exp(-i / 50.0) for i in range(5000)]
    assert (_first(TempTermination(100, 300.0, threshold=10.0), decay) or 0) > 0
    assert _first(TempTermination(50, 1000.0, threshold=10.0), [5.0] * 10000) is not None
    assert _first(TempTermination(100, 500.0, n_cull=3, threshold=10.0), decay) is not None
    t5 = _first(TempTermination(100, 300.0, threshold=5.0), decay)
    t15 = _first(TempTermination(100, 300.0, threshold=15.0), decay)
    assert t5 < t15


def test_temp_state_and_message():
    t = TempTermination(n_walkers=100, target_temp=300.0)
    assert t.log_Z_term_max == -math.inf and t.log_Z_term_last == -math.inf
    t.check(0, 5.0)
    assert t.log_Z_term_max > -math.inf and t.log_Z_term_last > -math.inf
    _first(t, [10.0 * math.exp(-i / 50.0) for i in range(1, 3000)])
    assert "300.0" in t.message() and "converged" in t.message()


def test_temp_fires_exactly_when_the_gap_passes_the_threshold():
    """A term sequence built from the weights: with h_i chosen so that the
    term falls by exactly 1 per iteration after a peak at i = 3, a threshold
    of 4.5 fires at i = 8, not before."""
    K, T = 20, 1000.0
    t = TempTermination(K, T, threshold=4.5)
    la, w = pm.log_alpha(K, 1), pm.iteration_offsets(K, 1)[1][-1]
    want = [0, 1, 2, 3, 2, 1, 0, -1, -2, -3]
    fired = [t.check(i, (i * la + w - want[i]) / t.beta) for i in range(len(want))]
    assert fired.index(True) == 8


def test_check_any_order_and_fallback():
    it, en = IterationTermination(1000), EnergyTermination(5.0)
    assert check_any([it, en], 10, 3.0) == (True, en.message())
    assert check_any([it, EnergyTermination(-100.0)], 10, 0.0) == (False, "")
    assert check_any([], 0, 0.0)[0] is False
    it5 = IterationTermination(5)
    assert check_any([it5, EnergyTermination(10.0)], 10, 0.0) == (True, it5.message())
    temp, it100 = TempTermination(100, 0.001, threshold=50.0), IterationTermination(100)
    for i in range(200):
        done, msg = check_any([temp, it100], i, 10.0 - i * 0.001)
        if done:
            break
    assert i == 99 and "iteration" in msg


# ------------------------------------------------------- step-size rule
@pytest.mark.parametrize("row", CASES["process_rate"]["rows"])
def test_bisection_rule_hand_cases(row):
    w = CASES["process_rate"]["window"]
    ss, conv, cap, floor, high, low = process_rate(*row["inp"], w["min_rate"], w["max_rate"],
                                                   w["adjust_factor"], w["max_step_size"])
    assert float(ss) == pytest.approx(row["step"], rel=1e-15)
    assert bool(conv) is row["converged"]
    for k, v in (("cap", cap), ("floor", floor), ("high", high), ("low", low)):
        if k in row:
            assert bool(v) is row[k], k


def test_bisection_window_is_open():
    """A rate on the lower edge is neither inside nor below the window: the
    step is kept and the replica is not converged."""
    ss, conv, cap, floor, high, low = process_rate(0.2, 0.3, 0.3, -1.0, 0.2, 0.7, 1.5, 10.0)
    assert float(ss) == 0.3 and not conv and not high and not low
    ss, conv, *_ = process_rate(0.7, 0.3, 0.3, -1.0, 0.2, 0.7, 1.5, 10.0)
    assert float(ss) == pytest.approx(0.45) and not conv
    assert FLOOR_THRESHOLD == 1e-7


def test_bisection_is_vectorised_per_replica():
    rate = np.array([0.4, 0.1, 0.9, 0.8])
    ss, conv, *_ = process_rate(rate, np.full(4, 0.5), np.full(4, 0.3),
                                np.array([-1.0, -1.0, -1.0, 0.15]), 0.2, 0.7, 1.5, 10.0)
    np.testing.assert_allclose(ss, [0.5, 0.5 / 1.5, 0.75, 0.3])
    assert conv.tolist() == [True, False, False, True]


# ------------------------------------------------------- initial cells
def test_cell_shape_walk_keeps_volume_and_aspect():
    from jaxrens_amd.engine import _aspect, cell_shape_walk
    rng = np.random.default_rng(5)
    c0 = np.eye(3) * 6.0
    for _ in range(20):
        c = cell_shape_walk(rng, c0, 50, 0.05, 0.6)
        v = abs(np.linalg.det(c))
        assert v == pytest.approx(216.0, rel=1e-12)
        assert _aspect(c, v) >= 0.6
        assert not np.allclose(c, c0)  # the shape moved
    # a cube has aspect 1; a flat cell is refused
    assert _aspect(c0, 216.0) == pytest.approx(1.0)
    assert cell_shape_walk(rng, c0, 200, 5.0, 0.999) is not None

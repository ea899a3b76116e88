"""Writes tests/golden/host_logic_cases.json: known-answer cases for the host
logic of the outer loop (prior-mass weights, stopping rules, the step-size
bisection rule).  Nothing here imports or runs the reference.

* Prior mass: the dead-point weights of pymatnest's convention (Martiniani
  et al., PRX 4, 031034 (2014), Eq. 3) evaluated in exact rational
  arithmetic, X_n = prod_{m <= n} (K - m mod P) / (K + 1 - m mod P) and
  a_n = X_n / (K + 1 - (n + 1) mod P), then logged; for the (K, P) pairs
  and ladder lengths the reference's own tests use
  (tests/test_prior_mass.py:35-47).
* Stopping rules and the bisection rule: the hand-calculated cases of the
  reference's tests (tests/sampling/test_termination.py:32-104,
  tests/sampling/adaptation/test_stepsize_handler.py:74-200) transcribed as
  inputs and expected outputs; where those tests assert only a relation
  (e.g. "the step shrinks"), the value the rule fixes is added next to it.

Run from the repository root: python tests/golden/make_host_vectors.py
"""

from __future__ import annotations

import json
import math
from fractions import Fraction
from pathlib import Path

OUT = Path(__file__).resolve().parent / "host_logic_cases.json"


def exact_ladder(K: int, P: int, n_dead: int):
    """log X_n and log a_n, n < n_dead, from exact rationals."""
    x = Fraction(1)
    log_x, log_a = [], []
    for n in range(n_dead):
        j = n % P
        x *= Fraction(K - j, K + 1 - j)
        a = x / (K + 1 - (n + 1) % P)
        log_x.append(math.log(x.numerator) - math.log(x.denominator))
        log_a.append(math.log(a.numerator) - math.log(a.denominator))
    return log_x, log_a


def prior_mass_cases():
    cases = []
    for K, P in [(10, 1), (600, 1), (50, 3), (600, 10), (7, 3)]:
        n_dead = 41 * P
        lx, la = exact_ladder(K, P, n_dead)
        cases.append({"n_live": K, "n_cull": P, "n_dead": n_dead,
                      "log_prior_mass": lx, "log_dead_weights": la,
                      "log_alpha": math.log(K + 1 - P) - math.log(K + 1)})
    return cases


def termination_cases():
    """Each case: a rule, optional evidence, and (iteration, emax, expected)
    checks applied in order to one instance."""
    return [
        {"rule": "iteration", "args": {"max_iterations": 100},
         "checks": [[98, 0.0, False], [99, 0.0, True]]},
        {"rule": "prior_mass", "args": {"n_live": 100},
         "checks": [[0, 0.0, False], [50, 0.0, False]]},
        {"rule": "prior_mass", "args": {"n_live": 50, "threshold": 0.1},
         "log_evidence": -5.0, "checks": [[1000, 0.0, True]]},
        # guard i * n_cull <= K holds at i = 10; at i = 11 the bound
        # 12 log(46/51) < 0 = log Z + log 1 fires
        {"rule": "prior_mass", "args": {"n_live": 50, "threshold": 1.0, "n_cull": 5},
         "log_evidence": 0.0, "checks": [[10, 0.0, False], [11, 0.0, True]]},
        {"rule": "energy", "args": {"min_energy": -1.0},
         "checks": [[0, 0.5, False], [1, -0.5, False], [2, -1.5, True]]},
        {"rule": "energy", "args": {"min_energy": 0.0},
         "checks": [[0, 0.0, False], [1, 0.001, False], [2, -0.001, True]]},
    ]


def first_trigger_cases():
    """Sequences run until the rule first fires: the reference asserts the
    relations; ``first`` is the iteration the rule's inequality fixes."""
    out = []
    # prior mass, constant contour 0, log Z = -1, threshold 1e-3: fires at the
    # first i with i n_cull > K and (i + 1) log alpha < -1 + log 1e-3
    for P in (1, 10):
        K = 100
        la = math.log(K + 1 - P) - math.log(K + 1)
        bound = -1.0 + math.log(1e-3)
        i = K // P + 1
        i = max(i, math.floor(bound / la - 1.0) + 1)
        while not ((i + 1) * la < bound):
            i += 1
        out.append({"rule": "prior_mass", "args": {"n_live": K, "threshold": 1e-3, "n_cull": P},
                    "log_evidence": -1.0, "emax": "zero", "first": i})
    return out


def process_rate_cases():
    """(rate, step, step_prev, rate_prev), window (0.2, 0.7), factor 1.5,
    cap 10 (the reference test's arguments), expected outputs."""
    w = {"min_rate": 0.2, "max_rate": 0.7, "adjust_factor": 1.5, "max_step_size": 10.0}
    rows = [
        # within the window: keep, converged, no flags
        dict(inp=[0.4, 0.1, 0.1, -1.0], step=0.1, converged=True, cap=False,
             floor=False, high=False, low=False),
        # below the window: shrink by 1 / 1.5 (the reference asserts step < 1)
        dict(inp=[0.1, 1.0, 1.0, -1.0], step=1.0 / 1.5, converged=False, cap=False,
             floor=False, high=False, low=True),
        # above: grow by 1.5 (the reference asserts step > 0.1)
        dict(inp=[0.9, 0.1, 0.1, -1.0], step=0.1 * 1.5, converged=False, cap=False,
             floor=False, high=True, low=False),
        # 9 * 1.5 = 13.5 reaches the cap 10
        dict(inp=[0.95, 9.0, 9.0, -1.0], step=10.0, converged=False, cap=True,
             floor=False, high=True, low=False),
        # 1e-20 / 1.5 falls below the floor 1e-7
        dict(inp=[0.01, 1e-20, 1e-20, -1.0], step=1e-7, converged=False, cap=False,
             floor=True, high=False, low=True),
        # bracket: the previous rate 0.15 is nearer the centre 0.45 -> its step 0.3
        dict(inp=[0.8, 0.5, 0.3, 0.15], step=0.3, converged=True),
        # bracket: the current rate 0.75 is nearer -> the current step 0.5
        dict(inp=[0.75, 0.5, 0.3, 0.05], step=0.5, converged=True),
    ]
    return {"window": w, "rows": rows}


def main():
    data = {"prior_mass": prior_mass_cases(), "termination": termination_cases(),
            "first_trigger": first_trigger_cases(), "process_rate": process_rate_cases()}
    OUT.write_text(json.dumps(data, indent=1) + "\n")
    print(f"wrote {OUT}")


if __name__ == "__main__":
    main()

"""Stopping rules of the outer nested-sampling loop.

Each rule is called once per iteration as ``check(i, h)`` with the 0-based
iteration index i and that iteration's contour h (the enthalpy of its last,
lowest-energy dead point), returns True when the run should stop, and
explains itself through ``message()``.  ``check_any`` asks the rules in
order and reports the first that fires.  The rules and their conditions are
the reference's (``sampling/termination.py:55-246``):

* iteration limit: iteration i is the last when i + 1 >= max_iterations;
* prior mass (Skilling's bound): the evidence still enclosed by the contour
  is at most X_(i+1) L_max with L_max <= exp(-h); stop once
  log X + (-h) < log Z + log(threshold), with log X = (i + 1) log alpha,
  but only after every live walker has on average been replaced
  (i n_cull > K);
* temperature (Baldock et al. 2017): the partition-function terms
  w_i exp(-beta h_i) at a target temperature, beta = 1 / (k_B T); stop once
  the newest term is more than ``threshold`` (natural-log units) below the
  largest seen, w_i being the weight of the iteration's last dead point;
* energy floor: stop once h < min_energy.
"""

from __future__ import annotations

import math

from .prior_mass import iteration_offsets, log_alpha

# Boltzmann constant in eV/K (the reference's value)
KB_EV = 8.6173324e-5


class IterationTermination:
    """Stop at the last of ``max_iterations`` iterations."""

    def __init__(self, max_iterations: int):
        self.max_iterations = int(max_iterations)

    def check(self, iteration: int, emax: float) -> bool:
        return iteration + 1 >= self.max_iterations

    def message(self) -> str:
        return f"iteration limit reached: max_iterations {self.max_iterations}"


class PriorMassTermination:
    """Stop when the prior mass left inside the contour can add at most a
    fraction ``threshold`` of the evidence accumulated so far (the caller
    passes log Z through ``update_evidence`` before each check)."""

    def __init__(self, n_live: int, threshold: float = 1e-3, n_cull: int = 1):
        self.n_live, self.n_cull, self.threshold = int(n_live), int(n_cull), threshold
        self._la = log_alpha(n_live, n_cull)
        self._log_thr = math.log(threshold)
        self._log_evidence = -math.inf

    def update_evidence(self, log_evidence: float) -> None:
        self._log_evidence = float(log_evidence)

    def remaining_bound(self, iteration: int, emax: float) -> float:
        """log of X_(i+1) exp(-h), the bound on the evidence still to come."""
        return (iteration + 1) * self._la - float(emax)

    def check(self, iteration: int, emax: float) -> bool:
        if iteration * self.n_cull <= self.n_live:
            return False
        return self.remaining_bound(iteration, emax) < self._log_evidence + self._log_thr

    def message(self) -> str:
        return (f"remaining prior mass negligible: bound below {self.threshold:g} "
                f"of the evidence")


class TempTermination:
    """Stop when the partition function at ``target_temp`` (kelvin) has
    converged: its newest term lies more than ``threshold`` below the
    largest term so far (natural-log units)."""

    def __init__(self, n_walkers: int, target_temp: float, n_cull: int = 1,
                 threshold: float = 10.0):
        self.n_walkers, self.n_cull = int(n_walkers), int(n_cull)
        self.target_temp, self.threshold = float(target_temp), float(threshold)
        self.beta = 1.0 / (KB_EV * self.target_temp)
        self._la = log_alpha(n_walkers, n_cull)
        # weight offset of the iteration's last (contour) dead point
        self._w_last = float(iteration_offsets(n_walkers, n_cull)[1][-1])
        self.log_Z_term_max = -math.inf
        self.log_Z_term_last = -math.inf

    def term(self, iteration: int, emax: float) -> float:
        """log(w_i exp(-beta h_i)) of the iteration's contour dead point."""
        return iteration * self._la + self._w_last - self.beta * float(emax)

    def check(self, iteration: int, emax: float) -> bool:
        t = self.term(iteration, emax)
        self.log_Z_term_last = t
        self.log_Z_term_max = max(self.log_Z_term_max, t)
        return t < self.log_Z_term_max - self.threshold

    def message(self) -> str:
        gap = self.log_Z_term_max - self.log_Z_term_last
        return (f"Z(T={self.target_temp:.1f} K) converged: newest term {gap:.1f} "
                f"below the largest")


class EnergyTermination:
    """Stop once the contour falls below ``min_energy``."""

    def __init__(self, min_energy: float):
        self.min_energy = float(min_energy)

    def check(self, iteration: int, emax: float) -> bool:
        return float(emax) < self.min_energy

    def message(self) -> str:
        return f"contour below min_energy {self.min_energy}"


def check_any(criteria, iteration: int, emax: float) -> tuple[bool, str]:
    """(True, message of the first rule that fires) or (False, "")."""
    fired = next((c for c in criteria if c.check(iteration, emax)), None)
    return (False, "") if fired is None else (True, fired.message())

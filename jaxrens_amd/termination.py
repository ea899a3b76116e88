"""Termination criteria of the outer loop (``sampling/termination.py:55-246``)."""

from __future__ import annotations

import math

from .prior_mass import log_alpha


class IterationTermination:
    def __init__(self, max_iterations: int):
        self.max_iterations = max_iterations

    def check(self, iteration: int, emax: float) -> bool:
        return iteration >= self.max_iterations - 1

    def message(self) -> str:
        return f"Reached max iterations ({self.max_iterations})"


class PriorMassTermination:
    """Stop when ``X_i * exp(-emax) < threshold * Z`` (after ``i*n_cull > K``)."""

    def __init__(self, n_live: int, threshold: float = 1e-3, n_cull: int = 1):
        self.n_live = n_live
        self.n_cull = n_cull
        self.threshold = threshold
        self._log_threshold = math.log(threshold)
        self._log_alpha = log_alpha(n_live, n_cull)
        self._log_evidence = -math.inf

    def update_evidence(self, log_evidence: float) -> None:
        self._log_evidence = float(log_evidence)

    def check(self, iteration: int, emax: float) -> bool:
        if iteration * self.n_cull <= self.n_live:
            return False
        log_X = (iteration + 1) * self._log_alpha
        return log_X - float(emax) < self._log_evidence + self._log_threshold

    def message(self) -> str:
        return "Prior mass negligible compared to evidence"


class EnergyTermination:
    def __init__(self, min_energy: float):
        self.min_energy = min_energy

    def check(self, iteration: int, emax: float) -> bool:
        return float(emax) < self.min_energy

    def message(self) -> str:
        return f"E_max below target energy ({self.min_energy})"


def check_any(criteria, iteration: int, emax: float) -> tuple[bool, str]:
    for c in criteria:
        if c.check(iteration, emax):
            return True, c.message()
    return False, ""

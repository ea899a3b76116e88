"""Step-size bisection of the outer loop.

``process_rate`` restates ``sampling/adaptation/stepsize_handler.py:35-102``
(``_process_rate_jax``) and ``adjust_step_size`` the loop of :110-434, with
the trial moves run on the device by ``jr_trial_moves`` (one trial move per
walker sampled with replacement, population untouched).
"""

from __future__ import annotations

from dataclasses import dataclass

import numpy as np

FLOOR_THRESHOLD = 1.0e-7


def process_rate(rate, step_size, step_size_prev, rate_prev, min_rate,
                 max_rate, adjust_factor, max_step_size):
    """Vectorised over replicas.  ``rate_prev < 0`` means no previous rate.

    Returns ``(new_ss, converged, cap_hit, floor_hit, too_high, too_low)``.
    """
    rate = np.asarray(rate, dtype=np.float64)
    step_size = np.asarray(step_size, dtype=np.float64)
    step_size_prev = np.asarray(step_size_prev, dtype=np.float64)
    rate_prev = np.asarray(rate_prev, dtype=np.float64)
    target = 0.5 * (min_rate + max_rate)
    in_window = (rate > min_rate) & (rate < max_rate)
    has_prev = rate_prev >= 0.0
    brackets = (has_prev & (np.minimum(rate, rate_prev) < min_rate)
                & (np.maximum(rate, rate_prev) > max_rate))
    prev_closer = np.abs(rate_prev - target) < np.abs(rate - target)
    bracket_ss = np.where(prev_closer, step_size_prev, step_size)
    too_high = rate >= max_rate
    too_low = rate < min_rate
    scale = np.where(too_low, 1.0 / adjust_factor,
                     np.where(too_high, adjust_factor, 1.0))
    proposed = step_size * scale
    cap_hit = proposed >= max_step_size
    floor_hit = proposed <= FLOOR_THRESHOLD
    adjusted = np.clip(proposed, FLOOR_THRESHOLD, max_step_size)
    new_ss = np.where(in_window, step_size, np.where(brackets, bracket_ss, adjusted))
    return new_ss, in_window | brackets, cap_hit, floor_hit, too_high, too_low


@dataclass
class AdaptationConfig:
    """``adaptation`` section defaults (``lj8_npt.config.snapshot.yaml``)."""

    adjust_interval: int = 20
    adjust_n_samples: int = 50
    adjust_max_rounds: int = 15
    min_rate: float = 0.3
    max_rate: float = 0.5
    adjust_factor: float = 1.5
    step_size_max: float = 0.5


def adjust_step_size(engine, move_idx: int, step_size, emax, rng, cfg: AdaptationConfig):
    """Bisection for one move on every replica at once.

    Each round samples ``n_samples`` walkers per replica with replacement,
    runs one trial move at the current trial step size, and updates replicas
    that have not converged (the reference's while-loop runs until all
    replicas converge or ``max_rounds``; converged replicas keep their step).
    Returns ``(new_step_size (R,), info dict)``.
    """
    R, K = engine.n_replicas, engine.n_walkers
    ss = np.asarray(step_size, dtype=np.float64).copy()
    ss_prev = ss.copy()
    rate_prev = np.full(R, -1.0)
    converged = np.zeros(R, dtype=bool)
    counts = np.zeros((R, 4), dtype=np.int64)
    cap_hits = np.zeros(R, np.int64)
    floor_hits = np.zeros(R, np.int64)
    saw_high = np.zeros(R, bool)
    saw_low = np.zeros(R, bool)
    rounds = np.zeros(R, np.int64)
    n_evals = 0
    rate = np.zeros(R)
    for _ in range(cfg.adjust_max_rounds):
        if converged.all():
            break
        idx = rng.integers(0, K, size=(R, cfg.adjust_n_samples), dtype=np.int32)
        acc, reason = engine.trial_moves(move_idx, ss, idx, emax=emax)
        active = ~converged
        r_rate = acc.mean(axis=1)
        new_ss, conv, cap, flo, hi, lo = process_rate(
            r_rate, ss, ss_prev, rate_prev, cfg.min_rate, cfg.max_rate,
            cfg.adjust_factor, cfg.step_size_max)
        for c in range(4):
            counts[active, c] = (reason[active] == c).sum(axis=1)
        cap_hits += (cap & active)
        floor_hits += (flo & active)
        saw_high |= (hi & active)
        saw_low |= (lo & active)
        rounds += active
        rate = np.where(active, r_rate, rate)
        ss_prev = np.where(active, ss, ss_prev)
        rate_prev = np.where(active, r_rate, rate_prev)
        ss = np.where(active, new_ss, ss)
        converged |= (conv & active)
        n_evals += int(active.sum()) * cfg.adjust_n_samples
    info = {
        "final_rate": rate,
        "final_counts": counts,
        "n_rounds": rounds,
        "converged": converged,
        "cap_hits": cap_hits,
        "floor_hits": floor_hits,
        "bracket_detected": saw_high & saw_low,
        "n_trial_moves": n_evals,
    }
    return ss, info

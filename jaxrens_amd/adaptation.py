"""Step-size bisection of the outer loop.

``process_rate`` follows the decision rule of
``sampling/adaptation/stepsize_handler.py:35-102`` (``_process_rate_jax``) and
``adjust_step_size`` the loop of :110-434, with
the trial moves run on the device by ``jr_trial_moves`` (one trial move per
walker sampled with replacement, population untouched).
"""

from __future__ import annotations

from dataclasses import dataclass

import numpy as np

FLOOR_THRESHOLD = 1.0e-7


def process_rate(rate, step_size, step_size_prev, rate_prev, min_rate,
                 max_rate, adjust_factor, max_step_size):
    """One bisection decision per replica from the acceptance rate measured
    at ``step_size`` (``rate_prev`` < 0: no earlier round).  The rules, the
    first that applies wins, fix the adapted step-size sequence:

    1. the rate lies strictly inside (min_rate, max_rate): keep the step;
       converged;
    2. the current and previous rates straddle the window (the lower one
       below min_rate, the higher one above max_rate): converged at the step
       whose rate lies nearer the window centre (the current one on a tie);
    3. otherwise scale the step by ``adjust_factor`` when the rate is at or
       above max_rate, by its inverse when below min_rate (unchanged on
       rate == min_rate) and clip to [FLOOR_THRESHOLD, max_step_size].

    Flags: cap_hit / floor_hit when the unclipped step of rule 3 reaches the
    cap / the floor, too_high = rate >= max_rate, too_low = rate < min_rate.
    Returns ``(new_step, converged, cap_hit, floor_hit, too_high, too_low)``,
    vectorised over replicas."""
    r = np.asarray(rate, dtype=np.float64)
    s_now = np.asarray(step_size, dtype=np.float64)
    s_old = np.asarray(step_size_prev, dtype=np.float64)
    r_old = np.asarray(rate_prev, dtype=np.float64)
    centre = 0.5 * (min_rate + max_rate)
    too_high, too_low = r >= max_rate, r < min_rate
    inside = (r > min_rate) & (r < max_rate)
    lo_rate, hi_rate = np.minimum(r, r_old), np.maximum(r, r_old)
    straddle = (r_old >= 0.0) & (lo_rate < min_rate) & (hi_rate > max_rate)
    nearer_old = np.abs(r_old - centre) < np.abs(r - centre)
    factor = np.where(too_low, 1.0 / adjust_factor, np.where(too_high, adjust_factor, 1.0))
    scaled = s_now * factor
    new_step = np.select([inside, straddle],
                         [s_now, np.where(nearer_old, s_old, s_now)],
                         default=np.clip(scaled, FLOOR_THRESHOLD, max_step_size))
    return (new_step, inside | straddle, scaled >= max_step_size,
            scaled <= FLOOR_THRESHOLD, too_high, too_low)


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

"""Prior-mass ladder and dead-point weights (pymatnest convention).

Restates ``prior_mass.py:38-108`` of the reference: with K live walkers and
``n_cull`` culled per iteration, dead point n (0-based) shrinks the prior mass
by ``(K - n % n_cull) / (K + 1 - n % n_cull)``.
"""

from __future__ import annotations

import numpy as np


def _shrinkage_terms(n_terms: int, n_live: int, n_cull: int) -> np.ndarray:
    j = np.mod(np.arange(n_terms, dtype=np.int64), n_cull).astype(np.float64)
    n = float(n_live)
    return np.log(n - j) - np.log(n + 1.0 - j)


def check_n_cull(n_live: int, n_cull: int) -> None:
    if int(n_cull) < 1:
        raise ValueError(f"n_cull must be >= 1, got {n_cull}")
    if int(n_cull) >= int(n_live):
        raise ValueError(f"n_cull={n_cull} must be smaller than n_live={n_live}")


def log_alpha(n_live: int, n_cull: int = 1) -> float:
    check_n_cull(n_live, n_cull)
    return float(np.log(n_live + 1.0 - n_cull) - np.log(n_live + 1.0))


def log_prior_mass(n_dead, n_live: int, n_cull: int = 1) -> np.ndarray:
    check_n_cull(n_live, n_cull)
    return np.cumsum(_shrinkage_terms(int(n_dead), n_live, n_cull))


def log_dead_weights(n_dead, n_live: int, n_cull: int = 1) -> np.ndarray:
    n_dead = int(n_dead)
    log_x = log_prior_mass(n_dead, n_live, n_cull)
    j_next = np.mod(np.arange(1, n_dead + 1, dtype=np.int64), n_cull)
    return log_x - np.log(float(n_live) + 1.0 - j_next.astype(np.float64))


def log_live_weight(n_dead, n_live: int, n_cull: int = 1) -> np.ndarray:
    check_n_cull(n_live, n_cull)
    log_x = np.sum(_shrinkage_terms(int(n_dead) + 1, n_live, n_cull))
    return np.asarray(log_x - np.log(float(n_live)), dtype=np.float64)


def iteration_offsets(n_live: int, n_cull: int = 1):
    x_offset = log_prior_mass(n_cull, n_live, n_cull)
    w_offset = log_dead_weights(n_cull, n_live, n_cull)
    return x_offset, w_offset


def dead_log_weight(iteration, n_live: int, n_cull: int = 1) -> np.ndarray:
    """``sampling/nested_sampling.py:215-226`` in float64."""
    _, w_off = iteration_offsets(n_live, n_cull)
    return np.asarray(iteration, dtype=np.float64)[..., None] * log_alpha(
        n_live, n_cull) + w_off


def logaddexp_accumulate(log_z: np.ndarray, terms: np.ndarray) -> np.ndarray:
    """``log Z <- logaddexp(log Z, logsumexp(terms))`` over the last axis."""
    m = np.max(terms, axis=-1)
    with np.errstate(invalid="ignore"):
        lse = m + np.log(np.sum(np.exp(terms - m[..., None]), axis=-1))
    lse = np.where(np.isneginf(m), -np.inf, lse)
    return np.logaddexp(log_z, lse)


def final_log_evidence(log_evidence, iteration, live_potentials, n_live: int,
                       n_cull: int = 1) -> np.ndarray:
    """Live walkers' share (``sampling/nested_sampling.py:229-252``)."""
    log_x = (np.asarray(iteration, dtype=np.float64) * log_alpha(n_live, n_cull)
             + (np.log(n_live) - np.log(n_live + 1.0)))
    lp = -np.asarray(live_potentials, dtype=np.float64)
    m = lp.max(axis=-1)
    lse = m + np.log(np.exp(lp - m[..., None]).sum(-1))
    return np.logaddexp(log_evidence, log_x - np.log(n_live) + lse)

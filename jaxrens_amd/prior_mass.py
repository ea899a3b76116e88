"""Prior-mass ladder and dead-point weights of nested sampling with K live
walkers and n_cull walkers culled per iteration (pymatnest convention,
Martiniani et al., PRX 4, 031034 (2014), Eq. 3; the reference's
``prior_mass.py`` implements the same convention).

The mathematics.  Of K points uniform in prior mass, the k-th largest has
expected enclosed mass (K + 1 - k) / (K + 1).  Removing the culls of one
iteration one at a time in descending energy, cull j (j = 0 .. n_cull - 1)
therefore shrinks the enclosed mass by

    t_j = (K - j) / (K + 1 - j) = 1 - 1 / (K + 1 - j),

and a whole iteration by alpha = prod_j t_j = (K + 1 - n_cull) / (K + 1).
Dead point n (global index, n = i n_cull + j) is the contour of
X_n = prod_{m <= n} t_(m mod n_cull) = alpha^i X_j, and its weight is the
mass between its contour and the next one,

    a_n = X_n - X_(n+1) = X_n / (K + 1 - (n + 1) mod n_cull).

After D dead points the K live walkers share X_D (the product over m <= D)
equally, so dead and live weights always sum to X_0 = K / (K + 1).

Everything is float64 and evaluated per index in closed form
(i log alpha + offset), with log t_j = log1p(-1 / (K + 1 - j)), so long runs
accumulate no rounding along the ladder.
"""

from __future__ import annotations

import numpy as np


def check_n_cull(n_live: int, n_cull: int) -> None:
    """A ladder needs 1 <= n_cull < n_live (the clones come from survivors)."""
    if not 1 <= int(n_cull) < int(n_live):
        raise ValueError(f"need 1 <= n_cull < n_live; got n_cull={n_cull}, "
                         f"n_live={n_live}")


def _log_t(j, n_live: int) -> np.ndarray:
    """log t_j for cull positions j within an iteration."""
    return np.log1p(-1.0 / (float(n_live) + 1.0 - np.asarray(j, dtype=np.float64)))


def log_alpha(n_live: int, n_cull: int = 1) -> float:
    """log of the per-iteration contraction (K + 1 - n_cull) / (K + 1)."""
    check_n_cull(n_live, n_cull)
    return float(np.log1p(-float(n_cull) / (float(n_live) + 1.0)))


def iteration_offsets(n_live: int, n_cull: int = 1) -> tuple[np.ndarray, np.ndarray]:
    """(x_off, w_off), each [n_cull]: for dead point j of iteration i,
    log X = i log alpha + x_off[j] and log a = i log alpha + w_off[j]."""
    check_n_cull(n_live, n_cull)
    j = np.arange(n_cull)
    x_off = np.cumsum(_log_t(j, n_live))
    # the next contour's cull position: j + 1 within the iteration, or 0 of
    # the next one (whose t_0 = K / (K + 1) gives 1 - t_0 = 1 / (K + 1))
    w_off = x_off - np.log(float(n_live) + 1.0 - ((j + 1) % n_cull))
    return x_off, w_off


def _ladder(n, n_live: int, n_cull: int, which: int) -> np.ndarray:
    n = np.asarray(n, dtype=np.int64)
    offs = iteration_offsets(n_live, n_cull)[which]
    return (n // n_cull).astype(np.float64) * log_alpha(n_live, n_cull) + offs[n % n_cull]


def log_prior_mass(n_dead, n_live: int, n_cull: int = 1) -> np.ndarray:
    """log X_n of dead points n = 0 .. n_dead - 1."""
    return _ladder(np.arange(int(n_dead)), n_live, n_cull, 0)


def log_dead_weights(n_dead, n_live: int, n_cull: int = 1) -> np.ndarray:
    """log a_n of dead points n = 0 .. n_dead - 1."""
    return _ladder(np.arange(int(n_dead)), n_live, n_cull, 1)


def log_live_weight(n_dead, n_live: int, n_cull: int = 1) -> np.ndarray:
    """log(X_D / K): the weight each live walker carries after D = n_dead
    dead points (0-d array)."""
    x_d = _ladder(int(n_dead), n_live, n_cull, 0)
    return np.asarray(x_d - np.log(float(n_live)), dtype=np.float64)


def dead_log_weight(iteration, n_live: int, n_cull: int = 1) -> np.ndarray:
    """log a of the n_cull dead points of each iteration in ``iteration``
    (shape iteration.shape + (n_cull,)); the reference forms the same weights
    in float32 (``sampling/nested_sampling.py:215-226``)."""
    it = np.asarray(iteration, dtype=np.float64)
    return it[..., None] * log_alpha(n_live, n_cull) + iteration_offsets(n_live, n_cull)[1]


def logaddexp_accumulate(log_z: np.ndarray, terms: np.ndarray) -> np.ndarray:
    """log(exp(log_z) + sum(exp(terms))) over the last axis of ``terms``;
    an all -inf row leaves log_z unchanged."""
    top = np.max(terms, axis=-1)
    safe = np.where(np.isfinite(top), top, 0.0)
    with np.errstate(divide="ignore"):
        s = safe + np.log(np.sum(np.exp(terms - safe[..., None]), axis=-1))
    return np.logaddexp(log_z, np.where(np.isneginf(top), -np.inf, s))


def final_log_evidence(log_evidence, iteration, live_potentials, n_live: int,
                       n_cull: int = 1) -> np.ndarray:
    """log Z with the live walkers' share added after ``iteration`` full
    iterations (``sampling/nested_sampling.py:229-252``): each live walker
    carries X_D / K with D = iteration n_cull."""
    d = np.asarray(iteration, dtype=np.int64) * int(n_cull)
    log_w = _ladder(d, n_live, n_cull, 0) - np.log(float(n_live))
    return logaddexp_accumulate(np.asarray(log_evidence, dtype=np.float64),
                                np.asarray(log_w)[..., None]
                                - np.asarray(live_potentials, dtype=np.float64))

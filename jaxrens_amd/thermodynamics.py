"""Thermodynamics from a nested-sampling run: evidence, partition function,
heat capacity, expectations.

Restates ``postprocess/thermodynamics.py:44-295`` in float64 NumPy on the
dead points (``.energies`` row order) and the final live walkers, with the
pymatnest weights of :mod:`jaxrens_amd.prior_mass`:

  log Z(beta)  = logsumexp_i (log w_i - beta E_i) over dead + live points,
  <O>(beta)    = sum_i p_i O_i,  p_i = w_i exp(-beta E_i) / Z(beta),
  C_v(beta)    = beta^2 sum_i p_i (E_i - <E>)^2   (deviation form, :179-230),
  F(beta)      = -log Z(beta) / beta.

Energies of an NPT run are the enthalpies H = U + P V the engine streams.
"""

from __future__ import annotations

import numpy as np

from .prior_mass import log_dead_weights, log_live_weight


def _logsumexp(a: np.ndarray, axis: int = -1, keepdims: bool = False) -> np.ndarray:
    m = np.max(a, axis=axis, keepdims=True)
    m = np.where(np.isfinite(m), m, 0.0)
    s = np.log(np.sum(np.exp(a - m), axis=axis, keepdims=True)) + m
    return s if keepdims else np.squeeze(s, axis=axis)


def _merged(dead_energies, live_energies, n_live: int, n_cull: int):
    dead = np.asarray(dead_energies, dtype=np.float64)
    live = np.asarray(live_energies, dtype=np.float64)
    n_dead = dead.shape[-1]
    lw_dead = log_dead_weights(n_dead, n_live, n_cull)
    lw_live = np.full(live.shape[-1], float(log_live_weight(n_dead, n_live, n_cull)))
    live = np.broadcast_to(live, dead.shape[:-1] + live.shape[-1:])
    return (np.concatenate([dead, live], axis=-1),
            np.concatenate([lw_dead, lw_live]))


def log_evidence(dead_energies, live_energies, n_live: int, n_cull: int = 1):
    """``log Z`` at beta = 1 (``thermodynamics.py:89-122``)."""
    return partition_function(1.0, dead_energies, live_energies, n_live, n_cull)


def partition_function(beta, dead_energies, live_energies, n_live: int,
                       n_cull: int = 1):
    """``log Z(beta)`` over the merged dead and live points (``:125-175``)."""
    E, lw = _merged(dead_energies, live_energies, n_live, n_cull)
    return _logsumexp(lw - float(beta) * E, axis=-1)


def _probabilities(beta, E, lw):
    lu = lw - float(beta) * E
    return np.exp(lu - _logsumexp(lu, axis=-1, keepdims=True))


def heat_capacity(beta, dead_energies, live_energies, n_live: int, n_cull: int = 1):
    """``C_v(beta) = beta^2 Var(E)`` in the deviation form (``:178-230``)."""
    E, lw = _merged(dead_energies, live_energies, n_live, n_cull)
    p = _probabilities(beta, E, lw)
    mean = np.sum(p * E, axis=-1, keepdims=True)
    return float(beta) ** 2 * np.sum(p * (E - mean) ** 2, axis=-1)


def expectation(observable_values, beta, dead_energies, live_energies, n_live: int,
                n_cull: int = 1):
    """Thermal average of an observable given per dead + live point (``:233-282``)."""
    E, lw = _merged(dead_energies, live_energies, n_live, n_cull)
    p = _probabilities(beta, E, lw)
    return np.sum(p * np.broadcast_to(np.asarray(observable_values, np.float64), E.shape),
                  axis=-1)


def free_energy(beta, log_z_beta):
    """``F = -log Z(beta) / beta`` (``:285-295``)."""
    return -np.asarray(log_z_beta, dtype=np.float64) / np.asarray(beta, dtype=np.float64)


def heat_capacity_curve(temperatures, dead_energies, live_energies, n_live: int,
                        n_cull: int = 1) -> np.ndarray:
    """``C_v`` at each temperature (k_B = 1)."""
    return np.array([heat_capacity(1.0 / t, dead_energies, live_energies, n_live, n_cull)
                     for t in np.asarray(temperatures, dtype=np.float64)])

"""Host-side draw tapes (explicit random inputs of one or more iterations).

A tape holds every random number an NS iteration consumes, in the layouts
of ``jr_tape`` (include/jaxrens_amd.h).  The device draws the same layouts
itself with Philox4x32-10 when no tape is given; a host tape lets the HIP
path and the CPU oracle consume identical inputs.
"""

from __future__ import annotations

from dataclasses import dataclass

import numpy as np


@dataclass
class Tape:
    sel_u: np.ndarray      # (n_iter, R, K)
    clone_u: np.ndarray    # (n_iter, R)
    extra_u: np.ndarray    # (n_iter, R, K)
    schedule: np.ndarray   # (n_iter, n_steps) int32
    normals: np.ndarray    # (n_iter, R, n_walk, n_steps, 3, Nmax), n_walk = n_cull + n_extra
    uniforms: np.ndarray   # (n_iter, R, n_walk, n_steps, 4)
    re_u: np.ndarray       # (n_iter, n_cycles, 2, max(R//2, 1), 2)

    @property
    def n_iter(self) -> int:
        return self.sel_u.shape[0]

    def iteration(self, it: int) -> "Tape":
        s = slice(it, it + 1)
        return Tape(self.sel_u[s], self.clone_u[s], self.extra_u[s],
                    self.schedule[s], self.normals[s], self.uniforms[s],
                    self.re_u[s])


def make_tape(rng: np.random.Generator, n_iter: int, R: int, K: int, Nmax: int,
              n_steps: int, n_extra: int, move_probs, n_swap_cycles: int = 1,
              ties: bool = False, n_cull: int = 1) -> Tape:
    """``ties``: round the selection noise to quarters so that equal noise
    occurs; with ``n_cull > 1`` also the energies' tie-break then reaches
    the index rule when energies are equal too."""
    n_walk = n_cull + n_extra
    probs = np.asarray(move_probs, dtype=np.float64)
    sched = rng.choice(len(probs), size=(n_iter, n_steps), p=probs / probs.sum())
    sel = rng.random((n_iter, R, K))
    if ties:  # exercise the first-index rule on equal noise
        sel = np.round(sel * 4) / 4
    return Tape(
        sel_u=sel,
        clone_u=rng.random((n_iter, R)),
        extra_u=rng.random((n_iter, R, K)),
        schedule=np.ascontiguousarray(sched, dtype=np.int32),
        normals=rng.standard_normal((n_iter, R, n_walk, n_steps, 3, Nmax)),
        uniforms=rng.random((n_iter, R, n_walk, n_steps, 4)),
        re_u=rng.random((n_iter, max(n_swap_cycles, 1), 2, max(R // 2, 1), 2)),
    )

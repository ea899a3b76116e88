"""Fixed-contour burn-in of the live walkers before nested sampling.

``initial_walk`` restates ``init/burn_in.py:154-407``: the contour of
replica r is held at ``Emax_r = max_k E_rk + emax_offset_per_atom * N_r``
(``:255-258``, ``reduce_emax`` per replica) for ``n_walks`` walks of
``walklength`` MWG sub-steps of every walker, each walk one
``jr_initial_walk`` launch on the device.  Before walk ``i`` with
``i % adjust_interval == 0`` the step sizes are re-tuned by the outer
loop's bisection at that fixed contour (``:347-372``); alchemical moves keep
their step, as in the outer loop.  The population is walked in place; the
NS state (contour, evidence, iteration) is left unchanged.
"""

from __future__ import annotations

import numpy as np

from .adaptation import AdaptationConfig, adjust_step_size
from .moves import STEPLESS


def burn_in_emax(engine, emax_offset_per_atom: float) -> np.ndarray:
    """Per-replica contour of the burn-in (``burn_in.py:255-258``)."""
    energy = engine.get_population()["energy"]
    return energy.max(axis=1) + float(emax_offset_per_atom) * engine.n_atoms_rep


def initial_walk(engine, *, n_walks: int, walklength: int, adjust_interval: int,
                 emax_offset_per_atom: float, adaptation: AdaptationConfig | None = None,
                 rng: np.random.Generator | None = None, tapes=None,
                 max_chunk: int = 0) -> dict:
    """Walk every live walker at a fixed contour.

    ``tapes``: optional list of ``(schedule, normals, uniforms)`` host draws,
    one per walk (``jr_initial_walk`` layouts), for replays; device draws
    otherwise.  Returns ``{"emax", "counters", "adaptation"}``.
    """
    if n_walks < 0 or walklength < 1:
        raise ValueError("need n_walks >= 0 and walklength >= 1")
    R = engine.n_replicas
    counters = np.zeros((R, engine.n_moves, 7), np.int64)
    emax = burn_in_emax(engine, emax_offset_per_atom)
    log = []
    if n_walks == 0:
        return {"emax": emax, "counters": counters, "adaptation": log}
    if tapes is not None and len(tapes) < n_walks:
        raise ValueError("one tape per walk is needed")
    adapt = adaptation is not None and adjust_interval > 0
    rng = rng if rng is not None else np.random.default_rng(0)
    adaptable = [m for m, spec in enumerate(engine.mwg.moves)
                 if spec.type not in STEPLESS]
    for i in range(n_walks):
        if adapt and i % adjust_interval == 0:
            ss = engine.step_sizes.copy()
            rec = {"walk": i}
            for m in adaptable:
                new, info = adjust_step_size(engine, m, ss[:, m], emax, rng, adaptation)
                ss[:, m] = new
                rec[engine.mwg.names[m]] = info
            engine.set_run_params(step_sizes=ss)
            log.append(rec)
        if tapes is None:
            counters += engine.initial_walk(walklength, emax, max_chunk=max_chunk)
        else:
            sc, nn, uu = tapes[i]
            counters += engine.initial_walk(walklength, emax, sc, nn, uu,
                                            max_chunk=max_chunk)
    return {"emax": emax, "counters": counters, "adaptation": log}

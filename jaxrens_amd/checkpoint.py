"""Checkpoint and resume of a nested-sampling run on the device engine.

Restates ``io/checkpoint.py:76-389`` with NumPy ``.npz`` files in place of
HDF5 (h5py is not part of this stack).  The field names follow the
reference's: ``positions``, ``types``, ``energies``, ``cells``,
``step_sizes``, ``log_evidence``, ``iteration``, ``emax``, ``n_dead``,
``n_walkers``.  Added for an exact resume of this engine: the per-walker
``n_atoms`` and GMC ``direction`` (with the persistent directions of the
species-scoped GMC moves, ``direction_slots``), the run parameters ``pressure`` / ``mu``,
the device draw counters ``rng_state`` (``jr_get_rng_state``; the reference
stores its PRNG key, ``:187-197``) with the ``seed``, and optionally the
host generator of the adaptation bisection as JSON.  Files are written and
read without pickling (``allow_pickle=False``).

A resumed engine redraws exactly the device stream an uninterrupted run
draws, and its population energies are recomputed by the same kernels, so
a run split at a checkpoint reproduces the uninterrupted run bit for bit.
"""

from __future__ import annotations

import ctypes as C
import json
from pathlib import Path

import numpy as np

from . import _lib

FORMAT = "jaxrens_amd-checkpoint-1"


def rng_state(engine) -> np.ndarray:
    st = (C.c_int64 * 3)()
    engine._check(engine.lib.jr_get_rng_state(engine.h, st))
    return np.array(list(st), dtype=np.int64)


def set_rng_state(engine, state) -> None:
    st = (C.c_int64 * 3)(*(int(x) for x in state))
    engine._check(engine.lib.jr_set_rng_state(engine.h, st))


def save_checkpoint(path, engine, host_rng: np.random.Generator | None = None) -> Path:
    """Write the engine's full run state to ``path`` (``.npz``)."""
    path = Path(path)
    path.parent.mkdir(parents=True, exist_ok=True)
    pop = engine.get_population()
    fields = {
        "format": np.array(FORMAT),
        "positions": pop["positions"],
        "types": pop["types"],
        "n_atoms": pop["n_atoms"],
        "cells": pop["cell"],
        "energies": pop["energy"],
        "direction": pop["direction"],
        "direction_slots": pop.get("direction_slots", np.zeros((0,) + pop["direction"].shape)),
        "step_sizes": engine.step_sizes,
        "pressure": engine.pressure,
        "mu": engine.mu,
        "log_evidence": engine.log_evidence,
        "iteration": np.array(engine.iteration, dtype=np.int64),
        "n_dead": np.array(engine.iteration * engine.run.n_cull, dtype=np.int64),
        "n_walkers": np.array(engine.n_walkers, dtype=np.int64),
        "emax": engine.emax(),
        "rng_state": rng_state(engine),
        "seed": np.array(engine.run.seed, dtype=np.uint64),
    }
    if host_rng is not None:
        fields["host_rng"] = np.array(json.dumps(host_rng.bit_generator.state))
    tmp = path.with_name(path.name + ".tmp.npz")
    np.savez(tmp, **fields)
    tmp.replace(path)
    return path


def load_checkpoint(path) -> dict:
    """Read a checkpoint written by :func:`save_checkpoint` (no pickling)."""
    with np.load(Path(path), allow_pickle=False) as f:
        d = {k: f[k] for k in f.files}
    if str(d.get("format", "")) != FORMAT:
        raise ValueError(f"{path}: not a {FORMAT} file")
    for k in ("iteration", "n_dead", "n_walkers"):
        d[k] = int(d[k])
    d["seed"] = int(d["seed"])
    if "host_rng" in d:
        d["host_rng"] = json.loads(str(d["host_rng"]))
    return d


def restore(engine, state: dict) -> np.random.Generator | None:
    """Put a loaded checkpoint into ``engine`` (same configuration and seed).

    Returns the adaptation bisection's host generator when the checkpoint
    holds one."""
    R, K = engine.n_replicas, engine.n_walkers
    if state["positions"].shape[:2] != (R, K) or state["n_walkers"] != K:
        raise ValueError("checkpoint population does not match the engine's R x K")
    if state["seed"] != (engine.run.seed & 0xFFFFFFFFFFFFFFFF):
        raise ValueError("checkpoint was written under another seed")
    engine.set_population(state["positions"], state["types"], state["cells"],
                          n_atoms=state["n_atoms"], direction=state["direction"])
    slots = state.get("direction_slots")
    if slots is not None and len(slots) != engine.mwg.n_dir_slots - 1:
        raise ValueError("checkpoint direction slots do not match the engine's moves")
    if slots is not None and len(slots):
        engine.set_direction_slots(slots)
    engine.set_run_params(pressure=state["pressure"], mu=state["mu"],
                          step_sizes=state["step_sizes"])
    engine.iteration = state["iteration"]
    engine.log_evidence = _lib.f64(state["log_evidence"]).copy()
    # the contour: the first adaptation after resume reads it (as the
    # reference restores NSState.emax, io/checkpoint.py:347-365)
    engine.set_emax(state["emax"])
    set_rng_state(engine, state["rng_state"])
    if "host_rng" in state:
        g = np.random.default_rng()
        g.bit_generator.state = state["host_rng"]
        return g
    return None

"""Configuration constraints: the minimum-distance floors.

``min_distance_floors`` builds the ``[S][S]`` threshold table of
``cli/schema/constraints.py:104-138`` (one value for every pair, or
``"A-B"`` keys plus ``"default"``, stored in float32 as the reference
stores it) and reports whether it depends on the species.
``constraint_tables`` combines several ``minimum_distance`` constraints
the way ``constraints/base.py:93-140`` gates moves: every constraint gates
the moves that move atoms or the cell, and only species-dependent ones gate
``alchemical_morph`` and ``species_swap``.  Since all constraints must
hold, combining them is an elementwise maximum of their floors.
"""

from __future__ import annotations

import numpy as np


def min_distance_floors(d_min, n_species: int, symbol_map: dict[int, str] | None = None):
    """``(floors [S][S] float64 of float32 values, type_dependent)``."""
    S = int(n_species)
    m = np.zeros((S, S), np.float32)
    if isinstance(d_min, dict):
        if symbol_map is None:
            raise ValueError("per-pair d_min needs a symbol_map (type code -> element)")
        default = float(d_min.get("default", 0.0))
        pairs = {}
        for key, v in d_min.items():
            if key == "default":
                continue
            a, sep, b = key.partition("-")
            if not sep or not a or not b:
                raise ValueError(f"minimum_distance d_min key {key!r} must be a 'A-B' pair")
            pairs[frozenset({a, b})] = float(v)
        unknown = {s for p in pairs for s in p} - set(symbol_map.values())
        if unknown:
            raise ValueError(f"minimum_distance d_min references element(s) {sorted(unknown)} "
                             "not present in the system")
        for i in range(S):
            for j in range(S):
                m[i, j] = pairs.get(frozenset({symbol_map[i], symbol_map[j]}), default)
    else:
        if float(d_min) < 0.0:
            raise ValueError(f"minimum_distance d_min must be >= 0, got {d_min}")
        m[:, :] = float(d_min)
    if np.any(m < 0):
        raise ValueError("minimum distances must be >= 0")
    return m.astype(np.float64), bool(m.min() != m.max())


def constraint_tables(constraints, n_species: int, symbol_map=None):
    """``(floors for atom/cell moves or None, floors for species moves or
    None)`` from a list of ``{"type": "minimum_distance", "d_min": ...}``."""
    pos, types = None, None
    for c in constraints or ():
        if c.get("type") != "minimum_distance":
            raise ValueError(f"unknown constraint type {c.get('type')!r}")
        m, dep = min_distance_floors(c["d_min"], n_species, symbol_map)
        pos = m if pos is None else np.maximum(pos, m)
        if dep:
            types = m if types is None else np.maximum(types, m)
    return pos, types

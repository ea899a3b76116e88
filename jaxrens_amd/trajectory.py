"""Dead-point trajectory in extended XYZ, the reference's ``.traj.extxyz``.

``io/trajectory.py:84-139`` (``ExtxyzTrajectoryWriter``) writes one frame
per recorded dead point through ASE; ASE is not part of this stack, so the
same text is written here directly, in the layout of the reference's
committed run output (``tests/_assets/data/lj8_npt/output``):

    <N>
    Lattice="<9 cell entries, rows = lattice vectors>" Properties=species:S:1:pos:R:3 ns_energy=<E> iter=<i> pbc="T T T"
    <symbol> <x> <y> <z>          (symbol left in 2 columns, coordinates %16.8f)

Floats of the header use Python's shortest round-trip form, as ASE's writer
does.  Positions are wrapped into the cell first (``wrap=True``, the
reference default): fractional coordinates ``f -> ((f - eps) mod 1) + eps``
with ``eps = 1e-7``, the convention of ``Atoms.wrap`` that the reference
calls.  A resume at ``restart_iteration`` drops the frames at or past it
(``io/restart_truncate.py:49-71``).
"""

from __future__ import annotations

import re
from pathlib import Path

import numpy as np

WRAP_EPS = 1e-7


def wrap_positions(positions, cell, eps: float = WRAP_EPS) -> np.ndarray:
    """Positions moved into the cell: fractional ``((f - eps) mod 1) + eps``;
    a degenerate cell leaves them unchanged."""
    pos = np.asarray(positions, dtype=np.float64)
    cell = np.asarray(cell, dtype=np.float64)
    if cell.shape != (3, 3) or abs(np.linalg.det(cell)) < 1e-12:
        return pos.copy()
    frac = np.linalg.solve(cell.T, pos.T).T
    frac = np.mod(frac - eps, 1.0) + eps
    return frac @ cell


def format_frame(positions, types, cell, energy: float, iteration: int,
                 symbol_map: dict[int, str], cull_rank: int | None = None) -> str:
    pos = np.asarray(positions, dtype=np.float64)
    lat = " ".join(repr(float(x)) for x in np.asarray(cell, dtype=np.float64).reshape(9))
    rank = "" if cull_rank is None else f" cull_rank={int(cull_rank)}"
    lines = [f"{len(pos)}",
             f'Lattice="{lat}" Properties=species:S:1:pos:R:3 '
             f'ns_energy={float(energy)!r} iter={int(iteration)}{rank} pbc="T T T"']
    for t, (x, y, z) in zip(np.asarray(types).reshape(-1), pos):
        lines.append(f"{symbol_map[int(t)]:<2s} {x:16.8f} {y:16.8f} {z:16.8f}")
    return "\n".join(lines) + "\n"


def read_extxyz(path) -> list[dict]:
    """Frames of a file in the layout above: symbols, positions, cell,
    ns_energy, iter."""
    lines = Path(path).read_text().splitlines()
    frames, i = [], 0
    while i < len(lines) and lines[i].strip():
        n = int(lines[i])
        hdr = lines[i + 1]
        lat = np.array([float(x) for x in
                        re.search(r'Lattice="([^"]+)"', hdr).group(1).split()])
        rows = [lines[i + 2 + a].split() for a in range(n)]
        frames.append({
            "symbols": [r[0] for r in rows],
            "positions": np.array([[float(v) for v in r[1:4]] for r in rows]).reshape(n, 3),
            "cell": lat.reshape(3, 3),
            "ns_energy": float(re.search(r"ns_energy=(\S+)", hdr).group(1)),
            "iter": int(re.search(r"iter=(-?\d+)", hdr).group(1)),
        })
        rank = re.search(r"cull_rank=(\d+)", hdr)
        if rank:
            frames[-1]["cull_rank"] = int(rank.group(1))
        i += 2 + n
    return frames


class ExtxyzTrajectoryWriter:
    """One frame per dead point (``io/trajectory.py:84-139``)."""

    def __init__(self, path, symbol_map: dict[int, str], wrap: bool = True,
                 restart_iteration: int = 0):
        self.path = Path(path)
        self.path.parent.mkdir(parents=True, exist_ok=True)
        self.symbol_map = symbol_map
        self.wrap = wrap
        if restart_iteration > 0 and self.path.exists():
            lines = self.path.read_text().splitlines(keepends=True)
            kept, i = [], 0
            while i < len(lines) and lines[i].strip():
                n = int(lines[i])
                it = int(re.search(r"iter=(-?\d+)", lines[i + 1]).group(1))
                if it < restart_iteration:
                    kept.extend(lines[i: i + 2 + n])
                i += 2 + n
            self.path.write_text("".join(kept))
            self._f = open(self.path, "a")
        else:
            self._f = open(self.path, "w")

    def write_dead_point(self, iteration: int, positions, types, cell, energy: float,
                         cull_rank: int | None = None):
        """One dead point; ``cull_rank`` (0 = highest energy among the
        iteration's culls, ``io/trajectory.py:118-139``) is written only
        when n_cull > 1."""
        pos = wrap_positions(positions, cell) if self.wrap else np.asarray(positions)
        self._f.write(format_frame(pos, types, cell, energy, iteration, self.symbol_map,
                                   cull_rank))

    def flush(self):
        self._f.flush()

    def close(self):
        if self._f is not None:
            self._f.close()
            self._f = None

"""Moves-YAML API of the reference, translated to the device move table.

Field names and defaults follow ``cli/schema/moves.py`` (``BaseMoveSpec``
:60-120, ``GMCMoveSpec`` :249-406 with the legacy ``galilean`` alias of
``cli/schema/root.py:29-40``, ``HMCMoveSpec`` :409-447, ``VolumeMoveSpec``
:628-682, ``ShearMoveSpec`` :685-736, ``StretchMoveSpec`` :739-788,
``SpeciesSwapMoveSpec`` :514-625, ``AlchemicalMorphMoveSpec`` :791-828) and
``cli/schema/cell.py`` for the
cell-geometry guard of the cell moves.  ``build_mwg`` mirrors
``sampling/mwg.py:40-259``: weights are normalised to probabilities and each
move carries its own step size.
"""

from __future__ import annotations

from typing import Annotated, Literal, Union

import numpy as np
from pydantic import BaseModel, ConfigDict, Field, TypeAdapter, field_validator

from . import _lib


class CellSpec(BaseModel):
    """``cli/schema/cell.py``: volume-per-atom band, aspect guard, V prior."""

    model_config = ConfigDict(extra="forbid", frozen=True)
    max_volume_per_atom: float = 1e4
    min_volume_per_atom: float = 1.0
    initial_min_volume_per_atom: float | None = None
    min_aspect_ratio: float = 0.8
    flat_V_prior: bool = False


class BaseMoveSpec(BaseModel):
    model_config = ConfigDict(extra="forbid", frozen=True)
    step_size: float = 0.1
    weight: float = 1.0
    adaptation_warmup: int = 100
    target_acceptance: float = 0.5
    name: str | None = None

    @property
    def move_type(self) -> str:
        return self.type  # type: ignore[attr-defined]

    def effective_name(self) -> str:
        return self.name if self.name is not None else self.type  # type: ignore[attr-defined]


class GMCMoveSpec(BaseMoveSpec):
    """Galilean MC with force reflection (``sampling/moves/galilean.py``).

    ``species`` scopes the move to the atoms of the listed type codes (the
    reference names elements, ``cli/schema/moves.py`` GMCMoveSpec; here the
    integer codes of the pair table); every other atom stays fixed, the
    direction and the reflection normal are projected onto the moving
    atoms, and the move keeps a persistent direction of its own (auto-named
    ``gmc_<codes>``).  ``refresh_direction`` draws a fresh direction on
    every call instead of perturbing the stored one."""

    type: Literal["gmc"] = "gmc"
    n_reflect: int = 5
    species: tuple[int, ...] | None = None
    refresh_direction: bool = False

    @field_validator("species", mode="before")
    @classmethod
    def _wrap_bare_code(cls, v):
        return (v,) if isinstance(v, (int, str)) else v

    def effective_name(self) -> str:
        if self.name is not None:
            return self.name
        if self.species:
            return "gmc_" + "_".join(str(c) for c in self.species)
        return self.type

    @property
    def scoped(self) -> bool:
        return bool(self.species)

    def device_move(self, cell: CellSpec) -> _lib.JrMove:
        if self.n_reflect < 1:
            raise ValueError("n_reflect must be >= 1")
        mask = 0
        if self.species is not None:
            codes = set(int(c) for c in self.species)
            if not codes:
                raise ValueError("gmc species must name at least one type code")
            if min(codes) < 0 or max(codes) >= _lib.JR_MAX_SPECIES:
                raise ValueError(f"species codes must be in [0, {_lib.JR_MAX_SPECIES})")
            mask = sum(1 << c for c in codes)
        return _lib.JrMove(type=_lib.JR_MOVE_GMC, n_reflect=self.n_reflect,
                           perturb_angle=0.1, species_mask=mask,
                           refresh_direction=int(self.refresh_direction))


class HMCMoveSpec(BaseMoveSpec):
    """Leapfrog HMC with the hard ``dH < 1`` cut (``sampling/moves/hmc.py``)."""

    type: Literal["hmc"] = "hmc"
    n_leapfrog: int = 10

    def device_move(self, cell: CellSpec) -> _lib.JrMove:
        if self.n_leapfrog < 0:
            raise ValueError("n_leapfrog must be >= 0")
        return _lib.JrMove(type=_lib.JR_MOVE_HMC, n_leapfrog=self.n_leapfrog)


class VolumeMoveSpec(BaseMoveSpec):
    """Isotropic volume move (``sampling/moves/volume.py``)."""

    type: Literal["volume"] = "volume"

    def device_move(self, cell: CellSpec) -> _lib.JrMove:
        return _lib.JrMove(type=_lib.JR_MOVE_VOLUME,
                           flat_v_prior=int(cell.flat_V_prior),
                           max_vpa=cell.max_volume_per_atom,
                           min_vpa=cell.min_volume_per_atom,
                           min_aspect=cell.min_aspect_ratio)


class ShearMoveSpec(BaseMoveSpec):
    """Volume-preserving shear of one cell vector (``sampling/moves/shear.py``)."""

    type: Literal["shear"] = "shear"

    def device_move(self, cell: CellSpec) -> _lib.JrMove:
        return _lib.JrMove(type=_lib.JR_MOVE_SHEAR,
                           max_vpa=cell.max_volume_per_atom,
                           min_vpa=cell.min_volume_per_atom,
                           min_aspect=cell.min_aspect_ratio)


class StretchMoveSpec(BaseMoveSpec):
    """Volume-preserving stretch of two Cartesian axes (``sampling/moves/stretch.py``)."""

    type: Literal["stretch"] = "stretch"

    def device_move(self, cell: CellSpec) -> _lib.JrMove:
        return _lib.JrMove(type=_lib.JR_MOVE_STRETCH,
                           max_vpa=cell.max_volume_per_atom,
                           min_vpa=cell.min_volume_per_atom,
                           min_aspect=cell.min_aspect_ratio)


class AlchemicalMorphMoveSpec(BaseMoveSpec):
    """Random atom to a random other species (``sampling/moves/alchemical.py``)."""

    type: Literal["alchemical_morph"] = "alchemical_morph"
    n_species: int = 2

    def device_move(self, cell: CellSpec) -> _lib.JrMove:
        if self.n_species < 2:
            raise ValueError("alchemical_morph needs n_species >= 2")
        return _lib.JrMove(type=_lib.JR_MOVE_MORPH, n_species=self.n_species)


class SpeciesSwapMoveSpec(BaseMoveSpec):
    """Exchange the types of two atoms of different species
    (``sampling/moves/swap.py``).  ``species`` restricts the exchange to the
    listed type codes (the reference names elements; here the integer codes
    of the pair table).  ``step_size`` is not used."""

    type: Literal["species_swap"] = "species_swap"
    species: tuple[int, ...] | None = None

    def device_move(self, cell: CellSpec) -> _lib.JrMove:
        mask = 0
        if self.species is not None:
            codes = set(int(c) for c in self.species)
            if len(codes) < 2:
                raise ValueError("species_swap needs at least two distinct species")
            if min(codes) < 0 or max(codes) >= _lib.JR_MAX_SPECIES:
                raise ValueError(f"species codes must be in [0, {_lib.JR_MAX_SPECIES})")
            mask = sum(1 << c for c in codes)
        return _lib.JrMove(type=_lib.JR_MOVE_SWAP, species_mask=mask)


MoveSpec = Annotated[
    Union[GMCMoveSpec, HMCMoveSpec, VolumeMoveSpec, ShearMoveSpec,
          StretchMoveSpec, SpeciesSwapMoveSpec, AlchemicalMorphMoveSpec],
    Field(discriminator="type"),
]
_MOVE_ADAPTER = TypeAdapter(MoveSpec)

_UNBUILT = {"random_walk", "single_atom", "single_atom_sweep"}

# moves without a continuous step (adaptation leaves them alone)
STEPLESS = {"alchemical_morph", "species_swap"}


def parse_move(spec) -> BaseMoveSpec:
    """Validate one move dict (``type: galilean`` is accepted as ``gmc``)."""
    if isinstance(spec, BaseMoveSpec):
        return spec
    d = dict(spec)
    if d.get("type") == "galilean":
        d["type"] = "gmc"
    if d.get("type") in _UNBUILT:
        raise NotImplementedError(f"move type {d['type']!r} is not built")
    return _MOVE_ADAPTER.validate_python(d)


def parse_moves(specs) -> list[BaseMoveSpec]:
    moves = [parse_move(s) for s in specs]
    if not moves:
        raise ValueError("at least one move is required")
    if len(moves) > _lib.JR_MAX_MOVES:
        raise ValueError(f"at most {_lib.JR_MAX_MOVES} moves")
    names = [m.effective_name() for m in moves]
    if len(set(names)) != len(names):
        raise ValueError(f"move names must be distinct, got {names}")
    return moves


class MWG:
    """Device move table + normalised weights (``sampling/mwg.py:40-259``)."""

    def __init__(self, moves, cell: CellSpec | dict | None = None):
        self.moves = parse_moves(moves)
        self.cell = (cell if isinstance(cell, CellSpec)
                     else CellSpec(**(cell or {})))
        w = np.array([m.weight for m in self.moves], dtype=np.float64)
        if (w < 0).any() or not (w.sum() > 0):
            raise ValueError("move weights must be >= 0 and not all zero")
        self.weights = w
        self.move_probs = w / w.sum()
        self.default_step_sizes = np.array([m.step_size for m in self.moves])
        self.names = [m.effective_name() for m in self.moves]
        self.n_moves = len(self.moves)
        # persistent-direction slots: unscoped GMC moves share the population's
        # direction (slot 0), each species-scoped one keeps its own
        # (cli/schema/moves.py GMCMoveSpec._direction_field)
        self.dir_slots, n = [], 0
        for m in self.moves:
            if getattr(m, "scoped", False):
                n += 1
                self.dir_slots.append(n)
            else:
                self.dir_slots.append(0)
        self.n_dir_slots = n + 1

    def _device_move(self, i):
        mv = self.moves[i].device_move(self.cell)
        mv.dir_slot = self.dir_slots[i]
        return mv

    def device_table(self):
        arr = (_lib.JrMove * self.n_moves)()
        for i in range(self.n_moves):
            arr[i] = self._device_move(i)
        return arr

    def as_oracle_moves(self):
        """Plain tuples for the CPU oracle (same fields as the device table)."""
        return [self._device_move(i) for i in range(self.n_moves)]


def build_mwg(backend, move_descriptors, cell=None) -> MWG:
    """Reference-shaped entry point; ``backend`` is held by the engine."""
    return MWG(move_descriptors, cell)

"""MI355X (gfx950) nested-sampling inner walk with the jaxrens energy/moves API."""

from .backends import (BackendResult, EnergyBackend, EnergyTerm, EnsembleBackend,
                       EnsembleTerm, Hamiltonian, HipLJBackend, LJBackend,
                       create_lj, ensemble_energy, eval_energy_and_forces,
                       make_ensemble_params, rebase_ensemble_energy, unwrap_model)
from .engine import NSEngine, RunConfig
from .moves import CellSpec, MWG, build_mwg, parse_moves

__all__ = [
    "BackendResult", "EnergyBackend", "EnergyTerm", "EnsembleBackend", "EnsembleTerm",
    "Hamiltonian", "HipLJBackend", "LJBackend", "create_lj", "ensemble_energy",
    "eval_energy_and_forces", "make_ensemble_params", "rebase_ensemble_energy",
    "unwrap_model", "NSEngine", "RunConfig", "CellSpec", "MWG", "build_mwg",
    "parse_moves",
]

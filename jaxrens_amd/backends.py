"""Energy API kept from the reference: ``BackendResult``, the
``EnergyBackend`` protocol, ``Hamiltonian`` / ``EnergyTerm`` /
``EnsembleTerm`` and the Lennard-Jones backend, computed by the HIP library.

Reference: ``backends/base.py:33-212``, ``backends/hamiltonian.py:54-160``,
``backends/ensemble.py:34-161``, ``backends/lj.py:41-228``.

Arrays are NumPy (host); each call is one batched launch of the FP64 pair
kernel through ``jr_eval_energy_forces``.  Leading batch axes are allowed on
``positions`` ``(*B, N, 3)`` / ``species`` ``(*B, N)`` / ``cell``
``(*B, 3, 3)``.
"""

from __future__ import annotations

import math
from collections.abc import Sequence
from typing import Any, NamedTuple, Protocol, runtime_checkable

import numpy as np

from . import _lib


class BackendResult(NamedTuple):
    """Return value of every backend call (``backends/base.py:33-71``)."""

    energy: np.ndarray | float
    forces: np.ndarray | None = None
    max_neighbor_count: int = 0
    overflow: bool = False
    energy_members: np.ndarray | None = None
    forces_members: np.ndarray | None = None


@runtime_checkable
class EnergyBackend(Protocol):
    """``backends/base.py:165-212``: ``backend(positions, species, cell,
    max_neighbors, ensemble_params=None) -> BackendResult``."""

    r_cutoff: float

    def __call__(self, positions, species, cell, max_neighbors: int = 0,
                 ensemble_params: dict | None = None) -> BackendResult: ...


class PairTable(NamedTuple):
    """Per-species-pair LJ parameters handed to the kernels."""

    n_species: int
    eps4: np.ndarray   # (S, S) 4 * eps_ab
    eps24: np.ndarray  # (S, S) 24 * eps_ab
    sig2: np.ndarray   # (S, S) sigma_ab ** 2
    rc2: float
    images: tuple[int, int, int] = (0, 0, 0)  # supercell_trafo // 2
    eval_chunks: int = 1  # evaluation chunks (summation order, team size)

    def as_struct(self, precision: int = _lib.JR_PREC_FP64):
        keep = (_lib.f64(self.eps4), _lib.f64(self.eps24), _lib.f64(self.sig2))
        pot = _lib.JrPot(self.n_species, precision, float(self.rc2),
                         _lib.dptr(keep[0]), _lib.dptr(keep[1]),
                         _lib.dptr(keep[2]))
        for d in range(3):
            pot.images[d] = int(self.images[d])
        pot.eval_chunks = int(self.eval_chunks)
        return pot, keep


def _f32(x) -> np.ndarray:
    return np.asarray(x, dtype=np.float32)


class LJBackend:
    """All-pairs Lennard-Jones with minimum-image PBC (``backends/lj.py:41-204``).

    ``epsilon``/``sigma`` scalar (one species) or 1-D per-species with
    Lorentz-Berthelot mixing, both rounded to float32 as the reference does
    (``lj.py:71-72,175-185``).  ``pair_epsilon``/``pair_sigma`` give an
    explicit symmetric (S, S) table in float64 instead of mixing (the binary
    semi-grand workload's eps_AB / sigma_AB; the reference has no such
    table).  ``cutoff=None`` means no cutoff; the cutoff is strict
    ``r^2 < rc^2`` and unshifted.  ``supercell_trafo`` adds the lattice
    translations ``-(s//2) .. s//2`` per cell vector on top of the minimum
    image, as ``lj.py:106-113,152-166`` does (entries up to 9, FP64 only).

    ``eval_chunks`` (1, 2, 4 or 8) fixes the summation order of every energy
    and force evaluation: the ring steps of the pair sum are split into that
    many consecutive chunks, each summed on its own and the chunk sums added
    in order.  With more than one chunk a team of that many wavefronts
    walks each chain, one chunk each, which cuts the latency of a walk with
    few chains per iteration (``n_extra = 0``).  1 (the default) is one
    running sum per atom.  FP64 minimum image only.
    """

    def __init__(
        self,
        epsilon: float | Sequence[float] = 1.0,
        sigma: float | Sequence[float] = 1.0,
        cutoff: float | None = None,
        supercell_trafo: tuple[int, int, int] = (1, 1, 1),
        *,
        pair_epsilon=None,
        pair_sigma=None,
        device: int = 0,
        precision: str = "fp64",
        eval_chunks: int = 1,
    ):
        sc = tuple(int(x) for x in supercell_trafo)
        if min(sc) < 1:
            raise ValueError(
                f"supercell_trafo entries must be >= 1, got {supercell_trafo}")
        images = tuple(s // 2 for s in sc)
        if max(images) > _lib.JR_MAX_IMAGES:
            raise ValueError(
                f"supercell_trafo entries must be <= {2 * _lib.JR_MAX_IMAGES + 1}")
        if max(images) > 0 and precision != "fp64":
            raise NotImplementedError("supercell images are built for fp64 only")
        self.supercell_trafo = sc
        self.cutoff = cutoff
        self.r_cutoff = cutoff if cutoff is not None else 0.0
        self.device = int(device)
        if precision not in ("fp64", "fp32_pairs"):
            raise ValueError("precision must be 'fp64' or 'fp32_pairs'")
        self.precision = precision
        self._prec_code = (_lib.JR_PREC_FP64 if precision == "fp64"
                           else _lib.JR_PREC_FP32_PAIRS)
        rc2 = math.inf if cutoff is None else float(cutoff) ** 2

        if pair_epsilon is not None or pair_sigma is not None:
            pe = np.asarray(pair_epsilon, dtype=np.float64)
            ps = np.asarray(pair_sigma, dtype=np.float64)
            if pe.ndim != 2 or pe.shape != ps.shape or pe.shape[0] != pe.shape[1]:
                raise ValueError("pair_epsilon/pair_sigma must be matching (S, S)")
            if not (np.allclose(pe, pe.T, rtol=0, atol=0)
                    and np.allclose(ps, ps.T, rtol=0, atol=0)):
                raise ValueError("pair tables must be symmetric")
            self._per_species = True
            self.epsilon = self.sigma = None
            self.n_species = pe.shape[0]
            eps, sig2 = pe, ps * ps
        else:
            eps_arr = _f32(epsilon)
            sig_arr = _f32(sigma)
            if eps_arr.ndim != sig_arr.ndim:
                raise ValueError(
                    "epsilon and sigma must have the same rank, got "
                    f"epsilon.ndim={eps_arr.ndim}, sigma.ndim={sig_arr.ndim}")
            if eps_arr.ndim > 1:
                raise ValueError(
                    "epsilon/sigma must be scalar or 1-D (per-species), got "
                    f"rank {eps_arr.ndim}")
            if eps_arr.ndim == 1:
                if eps_arr.shape != sig_arr.shape:
                    raise ValueError(
                        "per-species epsilon and sigma must have matching "
                        f"length, got {eps_arr.shape} vs {sig_arr.shape}")
                if eps_arr.shape[0] == 0:
                    raise ValueError("per-species epsilon/sigma must be non-empty")
                self._per_species = True
                self.epsilon = self.sigma = None
                self.n_species = int(eps_arr.shape[0])
                # float32 Lorentz-Berthelot mixing, lj.py:176-184
                eps_ij = np.sqrt(eps_arr[:, None] * eps_arr[None, :])
                sig_ij = np.float32(0.5) * (sig_arr[:, None] + sig_arr[None, :])
                eps = eps_ij.astype(np.float64)
                sig2 = (sig_ij ** 2).astype(np.float64)
            else:
                self._per_species = False
                self.epsilon = float(eps_arr)
                self.sigma = float(sig_arr)
                self.n_species = 1
                eps = np.array([[self.epsilon]])
                sig2 = np.array([[self.sigma ** 2]])
        if self.n_species > _lib.JR_MAX_SPECIES:
            raise ValueError(f"at most {_lib.JR_MAX_SPECIES} species")
        eval_chunks = int(eval_chunks)
        if eval_chunks not in (1, 2, 4, 8):
            raise ValueError("eval_chunks must be 1, 2, 4 or 8")
        if eval_chunks > 1 and (max(images) > 0 or precision != "fp64"):
            raise NotImplementedError(
                "eval_chunks > 1 is built for the fp64 minimum-image potential")
        self.eval_chunks = eval_chunks
        self.table = PairTable(self.n_species, 4.0 * eps, 24.0 * eps, sig2, rc2,
                               images, eval_chunks)

    # -- evaluation ---------------------------------------------------------
    def _eval(self, positions, species, cell, pressure=None, mu=None,
              forces=False):
        pos = np.asarray(positions, dtype=np.float64)
        batch = pos.shape[:-2]
        N = pos.shape[-2]
        if pos.shape[-1] != 3:
            raise ValueError("positions must have shape (*B, N, 3)")
        B = int(np.prod(batch)) if batch else 1
        sp = np.broadcast_to(np.asarray(species), batch + (N,))
        cl = np.broadcast_to(np.asarray(cell, dtype=np.float64), batch + (3, 3))
        if N > 256:
            raise ValueError("at most 256 atoms per configuration")
        if sp.size and (sp.min() < 0 or sp.max() >= self.n_species):
            raise ValueError(
                f"species codes must be in [0, {self.n_species}) for this table")
        Nmax = max(N, 1)
        p = _lib.f64(pos.reshape(B, N, 3))
        if N == 0:
            p = np.zeros((B, 1, 3))
        t = _lib.i32(sp.reshape(B, N)) if N else np.zeros((B, 1), np.int32)
        n_atoms = np.full(B, N, dtype=np.int32)
        c = _lib.f64(cl.reshape(B, 9))
        pr = None if pressure is None else _lib.f64(
            np.broadcast_to(np.asarray(pressure, dtype=np.float64), batch).reshape(B))
        m = None
        if mu is not None:
            mu = np.asarray(mu, dtype=np.float64)
            mu_full = np.zeros(batch + (self.n_species,))
            k = min(mu.shape[-1], self.n_species)
            mu_full[..., :k] = np.broadcast_to(mu, batch + mu.shape[-1:])[..., :k]
            m = _lib.f64(mu_full.reshape(B, self.n_species))
        U = np.empty(B)
        H = np.empty(B)
        F = np.empty((B, Nmax, 3)) if forces else None
        lib = _lib.load_library()
        pot, keep = self.table.as_struct(self._prec_code)
        rc = lib.jr_eval_energy_forces(
            pot, self.device, B, Nmax, _lib.dptr(p), _lib.iptr(t),
            _lib.iptr(n_atoms), _lib.dptr(c), _lib.dptr(pr), _lib.dptr(m),
            _lib.dptr(U), _lib.dptr(H), _lib.dptr(F))
        del keep
        _lib.check(rc)
        shape = batch
        U = U.reshape(shape) if shape else float(U[0])
        H = H.reshape(shape) if shape else float(H[0])
        if forces:
            F = F[:, :N].reshape(batch + (N, 3))
        return U, H, F

    def __call__(self, positions, species, cell, max_neighbors: int = 0,
                 ensemble_params: dict | None = None) -> BackendResult:
        U, _, _ = self._eval(positions, species, cell)
        return BackendResult(energy=U)

    def energy_and_forces(self, positions, species, cell, max_neighbors: int = 0,
                          ensemble_params: dict | None = None) -> BackendResult:
        U, _, F = self._eval(positions, species, cell, forces=True)
        return BackendResult(energy=U, forces=F)


HipLJBackend = LJBackend


def create_lj(epsilon=1.0, sigma=1.0, cutoff=None, supercell_trafo=(1, 1, 1),
              **kw) -> LJBackend:
    """``backends/lj.py:207-228``."""
    return LJBackend(epsilon=epsilon, sigma=sigma, cutoff=cutoff,
                     supercell_trafo=supercell_trafo, **kw)


def get_volume(cell) -> np.ndarray:
    """|det(cell)| (``utils/cell.py:12-14``)."""
    return np.abs(np.linalg.det(np.asarray(cell, dtype=np.float64)))


def ensemble_energy(cell, species, pressure=None, chemical_potentials=None):
    """``P*V - mu.N`` (``backends/ensemble.py:34-59``), host float64."""
    E = 0.0
    if pressure is not None and cell is not None:
        E = np.asarray(pressure, dtype=np.float64) * get_volume(cell)
    if chemical_potentials is not None:
        mu = np.asarray(chemical_potentials, dtype=np.float64)
        if mu.shape[-1] > 0:
            sp = np.asarray(species)
            counts = np.stack([(sp == s).sum(-1) for s in range(mu.shape[-1])],
                              axis=-1).astype(np.float64)
            E = E - (mu * counts).sum(-1)
    return E


def rebase_ensemble_energy(energy, cell, species, params_from, params_to):
    """``backends/ensemble.py:62-87``."""
    u = energy - ensemble_energy(cell, species, params_from.get("pressure"),
                                 params_from.get("chemical_potentials"))
    return u + ensemble_energy(cell, species, params_to.get("pressure"),
                               params_to.get("chemical_potentials"))


def eval_energy_and_forces(backend, positions, species, cell, max_neighbors=0,
                           ensemble_params=None) -> BackendResult:
    """``backends/base.py:74-113``: native ``energy_and_forces`` path."""
    native = getattr(backend, "energy_and_forces", None)
    if native is None:
        raise TypeError(
            f"{type(backend).__name__} has no energy_and_forces; the HIP "
            "path needs analytic forces (no autodiff fallback)")
    return native(positions, species, cell, max_neighbors,
                  ensemble_params=ensemble_params)


class EnergyTerm:
    """Additive, model-independent term (``backends/hamiltonian.py:54-86``)."""

    def energy(self, positions, species, cell, ensemble_params=None):
        raise NotImplementedError

    def energy_and_forces(self, positions, species, cell, ensemble_params=None):
        raise NotImplementedError


class EnsembleTerm(EnergyTerm):
    """``P*V - mu.N`` with zero forces (``backends/ensemble.py:90-133``)."""

    def __init__(self, pressure: float = 0.0, chemical_potentials=None):
        self.pressure = pressure
        self.chemical_potentials = chemical_potentials

    def params(self, ensemble_params: dict | None = None):
        pressure, mu = self.pressure, self.chemical_potentials
        if ensemble_params is not None:
            pressure = ensemble_params.get("pressure", pressure)
            mu = ensemble_params.get("chemical_potentials", mu)
        return pressure, mu

    def energy(self, positions, species, cell, ensemble_params=None):
        pressure, mu = self.params(ensemble_params)
        return ensemble_energy(cell, species, pressure, mu)

    def energy_and_forces(self, positions, species, cell, ensemble_params=None):
        e = self.energy(positions, species, cell, ensemble_params)
        return e, np.zeros_like(np.asarray(positions, dtype=np.float64))


class Hamiltonian:
    """Flat ``model + sum(terms)`` (``backends/hamiltonian.py:89-153``).

    With an LJ model and only ensemble terms the whole Hamiltonian is
    evaluated by one kernel launch (``P*V - mu.N`` added on the device).
    """

    def __init__(self, model, terms: Sequence[EnergyTerm] = ()):
        if isinstance(model, Hamiltonian):
            terms = (*model.terms, *terms)
            model = model.model
        self.model = model
        self.terms = tuple(terms)
        self.r_cutoff = model.r_cutoff

    def with_terms(self, *terms: EnergyTerm) -> "Hamiltonian":
        return Hamiltonian(self.model, (*self.terms, *terms))

    def ensemble_params_of(self, ensemble_params=None):
        """(pressure, mu) summed over ensemble terms, or None if other terms."""
        pressure, mu = 0.0, None
        for t in self.terms:
            if not isinstance(t, EnsembleTerm):
                return None
            p, m = t.params(ensemble_params)
            pressure = pressure + (0.0 if p is None else p)
            if m is not None:
                mu = np.asarray(m, dtype=np.float64) if mu is None else mu + m
        return pressure, mu

    def _fused(self, positions, species, cell, ensemble_params, forces):
        pm = self.ensemble_params_of(ensemble_params)
        if pm is None or not isinstance(self.model, LJBackend):
            return None
        return self.model._eval(positions, species, cell, pressure=pm[0],
                                mu=pm[1], forces=forces)

    def __call__(self, positions, species, cell, max_neighbors: int = 0,
                 ensemble_params=None) -> BackendResult:
        fused = self._fused(positions, species, cell, ensemble_params, False)
        if fused is not None:
            return BackendResult(energy=fused[1])
        res = self.model(positions, species, cell, max_neighbors)
        energy = res.energy
        for term in self.terms:
            energy = energy + term.energy(positions, species, cell, ensemble_params)
        return res._replace(energy=energy)

    def energy_and_forces(self, positions, species, cell, max_neighbors: int = 0,
                          ensemble_params=None) -> BackendResult:
        fused = self._fused(positions, species, cell, ensemble_params, True)
        if fused is not None:
            return BackendResult(energy=fused[1], forces=fused[2])
        res = eval_energy_and_forces(self.model, positions, species, cell,
                                     max_neighbors)
        energy, forces = res.energy, res.forces
        for term in self.terms:
            e, f = term.energy_and_forces(positions, species, cell, ensemble_params)
            energy = energy + e
            forces = forces + f
        return res._replace(energy=energy, forces=forces)


def unwrap_model(backend):
    """``backends/hamiltonian.py:156-160``."""
    return backend.model if isinstance(backend, Hamiltonian) else backend


def EnsembleBackend(base, pressure: float = 0.0, chemical_potentials=None):
    """``backends/ensemble.py:136-146``."""
    return Hamiltonian(base, [EnsembleTerm(pressure, chemical_potentials)])


def make_ensemble_params(pressure: float = 0.0, chemical_potentials=None) -> dict:
    """``backends/ensemble.py:149-161``."""
    params: dict[str, Any] = {"pressure": np.asarray(pressure)}
    if chemical_potentials is not None:
        params["chemical_potentials"] = np.asarray(chemical_potentials)
    return params

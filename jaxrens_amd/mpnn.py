"""Seeded random-init invariant message-passing potential: the expensive
exact judge of the surrogate-screened walk (no reference counterpart; the
reference's ML backends live in ``backends/`` and are out of scope here).

Architecture (FP32 weights and arithmetic on the device):
  radial basis  b_n(r) = sqrt(2/rc) sin(n pi r / rc) / r * u(r / rc),
                n = 1..8, u(x) = 1 - 28x^6 + 48x^7 - 21x^8, rc = 5 sigma
  layer 1       h1_i = emb[t_i] + silu(m1_i U1 + c1),
                m1_i[c] = sum_n W1[n, c] sum_{j in N(i)} b_n(r_ij) emb[t_j, c]
  layer 2       h2_i = h1_i + silu(m2_i U2 + c2),
                m2_i[c] = sum_n W2[n, c] sum_{j in N(i)} b_n(r_ij) h1_j[c]
  energy        E = sum_i wo . h2_i + core repulsion
                1/2 sum_{i, j in N(i), r < 1.2} ((0.85/r)^12 - (0.85/1.2)^12)
with 64 channels, N(i) every periodic image of every atom within rc, SiLU
activations.  Cubic cells only (the walk's volume moves are isotropic).
"""

from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np


@dataclass
class MPNNParams:
    S: int = 2
    C: int = 64
    NB: int = 8
    rc: float = 5.0
    core_eps: float = 1.0
    core_sig: float = 0.85
    core_rc: float = 1.2
    emb: np.ndarray = field(default=None)
    W1: np.ndarray = field(default=None)
    U1: np.ndarray = field(default=None)
    c1: np.ndarray = field(default=None)
    W2: np.ndarray = field(default=None)
    U2: np.ndarray = field(default=None)
    c2: np.ndarray = field(default=None)
    wo: np.ndarray = field(default=None)

    def arrays(self):
        return [self.emb, self.W1, self.U1, self.c1, self.W2, self.U2, self.c2,
                self.wo]


def init_mpnn(seed: int = 0, S: int = 2, C: int = 64, NB: int = 8,
              rc: float = 5.0) -> MPNNParams:
    """Seeded random weights.  Filter weights are scaled by 1/32 so the
    neighbour sums (~470 images within 5 sigma at rho = 0.9) stay O(1)."""
    rng = np.random.default_rng(seed)

    def f32(a):
        return np.ascontiguousarray(a, dtype=np.float32)

    p = MPNNParams(S=S, C=C, NB=NB, rc=rc)
    p.emb = f32(rng.standard_normal((S, C)))
    p.W1 = f32(rng.standard_normal((NB, C)) / 32.0)
    p.U1 = f32(rng.standard_normal((C, C)) / np.sqrt(C))
    p.c1 = f32(0.1 * rng.standard_normal(C))
    p.W2 = f32(rng.standard_normal((NB, C)) / 32.0)
    p.U2 = f32(rng.standard_normal((C, C)) / np.sqrt(C))
    p.c2 = f32(0.1 * rng.standard_normal(C))
    p.wo = f32(0.5 * rng.standard_normal(C) / np.sqrt(C))
    return p


class MPNNBackend:
    """Energy backend protocol over the HIP judge (``jr_mpnn_eval``):
    ``backend(positions, species, cell) -> BackendResult`` and
    ``energy_and_forces``; forces are analytic (-dE/dx)."""

    def __init__(self, params: MPNNParams | None = None, seed: int = 0,
                 device: int = 0):
        from jaxrens_amd import _lib
        self.params = params if params is not None else init_mpnn(seed)
        self.device = device
        self.r_cutoff = self.params.rc
        self._lib = _lib
        self.lib = _lib.load_library()
        self._keep = [np.ascontiguousarray(a, dtype=np.float32)
                      for a in self.params.arrays()]
        ptr = [a.ctypes.data_as(_lib._fp) for a in self._keep]
        p = self.params
        self.struct = _lib.JrMPNN(p.S, p.C, p.NB, 0, p.rc, p.core_eps, p.core_sig,
                                  p.core_rc, *ptr)

    def _eval(self, pos, types, cell, pressure=None, forces=True):
        """Batched: pos (B, N, 3), types (B, N), cell (B, 3, 3)."""
        _lib = self._lib
        pos = _lib.f64(pos)
        if pos.ndim == 2:
            pos = pos[None]
        B, N = pos.shape[:2]
        ty = _lib.i32(np.broadcast_to(np.asarray(types), (B, N)))
        cell = _lib.f64(np.broadcast_to(np.asarray(cell, dtype=np.float64), (B, 3, 3)))
        n = np.full(B, N, np.int32)
        pr = None if pressure is None else _lib.f64(np.broadcast_to(pressure, (B,)))
        U = np.empty(B)
        H = np.empty(B)
        F = np.zeros((B, N, 3)) if forces else None
        _lib.check(self.lib.jr_mpnn_eval(
            self.struct, self.device, B, N, _lib.dptr(pos), _lib.iptr(ty),
            _lib.iptr(n), _lib.dptr(cell), _lib.dptr(pr), _lib.dptr(U),
            _lib.dptr(H), _lib.dptr(F)))
        return U, H, F

    def __call__(self, positions, species, cell, max_neighbors=0,
                 ensemble_params=None):
        from jaxrens_amd.backends import BackendResult
        U, _, _ = self._eval(positions, species, cell, forces=False)
        return BackendResult(energy=float(U[0]))

    def energy_and_forces(self, positions, species, cell, max_neighbors=0,
                          ensemble_params=None):
        from jaxrens_amd.backends import BackendResult
        U, _, F = self._eval(positions, species, cell, forces=True)
        return BackendResult(energy=float(U[0]), forces=F[0])

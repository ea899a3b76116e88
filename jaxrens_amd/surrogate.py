"""Radial-descriptor MLP surrogate of the exact judge (no reference
counterpart; the reference has no surrogate).

Per atom i:
  descriptor  G_i[s][k] = 1/4 sum_{j within 3 sigma, t_j = s}
                  exp(-(r_ij - mu_k)^2 / (2 * 0.1^2)) u(r_ij / 3),
              mu_k = 0.8 + 1.7 k / 31, k = 0..31 (32 Gaussians, 0.8-2.5 sigma),
              u(x) = 1 - 28x^6 + 48x^7 - 21x^8 (smooth cutoff)
  input       x_i = [G_i[0], G_i[1], onehot(t_i), zeros] (80 values)
  network     h1 = silu(x W1 + b1), h2 = silu(h1 W2 + b2) (128 units each),
              e_i = h2 . w3 + b3
  energy      E = sum_i e_i,  forces -dE/dx (analytic, through the descriptor)
On the device the matrix products run on BF16 matrix cores with FP32
accumulation; the FP32 master parameters are trained with Adam.
"""

from __future__ import annotations

from dataclasses import dataclass

import numpy as np

D_IN = 80
HID = 128
NGAUSS = 32
# offsets into the FP32 parameter vector
OFF_W1 = 0
OFF_B1 = OFF_W1 + D_IN * HID
OFF_W2 = OFF_B1 + HID
OFF_B2 = OFF_W2 + HID * HID
OFF_W3 = OFF_B2 + HID
OFF_B3 = OFF_W3 + HID
N_PARAMS = OFF_B3 + 1


def init_surrogate(seed: int = 0, S: int = 2) -> np.ndarray:
    """Seeded initial FP32 parameters (scaled normal, zero biases)."""
    rng = np.random.default_rng(seed)
    p = np.zeros(N_PARAMS, np.float32)
    W1 = np.zeros((D_IN, HID), np.float32)
    n_in = NGAUSS * S + S
    W1[:n_in] = rng.standard_normal((n_in, HID)) / np.sqrt(n_in)
    p[OFF_W1:OFF_B1] = W1.ravel()
    p[OFF_W2:OFF_B2] = (rng.standard_normal((HID, HID)) / np.sqrt(HID)).ravel()
    p[OFF_W3:OFF_B3] = rng.standard_normal(HID) / np.sqrt(HID)
    return p


def _lib_mod():
    from jaxrens_amd import _lib
    return _lib


def sur_eval(params, pos, types, cell, S: int = 2, forces: bool = True,
             descriptors: bool = False, device: int = 0, base=None):
    """Surrogate energies (B,), forces (B, N, 3) and optionally the BF16
    descriptor rows (B, N, 80) as uint16 bit patterns, on the GPU.
    ``base`` = (eps, sig, rc) adds the pair baseline
    eps ((sig/r)^12 - (sig/rc)^12) for r < rc."""
    import ctypes as C
    _lib = _lib_mod()
    lib = _lib.load_library()
    pos = _lib.f64(pos)
    if pos.ndim == 2:
        pos = pos[None]
    B, N = pos.shape[:2]
    ty = _lib.i32(np.broadcast_to(np.asarray(types), (B, N)))
    cell = _lib.f64(np.broadcast_to(np.asarray(cell, dtype=np.float64), (B, 3, 3)))
    n = np.full(B, N, np.int32)
    p = np.ascontiguousarray(params, dtype=np.float32)
    U = np.empty(B)
    F = np.zeros((B, N, 3)) if forces else None
    X = np.zeros((B, N, D_IN), np.uint16) if descriptors else None
    xp = X.ctypes.data_as(C.POINTER(C.c_uint16)) if X is not None else None
    b = None if base is None else _lib.f64(base)
    _lib.check(lib.jr_sur_eval(p.ctypes.data_as(_lib._fp), _lib.dptr(b), device, B, N,
                               _lib.dptr(pos),
                               _lib.iptr(ty), _lib.iptr(n), _lib.dptr(cell), S,
                               _lib.dptr(U), _lib.dptr(F), xp))
    return U, F, X


def sur_grad(params, x_bits, c=None, device: int = 0):
    """Row energies e (n,) and, given c (n,), grad of sum_r c_r e_r."""
    import ctypes as C
    _lib = _lib_mod()
    lib = _lib.load_library()
    x = np.ascontiguousarray(x_bits, dtype=np.uint16)
    n = x.shape[0]
    p = np.ascontiguousarray(params, dtype=np.float32)
    e = np.empty(n, np.float32)
    g = np.empty(N_PARAMS, np.float32) if c is not None else None
    cc = np.ascontiguousarray(c, dtype=np.float32) if c is not None else None
    fp = _lib._fp
    _lib.check(lib.jr_sur_grad(p.ctypes.data_as(fp), device, n,
                               x.ctypes.data_as(C.POINTER(C.c_uint16)),
                               cc.ctypes.data_as(fp) if cc is not None else None,
                               e.ctypes.data_as(fp),
                               g.ctypes.data_as(fp) if g is not None else None))
    return e, g


def sur_adam(params, m, v, grad, lr: float, t: int, device: int = 0):
    _lib = _lib_mod()
    lib = _lib.load_library()
    p, m, v = (np.array(a, dtype=np.float32) for a in (params, m, v))
    g = np.ascontiguousarray(grad, dtype=np.float32)
    fp = _lib._fp
    _lib.check(lib.jr_sur_adam(device, p.ctypes.data_as(fp), m.ctypes.data_as(fp),
                               v.ctypes.data_as(fp), g.ctypes.data_as(fp), lr, t))
    return p, m, v


_MODES = {"plain": 0, "auto": 1, "screened": 2}


@dataclass
class LearnedConfig:
    """Settings of the surrogate-screened walk (``jr_learned``).

    mode: "plain" (the judge's forces steer and it decides every move; no
    surrogate), "auto" (two-stage delayed acceptance in an iteration whose
    held-out per-atom RMSE after the error model is <= rmse_bound and whose
    contours admit no state with U/N above train_max_epa, the range the
    surrogate trains on; plain otherwise) or "screened" (delayed acceptance
    in every iteration).  The surrogate
    trains online in "auto" and "screened" on the last ``ring_capacity``
    exactly evaluated configurations (those with U/N <= train_max_epa).
    delta_k: stage-1 slack in residual spreads of the error model.
    baseline: (eps, sig, rc) of the surrogate's pair term
    eps ((sig/r)^12 - (sig/rc)^12), r < rc <= 3; the network learns the
    judge's energy minus it (None: no baseline)."""

    mode: str = "auto"
    ring_capacity: int = 100_000
    train_batch: int = 32
    train_steps: int = 8
    min_train: int = 64
    min_heldout: int = 16
    lr: float = 1e-3
    rmse_bound: float = 0.05
    delta_k: float = 2.0
    train_max_epa: float = 50.0
    baseline: tuple[float, float, float] | None = None
    params: np.ndarray | None = None
    seed: int = 0

    def as_struct(self, n_species: int):
        from jaxrens_amd import _lib
        if self.mode not in _MODES:
            raise ValueError(f"learned mode must be one of {sorted(_MODES)}")
        p = self.params if self.params is not None else init_surrogate(self.seed, n_species)
        keep = np.ascontiguousarray(p, dtype=np.float32)
        if keep.shape != (N_PARAMS,):
            raise ValueError(f"surrogate parameters must have {N_PARAMS} values")
        be, bs, br = self.baseline if self.baseline is not None else (0.0, 0.0, 0.0)
        st = _lib.JrLearned(_MODES[self.mode], self.ring_capacity, self.train_batch,
                            self.train_steps, self.min_train, self.min_heldout, self.lr,
                            self.rmse_bound, self.delta_k, self.train_max_epa, be, bs, br,
                            keep.ctypes.data_as(_lib._fp))
        return st, keep

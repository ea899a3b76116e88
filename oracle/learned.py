"""ctypes wrapper of the CPU oracle for the learned part (test
infrastructure only): the message-passing judge in FP64."""

from __future__ import annotations

import ctypes as C

import numpy as np

from oracle.oracle import _d, _i, f64, i32, lib as _base_lib

_fp = C.POINTER(C.c_float)


class OrMPNN(C.Structure):
    _fields_ = [("S", C.c_int), ("C", C.c_int), ("NB", C.c_int), ("pad_", C.c_int),
                ("rc", C.c_double), ("core_eps", C.c_double),
                ("core_sig", C.c_double), ("core_rc", C.c_double),
                ("emb", _fp), ("W1", _fp), ("U1", _fp), ("c1", _fp),
                ("W2", _fp), ("U2", _fp), ("c2", _fp), ("wo", _fp)]


_ready = False


def lib():
    global _ready
    L = _base_lib()
    if not _ready:
        L.or_mpnn_energy_forces.restype = C.c_long
        L.or_mpnn_energy_forces.argtypes = [C.POINTER(OrMPNN), C.c_int,
                                            C.POINTER(C.c_double),
                                            C.POINTER(C.c_int32),
                                            C.POINTER(C.c_double),
                                            C.POINTER(C.c_double),
                                            C.POINTER(C.c_double)]
        _ready = True
    return L


def mpnn_struct(p) -> tuple[OrMPNN, list]:
    keep = [np.ascontiguousarray(a, dtype=np.float32) for a in p.arrays()]
    ptrs = [a.ctypes.data_as(_fp) for a in keep]
    s = OrMPNN(p.S, p.C, p.NB, 0, p.rc, p.core_eps, p.core_sig, p.core_rc, *ptrs)
    return s, keep


def mpnn_energy_forces(p, pos, types, cell, forces=True):
    """Energy (and forces) of one configuration; also the edge count."""
    s, keep = mpnn_struct(p)
    pos = f64(pos)
    N = pos.shape[0]
    ty = i32(types)
    cell = f64(np.asarray(cell).reshape(9))
    E = C.c_double()
    F = np.zeros((N, 3)) if forces else None
    ne = lib().or_mpnn_energy_forces(C.byref(s), N, _d(pos), _i(ty), _d(cell),
                                     C.byref(E), _d(F))
    return E.value, F, ne


_sur_ready = False


def _sur_lib():
    global _sur_ready
    L = lib()
    if not _sur_ready:
        dp, ip, fp = C.POINTER(C.c_double), C.POINTER(C.c_int32), _fp
        L.or_sur_descriptors.restype = None
        L.or_sur_descriptors.argtypes = [C.c_int, dp, ip, dp, C.c_int, dp]
        L.or_sur_energy_forces.restype = C.c_double
        L.or_sur_energy_forces.argtypes = [fp, C.c_int, C.c_int, dp, ip, dp, C.c_int, dp, dp]
        L.or_sur_loss_grad.restype = C.c_double
        L.or_sur_loss_grad.argtypes = [fp, C.c_int, C.c_int, ip, dp, dp, dp]
        L.or_adam_step.restype = None
        L.or_adam_step.argtypes = [C.c_int, fp, fp, fp, fp, C.c_double, C.c_int]
        _sur_ready = True
    return L


def _f32(a):
    return np.ascontiguousarray(a, dtype=np.float32)


def sur_descriptors(pos, types, cell, S=2):
    pos = f64(pos)
    N = pos.shape[0]
    x = np.zeros((N, 80))
    _sur_lib().or_sur_descriptors(N, _d(pos), _i(i32(types)),
                                  _d(f64(np.asarray(cell).reshape(9))), S, _d(x))
    return x


def sur_energy_forces(params, pos, types, cell, S=2, bf16=True, forces=True, base=None):
    """Surrogate energy (and forces); ``base`` = (eps, sig, rc) adds the
    pair baseline eps ((sig/r)^12 - (sig/rc)^12) for r < rc."""
    pos = f64(pos)
    N = pos.shape[0]
    p = _f32(params)
    F = np.zeros((N, 3)) if forces else None
    b = None if base is None else f64(base)
    E = _sur_lib().or_sur_energy_forces(p.ctypes.data_as(_fp), int(bf16), N, _d(pos),
                                        _i(i32(types)),
                                        _d(f64(np.asarray(cell).reshape(9))), S, _d(b),
                                        _d(F))
    return E, F


def sur_loss_grad(params, n_atoms, x, E, bf16=True):
    p = _f32(params)
    n_atoms = i32(n_atoms)
    x = f64(x)
    E = f64(E)
    g = np.zeros(p.size)
    loss = _sur_lib().or_sur_loss_grad(p.ctypes.data_as(_fp), int(bf16), len(n_atoms),
                                       _i(n_atoms), _d(x), _d(E), _d(g))
    return loss, g


def adam_step(p, m, v, g, lr, t):
    p, m, v, g = (_f32(a).copy() for a in (p, m, v, g))
    _sur_lib().or_adam_step(p.size, p.ctypes.data_as(_fp), m.ctypes.data_as(_fp),
                            v.ctypes.data_as(_fp), g.ctypes.data_as(_fp), lr, t)
    return p, m, v

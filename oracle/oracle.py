"""ctypes wrapper of the CPU oracle (test infrastructure only).

Imported by tests/, __graft_entry__.smoke() and bench.py's CPU-baseline leg;
never by the product package.  Build with ``make -C oracle``.
"""

from __future__ import annotations

import ctypes as C
import subprocess
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
LIB = HERE / "build" / "libjroracle.so"

_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int32)


class OrPot(C.Structure):
    _fields_ = [("n_species", C.c_int), ("rc2", C.c_double), ("eps4", _dp),
                ("eps24", _dp), ("sig2", _dp), ("img", C.c_int * 3),
                ("dmin_pos", _dp), ("dmin_types", _dp), ("chunks", C.c_int)]


class OrMove(C.Structure):
    _fields_ = [("type", C.c_int), ("n_reflect", C.c_int), ("n_leapfrog", C.c_int),
                ("flat_v_prior", C.c_int), ("n_species", C.c_int), ("species_mask", C.c_int),
                ("perturb_angle", C.c_double), ("max_vpa", C.c_double),
                ("min_vpa", C.c_double), ("min_aspect", C.c_double),
                ("refresh_direction", C.c_int), ("dir_slot", C.c_int)]


class OrPop(C.Structure):
    _fields_ = [("R", C.c_int), ("K", C.c_int), ("Nmax", C.c_int), ("S", C.c_int),
                ("pos", _dp), ("types", _ip), ("n_atoms", _ip), ("cell", _dp),
                ("energy", _dp), ("dir", _dp), ("dirx", _dp), ("n_xslots", C.c_int)]


class OrRun(C.Structure):
    _fields_ = [("pressure", _dp), ("mu", _dp), ("step_sizes", _dp)]


class OrDraws(C.Structure):
    _fields_ = [("sel_u", _dp), ("clone_u", _dp), ("extra_u", _dp), ("schedule", _ip),
                ("normals", _dp), ("uniforms", _dp), ("re_u", _dp)]


class OrIterOut(C.Structure):
    _fields_ = [("emax", _dp), ("worst", _ip), ("clone", _ip), ("walk", _ip),
                ("dead_energy", _dp), ("dead_volume", _dp), ("counters", _ip),
                ("re_acc", _ip), ("re_att", _ip)]


_lib = None


def build() -> None:
    subprocess.run(["make", "-C", str(HERE), "-s"], check=True)


def lib() -> C.CDLL:
    global _lib
    if _lib is None:
        if not LIB.exists():
            build()
        L = C.CDLL(str(LIB))
        L.or_det3.restype = C.c_double
        L.or_det3.argtypes = [_dp]
        L.or_cbrt.restype = C.c_double
        L.or_cbrt.argtypes = [C.c_double]
        L.or_exp.restype = C.c_double
        L.or_exp.argtypes = [C.c_double]
        L.or_ipow.restype = C.c_double
        L.or_ipow.argtypes = [C.c_double, C.c_int]
        L.or_check_cell_shape.restype = C.c_int
        L.or_check_cell_shape.argtypes = [_dp, C.c_int, C.c_double, C.c_double, C.c_double]
        L.or_ensemble_energy.restype = C.c_double
        L.or_ensemble_energy.argtypes = [C.c_int, _ip, _dp, C.c_double, _dp, C.c_int]
        L.or_energy_forces.restype = None
        L.or_energy_forces.argtypes = [C.c_int, _dp, _ip, _dp, C.POINTER(OrPot),
                                       C.c_double, _dp, _dp, _dp, _dp]
        L.or_walk_chain.restype = None
        L.or_walk_chain.argtypes = [C.c_int, C.c_int, _dp, _ip, _dp, _dp, _dp, _dp,
                                    C.c_size_t, C.POINTER(OrPot), C.c_double, _dp, C.c_double,
                                    _dp, C.POINTER(OrMove), C.c_int, _ip, C.c_int,
                                    _dp, _dp, _ip]
        L.or_ns_iteration.restype = C.c_int
        L.or_ns_iteration.argtypes = [C.POINTER(OrPop), C.POINTER(OrPot),
                                      C.POINTER(OrRun), C.POINTER(OrMove), C.c_int,
                                      C.c_int, C.c_int, C.c_int, C.c_int, _ip,
                                      C.POINTER(OrDraws), C.POINTER(OrIterOut), C.c_int]
        L.or_replica_exchange.restype = None
        L.or_replica_exchange.argtypes = [C.POINTER(OrPop), C.POINTER(OrRun), _dp,
                                          C.c_int, _ip, _dp, _ip, _ip]
        L.or_initial_walk.restype = None
        L.or_initial_walk.argtypes = [C.POINTER(OrPop), C.POINTER(OrPot),
                                      C.POINTER(OrRun), C.POINTER(OrMove), C.c_int,
                                      _dp, _ip, C.c_int, _dp, _dp, _ip, C.c_int]
        L.or_trial_moves.restype = None
        L.or_trial_moves.argtypes = [C.POINTER(OrPop), C.POINTER(OrPot),
                                     C.POINTER(OrRun), C.POINTER(OrMove), C.c_int,
                                     C.c_int, _dp, _dp, C.c_int, _ip, _dp, _dp, _ip,
                                     _ip, C.c_int]
        _lib = L
    return _lib


def _d(a):
    if a is None:
        return None
    assert a.dtype == np.float64 and a.flags.c_contiguous, (a.dtype, a.flags)
    return a.ctypes.data_as(_dp)


def _i(a):
    if a is None:
        return None
    assert a.dtype == np.int32 and a.flags.c_contiguous
    return a.ctypes.data_as(_ip)


def f64(a):
    return np.ascontiguousarray(np.asarray(a, dtype=np.float64))


def i32(a):
    return np.ascontiguousarray(np.asarray(a, dtype=np.int32))


class Pot:
    """Pair table as the oracle takes it (from an LJ backend's table)."""

    def __init__(self, n_species, eps4, eps24, sig2, rc2, images=(0, 0, 0), chunks=1):
        self.keep = (f64(eps4), f64(eps24), f64(sig2))
        self.s = OrPot(int(n_species), float(rc2), _d(self.keep[0]),
                       _d(self.keep[1]), _d(self.keep[2]),
                       (C.c_int * 3)(*(int(h) for h in images)), None, None,
                       int(chunks))

    @classmethod
    def from_table(cls, t):
        return cls(t.n_species, t.eps4, t.eps24, t.sig2, t.rc2,
                   getattr(t, "images", (0, 0, 0)), getattr(t, "eval_chunks", 1))

    def with_min_distance(self, floors_pos, floors_types=None):
        """Gate the walk with minimum-distance floors ``[S][S]`` (NULL when
        None): atom and cell moves by ``floors_pos``, species moves by
        ``floors_types``."""
        self.keep = self.keep + tuple(None if f is None else f64(f)
                                      for f in (floors_pos, floors_types))
        self.s.dmin_pos = None if floors_pos is None else _d(self.keep[-2])
        self.s.dmin_types = None if floors_types is None else _d(self.keep[-1])
        return self


def energy_forces(pot: Pot, pos, types, cell, pressure=0.0, mu=None, forces=True):
    pos = f64(pos)
    N = pos.shape[0]
    t = i32(types)
    c = f64(cell).reshape(9)
    m = None if mu is None else f64(mu)
    U = np.zeros(1)
    H = np.zeros(1)
    F = np.zeros((max(N, 1), 3)) if forces else None
    lib().or_energy_forces(N, _d(pos if N else np.zeros((1, 3))), _i(t if N else np.zeros(1, np.int32)),
                           _d(c), C.byref(pot.s), float(pressure), _d(m),
                           _d(U), _d(H), _d(F))
    return U[0], H[0], (F[:N] if forces else None)


def moves_array(jr_moves):
    arr = (OrMove * len(jr_moves))()
    for i, m in enumerate(jr_moves):
        arr[i] = OrMove(m.type, m.n_reflect, m.n_leapfrog, m.flat_v_prior,
                        m.n_species, m.species_mask, m.perturb_angle, m.max_vpa,
                        m.min_vpa, m.min_aspect, getattr(m, "refresh_direction", 0),
                        getattr(m, "dir_slot", 0))
    return arr


class Population:
    """Host copy of the population in the device layout."""

    def __init__(self, pos, types, n_atoms, cell, energy, direction, dirx=None):
        self.pos = f64(pos)
        self.types = i32(types)
        self.n_atoms = i32(n_atoms)
        self.cell = f64(np.asarray(cell).reshape(self.pos.shape[0], self.pos.shape[1], 9))
        self.energy = f64(energy)
        self.dir = f64(direction)
        # directions of the species-scoped GMC moves, (n_xslots, R, K, Nmax, 3)
        self.dirx = f64(np.zeros((0,) + self.dir.shape) if dirx is None else dirx)

    @classmethod
    def from_engine_dict(cls, d):
        return cls(d["positions"], d["types"], d["n_atoms"], d["cell"], d["energy"],
                   d["direction"], d.get("direction_slots"))

    def copy(self):
        return Population(self.pos.copy(), self.types.copy(), self.n_atoms.copy(),
                          self.cell.copy(), self.energy.copy(), self.dir.copy(),
                          self.dirx.copy())

    def struct(self, S):
        R, K, Nm = self.pos.shape[:3]
        nx = self.dirx.shape[0]
        return OrPop(R, K, Nm, S, _d(self.pos), _i(self.types), _i(self.n_atoms),
                     _d(self.cell), _d(self.energy), _d(self.dir),
                     _d(self.dirx) if nx else None, nx)


def ns_iteration(pop: Population, pot: Pot, pressure, mu, step_sizes, moves, n_steps,
                 n_extra, tape, it=0, n_swap_cycles=1, pair_valid=None, do_re=True,
                 n_threads=0, n_cull=1):
    """One iteration on ``pop`` (in place); ``tape`` is a draws.Tape.
    ``worst``, ``clone``, ``dead_energy`` and ``dead_volume`` are (R,) for
    n_cull = 1 and (R, n_cull) otherwise."""
    R, K, Nm = pop.pos.shape[:3]
    S = pot.s.n_species
    Cc = int(n_cull)
    n_walk = Cc + n_extra
    n_moves = len(moves)
    pr, m, ss = f64(pressure), f64(np.asarray(mu).reshape(R, S)), f64(step_sizes)
    run = OrRun(_d(pr), _d(m), _d(ss))
    keep = [f64(tape.sel_u[it]), f64(tape.clone_u[it]), f64(tape.extra_u[it]),
            i32(tape.schedule[it]), f64(tape.normals[it]), f64(tape.uniforms[it]),
            f64(tape.re_u[it])]
    dr = OrDraws(_d(keep[0]), _d(keep[1]), _d(keep[2]), _i(keep[3]), _d(keep[4]),
                 _d(keep[5]), _d(keep[6]) if (do_re and R > 1) else None)
    out = {
        "emax": np.zeros(R), "worst": np.zeros((R, Cc), np.int32),
        "clone": np.zeros((R, Cc), np.int32),
        "walk": np.zeros((R, n_walk), np.int32), "dead_energy": np.zeros((R, Cc)),
        "dead_volume": np.zeros((R, Cc)), "counters": np.zeros((R, n_moves, 7), np.int32),
        "re_acc": np.zeros(max(R - 1, 1), np.int32), "re_att": np.zeros(max(R - 1, 1), np.int32),
    }
    o = OrIterOut(_d(out["emax"]), _i(out["worst"]), _i(out["clone"]), _i(out["walk"]),
                  _d(out["dead_energy"]), _d(out["dead_volume"]), _i(out["counters"]),
                  _i(out["re_acc"]), _i(out["re_att"]))
    pv = i32(np.ones(max(R - 1, 1)) if pair_valid is None else pair_valid)
    ps = pop.struct(S)
    mv = moves_array(moves)
    rc = lib().or_ns_iteration(C.byref(ps), C.byref(pot.s), C.byref(run), mv, n_moves,
                               n_steps, n_extra, Cc, n_swap_cycles, _i(pv), C.byref(dr),
                               C.byref(o), n_threads)
    if rc != 0:
        raise RuntimeError(f"or_ns_iteration failed: {rc}")
    if Cc == 1:
        for k in ("worst", "clone", "dead_energy", "dead_volume"):
            out[k] = out[k][:, 0]
    return out


def trial_moves(pop: Population, pot: Pot, pressure, mu, step_sizes, moves, move_idx,
                trial_step, emax, sample_idx, normals, uniforms, n_threads=0):
    R, K, Nm = pop.pos.shape[:3]
    S = pot.s.n_species
    pr, m, ss = f64(pressure), f64(np.asarray(mu).reshape(R, S)), f64(step_sizes)
    run = OrRun(_d(pr), _d(m), _d(ss))
    idx = i32(sample_idx)
    n = idx.shape[1]
    acc = np.zeros((R, n), np.int32)
    rs = np.zeros((R, n), np.int32)
    ts, em, nn, uu = f64(trial_step), f64(emax), f64(normals), f64(uniforms)
    ps = pop.struct(S)
    lib().or_trial_moves(C.byref(ps), C.byref(pot.s), C.byref(run), moves_array(moves),
                         len(moves), move_idx, _d(ts), _d(em), n, _i(idx), _d(nn),
                         _d(uu), _i(acc), _i(rs), n_threads)
    return acc.astype(bool), rs


def initial_walk(pop: Population, pot: Pot, pressure, mu, step_sizes, moves, emax,
                 schedule, normals, uniforms, n_threads=0):
    """Fixed-contour walk of every walker of ``pop`` (in place); returns the
    counters [R, n_moves, 7]."""
    R, K, Nm = pop.pos.shape[:3]
    S = pot.s.n_species
    pr, m, ss = f64(pressure), f64(np.asarray(mu).reshape(R, S)), f64(step_sizes)
    run = OrRun(_d(pr), _d(m), _d(ss))
    sc = i32(schedule)
    em, nn, uu = f64(emax), f64(normals), f64(uniforms)
    ctr = np.zeros((R, len(moves), 7), np.int32)
    ps = pop.struct(S)
    lib().or_initial_walk(C.byref(ps), C.byref(pot.s), C.byref(run), moves_array(moves),
                          len(moves), _d(em), _i(sc), len(sc), _d(nn), _d(uu), _i(ctr),
                          n_threads)
    return ctr


def replica_exchange(pop: Population, pressure, mu, emax, re_u, n_swap_cycles=1,
                     pair_valid=None, S=1):
    R = pop.pos.shape[0]
    pr, m = f64(pressure), f64(np.asarray(mu).reshape(R, S))
    run = OrRun(_d(pr), _d(m), None)
    acc = np.zeros(max(R - 1, 1), np.int32)
    att = np.zeros(max(R - 1, 1), np.int32)
    pv = i32(np.ones(max(R - 1, 1)) if pair_valid is None else pair_valid)
    ps = pop.struct(S)
    ru = f64(re_u)
    em = f64(emax)
    lib().or_replica_exchange(C.byref(ps), C.byref(run), _d(em), n_swap_cycles, _i(pv),
                              _d(ru), _i(acc), _i(att))
    return acc, att


def cbrt(x):
    return lib().or_cbrt(float(x))


def exp(x):
    return lib().or_exp(float(x))


def det3(cell):
    return lib().or_det3(_d(f64(cell).reshape(9)))


def check_cell_shape(cell, n_atoms, max_vpa, min_vpa, min_aspect):
    return bool(lib().or_check_cell_shape(_d(f64(cell).reshape(9)), n_atoms, max_vpa,
                                          min_vpa, min_aspect))

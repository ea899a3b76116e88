"""Device-resident nested-sampling engine: the drop-in for the reference's
``ns_step`` / ``replica_exchange_step`` / step-size trials, on one GPU.

Reference call sites replaced: ``sampling/run_loop.py:419-551`` (jit'd
``ns_step`` per iteration, RE pass, adaptation), ``sampling/nested_sampling.py:
391-686`` (``ns_step``), ``sampling/moves/replica_exchange.py:318-436``.
Replica axis P of the reference's ``PmapVmapRuns(G=1, P)`` is the leading
axis R here; the population lives in HBM between calls.
"""

from __future__ import annotations

import ctypes as C
import math
from dataclasses import dataclass, field

import numpy as np

from . import _lib
from .backends import Hamiltonian, LJBackend, unwrap_model
from .draws import Tape
from .moves import MWG, CellSpec
from .prior_mass import dead_log_weight, final_log_evidence, logaddexp_accumulate


@dataclass
class RunConfig:
    n_replicas: int
    n_walkers: int
    n_atoms: int | list[int]          # per replica
    n_mcmc_steps: int = 20
    n_extra: int = 0
    n_cull: int = 1
    seed: int = 42
    pressures: list[float] | None = None
    chemical_potentials: list[list[float]] | None = None
    re_interval: int = 0
    n_swap_cycles: int = 1
    waves_per_block: int = 0
    device: int = 0


def _model_of(backend):
    """(model, default pressure, default mu) of an LJ or message-passing
    backend, bare or inside a Hamiltonian with ensemble terms."""
    from .mpnn import MPNNBackend
    model = unwrap_model(backend)
    if not isinstance(model, (LJBackend, MPNNBackend)):
        raise TypeError("the device walk needs an LJBackend or MPNNBackend model")
    pressure, mu = None, None
    if isinstance(backend, Hamiltonian):
        pm = backend.ensemble_params_of(None)
        if pm is None:
            raise TypeError("only EnsembleTerm terms are fused on the device")
        pressure, mu = pm
    return model, pressure, mu


def grid_positions_in_cell(rng: np.random.Generator, cell, n_atoms: int,
                           min_distance: float) -> np.ndarray:
    """``init/positions.py:34-100``: ``n_atoms`` distinct sites of the grid
    with floor(|a_i| / min_distance) points along cell vector i, at the
    centres of the grid cells (fractions stored in float32, as the reference
    stores them), drawn without replacement; Cartesian ``frac @ cell``.
    Sites are at least ``min_distance`` apart in orthogonal cells."""
    cell = np.asarray(cell, dtype=np.float64)
    n_ax = [max(1, int(float(np.linalg.norm(cell[i])) / min_distance)) for i in range(3)]
    n_grid = n_ax[0] * n_ax[1] * n_ax[2]
    if n_grid < n_atoms:
        raise ValueError(f"grid too coarse: {n_grid} sites available "
                         f"({n_ax[0]}x{n_ax[1]}x{n_ax[2]}), {n_atoms} atoms requested")
    axes = [(np.arange(m) + 0.5) / m for m in n_ax]
    g = np.stack([x.ravel() for x in np.meshgrid(*axes, indexing="ij")], axis=1)
    frac = g.astype(np.float32).astype(np.float64)
    return frac[rng.choice(n_grid, size=n_atoms, replace=False)] @ cell


def _aspect(cell: np.ndarray, volume: float) -> float:
    """Smallest height of the cell over V^(1/3), over the three faces
    (``utils/cell.py:80-102``)."""
    vc = volume ** (1.0 / 3.0)
    out = np.inf
    for i, j, k in ((0, 1, 2), (1, 2, 0), (0, 2, 1)):
        n = np.cross(cell[i], cell[j])
        out = min(out, abs(float(np.dot(n / np.linalg.norm(n), cell[k]))) / vc)
    return out


def cell_shape_walk(rng: np.random.Generator, cell, n_steps: int, step: float,
                    min_aspect: float) -> np.ndarray:
    """Volume-preserving random walk over the cell shape (``init/cells.py:
    140-230``): each step proposes, with equal odds, a stretch (axis pair
    (0,1), (0,2) or (1,2) scaled by exp(+-r), r ~ N(0, step)) or a shear (one
    cell vector moved by r1 u + r2 w in the plane of the other two, u along
    the first of them, w the unit part of the second orthogonal to u,
    r1, r2 ~ N(0, step)); the proposal is rescaled to the starting volume and
    kept when its aspect ratio is at least ``min_aspect``."""
    cell = np.array(cell, dtype=np.float64)
    v0 = abs(np.linalg.det(cell))
    for _ in range(n_steps):
        new = cell.copy()
        if rng.random() < 0.5:
            i, j = ((0, 1), (0, 2), (1, 2))[rng.integers(3)]
            r = step * rng.standard_normal()
            new[:, i] *= np.exp(r)
            new[:, j] *= np.exp(-r)
        else:
            idx = int(rng.integers(3))
            a, b = [c for c in range(3) if c != idx]
            u = cell[a] / np.linalg.norm(cell[a])
            w = cell[b] - np.dot(cell[b], u) * u
            w /= np.linalg.norm(w)
            r1, r2 = step * rng.standard_normal(2)
            new[idx] = cell[idx] + r1 * u + r2 * w
        new *= (v0 / abs(np.linalg.det(new))) ** (1.0 / 3.0)
        if _aspect(new, v0) >= min_aspect:
            cell = new
    return cell


class NSEngine:
    """One handle of ``libjrhip``: population, run parameters and draws."""

    def __init__(self, backend, moves, run: RunConfig, cell: CellSpec | dict | None = None,
                 learned=None, constraints=None, symbol_map: dict[int, str] | None = None):
        """``learned``: :class:`~jaxrens_amd.surrogate.LearnedConfig` of the
        surrogate-screened walk.  The judge is the backend's model (the
        message-passing potential, or the pair potential); with the
        message-passing backend and ``learned`` None the walk is plain.

        ``constraints``: the ``constraints:`` list of the reference's YAML
        (``minimum_distance`` entries, :mod:`jaxrens_amd.constraints`);
        ``symbol_map`` names the type codes for per-pair floors."""
        from .mpnn import MPNNBackend
        self.lib = _lib.load_library()
        self.model, p_default, mu_default = _model_of(backend)
        self.is_mpnn = isinstance(self.model, MPNNBackend)
        self.is_learned = self.is_mpnn or learned is not None
        self.mwg = MWG(moves, cell)
        self.run = run
        R, K = run.n_replicas, run.n_walkers
        n_atoms = run.n_atoms
        self.n_atoms_rep = (np.full(R, int(n_atoms), np.int32)
                            if np.isscalar(n_atoms) else np.asarray(n_atoms, np.int32))
        if self.n_atoms_rep.shape != (R,):
            raise ValueError("n_atoms must be an int or one value per replica")
        self.Nmax = int(self.n_atoms_rep.max())
        self.n_replicas, self.n_walkers = R, K
        self.n_walk = run.n_cull + run.n_extra
        self.n_moves = self.mwg.n_moves
        self._learned_keep = None
        if self.is_mpnn:
            self.S = self.model.params.S
            pot, self._pot_keep = _lib.JrPot(), None
        else:
            self.S = self.model.n_species
            pot, self._pot_keep = self.model.table.as_struct(self.model._prec_code)
        self._moves_arr = self.mwg.device_table()
        self._weights = _lib.f64(self.mwg.weights)
        cfg = _lib.JrConfig()
        cfg.n_replicas = R
        cfg.n_walkers = K
        cfg.n_atoms_max = self.Nmax
        cfg.n_moves = self.n_moves
        cfg.n_mcmc_steps = run.n_mcmc_steps
        cfg.n_extra = run.n_extra
        cfg.n_cull = run.n_cull
        cfg.n_swap_cycles = run.n_swap_cycles
        cfg.re_interval = run.re_interval
        cfg.device = run.device
        cfg.waves_per_block = run.waves_per_block
        cfg.seed = run.seed & 0xFFFFFFFFFFFFFFFF
        cfg.pot = pot
        cfg.moves = C.cast(self._moves_arr, C.POINTER(_lib.JrMove))
        cfg.move_weights = _lib.dptr(self._weights)
        if self.is_mpnn:
            cfg.mpnn = C.pointer(self.model.struct)
        if learned is not None:
            self._learned_st, self._learned_keep = learned.as_struct(self.S)
            cfg.learned = C.pointer(self._learned_st)
        from .constraints import constraint_tables
        self.min_distance, self.min_distance_types = constraint_tables(constraints, self.S,
                                                                       symbol_map)
        self._dmin_keep = [None if t is None else _lib.f64(t)
                           for t in (self.min_distance, self.min_distance_types)]
        if self._dmin_keep[0] is not None:
            cfg.min_distance = _lib.dptr(self._dmin_keep[0])
        if self._dmin_keep[1] is not None:
            cfg.min_distance_types = _lib.dptr(self._dmin_keep[1])
        h = C.c_void_p()
        _lib.check(self.lib.jr_create(C.byref(cfg), C.byref(h)))
        self.h = h
        self.iteration = 0
        self.log_evidence = np.full(R, -np.inf)
        pressures = (run.pressures if run.pressures is not None
                     else [0.0 if p_default is None else p_default] * R)
        mu = run.chemical_potentials
        if mu is None:
            mu = np.zeros((R, self.S))
            if mu_default is not None:
                mu[:, : len(mu_default)] = mu_default
        self.pressure = _lib.f64(pressures)
        self.mu = _lib.f64(np.asarray(mu, dtype=np.float64).reshape(R, self.S))
        self.step_sizes = _lib.f64(np.tile(self.mwg.default_step_sizes, (R, 1)))
        self.set_run_params()

    # ------------------------------------------------------------ lifetime
    def close(self):
        if getattr(self, "h", None):
            self.lib.jr_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _check(self, rc):
        _lib.check(rc, self.h)

    # ---------------------------------------------------------- parameters
    def set_run_params(self, pressure=None, mu=None, step_sizes=None):
        if pressure is not None:
            self.pressure = _lib.f64(pressure)
        if mu is not None:
            self.mu = _lib.f64(np.asarray(mu).reshape(self.n_replicas, self.S))
        if step_sizes is not None:
            self.step_sizes = _lib.f64(
                np.asarray(step_sizes).reshape(self.n_replicas, self.n_moves))
        self._check(self.lib.jr_set_run_params(
            self.h, _lib.dptr(self.pressure), _lib.dptr(self.mu),
            _lib.dptr(self.step_sizes)))

    # ---------------------------------------------------------- population
    def set_population(self, positions, types, cell, n_atoms=None, direction=None):
        R, K, Nm = self.n_replicas, self.n_walkers, self.Nmax
        pos = np.zeros((R, K, Nm, 3))
        p = np.asarray(positions, dtype=np.float64)
        pos[:, :, : p.shape[2]] = p
        ty = np.zeros((R, K, Nm), np.int32)
        t = np.asarray(types)
        ty[:, :, : t.shape[2]] = t
        na = (np.broadcast_to(self.n_atoms_rep[:, None], (R, K)) if n_atoms is None
              else np.asarray(n_atoms))
        na = _lib.i32(na)
        cl = _lib.f64(np.asarray(cell, dtype=np.float64).reshape(R, K, 9))
        d = None
        if direction is not None:
            d = np.zeros((R, K, Nm, 3))
            dd = np.asarray(direction, dtype=np.float64)
            d[:, :, : dd.shape[2]] = dd
            d = _lib.f64(d)
        self._check(self.lib.jr_set_population(
            self.h, _lib.dptr(_lib.f64(pos)), _lib.iptr(ty), _lib.iptr(na),
            _lib.dptr(cl), _lib.dptr(d)))

    def get_population(self) -> dict:
        R, K, Nm = self.n_replicas, self.n_walkers, self.Nmax
        pos = np.empty((R, K, Nm, 3))
        d = np.empty((R, K, Nm, 3))
        ty = np.empty((R, K, Nm), np.int32)
        na = np.empty((R, K), np.int32)
        cl = np.empty((R, K, 3, 3))
        e = np.empty((R, K))
        self._check(self.lib.jr_get_population(
            self.h, _lib.dptr(pos), _lib.iptr(ty), _lib.iptr(na), _lib.dptr(cl),
            _lib.dptr(e), _lib.dptr(d)))
        out = {"positions": pos, "types": ty, "n_atoms": na, "cell": cl,
               "energy": e, "direction": d}
        if self.mwg.n_dir_slots > 1:
            out["direction_slots"] = self.direction_slots()
        return out

    def direction_slots(self) -> np.ndarray:
        """Persistent directions of the species-scoped GMC moves,
        (n_slots - 1, R, K, Nmax, 3); slot s + 1 belongs to the moves with
        ``mwg.dir_slots == s + 1``."""
        R, K, Nm = self.n_replicas, self.n_walkers, self.Nmax
        n = self.mwg.n_dir_slots - 1
        out = np.empty((n, R, K, Nm, 3))
        for s in range(n):
            self._check(self.lib.jr_get_direction_slot(self.h, s + 1, _lib.dptr(out[s])))
        return out

    def set_direction_slots(self, dirs):
        dirs = _lib.f64(dirs)
        for s in range(self.mwg.n_dir_slots - 1):
            self._check(self.lib.jr_set_direction_slot(self.h, s + 1,
                                                       _lib.dptr(_lib.f64(dirs[s]))))

    def init_walkers(self, rng: np.random.Generator, v_lo_per_atom: float,
                     v_hi_per_atom: float, flat_V_prior: bool = True,
                     species_p=None, grid_distance: float | None = None,
                     shared_volume: bool = False, shape_walk_steps: int = 0,
                     shape_step: float = 0.05, min_aspect: float = 0.8):
        """Cubic cells from the volume prior and uniform positions
        (``init/cells.py:82-137``, ``init/positions.py:15-32``), or with
        ``grid_distance`` sites of a grid spaced by it
        (:func:`grid_positions_in_cell`); species Bernoulli/categorical per
        atom with probabilities ``species_p``.

        ``shared_volume``: the reference's ``random_initialise_cell`` path
        (``cli/resolve.py:285-343,1020-1035``): one cubic cell per replica
        from the volume prior, and each walker's cell from it by
        ``shape_walk_steps`` steps of :func:`cell_shape_walk` (step
        ``shape_step``, the run's ``min_aspect`` ratio); all walkers of a
        replica then share one volume."""
        R, K = self.n_replicas, self.n_walkers
        Nm = self.Nmax
        pos = np.zeros((R, K, Nm, 3))
        cell = np.zeros((R, K, 3, 3))
        ty = np.zeros((R, K, Nm), np.int32)
        for r in range(R):
            n = int(self.n_atoms_rep[r])
            u = np.full(K, rng.random()) if shared_volume else rng.random(K)
            v_lo, v_hi = n * v_lo_per_atom, n * v_hi_per_atom
            if flat_V_prior:
                vol = v_lo + u * (v_hi - v_lo)
            else:
                p = float(n + 1)
                a = (v_lo / v_hi) ** p if v_hi > 0 else 0.0
                vol = v_hi * (a + u * (1.0 - a)) ** (1.0 / p)
            lc = vol ** (1.0 / 3.0)
            cell[r] = lc[:, None, None] * np.eye(3)
            if shape_walk_steps > 0:
                for k in range(K):
                    cell[r, k] = cell_shape_walk(rng, cell[r, k], shape_walk_steps,
                                                 shape_step, min_aspect)
            if grid_distance is None and shape_walk_steps > 0:
                frac = rng.random((K, n, 3))
                pos[r, :, :n] = np.einsum("knd,kde->kne", frac, cell[r])
            elif grid_distance is None:
                frac = rng.random((K, n, 3))
                pos[r, :, :n] = frac * lc[:, None, None]
            else:
                for k in range(K):
                    pos[r, k, :n] = grid_positions_in_cell(rng, cell[r, k], n, grid_distance)
            if species_p is not None:
                ty[r, :, :n] = rng.choice(len(species_p), size=(K, n), p=species_p)
        self.set_population(pos, ty, cell)

    # ------------------------------------------------------------- draws
    @staticmethod
    def _tape_struct(t: Tape):
        keep = [_lib.f64(t.sel_u), _lib.f64(t.clone_u), _lib.f64(t.extra_u),
                _lib.i32(t.schedule), _lib.f64(t.normals), _lib.f64(t.uniforms),
                _lib.f64(t.re_u)]
        s = _lib.JrTape(_lib.dptr(keep[0]), _lib.dptr(keep[1]), _lib.dptr(keep[2]),
                        _lib.iptr(keep[3]), _lib.dptr(keep[4]), _lib.dptr(keep[5]),
                        _lib.dptr(keep[6]))
        return s, keep

    def last_tape(self) -> Tape:
        R, K, Nm, T = self.n_replicas, self.n_walkers, self.Nmax, self.run.n_mcmc_steps
        nc = max(self.run.n_swap_cycles, 1)
        t = Tape(np.zeros((1, R, K)), np.zeros((1, R)), np.zeros((1, R, K)),
                 np.zeros((1, T), np.int32),
                 np.zeros((1, R, self.n_walk, T, 3, Nm)),
                 np.zeros((1, R, self.n_walk, T, 4)),
                 np.zeros((1, nc, 2, max(R // 2, 1), 2)))
        self._check(self.lib.jr_get_last_tape(
            self.h, _lib.dptr(t.sel_u), _lib.dptr(t.clone_u), _lib.dptr(t.extra_u),
            _lib.iptr(t.schedule), _lib.dptr(t.normals), _lib.dptr(t.uniforms),
            _lib.dptr(t.re_u)))
        return t

    # -------------------------------------------------------- iterations
    def ns_iterations(self, n_iter: int, tape: Tape | None = None,
                      graph: bool = False, dead_records: bool = False) -> dict:
        """Run ``n_iter`` NS iterations; returns per-iteration outputs and
        updates ``log_evidence`` (float64 closed-form weights).

        The dead-point outputs (``worst``, ``clone``, ``dead_energy``,
        ``dead_volume`` and the ``dead_pos/types/cell`` records) carry a cull
        axis after the replica axis, in descending energy, when
        ``n_cull > 1``; with ``n_cull == 1`` it is dropped (``(n_iter, R)``)."""
        R, Nm, nm, Cc = self.n_replicas, self.Nmax, self.n_moves, self.run.n_cull
        out = {
            "emax": np.empty((n_iter, R)),
            "worst": np.empty((n_iter, R, Cc), np.int32),
            "clone": np.empty((n_iter, R, Cc), np.int32),
            "dead_energy": np.empty((n_iter, R, Cc)),
            "dead_volume": np.empty((n_iter, R, Cc)),
            "counters": np.empty((n_iter, R, nm, _lib.JR_NCOUNT), np.int32),
            "re_acc": np.empty((n_iter, max(R - 1, 1)), np.int32),
            "re_att": np.empty((n_iter, max(R - 1, 1)), np.int32),
            "walk": np.empty((n_iter, R, self.n_walk), np.int32),
        }
        if dead_records:
            out["dead_pos"] = np.empty((n_iter, R, Cc, Nm, 3))
            out["dead_types"] = np.empty((n_iter, R, Cc, Nm), np.int32)
            out["dead_cell"] = np.empty((n_iter, R, Cc, 3, 3))
        o = _lib.JrIterOut()
        for k in ("emax", "dead_energy", "dead_volume", "dead_pos", "dead_cell"):
            if k in out:
                setattr(o, k, _lib.dptr(out[k]))
        for k in ("worst", "clone", "counters", "re_acc", "re_att", "walk",
                  "dead_types"):
            if k in out:
                setattr(o, k, _lib.iptr(out[k]))
        self._check(self.lib.jr_set_iteration(self.h, self.iteration))
        if graph:
            if tape is not None:
                raise ValueError("graph replay draws on the device")
            self._check(self.lib.jr_ns_iterations_graph(self.h, n_iter, C.byref(o)))
        else:
            if tape is not None:
                ts, keep = self._tape_struct(tape)
                self._check(self.lib.jr_ns_iterations(self.h, n_iter, C.byref(ts),
                                                      C.byref(o)))
                del keep
            else:
                self._check(self.lib.jr_ns_iterations(self.h, n_iter, None,
                                                      C.byref(o)))
        its = self.iteration + np.arange(n_iter)
        w = dead_log_weight(its, self.n_walkers, Cc)  # (n_iter, n_cull)
        for i in range(n_iter):
            self.log_evidence = logaddexp_accumulate(
                self.log_evidence, w[i][None, :] - out["dead_energy"][i])
        if Cc == 1:
            for k in ("worst", "clone", "dead_energy", "dead_volume", "dead_pos",
                      "dead_types", "dead_cell"):
                if k in out:
                    out[k] = out[k][:, :, 0]
        self.iteration += n_iter
        if R == 1:
            out["re_acc"][:] = 0
            out["re_att"][:] = 0
        return out

    def ns_step(self, tape: Tape | None = None) -> list[dict]:
        """One iteration; one reference-shaped ``info`` dict per replica
        (keys of ``sampling/nested_sampling.py:653-685``)."""
        out = self.ns_iterations(1, tape=tape, dead_records=True)
        Cc, Nm = self.run.n_cull, self.Nmax
        infos = []
        for r in range(self.n_replicas):
            c = out["counters"][0, r]
            n_acc = c[:, 1]
            n_prop = c[:, 0]
            worst = np.reshape(out["worst"][0, r], (Cc,))
            clone = np.reshape(out["clone"][0, r], (Cc,))
            dead = {"positions": np.reshape(out["dead_pos"][0, r], (Cc, Nm, 3)),
                    "types": np.reshape(out["dead_types"][0, r], (Cc, Nm)),
                    "energy": np.reshape(out["dead_energy"][0, r], (Cc,)),
                    "cell": np.reshape(out["dead_cell"][0, r], (Cc, 3, 3))}
            infos.append({
                "emax": out["emax"][0, r],
                "hmax": out["emax"][0, r],
                "worst_idx": int(worst[-1]),
                "clone_idx": int(clone[-1]),
                "worst_indices": worst,
                "clone_indices": clone,
                "acceptance_rate": n_acc.sum() / max(self.n_walk * self.run.n_mcmc_steps, 1),
                "dead_walker": {k: v[-1] for k, v in dead.items()},
                "dead_walkers": dead,
                "n_accepted_per_move": n_acc,
                "n_proposed_per_move": n_prop,
                "reject_reason_counts_per_move": c[:, 1:5],
                "n_evaluations_per_move": c[:, 5],
                "n_grad_evaluations_per_move": c[:, 6],
                "inter_re_acc": out["re_acc"][0],
                "inter_re_att": out["re_att"][0],
            })
        return infos

    def final_log_evidence(self) -> np.ndarray:
        pop = self.get_population()
        return final_log_evidence(self.log_evidence, self.iteration, pop["energy"],
                                  self.n_walkers, self.run.n_cull)

    def emax(self) -> np.ndarray:
        e = np.empty(self.n_replicas)
        self._check(self.lib.jr_get_emax(self.h, _lib.dptr(e)))
        return e

    def set_emax(self, emax):
        """Set the resident contour (one value per replica)."""
        e = _lib.f64(np.broadcast_to(np.asarray(emax, dtype=np.float64), (self.n_replicas,)))
        self._check(self.lib.jr_set_emax(self.h, _lib.dptr(e)))

    def replica_exchange(self, re_u=None) -> dict:
        R = self.n_replicas
        acc = np.zeros(max(R - 1, 1), np.int32)
        att = np.zeros(max(R - 1, 1), np.int32)
        ru = None if re_u is None else _lib.f64(re_u)
        self._check(self.lib.jr_replica_exchange(self.h, _lib.dptr(ru),
                                                 _lib.iptr(acc), _lib.iptr(att)))
        return {"n_accepted": int(acc.sum()), "n_attempted": int(att.sum()),
                "n_energy_evals": 0, "n_accepted_per_pair": acc[: R - 1],
                "n_attempted_per_pair": att[: R - 1]}

    def trial_moves(self, move_idx: int, trial_step, sample_idx, emax=None,
                    normals=None, uniforms=None):
        R = self.n_replicas
        idx = _lib.i32(sample_idx)
        n = idx.shape[1]
        acc = np.empty((R, n), np.int32)
        rs = np.empty((R, n), np.int32)
        ts = _lib.f64(np.broadcast_to(np.asarray(trial_step, dtype=np.float64), (R,)))
        em = None if emax is None else _lib.f64(emax)
        nn = None if normals is None else _lib.f64(normals)
        uu = None if uniforms is None else _lib.f64(uniforms)
        self._check(self.lib.jr_trial_moves(
            self.h, int(move_idx), _lib.dptr(ts), _lib.dptr(em), n, _lib.iptr(idx),
            _lib.dptr(nn), _lib.dptr(uu), _lib.iptr(acc), _lib.iptr(rs)))
        return acc.astype(bool), rs

    def initial_walk(self, walklength: int, emax, schedule=None, normals=None,
                     uniforms=None, max_chunk: int = 0) -> np.ndarray:
        """One fixed-contour walk of every live walker (``init/burn_in.py:48-148``):
        ``walklength`` sub-steps at ``emax`` [R] held fixed.  Draws from the
        device, or the host arrays ``schedule`` [wl], ``normals``
        [R, K, wl, 3, Nmax], ``uniforms`` [R, K, wl, 4].  ``max_chunk`` > 0
        caps the walkers per replica of one launch.  Returns the walk's
        counters [R, n_moves, 7]."""
        R = self.n_replicas
        em = _lib.f64(np.broadcast_to(np.asarray(emax, dtype=np.float64), (R,)))
        ctr = np.zeros((R, self.n_moves, _lib.JR_NCOUNT), np.int32)
        sc = None if schedule is None else _lib.i32(schedule)
        nn = None if normals is None else _lib.f64(normals)
        uu = None if uniforms is None else _lib.f64(uniforms)
        self._check(self.lib.jr_initial_walk(
            self.h, int(walklength), _lib.dptr(em), _lib.iptr(sc), _lib.dptr(nn),
            _lib.dptr(uu), int(max_chunk), _lib.iptr(ctr)))
        return ctr

    def autotune(self, reps: int = 2) -> dict:
        """Pick the walk launch shape by timing candidates on this workload."""
        w, o, ms = C.c_int32(), C.c_int32(), C.c_double()
        self._check(self.lib.jr_autotune(self.h, int(reps), C.byref(w), C.byref(o),
                                         C.byref(ms)))
        return {"waves_per_block": w.value, "occupancy": o.value,
                "walk_ms": ms.value}

    def set_launch(self, waves_per_block: int, occupancy: int = 0):
        self._check(self.lib.jr_set_launch(self.h, int(waves_per_block), int(occupancy)))

    def set_walk_block_cap(self, max_blocks: int):
        """Cap the persistent walk grid at ``max_blocks`` blocks (0: none), so
        that each wave walks many chains in turn; results do not change."""
        self._check(self.lib.jr_set_walk_block_cap(self.h, int(max_blocks)))

    def walk_timing(self, reset: bool = False) -> tuple[float, int]:
        ms = C.c_double()
        n = C.c_int64()
        self._check(self.lib.jr_walk_timing(self.h, C.byref(ms), C.byref(n), int(reset)))
        return ms.value, n.value

    def launch_info(self) -> dict:
        a, b, c = C.c_int32(), C.c_int32(), C.c_int32()
        self._check(self.lib.jr_get_info(self.h, C.byref(a), C.byref(b), C.byref(c)))
        return {"n_cu": a.value, "blocks": b.value, "waves_per_block": c.value}

    # --------------------------------------------------- learned model
    def learned_stats(self, reset: bool = False) -> dict:
        """Counters of the learned-model walk since the last reset."""
        st = _lib.JrLearnedStats()
        self._check(self.lib.jr_learned_get_stats(self.h, C.byref(st), int(reset)))
        return st.as_dict()

    def surrogate_params(self) -> np.ndarray:
        from .surrogate import N_PARAMS
        p = np.empty(N_PARAMS, np.float32)
        self._check(self.lib.jr_learned_get_params(self.h, p.ctypes.data_as(_lib._fp)))
        return p

    def set_surrogate_params(self, params, reset_adam: bool = True):
        p = np.ascontiguousarray(params, dtype=np.float32)
        self._check(self.lib.jr_learned_set_params(self.h, p.ctypes.data_as(_lib._fp),
                                                   int(reset_adam)))

    def set_learned_mode(self, mode: str):
        from .surrogate import _MODES
        self._check(self.lib.jr_learned_set_mode(self.h, _MODES[mode]))

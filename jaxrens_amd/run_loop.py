"""Host outer loop of nested sampling over the device engine.

Per iteration, in the reference's order (``sampling/run_loop.py:486-625``):
adaptation when ``i % adjust_interval == 0`` (including ``i = 0`` with
``emax = +inf``), the NS step (select, clone, walk; replica exchange inside
the same device iteration when ``i > 0 and i % re_interval == 0``),
``.energies`` rows (``io/energy_log.py:57-67``), callbacks, then termination
on ``min(log Z)`` / ``max(emax)`` over replicas
(``batch_descriptor.py:415-426``) with ``check_any``.
"""

from __future__ import annotations

import math
from dataclasses import dataclass, field
from pathlib import Path

import numpy as np

from .adaptation import AdaptationConfig, adjust_step_size
from .moves import STEPLESS
from .termination import PriorMassTermination, check_any


class EnergyLogger:
    """``.energies`` file: header ``K n_cull n_dof 0.0 N``, rows
    ``iter emax volume`` (``io/energy_log.py:24-104``).  A resume at
    ``restart_iteration`` > 0 drops the rows at or past it and appends
    (``io/restart_truncate.py:28-50``)."""

    def __init__(self, path, n_walkers: int, n_cull: int = 1, n_dof: int = 0,
                 n_atoms: int = 0, restart_iteration: int = 0):
        self.path = Path(path)
        self.path.parent.mkdir(parents=True, exist_ok=True)
        if restart_iteration > 0 and self.path.exists() and self.path.stat().st_size > 0:
            with open(self.path) as f:
                lines = f.readlines()
            kept = [r for r in lines[1:]
                    if r.strip() and int(r.split()[0]) < restart_iteration]
            with open(self.path, "w") as f:
                f.write(lines[0])
                f.writelines(kept)
            self._f = open(self.path, "a")
        else:
            self._f = open(self.path, "w")
            self._f.write(f"{n_walkers} {n_cull} {n_dof} 0.0 {n_atoms}\n")

    def write_entry(self, iteration: int, energy: float, volume: float = 0.0):
        self._f.write(f"{iteration} {energy:.10e} {volume:.10e}\n")

    def flush(self):
        self._f.flush()

    def close(self):
        if self._f is not None:
            self._f.close()
            self._f = None

    @staticmethod
    def read(path):
        with open(path) as f:
            hdr = f.readline().split()
            rows = np.loadtxt(f, ndmin=2)
        return {"n_walkers": int(hdr[0]), "n_cull": int(hdr[1]),
                "n_dof": int(hdr[2]), "n_atoms": int(hdr[4]) if len(hdr) > 4 else 0,
                "iterations": rows[:, 0].astype(int), "energies": rows[:, 1],
                "volumes": rows[:, 2]}


@dataclass
class RunResult:
    iterations: int
    reason: str
    log_evidence: np.ndarray          # (R,) including the live walkers' share
    emax: np.ndarray                  # (iterations, R)
    dead_volume: np.ndarray           # (iterations, R[, n_cull])
    step_sizes: np.ndarray            # (R, n_moves) at the end
    counters: np.ndarray              # (R, n_moves, 7) summed over iterations
    re_acc: np.ndarray
    re_att: np.ndarray
    adaptation: list = field(default_factory=list)


def run_ns(engine, termination, adaptation: AdaptationConfig | None = None,
           energies_prefix: str | None = None, callbacks=(), seed: int = 0,
           block: int = 1, checkpoint_path=None, checkpoint_interval: int = 0,
           resume=None, traj_prefix: str | None = None, traj_interval: int = 1,
           symbol_map: dict[int, str] | None = None, traj_wrap: bool = True) -> RunResult:
    """Drive ``engine`` until a termination criterion fires.

    ``block`` > 1 runs that many iterations per device call when no
    adaptation falls inside the block and the criteria can only fire at the
    block's last iteration; with the default 1 every iteration syncs, which
    reproduces the reference's stopping iteration exactly.

    ``checkpoint_path`` / ``checkpoint_interval``: write the run state each
    time the count of finished iterations reaches a multiple of the interval
    (``cli/monitor.py:1225-1260``, ``CheckpointCallback``).  ``resume``: a
    checkpoint (path or loaded
    dict) to continue from; iterations keep their global numbers, and the
    ``.energies`` files are cut back to the checkpoint and appended to.

    With ``n_cull > 1`` every iteration writes its n_cull dead points, in
    descending energy, as n_cull ``.energies`` rows and frames (the frames
    tagged ``cull_rank``).

    ``traj_prefix``: write the dead walker of every iteration ``i`` with
    ``i % traj_interval == 0`` to ``<prefix>[.r].traj.extxyz`` per replica
    (``cli/monitor.py:1405-1515``, ``io/trajectory.py:84-139``), species
    named by ``symbol_map`` (default: type code t as ``"X<t>"``), positions
    wrapped into the cell unless ``traj_wrap`` is False (``output.wrap_atoms``,
    ``cli/schema/output.py:220-237``).
    """
    from .trajectory import ExtxyzTrajectoryWriter
    from .checkpoint import load_checkpoint, restore, save_checkpoint
    rng = np.random.default_rng(seed)
    if resume is not None:
        state = resume if isinstance(resume, dict) else load_checkpoint(resume)
        host = restore(engine, state)
        if host is not None:
            rng = host
    R = engine.n_replicas
    loggers = []
    if energies_prefix is not None:
        for r in range(R):
            suffix = "" if R == 1 else f".{r}"
            loggers.append(EnergyLogger(f"{energies_prefix}{suffix}.energies",
                                        engine.n_walkers, engine.run.n_cull, 0,
                                        int(engine.n_atoms_rep[r]),
                                        restart_iteration=engine.iteration))
    trajs = []
    if traj_prefix is not None:
        if traj_interval < 1:
            raise ValueError("traj_interval must be >= 1")
        smap = symbol_map if symbol_map is not None else {t: f"X{t}" for t in range(engine.S)}
        for r in range(R):
            suffix = "" if R == 1 else f".{r}"
            trajs.append(ExtxyzTrajectoryWriter(f"{traj_prefix}{suffix}.traj.extxyz", smap,
                                                wrap=traj_wrap,
                                                restart_iteration=engine.iteration))
    adaptable = [m for m, spec in enumerate(engine.mwg.moves)
                 if spec.type not in STEPLESS]
    emax_hist, vol_hist, adapt_log = [], [], []
    counters = np.zeros((R, engine.n_moves, 7), np.int64)
    re_acc = np.zeros(max(R - 1, 1), np.int64)
    re_att = np.zeros(max(R - 1, 1), np.int64)
    i = engine.iteration
    reason = ""
    try:
        while True:
            if adaptation is not None and i % adaptation.adjust_interval == 0:
                emax = engine.emax()
                ss = engine.step_sizes.copy()
                rec = {"iteration": i}
                for m in adaptable:
                    new, info = adjust_step_size(engine, m, ss[:, m], emax, rng,
                                                 adaptation)
                    ss[:, m] = new
                    rec[engine.mwg.names[m]] = info
                engine.set_run_params(step_sizes=ss)
                adapt_log.append(rec)
            n = 1
            if block > 1:
                n = block
                if adaptation is not None:
                    n = min(n, adaptation.adjust_interval
                            - (i % adaptation.adjust_interval))
                n = max(1, min(n, _iterations_before_check(termination, i)))
            out = engine.ns_iterations(n, dead_records=bool(trajs))
            R, Cc, Nm = engine.n_replicas, engine.run.n_cull, engine.Nmax
            for j in range(n):
                it = i + j
                emax_hist.append(out["emax"][j])
                vol_hist.append(out["dead_volume"][j])
                # one row (and frame) per dead point in descending energy; a
                # single cull records the contour (cli/monitor.py:1331-1378)
                de = out["emax"][j][:, None] if Cc == 1 else out["dead_energy"][j]
                dv = np.reshape(out["dead_volume"][j], (R, Cc))
                for r, lg in enumerate(loggers):
                    for c in range(Cc):
                        lg.write_entry(it, float(de[r, c]), float(dv[r, c]))
                if trajs and it % traj_interval == 0:
                    dp = np.reshape(out["dead_pos"][j], (R, Cc, Nm, 3))
                    dt = np.reshape(out["dead_types"][j], (R, Cc, Nm))
                    dc = np.reshape(out["dead_cell"][j], (R, Cc, 3, 3))
                    for r, tw in enumerate(trajs):
                        na = int(engine.n_atoms_rep[r])
                        for c in range(Cc):
                            tw.write_dead_point(it, dp[r, c, :na], dt[r, c, :na], dc[r, c],
                                                float(de[r, c]),
                                                cull_rank=c if Cc > 1 else None)
            counters += out["counters"].sum(0)
            re_acc += out["re_acc"].sum(0)
            re_att += out["re_att"].sum(0)
            for cb in callbacks:
                cb(i + n - 1, engine, out)
            i += n
            if (checkpoint_path is not None and checkpoint_interval > 0
                    and i // checkpoint_interval > (i - n) // checkpoint_interval):
                for lg in loggers + trajs:
                    lg.flush()
                save_checkpoint(checkpoint_path, engine, rng)
            log_z = float(np.min(engine.log_evidence))
            hmax = float(np.max(out["emax"][-1]))
            for c in termination:
                if isinstance(c, PriorMassTermination):
                    c.update_evidence(log_z)
            stop, reason = check_any(termination, i - 1, hmax)
            if stop:
                break
    finally:
        for lg in loggers + trajs:
            lg.close()
    return RunResult(
        iterations=i, reason=reason, log_evidence=engine.final_log_evidence(),
        emax=np.array(emax_hist), dead_volume=np.array(vol_hist),
        step_sizes=engine.step_sizes.copy(), counters=counters,
        re_acc=re_acc[: R - 1], re_att=re_att[: R - 1], adaptation=adapt_log)


def _iterations_before_check(termination, i: int) -> int:
    """Largest block that cannot skip a stopping iteration: iteration limits
    are exact; other criteria depend on the running evidence, so blocks are
    limited to one iteration for them."""
    from .termination import IterationTermination
    n = math.inf
    for c in termination:
        if isinstance(c, IterationTermination):
            n = min(n, c.max_iterations - i)
        else:
            return 1
    return int(n) if n != math.inf else 1

"""CPU restatement of one NS iteration of the learned-model walk (test
infrastructure only; the product path never imports this module).

Follows the reference for the parts it has:
  select / clone      sampling/nested_sampling.py:86-133,151-186,486-518
  galilean (GMC)      sampling/moves/galilean.py:65-95,142-300
  HMC                 sampling/moves/hmc.py:51-152
  volume              sampling/moves/volume.py:49-109, moves/_common.py:16-93,
                      utils/cell.py:105-125
  per-move counters   sampling/nested_sampling.py:533-609
and restates the surrogate screening, which has no reference counterpart:
with the corrected surrogate enthalpy H_s = U_s - (ca N + cv V) + P V, a
GMC trajectory is steered by the surrogate (reflect where
H_s >= emax + N bias_r); stage 1 keeps a proposal iff
[H_s(x) < thr] == [H_s(y) < thr], thr = emax + N (bias_r + k sig_r);
stage 2 accepts iff the judge's H(y) < emax.  Volume proposals pass the
cell and prior checks first.

Energies come from the FP64 judge oracle and the BF16-emulating surrogate
oracle (oracle/learned.py); draws come from a host tape in the jr_tape
layouts.
"""

from __future__ import annotations

import numpy as np

from oracle import learned as OL
from oracle import oracle as OR

GMC, VOLUME, HMC = 0, 1, 2


def select(E, sel_u, clone_u, extra_u, n_extra):
    """Worst (max energy, ties to the largest noise then the first index),
    clone = floor(u (K-1)) shifted past the worst, extras = the n_extra
    largest noises (stable, first index on ties) excluding both."""
    K = len(E)
    mx = np.max(E)
    cand = [k for k in range(K) if E[k] == mx]
    w = min(cand, key=lambda k: (-sel_u[k], k))
    cl = min(int(clone_u * (K - 1)), K - 2)
    if cl >= w:
        cl += 1
    keys = [(-np.inf if k in (w, cl) else extra_u[k], k) for k in range(K)]
    order = sorted(range(K), key=lambda k: (-keys[k][0], k))
    return w, cl, [w] + order[:n_extra], mx


def _ham(U, cell, P):
    h = U + P * abs(np.linalg.det(cell))
    return np.inf if np.isnan(h) else h


def _det(cell):
    return abs(OR.det3(cell))


def _ham_s(Us, cell, P, N, corr):
    """Corrected surrogate enthalpy U_s - (ca N + cv V) + P V."""
    V = abs(np.linalg.det(cell))
    h = Us + P * V - (corr[0] * N + corr[1] * V)
    return np.inf if np.isnan(h) else h


class Models:
    """The judge (message-passing parameters, or an oracle.oracle.Pot pair
    table) and the surrogate (parameters, species, pair baseline)."""

    def __init__(self, judge_params, sur_params, S=2, base=None):
        self.jp = judge_params
        self.sp = sur_params
        self.S = S
        self.base = base

    def judge(self, pos, ty, cell, forces):
        if isinstance(self.jp, OR.Pot):
            u, _, f = OR.energy_forces(self.jp, pos, ty, cell, forces=forces)
            return u, f
        e, f, _ = OL.mpnn_energy_forces(self.jp, pos, ty, cell, forces=forces)
        return e, f

    def sur(self, pos, ty, cell, forces):
        return OL.sur_energy_forces(self.sp, pos, ty, cell, S=self.S, bf16=True, forces=forces,
                                    base=self.base)


def gmc(ch, mv, step, xi, models, emax, P, thr, da, stats, records, use_sur, corr=(0.0, 0.0)):
    """One GMC move of chain ``ch`` (dict with pos, dir, cell, types, H, Hs)."""
    N = ch["n"]
    pos, d = ch["pos"][:N].copy(), ch["dir"][:N].copy()
    ty, cell = ch["types"][:N], ch["cell"]
    z = xi[:, :N].T
    fresh = np.sqrt(np.sum(d * d)) < 1e-8
    v = z.copy() if fresh else d + mv["perturb_angle"] * z
    d = v / max(np.sqrt(np.sum(v * v)), 1e-10)
    x0, d0 = pos.copy(), d.copy()
    nr = mv["n_reflect"]
    Hs_y = None
    for k in range(nr):
        pos = pos + step * d
        if da:
            Us, F = models.sur(pos, ty, cell, True)
            H = _ham_s(Us, cell, P, N, corr)
            Hs_y, Us_y = H, Us
            t = thr[0]
        else:
            U, F = models.judge(pos, ty, cell, True)
            H = _ham(U, cell, P)
            t = emax
        if k < nr - 1 and H >= t:
            fn = max(np.sqrt(np.sum(F * F)), 1e-10)
            d = d - 2.0 * (np.sum(F * d) / fn) * (F / fn)
    if da:
        stats["sur_e"] += nr
        code = 1 if (ch["Hs"] < thr[1]) == (Hs_y < thr[1]) else 0
        acc = False
        if code == 1:
            U = models.judge(pos, ty, cell, False)[0]
            Hy = _ham(U, cell, P)
            acc = Hy < emax
            stats["exact_e"] += 1
            stats["stage1_pass"] += 1
            records.append((U, Us_y, N, _det(cell)))
        n_eval, n_grad = code, 0
    else:
        Hy = H
        acc = Hy < emax
        code = 1
        stats["exact_fe"] += nr
        if use_sur:
            Us_y = models.sur(pos, ty, cell, False)[0]
            stats["sur_e"] += 1
            records.append((U, Us_y, N, _det(cell)))
        n_eval, n_grad = nr, nr
    if acc:
        ch["pos"][:N], ch["dir"][:N] = pos, d
        ch["H"] = Hy
        if da:
            ch["Hs"] = Hs_y
    else:
        ch["pos"][:N], ch["dir"][:N] = x0, -d0
    reason = 0 if acc else 1
    return reason, n_eval, n_grad


def _exp(x: float) -> float:
    """exp of the C oracle (IEEE operations only, as the device's)."""
    return float(OR.lib().or_exp(float(x)))


def hmc(ch, mv, step, xi, models, emax, P, thr, da, stats, records, use_sur, corr=(0.0, 0.0),
        u0=0.0):
    """One HMC move (sampling/moves/hmc.py:51-152): p = xi; n_leapfrog x
    {p += s/2 F; x += s p; F = F(x); p += s/2 F}.  Plain: accept iff
    H' < emax and (H' + K') - (H + K) < 1 (the reference's rule).
    That rule accepts almost every move at a small step and the exact flow
    keeps exp(-H - K), so it leans toward exp(-H) inside the contour.
    Screened: the surrogate's forces drive the leapfrog, its verdicts
    screen, and the move is the exact Metropolis step of the constrained
    prior uniform(x) N(p) under the leapfrog-and-flip involution: accept iff
    H' < emax and u0 < exp(K - K') (exp as the C oracle's or_exp)."""
    N = ch["n"]
    x0 = ch["pos"][:N].copy()
    ty, cell = ch["types"][:N], ch["cell"]
    p = xi[:, :N].T.copy()
    K0 = 0.5 * np.sum(p * p)
    h = 0.5 * step
    force = (lambda x: models.sur(x, ty, cell, True)) if da else \
        (lambda x: models.judge(x, ty, cell, True))
    x = x0.copy()
    E, F = force(x)
    for _ in range(mv["n_leapfrog"]):
        p = p + h * F
        x = x + step * p
        E, F = force(x)
        p = p + h * F
    K1 = 0.5 * np.sum(p * p)
    nl = mv["n_leapfrog"]
    code = 1
    if da:
        stats["sur_e"] += nl + 1
        Hs_y = _ham_s(E, cell, P, N, corr)
        code = 1 if (ch["Hs"] < thr[1]) == (Hs_y < thr[1]) else 0
        acc = False
        if code == 1:
            U = models.judge(x, ty, cell, False)[0]
            Hy = _ham(U, cell, P)
            acc = bool(Hy < emax) and bool(u0 < _exp(K0 - K1))
            stats["exact_e"] += 1
            stats["stage1_pass"] += 1
            records.append((U, E, N, _det(cell)))
        n_eval, n_grad = code, 0
    else:
        U = E
        Hy = _ham(U, cell, P)
        acc = (Hy < emax) and ((Hy + K1) - (ch["H"] + K0) < 1.0)
        stats["exact_fe"] += nl + 1
        if use_sur:
            records.append((U, models.sur(x, ty, cell, False)[0], N, _det(cell)))
            stats["sur_e"] += 1
        n_eval, n_grad = 2 * nl + 1, 2 * nl
    if acc:
        ch["pos"][:N] = x
        ch["H"] = Hy
        if da:
            ch["Hs"] = Hs_y
    return (0 if acc else 1), n_eval, n_grad


def volume(ch, mv, step, xi0, u0, models, emax, P, thr, da, stats, records, use_sur,
           corr=(0.0, 0.0)):
    N = ch["n"]
    pos, cell, ty = ch["pos"][:N], ch["cell"], ch["types"][:N]
    oldV = _det(cell)
    newV = abs(oldV + (step * N) * xi0)
    ratio = newV / oldV
    scale = OR.cbrt(ratio)
    nc = cell * scale
    npos = pos * scale
    cell_ok = OR.check_cell_shape(nc, N, mv["max_vpa"], mv["min_vpa"], mv["min_aspect"])
    p_acc = 1.0 if mv["flat_v_prior"] else min(1.0, ratio ** N)
    code = 2 if not cell_ok else (1 if u0 < p_acc else 3)
    Us = None
    if code == 1 and (da or use_sur):
        Us = models.sur(npos, ty, nc, False)[0]
        stats["sur_e"] += 1
    if da and code == 1:
        Hs_y = _ham_s(Us, nc, P, N, corr)
        code = 1 if (ch["Hs"] < thr[1]) == (Hs_y < thr[1]) else 0
        if code == 1:
            stats["stage1_pass"] += 1
    acc = False
    if code == 1:
        U = models.judge(npos, ty, nc, False)[0]
        Hy = _ham(U, nc, P)
        acc = Hy < emax
        stats["exact_e"] += 1
        if use_sur:
            records.append((U, Us, N, _det(nc)))
    if acc:
        ch["pos"][:N] = npos
        ch["cell"] = nc
        ch["H"] = Hy
        if da:
            ch["Hs"] = Hs_y
    reason = 0 if acc else (code if code in (2, 3) else 1)
    return reason, (1 if code == 1 else 0), 0


def ns_iteration(pop, models, pressure, step_sizes, moves, tape, it, n_extra, mode,
                 train_max_epa=np.inf, corr=(0.0, 0.0), bias=None, sig=None, k_slack=2.0):
    """One iteration (select -> clone -> walk) on ``pop`` (an
    oracle.oracle.Population, modified in place).  ``mode``: "plain" (judge
    steers; no surrogate), "plain+sur" (judge steers, the surrogate predicts
    the judged end points) or "screened" (surrogate steers, two-stage test).
    ``corr`` = (ca, cv), ``bias``/``sig`` [R] and ``k_slack`` form the surrogate's
    error model (zero: no held-out record yet).  Returns the per-iteration
    outputs, statistics and, in evaluation order, the (U, U_s, N, V) records
    of the exactly evaluated configurations that the training ring stores."""
    R, K, Nm = pop.pos.shape[:3]
    bias = np.zeros(R) if bias is None else np.asarray(bias, dtype=float)
    sig = np.zeros(R) if sig is None else np.asarray(sig, dtype=float)
    n_steps = tape.schedule.shape[1]
    n_moves = len(moves)
    da = mode == "screened"
    use_sur = mode != "plain"
    out = {"emax": np.zeros(R), "worst": np.zeros(R, int), "clone": np.zeros(R, int),
           "walk": np.zeros((R, 1 + n_extra), int),
           "counters": np.zeros((R, n_moves, 7), np.int64)}
    stats = dict(proposals=0, stage1_pass=0, accepted=0, exact_fe=0, exact_e=0, sur_e=0)
    chains = []
    for r in range(R):
        w, cl, walk, mx = select(pop.energy[r], tape.sel_u[it, r], tape.clone_u[it, r],
                                 tape.extra_u[it, r], n_extra)
        out["emax"][r], out["worst"][r], out["clone"][r] = mx, w, cl
        out["walk"][r] = walk
        for a in (pop.pos, pop.dir, pop.types, pop.cell, pop.energy, pop.n_atoms):
            a[r, w] = a[r, cl]
        for j, k in enumerate(walk):
            chains.append((r, j, k))
    state = []
    for (r, j, k) in chains:
        ch = {"pos": pop.pos[r, k].copy(), "dir": pop.dir[r, k].copy(),
              "types": pop.types[r, k], "cell": pop.cell[r, k].reshape(3, 3).copy(),
              "H": pop.energy[r, k], "n": int(pop.n_atoms[r, k]), "Hs": None}
        if da:
            N = ch["n"]
            Us = models.sur(ch["pos"][:N], ch["types"][:N], ch["cell"], False)[0]
            ch["Hs"] = _ham_s(Us, ch["cell"], pressure[r], N, corr)
            stats["sur_e"] += 1
        state.append(ch)
    records = []
    for t in range(n_steps):
        m = int(tape.schedule[it, t])
        mv = moves[m]
        step_records = []
        for c, (r, j, k) in enumerate(chains):
            ch = state[c]
            emax = out["emax"][r]
            thr = (emax + ch["n"] * bias[r], emax + ch["n"] * (bias[r] + k_slack * sig[r]))
            step = step_sizes[r, m]
            xi = tape.normals[it, r, j, t]
            if mv["type"] == GMC:
                rs, ne, ng = gmc(ch, mv, step, xi, models, emax, pressure[r], thr, da,
                                 stats, step_records, use_sur, corr)
            elif mv["type"] == HMC:
                rs, ne, ng = hmc(ch, mv, step, xi, models, emax, pressure[r], thr, da,
                                 stats, step_records, use_sur, corr,
                                 u0=tape.uniforms[it, r, j, t, 0])
            else:
                rs, ne, ng = volume(ch, mv, step, xi[0, 0], tape.uniforms[it, r, j, t, 0],
                                    models, emax, pressure[r], thr, da, stats,
                                    step_records, use_sur, corr)
            cnt = out["counters"][r, m]
            cnt[0] += 1
            cnt[1 + rs] += 1
            cnt[5] += ne
            cnt[6] += ng
            stats["proposals"] += 1
            stats["accepted"] += rs == 0
        records += [x for x in step_records
                    if np.isfinite(x[0]) and x[0] <= train_max_epa * x[2]]
    for c, (r, j, k) in enumerate(chains):
        ch = state[c]
        pop.pos[r, k], pop.dir[r, k] = ch["pos"], ch["dir"]
        pop.cell[r, k] = ch["cell"].reshape(-1)
        pop.energy[r, k] = ch["H"]
    return out, stats, records


def trial_moves(pop, models, pressure, moves, move_idx, trial_step, emax, sample_idx,
                normals, uniforms, mode, corr=(0.0, 0.0), bias=None, sig=None, k_slack=2.0):
    """Trial moves of the learned walk (the step-size adaptation's
    per-move function, sampling/adaptation/stepsize_handler.py:215-261):
    move ``move_idx`` once from each sampled walker (sample_idx [R, n]) at
    trial_step [R] under the contours emax [R], with the draws normals
    [R, n, 3, Nmax] and uniforms [R, n, 4], in ``mode`` as ns_iteration
    takes it; ``pop`` is not changed.  Returns the reasons [R, n]
    (0 = accepted)."""
    R = pop.pos.shape[0]
    bias = np.zeros(R) if bias is None else np.asarray(bias, dtype=float)
    sig = np.zeros(R) if sig is None else np.asarray(sig, dtype=float)
    da = mode == "screened"
    use_sur = mode != "plain"
    mv = moves[move_idx]
    n = sample_idx.shape[1]
    reason = np.zeros((R, n), int)
    stats = dict(proposals=0, stage1_pass=0, accepted=0, exact_fe=0, exact_e=0, sur_e=0)
    for r in range(R):
        for j in range(n):
            k = int(sample_idx[r, j])
            ch = {"pos": pop.pos[r, k].copy(), "dir": pop.dir[r, k].copy(),
                  "types": pop.types[r, k], "cell": pop.cell[r, k].reshape(3, 3).copy(),
                  "H": pop.energy[r, k], "n": int(pop.n_atoms[r, k]), "Hs": None}
            N = ch["n"]
            if da:
                Us = models.sur(ch["pos"][:N], ch["types"][:N], ch["cell"], False)[0]
                ch["Hs"] = _ham_s(Us, ch["cell"], pressure[r], N, corr)
            e = emax[r]
            thr = (e + N * bias[r], e + N * (bias[r] + k_slack * sig[r]))
            xi = normals[r, j]
            if mv["type"] == GMC:
                rs = gmc(ch, mv, trial_step[r], xi, models, e, pressure[r], thr, da, stats,
                         [], use_sur, corr)[0]
            elif mv["type"] == HMC:
                rs = hmc(ch, mv, trial_step[r], xi, models, e, pressure[r], thr, da, stats,
                         [], use_sur, corr, u0=uniforms[r, j, 0])[0]
            else:
                rs = volume(ch, mv, trial_step[r], xi[0, 0], uniforms[r, j, 0], models, e,
                            pressure[r], thr, da, stats, [], use_sur, corr)[0]
            reason[r, j] = rs
    return reason


def initial_walk(pop, models, pressure, step_sizes, moves, emax, schedule, normals, uniforms,
                 mode, corr=(0.0, 0.0), bias=None, sig=None, k_slack=2.0):
    """Burn-in walk of the learned walk (init/burn_in.py:48-148): every
    walker of replica r takes the sub-steps of ``schedule`` [wl] at the
    fixed contour emax[r] with the draws normals [R, K, wl, 3, Nmax] and
    uniforms [R, K, wl, 4], in ``mode`` as ns_iteration takes it; ``pop`` is
    walked in place.  Returns the counters [R, n_moves, 7]."""
    R, K = pop.pos.shape[:2]
    bias = np.zeros(R) if bias is None else np.asarray(bias, dtype=float)
    sig = np.zeros(R) if sig is None else np.asarray(sig, dtype=float)
    da = mode == "screened"
    use_sur = mode != "plain"
    counters = np.zeros((R, len(moves), 7), np.int64)
    stats = dict(proposals=0, stage1_pass=0, accepted=0, exact_fe=0, exact_e=0, sur_e=0)
    for r in range(R):
        for k in range(K):
            ch = {"pos": pop.pos[r, k].copy(), "dir": pop.dir[r, k].copy(),
                  "types": pop.types[r, k], "cell": pop.cell[r, k].reshape(3, 3).copy(),
                  "H": pop.energy[r, k], "n": int(pop.n_atoms[r, k]), "Hs": None}
            N = ch["n"]
            if da:
                Us = models.sur(ch["pos"][:N], ch["types"][:N], ch["cell"], False)[0]
                ch["Hs"] = _ham_s(Us, ch["cell"], pressure[r], N, corr)
            e = emax[r]
            thr = (e + N * bias[r], e + N * (bias[r] + k_slack * sig[r]))
            for t, m in enumerate(schedule):
                mv = moves[m]
                xi = normals[r, k, t]
                if mv["type"] == GMC:
                    rs, ne, ng = gmc(ch, mv, step_sizes[r, m], xi, models, e, pressure[r], thr,
                                     da, stats, [], use_sur, corr)
                elif mv["type"] == HMC:
                    rs, ne, ng = hmc(ch, mv, step_sizes[r, m], xi, models, e, pressure[r], thr,
                                     da, stats, [], use_sur, corr, u0=uniforms[r, k, t, 0])
                else:
                    rs, ne, ng = volume(ch, mv, step_sizes[r, m], xi[0, 0], uniforms[r, k, t, 0],
                                        models, e, pressure[r], thr, da, stats, [], use_sur,
                                        corr)
                cnt = counters[r, m]
                cnt[0] += 1
                cnt[1 + rs] += 1
                cnt[5] += ne
                cnt[6] += ng
            pop.pos[r, k], pop.dir[r, k] = ch["pos"], ch["dir"]
            pop.cell[r, k] = ch["cell"].reshape(-1)
            pop.energy[r, k] = ch["H"]
    return counters

"""Throughput of the nested-sampling inner walk on one MI355X.

Workload (the headline configuration): LJ-64 NPT replica stack, 64 identical
LJ atoms (eps = sigma = 1, strict unshifted cutoff 2.5, cubic PBC), P = 16
pressures log-spaced over 0.01-10, K = 1024 walkers each, 20 MCMC sub-steps
per iteration of galilean (n_reflect 10, step 0.1) and volume (step 0.3) moves
at weights 4:1, pressure replica exchange every iteration, FP64.  Initial
volumes uniform in [8, 4096] with uniform positions.  The whole population is
walked each iteration (n_extra = K - 2); a "step" is one NS iteration.

Prints one JSON line.  ``value`` = walker-steps/s = sum over replicas of
n_walk * n_mcmc_steps per iteration / wall time, inputs resident in HBM.
"""

from __future__ import annotations

import argparse
import json
import math
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

FP64_PEAK_TFLOPS = 78.6  # MI355X FP64 vector, dense
FP32_PEAK_TFLOPS = 157.3  # MI355X FP32 vector (packed issue)
METRIC = "walker-steps/s, LJ-64 NPT P×K=16×1024; surrogate vs plain at matched log Z"

# algorithmic FP64 FLOPs per unordered pair (cubic MIC): energy+forces, energy
FLOP_EF, FLOP_E = 38, 25

# measured HBM traffic of the walk kernel (rocprofv3 FETCH_SIZE / WRITE_SIZE
# passes scaled by a copy of known size); absent file -> traffic null
TRAFFIC_FILE = os.path.join(ROOT, "profiles", "walk_kernel_lj64_npt_16x1024_traffic.txt")


def measured_traffic() -> float | None:
    """Bytes per walk launch from the committed counter summary, if any."""
    try:
        with open(TRAFFIC_FILE) as f:
            for line in f:
                if line.startswith("traffic_bytes_per_launch:"):
                    return float(line.split(":", 1)[1])
    except OSError:
        return None
    return None


def copy_calibration(n_bytes: int = 1 << 30) -> dict:
    """A plain device copy of known size (float32 multiply by 1: reads and
    writes ``n_bytes`` each) for scaling the FETCH_SIZE / WRITE_SIZE
    counters; run it under the same counter pass as the bench."""
    import torch
    a = torch.ones(n_bytes // 4, dtype=torch.float32, device="cuda")
    b = torch.empty_like(a)
    for _ in range(2):
        torch.mul(a, 1.0, out=b)
    torch.cuda.synchronize()
    return {"copy_bytes_read": n_bytes, "copy_bytes_written": n_bytes,
            "kernel": "torch elementwise mul (float32, x1.0)"}


def build_engine(n_extra: int, R=16, K=1024, N=64, n_steps=20, seed=42,
                 precision="fp64", eval_chunks=1):
    from jaxrens_amd.backends import LJBackend
    from jaxrens_amd.engine import NSEngine, RunConfig
    lj = LJBackend(1.0, 1.0, 2.5, precision=precision, eval_chunks=eval_chunks)
    moves = [{"type": "galilean", "n_reflect": 10, "step_size": 0.1, "weight": 4},
             {"type": "volume", "step_size": 0.3, "weight": 1}]
    cell = {"min_volume_per_atom": 0.125, "max_volume_per_atom": 64.0,
            "flat_V_prior": True}
    P = [10 ** (-2 + 3 * i / (R - 1)) for i in range(R)]
    run = RunConfig(n_replicas=R, n_walkers=K, n_atoms=N, n_mcmc_steps=n_steps,
                    n_extra=n_extra, seed=seed, pressures=P, re_interval=1,
                    n_swap_cycles=1)
    eng = NSEngine(lj, moves, run, cell=cell)
    eng.init_walkers(np.random.default_rng(seed), 8.0 / N, 4096.0 / N, True)
    return eng


def launch_flops(counters: np.ndarray, eng) -> float:
    """Algorithmic FLOPs of the walk launches behind ``counters``
    (iterations, R, n_moves, 7): n_reflect E+F evaluations per galilean
    step, one energy evaluation per volume or morph step, N_r(N_r-1)/2 pairs
    each for replica r."""
    c = np.asarray(counters).reshape(-1, eng.n_replicas, eng.n_moves, 7)
    total = 0.0
    for r in range(eng.n_replicas):
        N = int(eng.n_atoms_rep[r])
        pairs = N * (N - 1) / 2
        for m, spec in enumerate(eng.mwg.moves):
            steps = float(c[:, r, m, 0].sum())
            if spec.type == "gmc":
                total += steps * spec.n_reflect * FLOP_EF * pairs
            elif spec.type == "hmc":
                total += steps * ((spec.n_leapfrog + 1) * FLOP_EF) * pairs
            else:
                total += steps * FLOP_E * pairs
    return float(total)


def iteration_distribution(eng, n_iter: int) -> dict:
    """Mean, p50 and p95 over ``n_iter`` single NS iterations, each its own
    call ending in a host synchronisation (wall clock of the whole
    iteration), and of the walk launch inside it (HIP events).  Run after
    the timed region, so ``value`` does not include the per-call syncs."""
    it_ms, walk_ms = [], []
    eng.walk_timing(reset=True)
    for _ in range(n_iter):
        t0 = time.perf_counter()
        eng.ns_iterations(1)
        it_ms.append((time.perf_counter() - t0) * 1e3)
        ms, n = eng.walk_timing(reset=True)
        walk_ms.append(ms / max(n, 1))

    def stats(v):
        v = np.asarray(v)
        return {"mean": float(v.mean()), "p50": float(np.percentile(v, 50)),
                "p95": float(np.percentile(v, 95))}
    return {"iterations": n_iter, "iteration_ms": stats(it_ms),
            "walk_launch_ms": stats(walk_ms),
            "note": "one ns_iterations(1) call per iteration, measured after the timed region"}


def config_throughput(n_iter_c1: int = 400, n_iter_c3: int = 20, cpu: bool = True) -> dict:
    """Throughput of the other synthetic configurations, whole iterations
    replayed from a captured hipGraph: LJ-8 NPT (P = 1, K = 128, n_extra = 0,
    the reference's defaults) and the binary semi-grand muPT batch (N in
    {32, 48, 64, 96, 128} over P = 20 replicas, K = 512, whole population
    walked, morph standing in for the identity swap, autotuned launch).
    FLOP rates are iteration-level (all kernels of an iteration)."""
    from jaxrens_amd.backends import LJBackend
    from jaxrens_amd.engine import NSEngine, RunConfig
    gv = [{"type": "galilean", "n_reflect": 10, "step_size": 0.1, "weight": 4},
          {"type": "volume", "step_size": 0.3, "weight": 1}]
    out = {}
    # LJ-8 NPT, the reference's CPU-runnable configuration
    run = RunConfig(n_replicas=1, n_walkers=128, n_atoms=8, n_mcmc_steps=20, n_extra=0,
                    seed=42, pressures=[0.1], re_interval=0)
    e1 = NSEngine(LJBackend(1.0, 1.0, 2.5), gv, run,
                  cell={"min_volume_per_atom": 0.125, "max_volume_per_atom": 64.0,
                        "flat_V_prior": True})
    e1.init_walkers(np.random.default_rng(42), 1.0 / 8, 64.0, True)
    e1.ns_iterations(5, graph=True)
    t0 = time.perf_counter()
    o1 = e1.ns_iterations(n_iter_c1, graph=True)
    d1 = time.perf_counter() - t0
    f1 = launch_flops(o1["counters"], e1) / d1 / 1e12
    out["lj8_npt"] = {
        "walker_steps_per_s": 1 * 1 * 20 * n_iter_c1 / d1,
        "ns_iterations_per_s": n_iter_c1 / d1, "iterations": n_iter_c1,
        "iteration_tflops": f1, "frac_of_fp64_peak": f1 / FP64_PEAK_TFLOPS,
        "note": "one chain of 8 atoms per iteration: latency-bound, 8 of 64 lanes busy"}
    if cpu:
        # the same configuration on one host core (its one chain per
        # iteration leaves nothing for more threads)
        out["lj8_npt"]["cpu_baseline"] = cpu_baseline(e1, 1, 1, n_iter=200)
    e1.close()
    # binary semi-grand muPT, uneven sizes in one persistent launch
    R, K = 20, 512
    sizes = [n for n in (32, 48, 64, 96, 128) for _ in range(4)]
    eps = np.array([[1.0, 1.5], [1.5, 0.5]])
    sig = np.array([[1.0, 0.8], [0.8, 0.88]])
    mu = np.stack([np.zeros(R), np.linspace(-2.0, 2.0, R)], axis=1)
    run = RunConfig(n_replicas=R, n_walkers=K, n_atoms=sizes, n_mcmc_steps=20,
                    n_extra=K - 2, seed=42, pressures=[1.0] * R, chemical_potentials=mu,
                    re_interval=1)
    moves = gv + [{"type": "alchemical_morph", "n_species": 2, "weight": 1}]
    e3 = NSEngine(LJBackend(pair_epsilon=eps, pair_sigma=sig, cutoff=2.5), moves, run,
                  cell={"min_volume_per_atom": 0.5, "max_volume_per_atom": 8.0,
                        "flat_V_prior": True})
    e3.init_walkers(np.random.default_rng(42), 1.0, 8.0, True, species_p=[0.5, 0.5])
    tune = e3.autotune(reps=1)
    e3.ns_iterations(2, graph=True)
    t0 = time.perf_counter()
    o3 = e3.ns_iterations(n_iter_c3, graph=True)
    d3 = time.perf_counter() - t0
    f3 = launch_flops(o3["counters"], e3) / d3 / 1e12
    out["binary_semigrand_mixed_n"] = {
        "walker_steps_per_s": R * (K - 1) * 20 * n_iter_c3 / d3,
        "ns_iterations_per_s": n_iter_c3 / d3, "iterations": n_iter_c3,
        "iteration_tflops": f3, "frac_of_fp64_peak": f3 / FP64_PEAK_TFLOPS,
        "launch": tune}
    e3.close()
    return out


def cpu_baseline(eng, n_replicas_sample: int, threads: int, n_iter: int = 1) -> dict:
    """The CPU oracle (FP64 C restatement, OpenMP) on ``n_iter`` iterations
    of a replica subset of the same workload, from the current population."""
    from oracle import oracle as O
    from jaxrens_amd.draws import make_tape
    O.build()
    d = eng.get_population()
    Rs = n_replicas_sample
    sub = {k: np.ascontiguousarray(v[:Rs]) for k, v in d.items()}
    pop = O.Population.from_engine_dict(sub)
    pot = O.Pot.from_table(eng.model.table)
    K, N, T = eng.n_walkers, eng.Nmax, eng.run.n_mcmc_steps
    tape = make_tape(np.random.default_rng(7), n_iter, Rs, K, N, T, eng.run.n_extra,
                     eng.mwg.move_probs)
    dt, flops = 0.0, 0.0
    for it in range(n_iter):
        t0 = time.perf_counter()
        out = O.ns_iteration(pop, pot, eng.pressure[:Rs], eng.mu[:Rs],
                             eng.step_sizes[:Rs], eng.mwg.as_oracle_moves(), T,
                             eng.run.n_extra, tape, it=it, do_re=Rs > 1,
                             n_threads=threads)
        dt += time.perf_counter() - t0
        flops += launch_flops(out["counters"][None], eng)
        print(f"[bench] cpu baseline iteration {it + 1}/{n_iter}: {dt:.1f} s",
              file=sys.stderr, flush=True)
    ws = Rs * eng.n_walk * T * n_iter
    return {"value": ws / dt, "unit": "walker-steps/s", "cores": threads,
            "kind": "port",
            "sample": (f"{n_iter} NS iteration(s) of {Rs} of {eng.n_replicas} replicas "
                       f"x {eng.n_walk} walked chains x {T} sub-steps = {ws} "
                       f"walker-steps in {dt:.3g} s; FP64 C oracle, OpenMP "
                       f"{threads} threads"),
            "flops_per_s": flops / dt}


# 256-atom message-passing workload of the surrogate-screened walk
SUR_P = [0.1, 0.5, 1.0, 2.0]
SUR_MOVES = [{"type": "galilean", "n_reflect": 10, "step_size": 0.1, "weight": 4},
             {"type": "volume", "step_size": 0.3, "weight": 1}]
# vol/atom >= 0.6 keeps the cubic edge above the judge's 5 sigma cutoff and
# its neighbour lists (<= 1024 images) in capacity
SUR_CELL = {"min_volume_per_atom": 0.6, "max_volume_per_atom": 64.0, "flat_V_prior": True}


def sur_population(R, K, N, seed=7, rho=0.9):
    """Jittered simple-cubic lattice at density rho, species 50/50 shuffled."""
    rng = np.random.default_rng(seed)
    L = (N / rho) ** (1 / 3)
    n = int(math.ceil(N ** (1 / 3)))
    g = (np.arange(n) + 0.5) / n * L
    lat = np.array([[x, y, z] for x in g for y in g for z in g])[:N]
    pos = lat[None, None] + rng.normal(scale=0.05 * L / n, size=(R, K, N, 3))
    ty = np.zeros((R, K, N), np.int32)
    ty[:] = np.arange(N) % 2
    ty = rng.permuted(ty, axis=2)
    cell = np.broadcast_to(np.eye(3) * L, (R, K, 3, 3)).copy()
    return pos, ty, cell


# The surrogate's own pair baseline: a generic r^-12 wall
# eps ((sig/r)^12 - (sig/rc)^12) below rc, not the judge's core term
# (eps 1, sig 0.85, rc 1.2); the network learns the judge minus it.
SUR_BASE = (1.0, 0.8, 1.0)
# committed record of the matched-log-Z runs (tests/test_gpu_learned_logz.py and
# tests/test_gpu_learned_walk.py)
LOGZ_FILE = os.path.join(ROOT, "profiles", "surrogate_log_z_mpnn256_lj128_lj8.txt")


def logz_record(config: str) -> dict | None:
    """The LOGZ_RESULT line for ``config`` from the committed record, if any."""
    try:
        with open(LOGZ_FILE) as f:
            for line in f:
                if line.startswith("LOGZ_RESULT "):
                    rec = json.loads(line.split(" ", 1)[1])
                    if rec.get("config") == config:
                        return {**rec, "source": os.path.relpath(LOGZ_FILE, ROOT)}
    except OSError:
        return None
    return None


def zero_readout(params: np.ndarray) -> np.ndarray:
    """The network with its per-atom readout (w3, b3) zeroed: the surrogate
    is then the pair baseline alone."""
    from jaxrens_amd.surrogate import N_PARAMS
    p = np.array(params, dtype=np.float32, copy=True)
    p[N_PARAMS - 129:] = 0.0
    return p


def sur_engine(mode, pop, R, K, N, T, params, train_steps=400):
    from jaxrens_amd.backends import EnsembleTerm, Hamiltonian
    from jaxrens_amd.engine import NSEngine, RunConfig
    from jaxrens_amd.mpnn import MPNNBackend, init_mpnn
    from jaxrens_amd.surrogate import LearnedConfig
    jp = init_mpnn(seed=11)
    judge = MPNNBackend(jp)
    run = RunConfig(n_replicas=R, n_walkers=K, n_atoms=N, n_mcmc_steps=T, n_extra=K - 2,
                    seed=42, pressures=SUR_P[:R], re_interval=1)
    lc = None
    if mode != "plain":
        lc = LearnedConfig(mode=mode, ring_capacity=100_000, train_batch=64,
                           train_steps=train_steps, min_train=256, min_heldout=64, lr=1e-3,
                           rmse_bound=0.05, delta_k=2.0, train_max_epa=50.0,
                           baseline=SUR_BASE, params=params)
    eng = NSEngine(Hamiltonian(judge, [EnsembleTerm(SUR_P[0])]), SUR_MOVES, run,
                   cell=SUR_CELL, learned=lc)
    eng.set_population(*pop)
    return eng


def surrogate_bench(n_plain=1, n_timed=2, n_train=3, R=4, K=1024, N=256, T=20) -> dict:
    """The 256-atom message-passing config walked three ways from one
    population: plain (the judge's forces steer every GMC trajectory), the
    two-stage screen with the surrogate reduced to its pair baseline (zeroed
    network, no training), and the screen with the network trained online
    (``n_train`` iterations of training before the timed ones).  Both screens
    use the generic wall ``SUR_BASE``."""
    from jaxrens_amd.surrogate import init_surrogate
    pop = sur_population(R, K, N)
    ws_it = R * (K - 1) * T
    out = {"workload": f"{N} atoms, 2 species 50/50, message-passing judge (rc 5, 8 Bessel, "
                       f"2x64 channels, SiLU, FP32), P={SUR_P[:R]}, K={K}, {T} sub-steps "
                       "gmc(n_reflect=10):volume 4:1, RE every iteration, n_extra=K-2",
           "initial": "jittered simple-cubic lattice at density 0.9",
           "surrogate_baseline": {"form": "eps ((sig/r)^12 - (sig/rc)^12), r < rc",
                                  "eps_sig_rc": list(SUR_BASE),
                                  "note": "a generic wall of its own, not the judge's core "
                                          "term (1, 0.85, 1.2)"}}

    def timed(eng, n):
        eng.learned_stats(reset=True)
        t = time.perf_counter()
        eng.ns_iterations(n)
        dt = time.perf_counter() - t
        st = eng.learned_stats()
        judge = st["exact_force_evals"] + st["exact_energy_evals"]
        return {"walker_steps_per_s": ws_it * n / dt, "s_per_iteration": dt / n,
                "accepted_moves_per_s": st["accepted"] / dt,
                "acceptance": st["accepted"] / max(st["proposals"], 1),
                "stage1_pass": st["stage1_pass"] / max(st["proposals"], 1),
                "accepted_per_judge_eval": st["accepted"] / max(judge, 1),
                "judge_evals_per_walker_step": judge / max(st["proposals"], 1),
                "judge_force_evals": st["exact_force_evals"],
                "judge_energy_evals": st["exact_energy_evals"],
                "surrogate_evals": st["surrogate_evals"],
                "screened_iterations": st["iterations_screened"],
                "heldout_rmse_per_atom": st["heldout_rmse"],
                "train_ms_per_iteration": st["train_ms"] / n, "iterations": n}

    eng = sur_engine("plain", pop, R, K, N, T, None)
    out["plain"] = timed(eng, n_plain)
    eng.close()
    print(f"[bench] surrogate: plain {out['plain']}", file=sys.stderr, flush=True)
    p0 = init_surrogate(1)
    eng = sur_engine("screened", pop, R, K, N, T, zero_readout(p0), train_steps=0)
    eng.ns_iterations(1)
    out["baseline_only"] = timed(eng, n_timed)
    eng.close()
    print(f"[bench] surrogate: baseline only {out['baseline_only']}", file=sys.stderr,
          flush=True)
    eng = sur_engine("screened", pop, R, K, N, T, p0)
    eng.ns_iterations(n_train)  # screened from the start, training online
    out["trained"] = timed(eng, n_timed)
    eng.close()
    print(f"[bench] surrogate: trained {out['trained']}", file=sys.stderr, flush=True)
    for k in ("baseline_only", "trained"):
        out[k]["ratio_walker_steps_vs_plain"] = (out[k]["walker_steps_per_s"] /
                                                 out["plain"]["walker_steps_per_s"])
    out["ratio_walker_steps"] = out["trained"]["ratio_walker_steps_vs_plain"]
    out["ratio_accepted_moves"] = (out["trained"]["accepted_moves_per_s"] /
                                   out["plain"]["accepted_moves_per_s"])
    out["trained_vs_baseline_only"] = {
        "walker_steps": out["trained"]["walker_steps_per_s"] /
        out["baseline_only"]["walker_steps_per_s"],
        "accepted_per_judge_eval": out["trained"]["accepted_per_judge_eval"] /
        max(out["baseline_only"]["accepted_per_judge_eval"], 1e-30)}
    out["log_z_check"] = logz_record("mpnn256")
    return out


def lj128_hmc_report(R=8, K=2048, N=128, T=20, n_timed=1) -> dict:
    """LJ-128 with HMC and surrogate delayed acceptance: the FP64 pair
    potential is the judge, HMC (10 leapfrog steps, dt 0.005) and galilean
    (n_reflect 10) at 1:1.  Reports the surrogate's held-out error against
    the FP64 judge and the stage-1 / stage-2 acceptance; the screened HMC is
    the exact constrained-prior move, its log Z tested against the plain
    random-walk moves at LJ-8.  The pair
    potential is far cheaper than the surrogate, so the screened walk is
    slower here; the configuration checks the method, not a speed-up."""
    from jaxrens_amd.backends import EnsembleTerm, Hamiltonian, LJBackend
    from jaxrens_amd.engine import NSEngine, RunConfig
    from jaxrens_amd.surrogate import LearnedConfig, init_surrogate
    P = list(np.geomspace(0.05, 5.0, R))
    moves = [{"type": "hmc", "n_leapfrog": 10, "step_size": 0.005, "weight": 1},
             {"type": "galilean", "n_reflect": 10, "step_size": 0.1, "weight": 1}]
    # 128 of the 6^3 simple-cubic sites at rho = 0.45 (spacing 1.10 sigma, near
    # the LJ minimum): the contours start inside the range the surrogate
    # trains on (U/N <= 0), so the auto walk can screen from its first records
    pos, ty, cell = sur_population(R, K, N, seed=9, rho=0.45)
    ty[:] = 0
    ws_it = R * (K - 1) * T
    out = {"workload": f"LJ-{N} (eps = sigma = 1, rc 2.5, FP64 judge), P = {R} pressures "
                       f"0.05-5, K = {K}, {T} sub-steps hmc(10 x 0.005):galilean(10) 1:1, "
                       "n_extra = K-2, initial jittered lattice at density 0.45 (site spacing 1.10)",
           "baseline": "pair term 4 (1/r)^12 below 2.5 (the r^-12 part of LJ)"}

    def run(mode, max_warm, n):
        lj = LJBackend(1.0, 1.0, 2.5)
        run = RunConfig(n_replicas=R, n_walkers=K, n_atoms=N, n_mcmc_steps=T, n_extra=K - 2,
                        seed=5, pressures=P, re_interval=1)
        # only configurations with U/N <= 0 are stored and scored: judged end
        # points with overlapping pairs (-4/r^6 below the Gaussians' reach)
        # are far outside any contour the walk samples
        # fallback bound 0.1 per atom: the network carries the whole -4/r^6 part
        # (about -7 per atom), so its BF16 operands alone leave ~1 % of that
        lc = LearnedConfig(mode=mode, ring_capacity=100_000, train_batch=64, train_steps=1000,
                           min_train=256, min_heldout=64, lr=2e-3, rmse_bound=0.1,
                           delta_k=2.0, train_max_epa=0.0, baseline=(4.0, 1.0, 2.5),
                           params=init_surrogate(2, S=1))
        eng = NSEngine(Hamiltonian(lj, [EnsembleTerm(P[0])]), moves, run, cell=None, learned=lc)
        eng.set_population(pos, ty, cell)
        warm = 0
        if mode == "plain" and max_warm:
            eng.ns_iterations(max_warm)  # the same contour as the screened run
            warm = max_warm
        # auto walks plain (training online) until the held-out error meets the bound
        while mode != "plain" and warm < max_warm:
            eng.ns_iterations(1)
            warm += 1
            st = eng.learned_stats()
            print(f"[bench] lj128 warm-up {warm}: held-out rmse {st['heldout_rmse']:.4f}, "
                  f"contour max {np.max(eng.emax()):.1f}, "
                  f"screened iterations {st['iterations_screened']}", file=sys.stderr,
                  flush=True)
            if st["iterations_screened"] > 0:
                break
        eng.learned_stats(reset=True)
        t = time.perf_counter()
        eng.ns_iterations(n)
        dt = time.perf_counter() - t
        st = eng.learned_stats()
        eng.close()
        r = {"walker_steps_per_s": ws_it * n / dt, "s_per_iteration": dt / n,
             "acceptance": st["accepted"] / max(st["proposals"], 1),
             "screened_iterations": st["iterations_screened"], "warm_up_iterations": warm}
        if mode != "plain":
            scr = st["iterations_screened"] > 0
            r.update({"heldout_rmse_per_atom_vs_fp64": st["heldout_rmse"],
                      "stage1_acceptance": (st["stage1_pass"] / max(st["proposals"], 1)
                                            if scr else None),
                      "stage2_acceptance": (st["accepted"] / max(st["stage1_pass"], 1)
                                            if scr else None),
                      "judge_energy_evals": st["exact_energy_evals"],
                      "surrogate_evals": st["surrogate_evals"]})
        return r

    out["screened"] = run("auto", 20, n_timed)
    out["plain"] = run("plain", out["screened"]["warm_up_iterations"], n_timed)
    # log Z: the screened HMC is the exact constrained-prior move, checked
    # against the plain pair-potential random walk (GMC + volume) at LJ-128
    # (screened in every iteration, 96 sub-steps) and at LJ-8, where the
    # reference's HMC rule is also compared
    out["log_z_check"] = {"lj128_screened_hmc_vs_random_walk": logz_record("lj128_hmc"),
                          "screened_hmc_vs_random_walk": logz_record("lj8_hmc"),
                          "reference_hmc_vs_random_walk": logz_record("lj8_reference_hmc")}
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-replicas", type=int, default=16)
    ap.add_argument("--cpu-iters", type=int, default=3)
    ap.add_argument("--skip-secondary", action="store_true")
    ap.add_argument("--skip-surrogate", action="store_true")
    ap.add_argument("--sur-timed", type=int, default=2,
                    help="timed iterations of the screened walk (256-atom config)")
    ap.add_argument("--sur-only", action="store_true",
                    help="run only the 256-atom surrogate comparison")
    ap.add_argument("--lj128-only", action="store_true",
                    help="run only the LJ-128 HMC surrogate report")
    ap.add_argument("--skip-lj128", action="store_true")
    ap.add_argument("--launch", default="1,3",
                    help="walk launch shape 'waves_per_block,occupancy' of the headline "
                         "(the fastest in a sweep of 1/2/4 waves x occupancy 0/3)")
    ap.add_argument("--tune", action="store_true",
                    help="pick the headline launch shape with the autotuner instead")
    ap.add_argument("--no-tune", action="store_true",
                    help="no autotuning for the secondary FP32-pair engine")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="run only a 1 GiB device copy (counter calibration)")
    args = ap.parse_args()
    if args.copy_calibration:
        print(json.dumps(copy_calibration()), flush=True)
        return
    if args.gpus != 1:
        print(f"bench.py measures one GPU only (got --gpus {args.gpus}); "
              "the multi-GPU replica sharding is a design, not a measured path.",
              file=sys.stderr)
        sys.exit(2)

    if args.sur_only:
        print(json.dumps({"surrogate": surrogate_bench(n_timed=args.sur_timed)}), flush=True)
        return
    if args.lj128_only:
        print(json.dumps({"lj128_hmc_surrogate": lj128_hmc_report()}), flush=True)
        return

    eng = build_engine(n_extra=1022)
    R, n_walk, T = eng.n_replicas, eng.n_walk, eng.run.n_mcmc_steps
    if args.tune:
        tune = eng.autotune(reps=2)
    else:
        wpb, occ = (int(x) for x in args.launch.split(","))
        eng.set_launch(wpb, occ)
        tune = {"waves_per_block": wpb, "occupancy": occ, "autotuned": False}
    print(f"[bench] engine ready: R={R} K={eng.n_walkers} N={eng.Nmax} "
          f"n_walk={n_walk} tune={tune} launch={eng.launch_info()}",
          file=sys.stderr, flush=True)
    eng.ns_iterations(args.warmup)
    eng.walk_timing(reset=True)
    t0 = time.perf_counter()
    out = eng.ns_iterations(args.steps)  # one host sync at the end
    dt = time.perf_counter() - t0
    walk_ms, launches = eng.walk_timing(reset=True)
    ws_per_iter = R * n_walk * T
    value = ws_per_iter * args.steps / dt
    flops = launch_flops(out["counters"], eng)
    avg_launch_ms = walk_ms / max(launches, 1)
    achieved = flops / launches / (avg_launch_ms * 1e-3) / 1e12
    print(f"[bench] timed {args.steps} iterations in {dt:.3f} s; walk kernel "
          f"avg {avg_launch_ms:.3f} ms", file=sys.stderr, flush=True)
    dist = iteration_distribution(eng, args.steps)

    secondary = None
    if not args.skip_secondary:
        secondary = {}
        for key, chunks in (("n_extra_0", 1), ("n_extra_0_teams", 4)):
            e0 = build_engine(n_extra=0, eval_chunks=chunks)
            e0.ns_iterations(5, graph=True)
            n0 = 200
            t1 = time.perf_counter()
            e0.ns_iterations(n0, graph=True)
            d0 = time.perf_counter() - t1
            secondary[key] = {
                "walker_steps_per_s": R * 1 * T * n0 / d0,
                "ns_iterations_per_s": n0 / d0, "iterations": n0,
                "mode": "hipGraph replay of one iteration",
                "eval_chunks": chunks, "launch": e0.launch_info()}
            if chunks > 1:
                secondary[key]["note"] = (
                    "evaluation teams: every chain walked by a block of 4 waves, each "
                    "evaluating one of 4 chunks of the ring steps (chunked FP64 order)")
            e0.close()
        secondary["n_extra_0_teams"]["speedup_vs_one_wave"] = (
            secondary["n_extra_0_teams"]["walker_steps_per_s"]
            / secondary["n_extra_0"]["walker_steps_per_s"])
        e32 = build_engine(n_extra=1022, precision="fp32_pairs")
        t32 = e32.autotune(reps=2) if not args.no_tune else None
        e32.ns_iterations(2)
        e32.walk_timing(reset=True)
        n32 = max(5, args.steps // 2)
        t2 = time.perf_counter()
        o32 = e32.ns_iterations(n32)
        d32 = time.perf_counter() - t2
        ms32, l32 = e32.walk_timing(reset=True)
        f32 = launch_flops(o32["counters"], e32) / l32 / (ms32 / l32 * 1e-3) / 1e12
        secondary["fp32_pairs"] = {
            "walker_steps_per_s": ws_per_iter * n32 / d32, "iterations": n32,
            "walk_avg_ms": ms32 / l32, "achieved_tflops": f32,
            "frac_of_fp32_peak": f32 / FP32_PEAK_TFLOPS,
            "frac_of_fp64_peak": f32 / FP64_PEAK_TFLOPS, "launch": t32,
            "note": "float32 pair terms, float64 minimum image and accumulation"}
        e32.close()
        secondary.update(config_throughput(cpu=not args.no_cpu_baseline))

    surrogate = None
    if not args.skip_surrogate:
        surrogate = surrogate_bench(n_timed=args.sur_timed)
    lj128 = None
    if not args.skip_lj128:
        lj128 = lj128_hmc_report()

    cpu = None
    if not args.no_cpu_baseline:
        # the box allots 16 host cores to one GPU (OMP_NUM_THREADS there)
        threads = min(int(os.environ.get("OMP_NUM_THREADS", "16") or 16), os.cpu_count() or 1)
        cpu = cpu_baseline(eng, args.cpu_replicas, threads, args.cpu_iters)
        cpu["host_cpus_visible"] = os.cpu_count()
        e0 = build_engine(n_extra=0)
        cpu["n_extra_0"] = cpu_baseline(e0, e0.n_replicas, threads, 200)
        e0.close()

    line = {
        "metric": METRIC,
        "value": value,
        "unit": "walker-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": dt / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (random initial walkers, device Philox draws)",
        "config": {
            "workload": "LJ-64 NPT replica stack: N=64, P=16 pressures 0.01-10, "
                        "K=1024, 20 MCMC sub-steps gmc(n_reflect=10):volume 4:1, "
                        "pressure RE every iteration, n_extra=K-2 (whole population)",
            "n_replicas": R, "n_walkers": eng.n_walkers, "n_atoms": eng.Nmax,
            "n_mcmc_steps": T, "n_extra": eng.run.n_extra,
            "walker_steps_per_iteration": ws_per_iter,
            "ns_iterations_per_s": args.steps / dt,
            "launch": {**eng.launch_info(), **(tune or {})},
            "iteration_distribution": dist,
        },
        "roofline": {
            "bound": "fp64_valu",
            "note": ("compute-bound on the FP64 vector ALUs; no MFMA instruction is "
                     "issued (the pair sums are not a dense contraction); peak is "
                     "MI355X's FP64 vector 78.6 TFLOP/s"),
            "kernel": "walk_kernel (persistent, FP64)",
            "achieved": achieved,
            "peak": FP64_PEAK_TFLOPS,
            "unit": "TFLOP/s",
            "frac": achieved / FP64_PEAK_TFLOPS,
            "traffic": measured_traffic(),
            "flops_per_launch": flops / launches,
            "avg_launch_ms": avg_launch_ms,
            "launches": launches,
        },
        "cpu_baseline": cpu,
        "secondary": secondary,
        "surrogate": surrogate,
        "lj128_hmc_surrogate": lj128,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()

"""Pins of the CPU restatement of the learned-model walk (oracle/learned_walk.py),
which the GPU parity tests compare against: its selection against the C
oracle's on the same draws, and its moves and two-stage test against closed
forms with stand-in models (``sampling/nested_sampling.py:86-212``,
``sampling/moves/galilean.py:142-300``, ``hmc.py:51-152``,
``volume.py:49-109``)."""

import numpy as np
import pytest

from jaxrens_amd.backends import LJBackend
from jaxrens_amd.draws import make_tape
from jaxrens_amd.moves import MWG
from oracle import learned_walk as OW

CELL = {"min_volume_per_atom": 0.125, "max_volume_per_atom": 64.0, "flat_V_prior": True}


class Flat:
    """Stand-in judge and surrogate: constant energies, zero forces."""

    def __init__(self, e_judge=0.0, e_sur=0.0):
        self.e_judge, self.e_sur = e_judge, e_sur
        self.judge_calls = 0

    def judge(self, pos, ty, cell, forces):
        self.judge_calls += 1
        return self.e_judge, (np.zeros_like(pos) if forces else None)

    def sur(self, pos, ty, cell, forces):
        return self.e_sur, (np.zeros_like(pos) if forces else None)


def chain(rng, N=5, L=6.0, H=0.0, Hs=0.0):
    return {"pos": rng.random((N, 3)) * L, "dir": np.zeros((N, 3)),
            "types": np.zeros(N, np.int32), "cell": np.eye(3) * L, "H": H, "n": N, "Hs": Hs}


def stats():
    return dict(proposals=0, stage1_pass=0, accepted=0, exact_fe=0, exact_e=0, sur_e=0)


GMC = {"type": 0, "n_reflect": 5, "perturb_angle": 0.1}
HMC = {"type": 2, "n_leapfrog": 4}


def test_select_matches_the_c_oracle(oracle):
    """Worst, clone, extras and the contour equal the C oracle's selection on
    the same draws, exact maxima and noise ties included."""
    rng = np.random.default_rng(3)
    R, K, N, n_extra = 3, 10, 4, 5
    L = 4.0
    pos = rng.random((R, K, N, 3)) * L
    pop = oracle.Population(pos, np.zeros((R, K, N), np.int32), np.full((R, K), N, np.int32),
                            np.tile(np.eye(3) * L, (R, K, 1, 1)), rng.normal(size=(R, K)),
                            np.zeros((R, K, N, 3)))
    pop.energy[0, [1, 4, 8]] = 9.0   # exact maxima
    pop.energy[1, [0, 9]] = 9.0
    mw = MWG([{"type": "gmc", "n_reflect": 1, "step_size": 0.0}], CELL)
    tape = make_tape(rng, 1, R, K, N, 1, n_extra, mw.move_probs, ties=True)
    E0 = pop.energy.copy()
    pot = oracle.Pot.from_table(LJBackend(1.0, 1.0, 2.5).table)
    out = oracle.ns_iteration(pop, pot, np.zeros(R), np.zeros((R, 1)), np.zeros((R, 1)),
                              mw.as_oracle_moves(), 1, n_extra, tape, do_re=False)
    for r in range(R):
        w, cl, walk, mx = OW.select(E0[r], tape.sel_u[0, r], tape.clone_u[0, r],
                                    tape.extra_u[0, r], n_extra)
        assert (w, cl, mx) == (out["worst"][r], out["clone"][r], out["emax"][r])
        assert walk == out["walk"][r].tolist()


def test_gmc_without_forces_is_a_straight_line():
    """Zero forces and Emax = +inf: the trajectory is x0 + n s d with
    d = xi / |xi| (fresh direction), accepted, direction kept."""
    rng = np.random.default_rng(1)
    ch = chain(rng)
    x0 = ch["pos"].copy()
    xi = rng.normal(size=(3, ch["n"]))
    st = stats()
    rs, ne, ng = OW.gmc(ch, GMC, 0.05, xi, Flat(), np.inf, 0.0, (np.inf, np.inf), False, st,
                        [], False)
    d = xi.T / np.sqrt(np.sum(xi * xi))
    assert (rs, ne, ng) == (0, 5, 5) and st["exact_fe"] == 5
    np.testing.assert_allclose(ch["pos"], x0 + 5 * 0.05 * d, rtol=0, atol=1e-14)
    np.testing.assert_allclose(ch["dir"], d, rtol=1e-15)


def test_gmc_reject_reverts_and_flips_the_direction():
    rng = np.random.default_rng(2)
    ch = chain(rng)
    ch["dir"] = rng.normal(size=(ch["n"], 3))
    x0, dprev = ch["pos"].copy(), ch["dir"].copy()
    xi = rng.normal(size=(3, ch["n"]))
    rs, _, _ = OW.gmc(ch, GMC, 0.05, xi, Flat(), -np.inf, 0.0, (-np.inf, -np.inf), False,
                      stats(), [], False)
    v = dprev + 0.1 * xi.T
    d0 = v / np.sqrt(np.sum(v * v))
    assert rs == 1
    np.testing.assert_array_equal(ch["pos"], x0)
    np.testing.assert_allclose(ch["dir"], -d0, rtol=1e-15)


def test_hmc_without_forces_drifts_and_conserves():
    """Zero forces: p stays xi, x moves by n_leapfrog s p, K is conserved, so
    the reference's test reduces to H' < emax."""
    rng = np.random.default_rng(4)
    ch = chain(rng, H=0.0)
    x0 = ch["pos"].copy()
    xi = rng.normal(size=(3, ch["n"]))
    rs, ne, ng = OW.hmc(ch, HMC, 0.01, xi, Flat(), 1.0, 0.0, (np.inf, np.inf), False, stats(),
                        [], False)
    assert (rs, ne, ng) == (0, 9, 8)
    np.testing.assert_allclose(ch["pos"], x0 + 4 * 0.01 * xi.T, rtol=0, atol=1e-14)
    ch2 = chain(rng, H=0.0)
    rs, _, _ = OW.hmc(ch2, HMC, 0.01, xi, Flat(e_judge=0.5), 0.25, 0.0, (np.inf, np.inf),
                      False, stats(), [], False)
    assert rs == 1  # H' = 0.5 >= emax


def test_hmc_energy_cut_uses_the_judge():
    """Accepted iff H' < emax and (H' + K') - (H + K) < 1 with the judge's H'."""
    rng = np.random.default_rng(5)
    xi = rng.normal(size=(3, 5))
    for e_judge, want in ((0.9, 0), (1.1, 1)):
        ch = chain(rng, H=0.0)
        rs, _, _ = OW.hmc(ch, HMC, 0.01, xi, Flat(e_judge=e_judge), 10.0, 0.0,
                          (np.inf, np.inf), False, stats(), [], False)
        assert rs == want


class ConstForce(Flat):
    """Stand-in surrogate with a uniform force f on every atom (energy
    linear in the positions, its value irrelevant here)."""

    def __init__(self, f, e_judge=0.0):
        super().__init__(e_judge=e_judge)
        self.f = np.asarray(f, dtype=float)

    def sur(self, pos, ty, cell, forces):
        return 0.0, (np.broadcast_to(self.f, pos.shape).copy() if forces else None)


def test_screened_hmc_is_metropolis_on_the_kinetic_energy():
    """Screened: the surrogate's forces drive the leapfrog and the move is
    accepted iff the judge's H' < emax and u < exp(K - K'), whatever H' is
    below the contour (the constrained prior is flat in x).  With a uniform
    surrogate force f each of the L leapfrog steps adds s f to every
    momentum, so K' = |p + L s f|^2 / 2 in closed form; u just below and
    just above exp(K - K') decide the move, a judge H' lower than H leaves
    the bar where it is, and a judge H' above the contour rejects whatever
    u.  (The plain rule's hard cut on H' + K' - H - K would accept the
    second case.)"""
    rng = np.random.default_rng(8)
    xi = rng.normal(size=(3, 5))
    p = xi.T
    f, s, L = 2.0 * p.sum(0) / np.linalg.norm(p.sum(0)), 0.01, HMC["n_leapfrog"]
    K0 = 0.5 * np.sum(p * p)
    K1 = 0.5 * np.sum((p + L * s * f) ** 2)
    bar = np.exp(K0 - K1)
    assert 0.3 < bar < 0.95  # the force raises K: a real test
    cases = ((bar * (1 - 1e-9), 0.0, 0), (bar * (1 + 1e-9), 0.0, 1),
             (bar * (1 - 1e-9), -0.5, 0), (bar * (1 + 1e-9), -0.5, 1),
             (0.0, 20.0, 1))
    for u, e_judge, want in cases:
        ch = chain(rng, H=0.0, Hs=0.0)
        rs, ne, ng = OW.hmc(ch, HMC, s, xi, ConstForce(f, e_judge), 10.0, 0.0,
                            (np.inf, np.inf), True, stats(), [], True, u0=u)
        assert rs == want, (u, e_judge)
        assert (ne, ng) == (1, 0)  # one judge energy, no judge force
        assert ch["H"] == (e_judge if want == 0 else 0.0)


@pytest.mark.parametrize("move", ["gmc", "hmc"])
def test_stage1_is_symmetric_and_screens_before_the_judge(move):
    """Screened: when the surrogate's verdicts at x and y differ the proposal
    is rejected without a judge evaluation; when they agree the judge
    decides (for either common verdict)."""
    rng = np.random.default_rng(6)
    xi = rng.normal(size=(3, 5))
    fn = OW.gmc if move == "gmc" else OW.hmc
    mv = GMC if move == "gmc" else HMC
    thr = (1.0, 1.0)
    cases = [  # (Hs at x, surrogate energy at y, judge energy, rejected, judge called)
        (0.0, 5.0, 0.0, True, False),   # inside -> outside by the surrogate
        (5.0, 0.0, 0.0, True, False),   # outside -> inside
        (0.0, 0.5, 0.0, False, True),   # both inside, judge inside
        (5.0, 6.0, 0.0, False, True),   # both outside, judge inside
        (0.0, 0.5, 3.0, True, True),    # both inside, judge outside
    ]
    for hs_x, es_y, ej, rejected, called in cases:
        ch = chain(rng, H=0.0, Hs=hs_x)
        m = Flat(e_judge=ej, e_sur=es_y)
        st = stats()
        rs, _, _ = fn(ch, mv, 0.01, xi, m, 2.0, 0.0, thr, True, st, [], True)
        assert (rs != 0) == rejected
        assert (m.judge_calls > 0) == called
        assert st["stage1_pass"] == (1 if called else 0)


def test_volume_codes_and_prior():
    """Cell-shape failures (code 2) and prior failures (code 3) are rejected
    before the judge; a passing proposal scales cell and positions by
    (V'/V)^(1/3)."""
    rng = np.random.default_rng(7)
    mv = {"type": 1, "flat_v_prior": 0, "max_vpa": 64.0, "min_vpa": 0.5, "min_aspect": 0.8}
    ch = chain(rng, N=4, L=3.0)
    m = Flat()
    # far below min_vpa: V' = |V + s N xi| with V = 27, N = 4 -> V' = 1
    rs, ne, _ = OW.volume(ch, mv, 1.0, -6.5, 0.5, m, np.inf, 0.0, (np.inf, np.inf), False,
                          stats(), [], False)
    assert rs == 2 and ne == 0 and m.judge_calls == 0
    # shrink by half with the V^N prior: p_acc = 0.5^4, u = 0.5 fails
    rs, ne, _ = OW.volume(ch, mv, 1.0, -27 / 8, 0.5, m, np.inf, 0.0, (np.inf, np.inf), False,
                          stats(), [], False)
    assert rs == 3 and ne == 0 and m.judge_calls == 0
    x0, c0 = ch["pos"].copy(), ch["cell"].copy()
    rs, ne, _ = OW.volume(ch, mv, 1.0, 27 / 4, 0.0, m, np.inf, 0.0, (np.inf, np.inf), False,
                          stats(), [], False)
    s = (54 / 27) ** (1 / 3)
    assert rs == 0 and ne == 1 and m.judge_calls == 1
    np.testing.assert_allclose(ch["cell"], c0 * s, rtol=1e-15)
    np.testing.assert_allclose(ch["pos"], x0 * s, rtol=1e-15)


def _flat_pop(oracle, rng, R=2, K=3, N=5, L=6.0, H=0.0):
    return oracle.Population(rng.random((R, K, N, 3)) * L, np.zeros((R, K, N), np.int32),
                             np.full((R, K), N, np.int32), np.tile(np.eye(3) * L, (R, K, 1, 1)),
                             np.full((R, K), H), np.zeros((R, K, N, 3)))


VOL = {"type": 1, "flat_v_prior": True, "max_vpa": 64.0, "min_vpa": 0.125, "min_aspect": 0.8}


def test_trial_moves_decide_without_touching_the_population(oracle):
    """Trial moves with a flat judge: every GMC proposal is accepted under a
    contour above the judge's energy and rejected (reason 1) under one below,
    a volume step that leaves the cell bounds is a cell reject (reason 2),
    and the population is left as it was; a walker sampled twice with the
    same draws gets the same reason."""
    rng = np.random.default_rng(5)
    pop = _flat_pop(oracle, rng)
    before = pop.pos.copy(), pop.dir.copy(), pop.energy.copy()
    R, n, N = 2, 4, 5
    idx = np.array([[0, 1, 1, 2], [2, 2, 0, 1]], np.int32)
    nrm = rng.standard_normal((R, n, 3, N))
    nrm[:, 2] = nrm[:, 1]
    uni = rng.random((R, n, 4))
    P = np.zeros(2)  # H = U: the flat judge sits at 0
    m = Flat()
    for emax, want in ((1.0, 0), (-1.0, 1)):
        rs = OW.trial_moves(pop, m, P, [GMC], 0, np.array([0.1, 0.2]),
                            np.array([emax, emax]), idx, nrm, uni, "plain")
        assert np.all(rs == want)
    # a volume step of 1e4 per atom leaves the cell bounds
    nrm[..., 0, 0] = 1.0
    rs = OW.trial_moves(pop, m, P, [GMC, VOL], 1, np.array([1e4, 1e4]),
                        np.array([1.0, 1.0]), idx, nrm, uni, "plain")
    assert np.all(rs == 2)
    assert rs[0, 1] == rs[0, 2]
    for a, b in zip(before, (pop.pos, pop.dir, pop.energy)):
        assert np.array_equal(a, b)


def test_initial_walk_is_the_moves_in_schedule_order(oracle):
    """The burn-in walk with a flat judge above the contour: every GMC step
    is the straight line of the normalised draw (fresh at a zero direction,
    then perturbed), applied to each walker in schedule order, and the
    counters count every proposal as accepted; one sub-step from a walker
    equals that walker's trial move from the same state and draws."""
    rng = np.random.default_rng(6)
    R, K, N, wl = 2, 3, 5, 3
    pop = _flat_pop(oracle, rng, R, K, N)
    pop0 = _flat_pop(oracle, np.random.default_rng(6), R, K, N)
    nrm = rng.standard_normal((R, K, wl, 3, N))
    uni = rng.random((R, K, wl, 4))
    steps = np.array([[0.1], [0.2]])
    emax = np.array([1.0, 1.0])
    P = np.zeros(2)  # H = U: the flat judge sits at 0
    ctr = OW.initial_walk(pop, Flat(), P, steps, [GMC], emax, np.zeros(wl, np.int32), nrm,
                          uni, "plain")
    assert np.array_equal(ctr[:, 0, :2], np.full((R, 2), K * wl))
    for r in range(R):
        for k in range(K):
            x, d = pop0.pos[r, k].copy(), np.zeros((N, 3))
            for t in range(wl):
                z = nrm[r, k, t].T
                v = z if np.sqrt(np.sum(d * d)) < 1e-8 else d + GMC["perturb_angle"] * z
                d = v / max(np.sqrt(np.sum(v * v)), 1e-10)
                for _ in range(GMC["n_reflect"]):
                    x = x + steps[r, 0] * d
            np.testing.assert_allclose(pop.pos[r, k], x, rtol=0, atol=1e-12)
    # one sub-step equals the trial move of each walker
    pop1 = _flat_pop(oracle, np.random.default_rng(6), R, K, N)
    idx = np.tile(np.arange(K, dtype=np.int32), (R, 1))
    rs = OW.trial_moves(pop1, Flat(), P, [GMC], 0, steps[:, 0], np.array([-1.0, 1.0]), idx,
                        nrm[:, :, 0], uni[:, :, 0], "plain")
    c1 = OW.initial_walk(pop1, Flat(), P, steps, [GMC], np.array([-1.0, 1.0]),
                         np.zeros(1, np.int32), nrm[:, :, :1], uni[:, :, :1], "plain")
    assert np.array_equal(c1[:, 0, 1], (rs == 0).sum(axis=1))
    assert np.array_equal(c1[:, 0, 2], (rs == 1).sum(axis=1))


class Tilted(Flat):
    """Stand-in judge E = sum of the x coordinates, zero forces."""

    def judge(self, pos, ty, cell, forces):
        self.judge_calls += 1
        return float(np.sum(pos[:, 0])), (np.zeros_like(pos) if forces else None)


def test_trial_moves_take_the_sampled_walker_and_its_draws(oracle):
    """With E = sum x and no forces each trial GMC move ends at
    x0 + n_reflect s d for the sampled walker's x0 and the sample's own draw,
    so the reason follows from that end point against the contour; a wrong
    walker or draw index changes it."""
    rng = np.random.default_rng(8)
    R, K, N, n = 2, 6, 5, 8
    pop = _flat_pop(oracle, rng, R, K, N)
    idx = rng.integers(0, K, size=(R, n)).astype(np.int32)
    nrm = rng.standard_normal((R, n, 3, N))
    uni = rng.random((R, n, 4))
    steps = np.array([0.3, 0.5])
    emax = np.array([np.median(pop.pos[0, :, :, 0].sum(-1)),
                     np.median(pop.pos[1, :, :, 0].sum(-1))])
    rs = OW.trial_moves(pop, Tilted(), np.zeros(2), [GMC], 0, steps, emax, idx, nrm, uni,
                        "plain")
    want = np.zeros((R, n), int)
    for r in range(R):
        for j in range(n):
            z = nrm[r, j].T
            d = z / max(np.sqrt(np.sum(z * z)), 1e-10)
            x = pop.pos[r, idx[r, j]].copy()
            for _ in range(GMC["n_reflect"]):
                x = x + steps[r] * d
            want[r, j] = 0 if np.sum(x[:, 0]) < emax[r] else 1
    assert 0 < (want == 0).sum() < want.size  # both outcomes occur
    assert np.array_equal(rs, want)

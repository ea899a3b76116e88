"""Pins the oracle's move kernels and NS bookkeeping by invariants and
closed forms (``sampling/moves/galilean.py:142-300``, ``volume.py:49-109``,
``_common.py:16-93``, ``hmc.py:51-152``, ``alchemical.py:40-79``,
``nested_sampling.py:86-212,391-686``, ``replica_exchange.py:318-436``)."""

import ctypes as C

import numpy as np
import pytest

from jaxrens_amd import _lib
from jaxrens_amd.backends import LJBackend
from jaxrens_amd.draws import make_tape
from jaxrens_amd.moves import MWG
from jaxrens_amd import prior_mass as pm

CELL = {"min_volume_per_atom": 0.125, "max_volume_per_atom": 64.0, "flat_V_prior": True}


def lattice(n_side, a):
    g = np.arange(n_side) * a
    return np.array([[x, y, z] for x in g for y in g for z in g], dtype=np.float64)


class Chain:
    """One walker driven through or_walk_chain."""

    def __init__(self, oracle, pos, cell, types=None, pressure=0.0, mu=None,
                 lj=None, moves=None):
        self.O = oracle
        self.lj = lj or LJBackend(1.0, 1.0, 2.5)
        self.pot = oracle.Pot.from_table(self.lj.table)
        self.N = len(pos)
        self.pos = oracle.f64(pos).copy()
        self.cell = oracle.f64(cell).reshape(9).copy()
        self.types = oracle.i32(np.zeros(self.N) if types is None else types).copy()
        self.dir = np.zeros((self.N, 3))
        self.pressure = pressure
        S = self.lj.n_species
        self.mu = oracle.f64(np.zeros(S) if mu is None else mu)
        U, H, _ = oracle.energy_forces(self.pot, self.pos, self.types, self.cell,
                                       pressure, self.mu)
        self.E = np.array([H])
        self.mwg = MWG(moves or [{"type": "gmc", "n_reflect": 3, "step_size": 0.05},
                                 {"type": "volume", "step_size": 0.3},
                                 {"type": "hmc", "n_leapfrog": 5, "step_size": 0.002},
                                 {"type": "alchemical_morph", "n_species": 2}], CELL)
        nx = self.mwg.n_dir_slots - 1
        self.dirx = np.zeros((nx, self.N, 3)) if nx else None

    def step(self, m, emax, xi, u=(0.5, 0.5, 0.5, 0.5), step=None):
        O = self.O
        ss = O.f64(self.mwg.default_step_sizes if step is None
                   else np.full(self.mwg.n_moves, step))
        xi = O.f64(xi).reshape(3, self.N)
        ctr = np.zeros((self.mwg.n_moves, 7), np.int32)
        mv = O.moves_array(self.mwg.as_oracle_moves())
        sched = np.array([m], np.int32)
        O.lib().or_walk_chain(self.N, self.N, O._d(self.pos), O._i(self.types),
                              O._d(self.cell), O._d(self.E), O._d(self.dir),
                              O._d(self.dirx) if self.dirx is not None else None,
                              self.N * 3, C.byref(self.pot.s), self.pressure, O._d(self.mu), emax,
                              O._d(ss), mv, self.mwg.n_moves, O._i(sched), 1,
                              O._d(xi), O._d(O.f64(u)), O._i(ctr))
        return ctr[m]

    def energy(self):
        return self.O.energy_forces(self.pot, self.pos, self.types, self.cell,
                                    self.pressure, self.mu)


def test_gmc_unconstrained_is_straight_line(oracle):
    pos = lattice(2, 1.3) + 2.0
    ch = Chain(oracle, pos, np.eye(3) * 6.0)
    xi = np.random.default_rng(0).standard_normal((3, 8))
    x0 = ch.pos.copy()
    c = ch.step(0, np.inf, xi)
    assert c[1] == 1  # accepted
    d = xi.T / np.linalg.norm(xi)  # fresh direction (stored one was zero)
    np.testing.assert_allclose(ch.pos, x0 + 3 * 0.05 * d, rtol=0, atol=1e-14)
    np.testing.assert_allclose(ch.dir, d, rtol=0, atol=1e-15)
    assert ch.E[0] == ch.energy()[1]
    assert c[5] == 3 and c[6] == 3  # n_reflect evaluations, all with forces


def test_gmc_reject_reverts_and_flips_perturbed_direction(oracle):
    pos = lattice(2, 1.3) + 2.0
    ch = Chain(oracle, pos, np.eye(3) * 6.0)
    rng = np.random.default_rng(1)
    ch.step(0, np.inf, rng.standard_normal((3, 8)))  # sets a direction
    d_prev = ch.dir.copy()
    x0, E0 = ch.pos.copy(), ch.E[0]
    xi = rng.standard_normal((3, 8))
    c = ch.step(0, -np.inf, xi)
    assert c[2] == 1  # energy reject
    assert np.array_equal(ch.pos, x0) and ch.E[0] == E0
    v = d_prev + 0.1 * xi.T
    np.testing.assert_allclose(ch.dir, -v / np.linalg.norm(v), atol=1e-15)


def test_gmc_reflection_is_reversible(oracle):
    """With zero perturbation noise GMC is a reversible, norm-preserving map:
    running back from (x_n, -d_n) returns to (x_0, -d_0)."""
    rng = np.random.default_rng(2)
    base = lattice(3, 1.12) + 0.3 + rng.uniform(-0.03, 0.03, (27, 3))
    ch = Chain(oracle, base, np.eye(3) * 3.36)
    ch.dir = rng.standard_normal((27, 3))
    ch.dir /= np.linalg.norm(ch.dir)
    # emax just above the current energy so that steps reflect
    E_now = ch.E[0]
    emax = E_now + 0.5
    zero = np.zeros((3, 27))
    x0 = ch.pos.copy()
    d0 = ch.dir / np.linalg.norm(ch.dir)
    ch.mwg = MWG([{"type": "gmc", "n_reflect": 12, "step_size": 0.01}], CELL)
    c = ch.step(0, emax, zero)
    if c[1] != 1:
        pytest.skip("trajectory not accepted for this seed")
    assert abs(np.linalg.norm(ch.dir) - 1.0) < 1e-12
    ch.dir = -ch.dir
    ch.step(0, emax, zero)
    np.testing.assert_allclose(ch.pos, x0, atol=1e-9)
    np.testing.assert_allclose(ch.dir, -d0, atol=1e-9)


def test_gmc_reflects_off_force_normal(oracle):
    """One violated step: the direction is mirrored in the plane normal to F."""
    pos = np.array([[0.0, 0, 0], [1.0, 0.2, 0.0]])
    ch = Chain(oracle, pos, np.eye(3) * 8.0,
               moves=[{"type": "gmc", "n_reflect": 2, "step_size": 1e-4}])
    xi = np.array([[1.0, -1.0], [0.3, 0.1], [0.0, 0.2]])
    d = xi.T / np.linalg.norm(xi)
    x1 = ch.pos + 1e-4 * d
    _, H1, F1 = oracle.energy_forces(ch.pot, x1, ch.types, ch.cell)
    emax = H1 - 1e-12  # first step violates; walk continues
    ch.step(0, emax, xi)
    f = F1 / np.linalg.norm(F1)
    d_ref = d - 2 * np.sum(f * d) * f
    x2 = x1 + 1e-4 * d_ref
    _, H2, _ = oracle.energy_forces(ch.pot, x2, ch.types, ch.cell)
    if H2 < emax:  # accepted: state is the end of the reflected path
        np.testing.assert_allclose(ch.pos, x2, atol=1e-13)
    else:  # rejected: flip of the initial direction
        np.testing.assert_allclose(ch.dir, -d, atol=1e-14)
    assert np.sum(d_ref * f) == pytest.approx(-np.sum(d * f), abs=1e-12)


def test_volume_move_closed_form_and_reasons(oracle):
    pos = lattice(2, 1.5) + 0.5
    cell = np.eye(3) * 4.0
    ch = Chain(oracle, pos, cell, pressure=0.1)
    V0 = 64.0
    c = ch.step(1, np.inf, np.full((3, 8), 0.7), u=(0.3, 0, 0, 0))
    assert c[1] == 1
    V1 = abs(np.linalg.det(ch.cell.reshape(3, 3)))
    assert V1 == pytest.approx(V0 + 0.3 * 8 * 0.7, rel=1e-14)
    s = (V1 / V0) ** (1 / 3)
    np.testing.assert_allclose(ch.pos, pos * s, rtol=1e-15)
    assert ch.E[0] == ch.energy()[1]
    # energy reject (reason 1), cell reject (reason 2)
    assert ch.step(1, -np.inf, np.full((3, 8), 0.7))[2] == 1
    big = np.full((3, 8), 1000.0)  # vpa far above 64
    assert ch.step(1, np.inf, big)[3] == 1


def test_volume_prior_reject(oracle):
    pos = lattice(2, 1.5) + 0.5
    ch = Chain(oracle, pos, np.eye(3) * 4.0,
               moves=[{"type": "volume", "step_size": 0.3}])
    ch.mwg = MWG([{"type": "volume", "step_size": 0.3}],
                 {**CELL, "flat_V_prior": False})
    # shrink: ratio < 1 so p_acc = ratio**N < 1; u close to 1 rejects
    c = ch.step(0, np.inf, np.full((3, 8), -0.5), u=(0.999999, 0, 0, 0))
    assert c[4] == 1  # prior reject (reason 3)
    c = ch.step(0, np.inf, np.full((3, 8), -0.5), u=(0.0, 0, 0, 0))
    assert c[1] == 1


def test_hmc_small_step_conserves_and_accepts(oracle):
    rng = np.random.default_rng(3)
    pos = lattice(3, 1.2) + rng.uniform(-0.05, 0.05, (27, 3))
    ch = Chain(oracle, pos, np.eye(3) * 3.6)
    x0 = ch.pos.copy()
    c = ch.step(2, ch.E[0] + 50.0, rng.standard_normal((3, 27)))
    assert c[1] == 1 and c[5] == 11 and c[6] == 10
    assert not np.array_equal(ch.pos, x0)
    c = ch.step(2, -np.inf, rng.standard_normal((3, 27)))
    assert c[2] == 1


def test_hmc_energy_error_cut(oracle):
    rng = np.random.default_rng(4)
    pos = lattice(3, 1.2) + rng.uniform(-0.05, 0.05, (27, 3))
    ch = Chain(oracle, pos, np.eye(3) * 3.6,
               moves=[{"type": "hmc", "n_leapfrog": 5, "step_size": 0.2}])
    x0 = ch.pos.copy()
    c = ch.step(0, np.inf, 3.0 * rng.standard_normal((3, 27)))
    assert c[2] == 1 and np.array_equal(ch.pos, x0)  # dH >= 1 rejects


def test_morph_changes_one_atom(oracle):
    lj = LJBackend(pair_epsilon=[[1.0, 1.5], [1.5, 0.5]],
                   pair_sigma=[[1.0, 0.8], [0.8, 0.88]], cutoff=2.5)
    pos = lattice(2, 1.4) + 0.5
    types = np.array([0, 1, 0, 1, 0, 1, 0, 1])
    ch = Chain(oracle, pos, np.eye(3) * 4.0, types=types, pressure=1.0,
               mu=[0.0, 0.7], lj=lj)
    E0 = ch.E[0]
    c = ch.step(3, np.inf, np.zeros((3, 8)), u=(0.3, 0.9, 0, 0))
    assert c[1] == 1
    assert (ch.types != types).sum() == 1
    assert ch.types[2] == 1  # atom floor(0.3*8)=2, species 0 -> 1
    assert ch.E[0] == ch.energy()[1]
    assert ch.E[0] != E0
    t1 = ch.types.copy()
    c = ch.step(3, -np.inf, np.zeros((3, 8)), u=(0.3, 0.2, 0, 0))
    assert c[2] == 1 and np.array_equal(ch.types, t1)


# --------------------------------------------------------------------- NS
def _pop(oracle, R, K, N, rng, energies=None):
    L = 4.0
    pos = rng.random((R, K, N, 3)) * L
    cell = np.tile(np.eye(3) * L, (R, K, 1, 1))
    types = np.zeros((R, K, N), np.int32)
    na = np.full((R, K), N, np.int32)
    pop = oracle.Population(pos, types, na, cell, np.zeros((R, K)), np.zeros((R, K, N, 3)))
    lj = LJBackend(1.0, 1.0, 2.5)
    pot = oracle.Pot.from_table(lj.table)
    for r in range(R):
        for k in range(K):
            pop.energy[r, k] = oracle.energy_forces(pot, pop.pos[r, k], types[r, k],
                                                    cell[r, k], forces=False)[1]
    if energies is not None:
        pop.energy[:] = energies
    return pop, pot


def test_select_tie_break_clone_and_extras(oracle):
    rng = np.random.default_rng(5)
    R, K, N = 2, 12, 4
    E = rng.normal(size=(R, K))
    E[0, [2, 7, 9]] = 50.0  # three exact maxima in replica 0
    pop, pot = _pop(oracle, R, K, N, rng, E)
    mw = MWG([{"type": "gmc", "n_reflect": 1, "step_size": 0.0}], CELL)
    tape = make_tape(rng, 1, R, K, N, 1, 6, mw.move_probs)
    tape.sel_u[0, 0, :] = 0.1
    tape.sel_u[0, 0, 7] = 0.9
    before = pop.copy()
    out = oracle.ns_iteration(pop, pot, np.zeros(R), np.zeros((R, 1)),
                              np.full((R, 1), 0.0), mw.as_oracle_moves(), 1, 6, tape,
                              do_re=False)
    assert out["worst"][0] == 7 and out["emax"][0] == 50.0
    assert out["emax"][1] == before.energy[1].max()
    for r in range(R):
        w, c = out["worst"][r], out["clone"][r]
        assert c != w
        walk = out["walk"][r]
        assert walk[0] == w and c not in walk[1:] and w not in walk[1:]
        assert len(set(walk.tolist())) == len(walk)
        # extras are the top-6 noise walkers excluding worst and clone
        noise = tape.extra_u[0, r].copy()
        noise[[w, c]] = -np.inf
        assert np.array_equal(np.sort(walk[1:]), np.sort(np.argsort(-noise, kind="stable")[:6]))
        # the culled slot holds a copy of the clone source
        assert np.array_equal(pop.types[r, w], before.types[r, c])


def test_clone_index_is_uniform_over_survivors(oracle):
    rng = np.random.default_rng(6)
    R, K, N = 1, 5, 2
    pop, pot = _pop(oracle, R, K, N, rng, np.arange(5.0)[None])
    mw = MWG([{"type": "gmc", "n_reflect": 1, "step_size": 0.0}], CELL)
    seen = []
    for u in np.linspace(0, 0.999, 40):
        p = pop.copy()
        tape = make_tape(rng, 1, R, K, N, 1, 0, mw.move_probs)
        tape.clone_u[0, 0] = u
        out = oracle.ns_iteration(p, pot, np.zeros(1), np.zeros((1, 1)), np.zeros((1, 1)),
                                  mw.as_oracle_moves(), 1, 0, tape, do_re=False)
        assert out["worst"][0] == 4
        seen.append(out["clone"][0])
    assert sorted(set(seen)) == [0, 1, 2, 3]
    assert np.bincount(seen).tolist() == [10, 10, 10, 10]


def test_prior_mass_matches_pymatnest_closed_form():
    K = 32
    n = 100
    # pymatnest (PRX 4, 031034 (2014), Eq. 3), n_cull = 1: the mass left
    # after dead point i is X_i = (K/(K+1))**(i+1) and a_i = X_i/(K+1)
    ref = (np.arange(n) + 1) * np.log(K / (K + 1.0)) - np.log(K + 1.0)
    np.testing.assert_allclose(pm.log_dead_weights(n, K, 1), ref, rtol=1e-13)
    # mass conservation at iteration boundaries: dead + live = X_0 = K/(K+1)
    for n_cull, nd in ((1, 100), (3, 99)):
        lw = pm.log_dead_weights(nd, K, n_cull)
        live = pm.log_live_weight(nd, K, n_cull)
        total = np.logaddexp.reduce(np.append(lw, np.log(K) + live))
        assert total == pytest.approx(np.log(K / (K + 1.0)), abs=1e-12)
    it = np.arange(5)
    np.testing.assert_allclose(pm.dead_log_weight(it, K, 1)[:, 0], ref[:5], rtol=1e-13)


def test_replica_exchange_rebase_and_schedule(oracle):
    rng = np.random.default_rng(8)
    R, K, N = 3, 2, 2
    pop, pot = _pop(oracle, R, K, N, rng)
    S = 2
    pop.types[:] = np.array([0, 1], np.int32)
    P = np.array([0.1, 0.2, 0.4])
    mu = np.array([[0.0, 0.5], [0.0, 1.0], [0.0, 2.0]])
    V = np.abs(np.linalg.det(pop.cell.reshape(R, K, 3, 3)))
    U = np.full((R, K), -3.0)
    pop.energy[:] = U + P[:, None] * V - mu[:, None, 1]
    emax = np.full(R, 1e9)
    re_u = np.zeros((1, 2, 1, 2))  # always walker 0
    before = pop.copy()
    acc, att = oracle.replica_exchange(pop, P, mu, emax, re_u, 1, S=S)
    assert att.tolist() == [1, 1]  # pairs (0,1) then (1,2)
    assert acc.tolist() == [1, 1]
    # walker from replica 0 moved to 1 then to 2; energies re-based by hand
    assert np.array_equal(pop.pos[2, 0], before.pos[0, 0])
    assert pop.energy[2, 0] == pytest.approx(-3.0 + 0.4 * V[0, 0] - 2.0, rel=1e-13)
    assert pop.energy[0, 0] == pytest.approx(-3.0 + 0.1 * V[1, 0] - 0.5, rel=1e-13)
    # an incoming energy above the receiving emax blocks the swap
    pop2 = before.copy()
    acc, att = oracle.replica_exchange(pop2, P, mu, np.full(R, -1e9), re_u, 1, S=S)
    assert acc.tolist() == [0, 0] and att.tolist() == [1, 1]
    assert np.array_equal(pop2.pos, before.pos)


BINARY = dict(pair_epsilon=[[1.0, 1.5], [1.5, 0.5]], pair_sigma=[[1.0, 0.8], [0.8, 0.88]],
              cutoff=2.5)
SCOPED = [{"type": "gmc", "species": 0, "n_reflect": 3, "step_size": 0.05},
          {"type": "gmc", "species": [1], "n_reflect": 3, "step_size": 0.05,
           "refresh_direction": True},
          {"type": "gmc", "n_reflect": 3, "step_size": 0.05}]


def _binary_chain(oracle):
    pos = lattice(2, 1.3) + 2.0
    types = np.array([0, 1, 0, 1, 1, 0, 0, 1])
    return Chain(oracle, pos, np.eye(3) * 6.0, types=types, lj=LJBackend(**BINARY),
                 moves=SCOPED)


def test_scoped_gmc_moves_only_its_species(oracle):
    """galilean.py:58-72,171-180: the direction is the normalised masked
    normal draw; atoms outside the scope never move, and the move keeps its
    own persistent direction (slot 1), leaving the shared one untouched."""
    ch = _binary_chain(oracle)
    assert ch.mwg.dir_slots == [1, 2, 0] and ch.mwg.names[:2] == ["gmc_0", "gmc_1"]
    xi = np.random.default_rng(3).standard_normal((3, 8))
    x0 = ch.pos.copy()
    c = ch.step(0, np.inf, xi)
    assert c[1] == 1
    m = (ch.types == 0)[:, None].astype(float)
    d = xi.T * m / np.linalg.norm(xi.T * m)
    np.testing.assert_allclose(ch.dirx[0], d, rtol=0, atol=1e-15)
    np.testing.assert_allclose(ch.pos, x0 + 3 * 0.05 * d, rtol=0, atol=1e-14)
    assert np.array_equal(ch.pos[ch.types == 1], x0[ch.types == 1])  # bit for bit
    assert not ch.dir.any() and not ch.dirx[1].any()
    # the second call perturbs the stored slot-1 direction (masked again)
    xi2 = np.random.default_rng(4).standard_normal((3, 8))
    ch.step(0, np.inf, xi2)
    v = (d + 0.1 * xi2.T) * m
    np.testing.assert_allclose(ch.dirx[0], v / np.linalg.norm(v), rtol=0, atol=1e-15)


def test_refresh_direction_ignores_the_stored_direction(oracle):
    """refresh_direction draws a fresh direction on every call
    (galilean.py:160-168), so two calls with the same normals give the same
    direction whatever the first call left behind."""
    ch = _binary_chain(oracle)
    xi = np.random.default_rng(5).standard_normal((3, 8))
    ch.step(1, np.inf, xi)
    m = (ch.types == 1)[:, None].astype(float)
    d = xi.T * m / np.linalg.norm(xi.T * m)
    np.testing.assert_allclose(ch.dirx[1], d, rtol=0, atol=1e-15)
    ch.dirx[1] = np.random.default_rng(6).standard_normal((8, 3))  # any stored value
    ch.step(1, np.inf, xi)
    np.testing.assert_allclose(ch.dirx[1], d, rtol=0, atol=1e-15)


def test_scoped_reflection_uses_the_projected_force(oracle):
    """On a violated step the direction is mirrored off the force projected
    onto the moving atoms; frozen rows of the direction stay exactly zero."""
    ch = _binary_chain(oracle)
    ch.mwg = MWG([{"type": "gmc", "species": 0, "n_reflect": 2, "step_size": 1e-4}], CELL)
    ch.dirx = np.zeros((1, 8, 3))
    xi = np.random.default_rng(7).standard_normal((3, 8))
    m = (ch.types == 0)[:, None].astype(float)
    d = xi.T * m / np.linalg.norm(xi.T * m)
    x1 = ch.pos + 1e-4 * d
    _, H1, F1 = oracle.energy_forces(ch.pot, x1, ch.types, ch.cell)
    emax = H1 - 1e-12  # the first step violates
    ch.step(0, emax, xi)
    fm = F1 * m
    f = fm / np.linalg.norm(fm)
    d_ref = d - 2 * np.sum(f * d) * f
    x2 = x1 + 1e-4 * d_ref
    _, H2, _ = oracle.energy_forces(ch.pot, x2, ch.types, ch.cell)
    if H2 < emax:
        np.testing.assert_allclose(ch.pos, x2, atol=1e-13)
        np.testing.assert_allclose(ch.dirx[0], d_ref, atol=1e-12)
    else:
        np.testing.assert_allclose(ch.dirx[0], -d, atol=1e-14)
    assert not ch.dirx[0][ch.types == 1].any()
    assert np.array_equal(ch.pos[ch.types == 1], (lattice(2, 1.3) + 2.0)[ch.types == 1])

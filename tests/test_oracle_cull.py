"""Selection with n_cull > 1 in the oracle (``sampling/nested_sampling.py:188-212``).

Pinned against numpy's stable ``lexsort``, which restates the reference's
``jnp.lexsort((-noise, -potentials))`` directly, and against a sort of the
survivors on their permutation keys (``jax.random.permutation`` sorts on
random keys); energy, noise and key ties reach every tie rule.  The
permutation is also checked exhaustively on a small case: every ordering of
the keys yields a distinct ordered pair of clone sources.
"""

import itertools

import numpy as np
import pytest

from jaxrens_amd.backends import LJBackend
from jaxrens_amd.draws import make_tape
from jaxrens_amd.moves import MWG

CELL = {"min_volume_per_atom": 0.125, "max_volume_per_atom": 64.0, "flat_V_prior": True}
# a zero-length gmc step leaves every walked configuration where it was
STILL = MWG([{"type": "gmc", "n_reflect": 1, "step_size": 0.0}], CELL)


def _pop(oracle, R, K, N, rng, energies):
    L = 4.0
    pos = rng.random((R, K, N, 3)) * L
    cell = np.tile(np.eye(3) * L, (R, K, 1, 1)) * (1.0 + 0.01 * rng.random((R, K, 1, 1)))
    types = np.zeros((R, K, N), np.int32)
    pop = oracle.Population(pos, types, np.full((R, K), N, np.int32), cell,
                            np.array(energies, dtype=np.float64), np.zeros((R, K, N, 3)))
    return pop, oracle.Pot.from_table(LJBackend(1.0, 1.0, 2.5).table)


def _expected(E, noise, keys, C, n_extra):
    """Reference selection restated with numpy's stable sorts."""
    K = len(E)
    order = np.lexsort((-noise, -E))  # primary -E, then -noise, then index
    worst = order[:C]
    surv = order[C:]
    perm = surv[np.lexsort((surv, keys[:K - C]))]  # keys ascending, then index
    return worst, perm[:C], np.concatenate([worst, perm[C:C + n_extra]]), E[worst[-1]]


def _run(oracle, pop, pot, tape, C, n_extra):
    R = pop.pos.shape[0]
    return oracle.ns_iteration(pop, pot, np.zeros(R), np.zeros((R, 1)), np.zeros((R, 1)),
                               STILL.as_oracle_moves(), 1, n_extra, tape, do_re=False,
                               n_cull=C)


@pytest.mark.parametrize("C,n_extra", [(2, 0), (3, 4), (5, 1), (8, 0)])
def test_cull_matches_numpy_lexsort(oracle, C, n_extra):
    rng = np.random.default_rng(40 + C)
    R, K, N = 2, 16, 3
    E = np.round(rng.normal(size=(R, K)) * 2) / 2  # many exact energy ties
    E[0, [1, 5, 9]] = 7.0
    tape = make_tape(rng, 1, R, K, N, 1, n_extra, STILL.move_probs, ties=True, n_cull=C)
    tape.sel_u[0, 0, [1, 5]] = 0.5  # equal energy and equal noise: index rule
    tape.sel_u[0, 0, 9] = 0.25
    tape.extra_u[0, 1, :4] = 0.75   # equal permutation keys: walker index rule
    pop, pot = _pop(oracle, R, K, N, rng, E)
    before = pop.copy()
    out = _run(oracle, pop, pot, tape, C, n_extra)
    for r in range(R):
        worst, clone, walk, contour = _expected(E[r], tape.sel_u[0, r], tape.extra_u[0, r],
                                                C, n_extra)
        assert np.array_equal(out["worst"][r], worst), r
        assert np.array_equal(out["clone"][r], clone), r
        assert np.array_equal(out["walk"][r], walk), r
        assert out["emax"][r] == contour
        assert np.array_equal(out["dead_energy"][r], E[r, worst])
        vol = np.abs(np.linalg.det(before.cell[r].reshape(K, 3, 3)[worst]))
        np.testing.assert_allclose(out["dead_volume"][r], vol, rtol=1e-14)
        for w, c in zip(worst, clone):
            assert np.array_equal(pop.pos[r, w], before.pos[r, c])
            assert np.array_equal(pop.cell[r, w], before.cell[r, c])
    assert list(out["worst"][0, :2]) == [1, 5]


def test_cull_invariants_random(oracle):
    rng = np.random.default_rng(7)
    for trial in range(30):
        K = int(rng.integers(6, 40))
        C = int(rng.integers(2, K // 2 + 1))
        n_extra = int(rng.integers(0, K - 2 * C + 1))
        E = rng.normal(size=(1, K))
        if trial % 3 == 0:
            E = np.round(E)
        pop, pot = _pop(oracle, 1, K, 2, rng, E)
        tape = make_tape(rng, 1, 1, K, 2, 1, n_extra, STILL.move_probs, n_cull=C)
        out = _run(oracle, pop, pot, tape, C, n_extra)
        worst, clone, walk = out["worst"][0], out["clone"][0], out["walk"][0]
        dead = out["dead_energy"][0]
        assert np.all(np.diff(dead) <= 0)  # descending
        survivors = np.setdiff1d(np.arange(K), worst)
        assert np.all(E[0, survivors] <= out["emax"][0])
        assert out["emax"][0] == dead[-1]
        assert len(set(walk.tolist())) == C + n_extra
        assert set(clone.tolist()).isdisjoint(walk.tolist())
        assert len(set(clone.tolist())) == C
        assert np.array_equal(walk[:C], worst)


def test_cull_permutation_is_exhaustive(oracle):
    """K = 5, C = 2: the three survivors take keys in every order; the
    clone pairs are then every ordered pair of distinct survivors once."""
    rng = np.random.default_rng(3)
    E = np.array([[0.0, 4.0, 1.0, 3.0, 2.0]])  # culled: 1, 3; survivors 4, 2, 0
    pairs = []
    for keys in itertools.permutations([0.1, 0.2, 0.3]):
        pop, pot = _pop(oracle, 1, 5, 2, rng, E)
        tape = make_tape(rng, 1, 1, 5, 2, 1, 0, STILL.move_probs, n_cull=2)
        tape.extra_u[0, 0, :3] = keys
        out = _run(oracle, pop, pot, tape, 2, 0)
        assert list(out["worst"][0]) == [1, 3] and out["emax"][0] == 3.0
        pairs.append(tuple(out["clone"][0].tolist()))
        # survivor position p (order 4, 2, 0) takes key p: the smallest key wins
        assert pairs[-1][0] == [4, 2, 0][int(np.argmin(keys))]
    assert sorted(pairs) == sorted(itertools.permutations([0, 2, 4], 2))


def test_cull_needs_room_for_distinct_sources(oracle):
    rng = np.random.default_rng(1)
    pop, pot = _pop(oracle, 1, 8, 2, rng, rng.normal(size=(1, 8)))
    tape = make_tape(rng, 1, 1, 8, 2, 1, 1, STILL.move_probs, n_cull=4)
    with pytest.raises(RuntimeError):
        _run(oracle, pop, pot, tape, 4, 1)  # 2*4 + 1 > 8

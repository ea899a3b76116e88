"""Pins of the oracle's species swap (``sampling/moves/swap.py:119-262``):
one unlike pair exchanged, composition and positions kept, the pair uniform
over all unlike pairs in scope (chi-square), scope respected, no-pair and
energy rejections leave the walker."""

import ctypes as C

import numpy as np
import pytest

from jaxrens_amd.backends import LJBackend
from jaxrens_amd.moves import MWG

CELL = {"min_volume_per_atom": 0.1, "max_volume_per_atom": 100.0, "flat_V_prior": True}


def swap_once(oracle, lj, pos, types, cell, u, emax=np.inf, species=None):
    O = oracle
    N = len(pos)
    mw = MWG([{"type": "species_swap", **({"species": species} if species else {})}], CELL)
    pot = O.Pot.from_table(lj.table)
    p, t, c = O.f64(pos).copy(), O.i32(types).copy(), O.f64(cell).reshape(9).copy()
    S = lj.n_species
    mu = O.f64(np.zeros(S))
    E = np.array([O.energy_forces(pot, p, t, c, 0.0, mu)[1]])
    ctr = np.zeros((1, 7), np.int32)
    O.lib().or_walk_chain(N, N, O._d(p), O._i(t), O._d(c), O._d(E), O._d(np.zeros((N, 3))),
                          None, 0, C.byref(pot.s), 0.0, O._d(mu), emax, O._d(O.f64([0.1])),
                          O.moves_array(mw.as_oracle_moves()), 1, O._i(np.zeros(1, np.int32)),
                          1, O._d(np.zeros((3, N))), O._d(O.f64(list(u) + [0.5])), O._i(ctr))
    return p, t, E[0], ctr[0], pot


def _config(rng, types):
    N = len(types)
    L = (N / 0.5) ** (1 / 3)
    return rng.random((N, 3)) * L, np.eye(3) * L


LJ3 = LJBackend(np.array([1.0, 0.8, 0.6], np.float32), np.array([1.0, 0.9, 1.1], np.float32), 2.5)


def test_swap_exchanges_one_unlike_pair(oracle):
    rng = np.random.default_rng(0)
    types = np.array([0, 1, 1, 2, 0, 2, 2, 1, 0, 1])
    pos, cell = _config(rng, types)
    p, t, E, ctr, pot = swap_once(oracle, LJ3, pos, types, cell, rng.random(3))
    assert ctr[1] == 1 and ctr[5] == 1
    changed = np.flatnonzero(t != types)
    assert len(changed) == 2 and t[changed[0]] == types[changed[1]] \
        and t[changed[1]] == types[changed[0]]
    assert np.array_equal(np.bincount(t, minlength=3), np.bincount(types, minlength=3))
    assert np.array_equal(p, pos)
    assert E == oracle.energy_forces(pot, p, t, cell)[1]


@pytest.mark.parametrize("species", [None, [0, 2]])
def test_swap_pair_is_uniform_over_unlike_pairs(oracle, species):
    rng = np.random.default_rng(1)
    types = np.array([0, 0, 1, 1, 1, 2, 2, 2, 2])
    pos, cell = _config(rng, types)
    scope = set(range(3)) if species is None else set(species)
    pairs = [(i, j) for i in range(9) for j in range(i + 1, 9)
             if types[i] != types[j] and types[i] in scope and types[j] in scope]
    index = {p: n for n, p in enumerate(pairs)}
    hits = np.zeros(len(pairs))
    draws = 60 * len(pairs)
    for _ in range(draws):
        _, t, _, ctr, _ = swap_once(oracle, LJ3, pos, types, cell, rng.random(3),
                                    species=species)
        ch = tuple(np.flatnonzero(t != types))
        hits[index[ch]] += 1
    expected = draws / len(pairs)
    chi2 = float(np.sum((hits - expected) ** 2 / expected))
    # 99.9% point of chi-square with 25 (all) or 7 (scoped) degrees of freedom
    assert chi2 < {26: 52.6, 8: 24.3}[len(pairs)], (chi2, hits)


def test_swap_without_unlike_pair_or_above_emax_rejects(oracle):
    rng = np.random.default_rng(2)
    same = np.array([1, 1, 1, 1, 1])
    pos, cell = _config(rng, same)
    p, t, _, ctr, _ = swap_once(oracle, LJ3, pos, same, cell, rng.random(3))
    assert ctr[2] == 1 and ctr[5] == 1 and np.array_equal(t, same)
    types = np.array([0, 1, 2, 0, 1])
    p, t, _, ctr, _ = swap_once(oracle, LJ3, pos, types, cell, rng.random(3), emax=-np.inf)
    assert ctr[2] == 1 and np.array_equal(t, types) and np.array_equal(p, pos)
    # a scope with one species present holds no unlike pair
    p, t, _, ctr, _ = swap_once(oracle, LJ3, pos, np.array([0, 0, 1, 1, 0]), cell,
                                rng.random(3), species=[0, 2])
    assert ctr[2] == 1 and np.array_equal(t, [0, 0, 1, 1, 0])

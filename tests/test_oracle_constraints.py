"""Minimum-distance constraint of the oracle (``constraints/min_distance.py``,
gated as in ``sampling/mwg.py:98-141``) and the host-side floor tables
(``cli/schema/constraints.py:104-138``).

The predicate is pinned against the reference's own distance routine
restated in numpy (``backends/geometry.py:18-70``: fractional coordinates,
``round``, back to Cartesian) and against a brute-force search over the 27
nearest images, with thresholds just either side of the true minimum.  The
gate is pinned on constructed moves: a volume compression and a species
change that break the floor are undone and counted in no reject bucket.
"""

import itertools

import numpy as np
import pytest

from jaxrens_amd.backends import LJBackend
from jaxrens_amd.constraints import constraint_tables, min_distance_floors
from jaxrens_amd.moves import MWG


def _ref_min_ratio(pos, types, cell, floors):
    """min over pairs of r_ij / floor_ij with the reference's distances."""
    frac = pos @ np.linalg.inv(cell)
    df = frac[:, None, :] - frac[None, :, :]
    df = df - np.round(df)
    r = np.sqrt(((df @ cell) ** 2).sum(-1))
    th = floors[types[:, None], types[None, :]]
    n = len(pos)
    mask = ~np.eye(n, dtype=bool) & (th > 0)
    return (r[mask] / th[mask]).min(), r[~np.eye(n, dtype=bool)].min()


def _brute_min(pos, cell):
    best = np.inf
    shifts = np.array(list(itertools.product((-1, 0, 1), repeat=3))) @ cell
    for i in range(len(pos)):
        for j in range(len(pos)):
            if i != j:
                best = min(best, np.sqrt(((pos[j] - pos[i] + shifts) ** 2).sum(-1)).min())
    return best


def _ok(oracle, pos, types, cell, floors):
    import ctypes as C
    from oracle.oracle import _d, _i, f64, i32
    L = oracle.lib()
    L.or_min_distance_ok.restype = C.c_int
    p, t, c, f = f64(pos), i32(types), f64(np.asarray(cell).reshape(9)), f64(floors)
    return bool(L.or_min_distance_ok(len(p), _d(p), _i(t), _d(c), _d(f), floors.shape[0]))


def test_floor_tables_as_the_reference_builds_them():
    m, dep = min_distance_floors(0.8, 2)
    assert not dep and np.all(m == np.float32(0.8)) and m.dtype == np.float64
    sm = {0: "Si", 1: "O"}
    m, dep = min_distance_floors({"default": 1.0, "Si-Si": 2.0, "Si-O": 1.6}, 2, sm)
    assert dep and m[0, 0] == np.float32(2.0) and m[0, 1] == m[1, 0] == np.float32(1.6)
    assert m[1, 1] == 1.0
    with pytest.raises(ValueError):
        min_distance_floors({"Si-C": 1.0}, 2, sm)
    with pytest.raises(ValueError):
        min_distance_floors(-0.1, 2)
    pos, typ = constraint_tables([{"type": "minimum_distance", "d_min": 0.5},
                                  {"type": "minimum_distance", "d_min": {"default": 0.2,
                                                                         "O-O": 0.9}}], 2, sm)
    assert np.array_equal(pos, [[0.5, 0.5], [0.5, np.float32(0.9)]])
    assert np.array_equal(typ, [[np.float32(0.2), np.float32(0.2)],
                                [np.float32(0.2), np.float32(0.9)]])


@pytest.mark.parametrize("triclinic", [False, True])
def test_predicate_matches_reference_distances(oracle, triclinic):
    rng = np.random.default_rng(11 + triclinic)
    for trial in range(40):
        N = int(rng.integers(2, 20))
        L = 3.0 + 2 * rng.random()
        cell = np.eye(3) * L
        if triclinic:
            cell = cell + rng.uniform(-0.2, 0.2, (3, 3)) * L
        pos = (rng.random((N, 3)) - 0.3) @ cell * 1.4  # some atoms outside the cell
        types = rng.integers(0, 2, N)
        floors = np.array([[1.0, 0.7], [0.7, 0.5]]) * rng.uniform(0.2, 1.5)
        ratio, rmin = _ref_min_ratio(pos, types, cell, floors)
        if not triclinic:
            assert rmin == pytest.approx(_brute_min(pos, cell), rel=1e-12)
        if abs(ratio - 1.0) < 1e-9:
            continue
        assert _ok(oracle, pos, types, cell, floors) == (ratio > 1.0), trial
        # thresholds just either side of the closest pair
        f_on = floors * ratio
        assert not _ok(oracle, pos, types, cell, f_on * (1 + 1e-9))
        assert _ok(oracle, pos, types, cell, f_on * (1 - 1e-9))
    assert _ok(oracle, pos, types, cell, np.zeros((2, 2)))  # zero floors disable


def _pair_walker(oracle, types=(0, 0), S=1, d=1.0, L=4.0):
    from oracle.oracle import Population
    pos = np.zeros((1, 1, 2, 3))
    pos[0, 0, 1, 0] = d
    cell = (np.eye(3) * L)[None, None]
    return Population(pos, np.array(types, np.int32)[None, None], np.full((1, 1), 2, np.int32),
                      cell, np.zeros((1, 1)), np.zeros((1, 1, 2, 3)))


def _trial(oracle, pop, pot, moves, move_idx, xi0=0.0, u=(0.0, 0.0, 0.0, 0.0), S=1):
    normals = np.zeros((1, 3, 2))
    normals[0, 0, 0] = xi0
    uniforms = np.array([u], dtype=np.float64)
    mw = MWG(moves, {"min_volume_per_atom": 0.01, "max_volume_per_atom": 1e4,
                     "flat_V_prior": True})
    return oracle.trial_moves(pop, pot, np.zeros(1), np.zeros((1, S)),
                              np.ones((1, len(moves))), mw.as_oracle_moves(), move_idx,
                              np.ones(1), np.full(1, np.inf), np.zeros((1, 1), np.int32),
                              normals, uniforms)


def test_gate_undoes_a_compression_below_the_floor(oracle):
    """Volume step V' = V + s N xi with V = 64, N = 2: xi = -15.616 shrinks the
    box by 0.8, bringing the pair from 1.0 to 0.8."""
    lj = LJBackend(1.0, 1.0, 2.5)
    xi0 = (64 * 0.512 - 64) / 2
    vol = [{"type": "volume", "step_size": 1.0, "weight": 1}]
    for floor, accepted, reason in [(None, True, 0), (0.9, False, 4), (0.7, True, 0)]:
        pot = oracle.Pot.from_table(lj.table)
        if floor is not None:
            pot.with_min_distance(np.full((1, 1), floor))
        acc, rs = _trial(oracle, _pair_walker(oracle), pot, vol, 0, xi0=xi0)
        assert bool(acc[0, 0]) == accepted and rs[0, 0] == reason, floor


def test_gate_on_species_moves_only_when_type_dependent(oracle):
    """Two species-0 atoms 1.0 apart; morphing one to species 1 makes the
    pair's floor 1.2."""
    lj = LJBackend(pair_epsilon=np.ones((2, 2)), pair_sigma=np.ones((2, 2)), cutoff=2.5)
    floors = np.array([[0.5, 1.2], [1.2, 0.5]])
    morph = [{"type": "alchemical_morph", "n_species": 2, "weight": 1}]
    for types_table, reason in [(None, 0), (floors, 4)]:
        pot = oracle.Pot.from_table(lj.table).with_min_distance(floors, types_table)
        acc, rs = _trial(oracle, _pair_walker(oracle, S=2), pot, morph, 0,
                         u=(0.7, 0.0, 0.0, 0.0), S=2)
        assert rs[0, 0] == reason and bool(acc[0, 0]) == (reason == 0)


def test_gate_restores_the_walker_and_counts_no_bucket(oracle):
    """Through a whole NS iteration: the compression is undone, the counters
    show one proposal, one evaluation and no reject bucket."""
    from jaxrens_amd.draws import make_tape
    lj = LJBackend(1.0, 1.0, 2.5)
    pot = oracle.Pot.from_table(lj.table).with_min_distance(np.full((1, 1), 0.9))
    pop = _pair_walker(oracle)
    # two walkers so that one is culled and the other walked
    from oracle.oracle import Population
    pop = Population(np.repeat(pop.pos, 2, 1), np.repeat(pop.types, 2, 1),
                     np.repeat(pop.n_atoms, 2, 1), np.repeat(pop.cell.reshape(1, 1, 3, 3), 2, 1),
                     np.array([[1.0e6, -1.0]]), np.zeros((1, 2, 2, 3)))
    mw = MWG([{"type": "volume", "step_size": 1.0, "weight": 1}],
             {"min_volume_per_atom": 0.01, "max_volume_per_atom": 1e4, "flat_V_prior": True})
    tape = make_tape(np.random.default_rng(0), 1, 1, 2, 2, 1, 0, mw.move_probs)
    tape.normals[:] = 0.0
    tape.normals[0, 0, 0, 0, 0, 0] = (64 * 0.512 - 64) / 2
    tape.uniforms[:] = 0.0
    before = pop.copy()
    out = oracle.ns_iteration(pop, pot, np.zeros(1), np.zeros((1, 1)), np.ones((1, 1)),
                              mw.as_oracle_moves(), 1, 0, tape, do_re=False)
    w = int(out["worst"][0])
    assert w == 0
    assert np.array_equal(out["counters"][0, 0], [1, 0, 0, 0, 0, 1, 0])
    assert np.array_equal(pop.pos[0, 0], before.pos[0, 1])  # the clone, unmoved
    assert np.array_equal(pop.cell[0, 0], before.cell[0, 1])
    assert pop.energy[0, 0] == before.energy[0, 1]


def test_grid_positions_respect_the_spacing():
    """``init/positions.py:34-100``: in an orthogonal cell every pair of grid
    sites is at least the spacing apart (minimum image included), sites are
    distinct cell-centred grid points, and a too-coarse grid is an error."""
    from jaxrens_amd.engine import grid_positions_in_cell
    rng = np.random.default_rng(3)
    for L, n, d in [(4.0, 20, 1.0), (6.0, 64, 1.3), (3.0, 1, 2.9)]:
        cell = np.diag([L, L * 1.1, L * 0.9])
        pos = grid_positions_in_cell(rng, cell, n, d)
        assert pos.shape == (n, 3)
        frac = pos @ np.linalg.inv(cell)
        assert np.all((frac > 0) & (frac < 1))
        if n > 1:
            df = frac[:, None] - frac[None]
            df -= np.round(df)
            r = np.sqrt(((df @ cell) ** 2).sum(-1))[~np.eye(n, dtype=bool)]
            assert r.min() >= d - 1e-6
        na = [max(1, int(np.linalg.norm(cell[i]) / d)) for i in range(3)]
        idx = np.rint(frac * na - 0.5)
        assert np.allclose(frac * na - 0.5, idx, atol=1e-6)  # cell-centred sites
        assert len({tuple(v) for v in idx.astype(int)}) == n
    with pytest.raises(ValueError, match="grid too coarse"):
        grid_positions_in_cell(rng, np.eye(3) * 2.0, 9, 1.0)

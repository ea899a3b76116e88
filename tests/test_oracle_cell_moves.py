"""Pins the oracle's supercell image sums and its volume-preserving cell moves
(``backends/lj.py:106-113,152-197``, ``sampling/moves/shear.py:19-120``,
``stretch.py:18-97``, ``_common.py:16-93``, ``utils/cell.py:80-152``).

Independent checks: the reference's image formula restated directly in NumPy
(all offsets, all ordered pairs, the central self pair excluded, halved), a
single-atom lattice sum, the full periodic sum when the images cover the
cutoff sphere, finite differences, and the invariants of the cell moves
(volume kept, fractional coordinates kept, step lengths)."""

import ctypes as C
import math

import numpy as np
import pytest

from jaxrens_amd.backends import LJBackend
from jaxrens_amd.moves import MWG

CELL = {"min_volume_per_atom": 0.125, "max_volume_per_atom": 64.0,
        "min_aspect_ratio": 0.5, "flat_V_prior": True}


def reference_image_energy(pos, cell, eps, sig2, cutoff, sc, types=None):
    """``lj.py:152-197`` as written: (K, N, N) image tensor over the centred
    offsets of ``_graph_neighbors.py:20-33``; ``sig2`` is the squared sigma
    table."""
    pos = np.asarray(pos, dtype=np.float64)
    n = len(pos)
    types = np.zeros(n, int) if types is None else np.asarray(types)
    eps = np.atleast_2d(eps)
    sig2 = np.atleast_2d(sig2)
    offs = np.array([(a, b, c)
                     for a in range(-(sc[0] // 2), sc[0] // 2 + 1)
                     for b in range(-(sc[1] // 2), sc[1] // 2 + 1)
                     for c in range(-(sc[2] // 2), sc[2] // 2 + 1)], dtype=np.float64)
    inv = np.linalg.inv(cell)
    dr = pos[None, :, :] - pos[:, None, :]
    dr = dr - np.round(dr @ inv) @ cell
    d = dr[None] + (offs @ cell)[:, None, None, :]
    r2 = np.sum(d * d, axis=-1)
    excl = np.all(offs == 0, axis=1)[:, None, None] & np.eye(n, dtype=bool)[None]
    r2s = np.where(excl, 1.0, r2)
    e_ij = eps[types[:, None], types[None, :]][None]
    s2_ij = sig2[types[:, None], types[None, :]][None]
    s6 = (s2_ij / r2s) ** 3
    pe = 4.0 * e_ij * (s6 * s6 - s6)
    if cutoff is not None:
        pe = np.where(r2 < cutoff ** 2, pe, 0.0)
    return 0.5 * float(np.sum(np.where(excl, 0.0, pe)))


def periodic_energy(pos, cell, cutoff, nmax=4):
    """Every periodic image within the cutoff (one species, eps = sig = 1)."""
    n = len(pos)
    rng = range(-nmax, nmax + 1)
    T = np.array([(a, b, c) for a in rng for b in rng for c in rng], float) @ cell
    e = 0.0
    for i in range(n):
        for j in range(n):
            d = pos[j] - pos[i] + T
            r2 = np.sum(d * d, axis=1)
            if i == j:
                r2 = r2[r2 > 0]
            r2 = r2[r2 < cutoff ** 2]
            s6 = (1.0 / r2) ** 3
            e += 0.5 * float(np.sum(4.0 * (s6 * s6 - s6)))
    return e


def spread(rng, n, cell, dmin=0.8):
    inv = np.linalg.inv(cell)
    pts = []
    while len(pts) < n:
        x = rng.random(3) @ cell
        ok = True
        for p in pts:
            d = x - p
            d = d - np.round(d @ inv) @ cell
            if d @ d < dmin ** 2:
                ok = False
                break
        if ok:
            pts.append(x)
    return np.array(pts)


def test_exp_matches_libm(oracle):
    rng = np.random.default_rng(0)
    xs = np.concatenate([rng.uniform(-1, 1, 2000), rng.uniform(-700, 700, 2000),
                         [0.0, -0.0, 1e-300, -1e-300, 0.5 * math.log(2), 1.0, -1.0,
                          709.0, -708.0]])
    for x in xs:
        got, want = oracle.exp(x), math.exp(x)
        assert abs(got - want) <= 2 * math.ulp(want), (x, got, want)
    assert oracle.exp(0.0) == 1.0
    assert oracle.exp(800.0) == math.inf and oracle.exp(-800.0) == 0.0
    assert math.isnan(oracle.exp(math.nan))
    # subnormal results: within one subnormal step
    for x in (-740.0, -744.0, -709.5):
        assert abs(oracle.exp(x) - math.exp(x)) <= 2 * math.ulp(math.exp(x))


@pytest.mark.parametrize("sc,N,tri,species", [
    ((2, 2, 2), 5, False, 1), ((4, 4, 4), 4, False, 1), ((3, 2, 1), 6, True, 1),
    ((4, 4, 4), 1, True, 1), ((2, 2, 2), 2, True, 1), ((5, 3, 4), 7, True, 2)])
def test_images_match_reference_formula(oracle, sc, N, tri, species):
    rng = np.random.default_rng(N + 10 * species)
    L = 2.2
    cell = np.eye(3) * L
    if tri:
        cell = cell + rng.uniform(-0.25, 0.25, (3, 3))
    pos = spread(rng, N, cell, 0.9) + rng.integers(-1, 2, (N, 3)) @ cell
    if species == 1:
        lj = LJBackend(1.0, 1.0, 2.5, supercell_trafo=sc)
        eps, sig2, ty = 1.0, 1.0, np.zeros(N, int)
    else:
        e1, s1 = np.array([1.0, 0.6], np.float32), np.array([1.0, 0.8], np.float32)
        lj = LJBackend(e1, s1, 2.5, supercell_trafo=sc)
        eps = np.sqrt(e1[:, None] * e1[None, :]).astype(np.float64)
        sig2 = (np.float32(0.5) * (s1[:, None] + s1[None, :])) ** 2  # squared in float32
        ty = rng.integers(0, 2, N)
    assert lj.table.images == tuple(s // 2 for s in sc)
    pot = oracle.Pot.from_table(lj.table)
    U, _, F = oracle.energy_forces(pot, pos, ty, cell)
    want = reference_image_energy(pos, cell, eps, sig2, 2.5, sc, ty)
    assert U == pytest.approx(want, rel=1e-12, abs=1e-12)
    # forces: -dU/dx by central differences
    h = 1e-6
    for a in range(min(N, 3)):
        for k in range(3):
            p = pos.copy()
            p[a, k] += h
            up = oracle.energy_forces(pot, p, ty, cell, forces=False)[0]
            p[a, k] -= 2 * h
            um = oracle.energy_forces(pot, p, ty, cell, forces=False)[0]
            assert F[a, k] == pytest.approx(-(up - um) / (2 * h), rel=1e-5, abs=1e-5)
    if N > 1:
        assert np.abs(F.sum(axis=0)).max() < 1e-9


def test_single_atom_lattice_sum(oracle):
    """One atom: U = 1/2 sum over lattice vectors T != 0 with |T| < rc."""
    L = 1.2
    lj = LJBackend(1.0, 1.0, 2.5, supercell_trafo=(5, 5, 5))
    U, _, F = oracle.energy_forces(oracle.Pot.from_table(lj.table), np.zeros((1, 3)),
                                   [0], np.eye(3) * L)
    want = 0.0
    for a in range(-2, 3):
        for b in range(-2, 3):
            for c in range(-2, 3):
                r2 = (a * a + b * b + c * c) * L * L
                if 0 < r2 < 2.5 ** 2:
                    want += 0.5 * 4.0 * (r2 ** -6 - r2 ** -3)
    assert U == pytest.approx(want, rel=1e-13)
    assert np.all(F == 0.0)


def test_images_reach_the_periodic_sum(oracle):
    """Cell edge 2 < 2 rc: minimum image misses neighbours, images (5,5,5)
    cover the cutoff sphere and give the full periodic energy."""
    rng = np.random.default_rng(3)
    cell = np.eye(3) * 2.0
    pos = spread(rng, 4, cell, 0.9)
    full = periodic_energy(pos, cell, 2.5)
    plain = oracle.energy_forces(oracle.Pot.from_table(LJBackend(1.0, 1.0, 2.5).table),
                                 pos, np.zeros(4), cell)[0]
    img = oracle.energy_forces(oracle.Pot.from_table(
        LJBackend(1.0, 1.0, 2.5, supercell_trafo=(5, 5, 5)).table), pos, np.zeros(4), cell)[0]
    assert img == pytest.approx(full, rel=1e-12)
    assert abs(plain - full) > 1e-3


def test_image_translation_invariance(oracle):
    rng = np.random.default_rng(4)
    cell = np.array([[2.0, 0.0, 0.0], [0.5, 2.1, 0.0], [-0.3, 0.4, 1.9]])
    pos = spread(rng, 6, cell, 0.9)
    pot = oracle.Pot.from_table(LJBackend(1.0, 1.0, 2.5, supercell_trafo=(4, 4, 4)).table)
    U0 = oracle.energy_forces(pot, pos, np.zeros(6), cell)[0]
    shift = rng.integers(-3, 4, (6, 3)) @ cell + rng.normal(size=3)
    U1 = oracle.energy_forces(pot, pos + shift, np.zeros(6), cell)[0]
    assert U1 == pytest.approx(U0, rel=1e-11)


class CellChain:
    """One walker driven through or_walk_chain with one cell move."""

    def __init__(self, oracle, pos, cell, move, sc=(1, 1, 1), pressure=0.0):
        self.O = oracle
        self.lj = LJBackend(1.0, 1.0, 2.5, supercell_trafo=sc)
        self.pot = oracle.Pot.from_table(self.lj.table)
        self.N = len(pos)
        self.pos = oracle.f64(pos).copy()
        self.cell = oracle.f64(cell).reshape(9).copy()
        self.types = oracle.i32(np.zeros(self.N))
        self.dir = np.zeros((self.N, 3))
        self.pressure = pressure
        self.mu = oracle.f64(np.zeros(1))
        self.E = np.array([oracle.energy_forces(self.pot, self.pos, self.types, self.cell,
                                                pressure, self.mu)[1]])
        self.mwg = MWG([move], CELL)

    def step(self, emax, xi0, xi1, u0, step):
        O = self.O
        xi = np.zeros((3, self.N))
        xi[0, 0], xi[1, 0] = xi0, xi1
        ctr = np.zeros((1, 7), np.int32)
        O.lib().or_walk_chain(self.N, self.N, O._d(self.pos), O._i(self.types),
                              O._d(self.cell), O._d(self.E), O._d(self.dir), None, 0,
                              C.byref(self.pot.s), self.pressure, O._d(self.mu), emax,
                              O._d(O.f64([step])), O.moves_array(self.mwg.as_oracle_moves()),
                              1, O._i(np.zeros(1, np.int32)), 1, O._d(O.f64(xi)),
                              O._d(O.f64([u0, 0.5, 0.5, 0.5])), O._i(ctr))
        return ctr[0]


@pytest.mark.parametrize("u0,idx", [(0.1, 0), (0.5, 1), (0.9, 2)])
def test_shear_keeps_volume_and_fractional_coordinates(oracle, u0, idx):
    rng = np.random.default_rng(idx)
    c0 = np.eye(3) * 3.0 + rng.uniform(-0.3, 0.3, (3, 3))
    pos = spread(rng, 6, c0, 0.9)
    ch = CellChain(oracle, pos, c0, {"type": "shear", "step_size": 0.2})
    ctr = ch.step(np.inf, 0.7, -1.3, u0, 0.2)
    assert list(ctr[:3]) == [1, 1, 0] and ctr[5] == 1 and ctr[6] == 0
    c1 = ch.cell.reshape(3, 3)
    assert abs(np.linalg.det(c1)) == pytest.approx(abs(np.linalg.det(c0)), rel=1e-13)
    others = [i for i in range(3) if i != idx]
    assert np.array_equal(c1[others], c0[others])
    delta = c1[idx] - c0[idx]
    assert np.dot(delta, delta) == pytest.approx(0.2 ** 2 * (0.7 ** 2 + 1.3 ** 2), rel=1e-13)
    normal = np.cross(c0[others[0]], c0[others[1]])
    assert abs(np.dot(delta, normal)) < 1e-13 * np.linalg.norm(normal)
    # the component along the first other vector is step * xi0
    e1 = c0[others[0]] / np.linalg.norm(c0[others[0]])
    assert np.dot(delta, e1) == pytest.approx(0.2 * 0.7, rel=1e-12)
    np.testing.assert_allclose(ch.pos.reshape(-1, 3) @ np.linalg.inv(c1),
                               pos @ np.linalg.inv(c0), atol=1e-13)
    assert ch.E[0] == oracle.energy_forces(ch.pot, ch.pos, ch.types, ch.cell)[1]


@pytest.mark.parametrize("u0,axes", [(0.1, (0, 1)), (0.5, (0, 2)), (0.9, (1, 2))])
def test_stretch_scales_two_axes(oracle, u0, axes):
    rng = np.random.default_rng(5)
    c0 = np.eye(3) * 3.0 + rng.uniform(-0.3, 0.3, (3, 3))
    pos = spread(rng, 6, c0, 0.9)
    ch = CellChain(oracle, pos, c0, {"type": "stretch", "step_size": 0.1})
    ch.step(np.inf, 0.8, 0.0, u0, 0.1)
    d = np.ones(3)
    d[axes[0]], d[axes[1]] = math.exp(0.08), math.exp(-0.08)
    np.testing.assert_allclose(ch.cell.reshape(3, 3), c0 * d[None, :], rtol=4e-16)
    np.testing.assert_allclose(ch.pos.reshape(-1, 3), pos * d[None, :], rtol=4e-16)
    assert abs(np.linalg.det(ch.cell.reshape(3, 3))) == pytest.approx(
        abs(np.linalg.det(c0)), rel=1e-14)


def test_cell_move_reject_reasons(oracle):
    """Energy (1) before cell (2); a rejected proposal leaves the walker."""
    rng = np.random.default_rng(6)
    c0 = np.eye(3) * 3.0
    pos = spread(rng, 6, c0, 0.9)
    for move, xi in (({"type": "shear", "step_size": 1.0}, 20.0),
                     ({"type": "stretch", "step_size": 1.0}, 3.0)):
        ch = CellChain(oracle, pos, c0, move)
        p0, e0 = ch.pos.copy(), ch.E.copy()
        # a large deformation fails the aspect guard (0.5): cell reject ...
        ctr = ch.step(np.inf, xi, xi, 0.2, 1.0)
        assert ctr[3] == 1
        # ... unless the energy also fails, which takes precedence
        ctr = ch.step(-np.inf, xi, xi, 0.2, 1.0)
        assert ctr[2] == 1
        assert np.array_equal(ch.pos, p0) and np.array_equal(ch.E, e0)
        assert np.array_equal(ch.cell, c0.reshape(9))


def test_shear_with_images_on_a_small_cell(oracle):
    """The tutorial's NPT setting: small cells with (4,4,4) images; an
    accepted shear stores the image energy of the new cell."""
    rng = np.random.default_rng(7)
    c0 = np.eye(3) * 2.1
    pos = spread(rng, 8, c0, 0.9)
    ch = CellChain(oracle, pos, c0, {"type": "shear", "step_size": 0.1}, sc=(4, 4, 4),
                   pressure=0.1)
    assert ch.step(np.inf, 0.3, 0.2, 0.4, 0.1)[1] == 1
    want = reference_image_energy(ch.pos.reshape(-1, 3), ch.cell.reshape(3, 3), 1.0, 1.0,
                                  2.5, (4, 4, 4))
    V = abs(np.linalg.det(ch.cell.reshape(3, 3)))
    assert ch.E[0] == pytest.approx(want + 0.1 * V, rel=1e-12)

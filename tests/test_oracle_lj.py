"""Pins the oracle's LJ energy / forces / ensemble term.

Independent checks: two-atom closed form, a brute-force periodic loop
(minimum image by fractional rounding, ``backends/lj.py:141-154``), finite
differences, lattice-translation invariance, the reference's committed LJ-8
NPT dead frames (``tests/_assets/data/lj8_npt``) and hand-computed
ensemble terms (``backends/ensemble.py:34-59``).
"""

import json
from pathlib import Path

import numpy as np
import pytest

from jaxrens_amd.backends import LJBackend

GOLD = Path(__file__).parent / "golden"


def brute_energy(pos, cell, eps, sig, cutoff, types=None):
    """Plain double loop over ordered pairs with MIC, halved."""
    pos = np.asarray(pos, dtype=np.float64)
    n = len(pos)
    types = np.zeros(n, int) if types is None else np.asarray(types)
    eps = np.atleast_2d(eps)
    sig = np.atleast_2d(sig)
    cell = np.asarray(cell, dtype=np.float64)
    det = np.linalg.det(cell)
    if det == 0.0:
        cell = 1e10 * np.eye(3)
    inv = np.linalg.inv(cell)
    e = 0.0
    for i in range(n):
        for j in range(n):
            if i == j:
                continue
            dr = pos[j] - pos[i]
            dr = dr - np.round(dr @ inv) @ cell
            r2 = float(dr @ dr)
            if cutoff is not None and r2 >= cutoff ** 2:
                continue
            a, b = types[i], types[j]
            s6 = (sig[a, b] ** 2 / r2) ** 3
            e += 0.5 * 4.0 * eps[a, b] * (s6 * s6 - s6)
    return e


def spread_config(rng, n, cell, min_dist=0.8):
    """Random positions in the cell with no pair closer than min_dist."""
    cell = np.asarray(cell, dtype=np.float64)
    pts = []
    while len(pts) < n:
        f = rng.random(3)
        x = f @ cell
        ok = True
        inv = np.linalg.inv(cell)
        for p in pts:
            d = x - p
            d = d - np.round(d @ inv) @ cell
            if d @ d < min_dist ** 2:
                ok = False
                break
        if ok:
            pts.append(x)
    return np.array(pts)


@pytest.fixture(scope="module")
def lj():
    return LJBackend(1.0, 1.0, 2.5)


def test_two_atom_closed_form(oracle, lj):
    pot = oracle.Pot.from_table(lj.table)
    for r in [0.9, 1.0, 2 ** (1 / 6), 1.2, 2.0, 2.49]:
        pos = np.array([[1.0, 2.0, 3.0], [1.0 + r, 2.0, 3.0]])
        U, H, F = oracle.energy_forces(pot, pos, [0, 0], np.eye(3) * 20)
        ex = 4.0 * (r ** -12 - r ** -6)
        assert U == pytest.approx(ex, rel=1e-14, abs=1e-14)
        f = 24.0 * (2 * r ** -12 - r ** -6) / r
        assert F[1, 0] == pytest.approx(f, rel=1e-13, abs=1e-13)
        assert F[0, 0] == pytest.approx(-f, rel=1e-13, abs=1e-13)
        assert np.all(F[:, 1:] == 0)


def test_cutoff_is_strict_and_unshifted(oracle, lj):
    pot = oracle.Pot.from_table(lj.table)
    for r, expect_zero in [(2.5, True), (2.5000001, True), (2.4999999, False)]:
        pos = np.array([[0.0, 0, 0], [r, 0, 0]])
        U, _, F = oracle.energy_forces(pot, pos, [0, 0], np.eye(3) * 20)
        if expect_zero:
            assert U == 0.0 and np.all(F == 0)
        else:
            assert U == pytest.approx(4 * (r ** -12 - r ** -6), rel=1e-12)


@pytest.mark.parametrize("n", [1, 2, 3, 5, 8, 13, 16, 31, 64, 65, 96, 130])
def test_brute_force_cubic(oracle, lj, n):
    rng = np.random.default_rng(100 + n)
    L = (n / 0.4) ** (1 / 3) + 1.0
    cell = np.eye(3) * L
    pos = spread_config(rng, n, cell) + rng.integers(-3, 4, size=(n, 3)) * L
    pot = oracle.Pot.from_table(lj.table)
    U, H, F = oracle.energy_forces(pot, pos, np.zeros(n, int), cell)
    ref = brute_energy(pos, cell, 1.0, 1.0, 2.5)
    assert U == pytest.approx(ref, rel=1e-12, abs=1e-12)
    assert np.abs(F.sum(0)).max() < 1e-9 * max(1.0, np.abs(F).max())


def test_brute_force_sheared_cell(oracle, lj):
    rng = np.random.default_rng(7)
    cell = np.array([[5.0, 0.5, 0.0], [0.6, 5.0, 0.3], [0.2, 0.4, 5.0]])
    pos = spread_config(rng, 8, cell) + rng.uniform(-1, 1, size=(8, 3))
    pot = oracle.Pot.from_table(lj.table)
    U, _, _ = oracle.energy_forces(pot, pos, np.zeros(8, int), cell)
    assert U == pytest.approx(brute_energy(pos, cell, 1.0, 1.0, 2.5), rel=1e-12)


def test_lattice_translation_invariance(oracle, lj):
    rng = np.random.default_rng(3)
    cell = np.eye(3) * 6.0034575
    pos = spread_config(rng, 8, cell)
    shifted = pos + rng.integers(-8, 9, size=(8, 3)) @ cell
    pot = oracle.Pot.from_table(lj.table)
    U0, _, F0 = oracle.energy_forces(pot, pos, np.zeros(8, int), cell)
    U1, _, F1 = oracle.energy_forces(pot, shifted, np.zeros(8, int), cell)
    assert U1 == pytest.approx(U0, rel=1e-12)
    np.testing.assert_allclose(F1, F0, rtol=1e-9, atol=1e-9)


def test_forces_are_minus_gradient(oracle, lj):
    rng = np.random.default_rng(11)
    cell = np.array([[6.0, 0.3, 0.0], [0.0, 6.2, 0.4], [0.5, 0.0, 5.8]])
    pos = spread_config(rng, 12, cell, 0.95)
    pot = oracle.Pot.from_table(lj.table)
    _, _, F = oracle.energy_forces(pot, pos, np.zeros(12, int), cell)
    h = 1e-6
    fd = np.zeros_like(pos)
    for i in range(12):
        for k in range(3):
            p = pos.copy()
            p[i, k] += h
            e1 = oracle.energy_forces(pot, p, np.zeros(12, int), cell, forces=False)[0]
            p[i, k] -= 2 * h
            e2 = oracle.energy_forces(pot, p, np.zeros(12, int), cell, forces=False)[0]
            fd[i, k] = -(e1 - e2) / (2 * h)
    np.testing.assert_allclose(F, fd, rtol=1e-6, atol=1e-6 * np.abs(F).max())


def test_lorentz_berthelot_value(oracle):
    be = LJBackend([1.0, 0.5], [1.0, 1.2])
    pot = oracle.Pot.from_table(be.table)
    r = 1.4
    pos = np.array([[0.0, 0, 0], [r, 0, 0]])
    U, _, _ = oracle.energy_forces(pot, pos, [0, 1], np.zeros((3, 3)))
    eps_ij = float(np.sqrt(np.float32(1.0) * np.float32(0.5)))
    sig_ij = 1.1
    expect = 4.0 * eps_ij * ((sig_ij / r) ** 12 - (sig_ij / r) ** 6)
    assert U == pytest.approx(expect, rel=1e-6)  # float32 tables (lj.py:71-72)


def test_per_species_reduces_to_scalar(oracle):
    pos = np.array([[0.0, 0, 0], [1.5, 0, 0], [0.3, 1.1, 0.2]])
    a = oracle.Pot.from_table(LJBackend(0.7, 0.9).table)
    b = oracle.Pot.from_table(LJBackend([0.7, 0.7], [0.9, 0.9]).table)
    Ua = oracle.energy_forces(a, pos, [0, 0, 0], np.zeros((3, 3)))[0]
    Ub = oracle.energy_forces(b, pos, [1, 0, 1], np.zeros((3, 3)))[0]
    # the table path squares sigma in float32 (lj.py:184), the scalar path in
    # float64 of the float32 sigma (lj.py:188): equal to float32 precision
    assert Ua == pytest.approx(Ub, rel=1e-6)


def test_explicit_pair_table_matches_brute(oracle):
    eps = np.array([[1.0, 1.5], [1.5, 0.5]])
    sig = np.array([[1.0, 0.8], [0.8, 0.88]])
    be = LJBackend(pair_epsilon=eps, pair_sigma=sig, cutoff=2.5)
    rng = np.random.default_rng(5)
    cell = np.eye(3) * 5.5
    pos = spread_config(rng, 20, cell, 0.75)
    ty = rng.integers(0, 2, 20)
    pot = oracle.Pot.from_table(be.table)
    U, _, _ = oracle.energy_forces(pot, pos, ty, cell)
    assert U == pytest.approx(brute_energy(pos, cell, eps, sig, 2.5, ty), rel=1e-12)


def test_nonperiodic_zero_cell(oracle, lj):
    pos = np.array([[0.0, 0, 0], [1.1, 0, 0], [0.0, 1.3, 0.0], [5.0, 5.0, 5.0]])
    pot = oracle.Pot.from_table(LJBackend(1.0, 1.0, None).table)
    U, _, _ = oracle.energy_forces(pot, pos, np.zeros(4, int), np.zeros((3, 3)))
    ref = 0.0
    for i in range(4):
        for j in range(i + 1, 4):
            r2 = ((pos[i] - pos[j]) ** 2).sum()
            ref += 4 * (r2 ** -6 - r2 ** -3)
    assert U == pytest.approx(ref, rel=1e-13)


def test_lj8_reference_dead_frames(oracle, lj):
    """The reference run's logged ns_energy = U + 0.1 V (float32 output)."""
    data = json.loads((GOLD / "lj8_dead_frames.json").read_text())
    pot = oracle.Pot.from_table(lj.table)
    for fr in data["frames"]:
        pos = np.array(fr["positions"])
        cell = np.array(fr["cell"])
        U, H, _ = oracle.energy_forces(pot, pos, np.zeros(len(pos), int), cell,
                                       pressure=data["pressure"])
        assert H == pytest.approx(fr["ns_energy"], rel=1e-6), fr["iter"]
        V = abs(np.linalg.det(cell))
        assert H - U == pytest.approx(0.1 * V, rel=1e-12)


def test_ensemble_term_hand_values(oracle):
    cell = np.diag([2.0, 3.0, 4.0])  # V = 24
    types = np.array([0, 1, 1, 0, 1], np.int32)
    mu = np.array([0.5, -1.25])
    e = oracle.lib().or_ensemble_energy(5, oracle._i(types), oracle._d(oracle.f64(cell).reshape(9)),
                                        0.1, oracle._d(oracle.f64(mu)), 2)
    # 0.1*24 - (0.5*2 + (-1.25)*3) = 2.4 - (1.0 - 3.75) = 5.15
    assert e == pytest.approx(5.15, rel=1e-14)


def test_cbrt_and_cell_shape(oracle):
    for x in [1e-3, 0.5, 1.0, 2.0, 27.0, 1234.5678, 3.7e5]:
        y = oracle.cbrt(x)
        assert y == pytest.approx(x ** (1 / 3), rel=2e-16)
    assert oracle.cbrt(0.0) == 0.0
    assert oracle.check_cell_shape(np.eye(3) * 4, 64, 64.0, 0.125, 0.8)
    assert not oracle.check_cell_shape(np.eye(3) * 1.0, 64, 64.0, 0.125, 0.8)
    assert not oracle.check_cell_shape(np.eye(3) * 17, 64, 64.0, 0.125, 0.8)
    flat = np.diag([10.0, 10.0, 1.0])  # aspect 1/cbrt(100) ~ 0.215
    assert not oracle.check_cell_shape(flat, 8, 1e4, 0.0, 0.8)
    assert oracle.check_cell_shape(flat, 8, 1e4, 0.0, 0.2)

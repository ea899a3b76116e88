"""Pins the oracle's chunked evaluation order (``eval_chunks`` C > 1).

The chunked order splits the ring steps of the pair sum into C consecutive
chunks, sums each on its own and adds the chunk sums in order.  It is the
same sum in another order, so: the energy matches a brute-force periodic
loop; energies and forces match the single running sum (C = 1) to rounding;
forces are minus the finite-difference gradient and sum to zero; and for
N <= 8 (fewer than four ring steps, so every chunk but the last is empty)
the chunked result equals C = 1 bit for bit.  A dropped or repeated group of
steps at a chunk boundary moves the energy by whole pair terms and fails the
brute-force and C = 1 comparisons.
"""

import numpy as np
import pytest

from jaxrens_amd.backends import LJBackend
from test_oracle_lj import brute_energy, spread_config


def _pot(oracle, lj, chunks):
    t = lj.table
    return oracle.Pot(t.n_species, t.eps4, t.eps24, t.sig2, t.rc2, chunks=chunks)


@pytest.fixture(scope="module")
def lj():
    return LJBackend(1.0, 1.0, 2.5)


@pytest.mark.parametrize("n", [1, 2, 3, 5, 8, 9, 13, 16, 31, 33, 64, 65, 96, 128, 130, 200, 256])
@pytest.mark.parametrize("chunks", [2, 4, 8])
def test_chunked_matches_brute_force_and_single_sum(oracle, lj, n, chunks):
    rng = np.random.default_rng(1000 + n)
    L = (n / 0.5) ** (1 / 3) + 1.0
    cell = np.eye(3) * L
    pos = spread_config(rng, n, cell) + rng.integers(-2, 3, size=(n, 3)) * L
    types = np.zeros(n, int)
    U1, H1, F1 = oracle.energy_forces(_pot(oracle, lj, 1), pos, types, cell, pressure=0.3)
    Uc, Hc, Fc = oracle.energy_forces(_pot(oracle, lj, chunks), pos, types, cell, pressure=0.3)
    ref = brute_energy(pos, cell, 1.0, 1.0, 2.5)
    assert Uc == pytest.approx(ref, rel=1e-12, abs=1e-12)
    assert Uc == pytest.approx(U1, rel=1e-13, abs=1e-13)
    assert Hc - Uc == pytest.approx(H1 - U1, rel=1e-12, abs=1e-12 * abs(H1))
    scale = max(1.0, float(np.abs(F1).max()))
    np.testing.assert_allclose(Fc, F1, rtol=0, atol=1e-12 * scale)
    assert np.abs(Fc.sum(0)).max() < 1e-9 * scale
    if n <= 8:
        assert Uc == U1 and np.array_equal(Fc, F1)


@pytest.mark.parametrize("chunks", [2, 8])
def test_chunked_triclinic_and_pair_table(oracle, chunks):
    rng = np.random.default_rng(5)
    cell = np.array([[6.5, 0.4, 0.0], [0.3, 6.2, 0.5], [0.1, 0.6, 6.8]])
    n = 70
    pos = spread_config(rng, n, cell, 0.85)
    types = rng.integers(0, 2, n)
    lj = LJBackend(pair_epsilon=[[1.0, 1.5], [1.5, 0.5]],
                   pair_sigma=[[1.0, 0.8], [0.8, 0.88]], cutoff=2.5)
    t = lj.table
    p1 = oracle.Pot(t.n_species, t.eps4, t.eps24, t.sig2, t.rc2, chunks=1)
    pc = oracle.Pot(t.n_species, t.eps4, t.eps24, t.sig2, t.rc2, chunks=chunks)
    mu = np.array([0.2, -0.4])
    U1, H1, F1 = oracle.energy_forces(p1, pos, types, cell, 0.7, mu)
    Uc, Hc, Fc = oracle.energy_forces(pc, pos, types, cell, 0.7, mu)
    eps = np.array([[1.0, 1.5], [1.5, 0.5]])
    sig = np.array([[1.0, 0.8], [0.8, 0.88]])
    assert Uc == pytest.approx(brute_energy(pos, cell, eps, sig, 2.5, types), rel=1e-12)
    assert Uc == pytest.approx(U1, rel=1e-13)
    assert Hc - Uc == pytest.approx(H1 - U1, rel=1e-12, abs=1e-12 * abs(H1))
    np.testing.assert_allclose(Fc, F1, rtol=0, atol=1e-12 * np.abs(F1).max())


def test_chunked_forces_are_minus_gradient(oracle, lj):
    rng = np.random.default_rng(21)
    cell = np.eye(3) * 5.3
    n = 40
    pos = spread_config(rng, n, cell, 0.95)
    pot = _pot(oracle, lj, 8)
    _, _, F = oracle.energy_forces(pot, pos, np.zeros(n, int), cell)
    h = 1e-6
    for a, d in [(0, 0), (7, 1), (19, 2), (39, 0)]:
        p = pos.copy()
        p[a, d] += h
        Up, _, _ = oracle.energy_forces(pot, p, np.zeros(n, int), cell, forces=False)
        p[a, d] -= 2 * h
        Um, _, _ = oracle.energy_forces(pot, p, np.zeros(n, int), cell, forces=False)
        assert F[a, d] == pytest.approx(-(Up - Um) / (2 * h), rel=1e-5, abs=1e-6)


def test_one_chunk_is_the_single_running_sum(oracle, lj):
    rng = np.random.default_rng(9)
    cell = np.eye(3) * 6.0
    pos = spread_config(rng, 50, cell)
    t = lj.table
    r0 = oracle.energy_forces(oracle.Pot(t.n_species, t.eps4, t.eps24, t.sig2, t.rc2, chunks=0),
                              pos, np.zeros(50, int), cell)
    r1 = oracle.energy_forces(_pot(oracle, lj, 1), pos, np.zeros(50, int), cell)
    assert r0[0] == r1[0] and np.array_equal(r0[2], r1[2])


def test_chunked_configurations_rejected():
    """Teams are built for the FP64 minimum-image pair walk without the
    minimum-distance constraint; other combinations fail at construction."""
    with pytest.raises(NotImplementedError):
        LJBackend(1.0, 1.0, 2.5, supercell_trafo=(3, 3, 3), eval_chunks=4)
    with pytest.raises(NotImplementedError):
        LJBackend(1.0, 1.0, 2.5, precision="fp32_pairs", eval_chunks=4)
    with pytest.raises(ValueError):
        LJBackend(1.0, 1.0, 2.5, eval_chunks=3)

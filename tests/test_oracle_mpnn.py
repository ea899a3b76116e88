"""Pins of the FP64 oracle of the message-passing judge (CPU only).

The judge is this project's own seeded random-init potential, so there is no
external fixture; the oracle is pinned by mathematics: forces against
central differences of the energy, translation / permutation / axis-swap
invariance, zero net force, continuity at both cutoffs, and an independent
NumPy restatement written from the documented weight layouts (z = m U + c,
emb [S][C], W [NB][C]) that a transposed matrix or a swapped index fails.
"""

import numpy as np
import pytest

from jaxrens_amd.mpnn import init_mpnn
from oracle import learned as OL


def _liquid(rng, N, rho=0.9, S=2):
    L = (N / rho) ** (1 / 3)
    n = int(np.ceil(N ** (1 / 3)))
    g = (np.arange(n) + 0.5) / n * L
    lat = np.array([[x, y, z] for x in g for y in g for z in g])[:N]
    pos = lat + rng.normal(scale=0.06 * L / n, size=(N, 3))
    ty = rng.integers(0, S, N)
    return pos, ty, np.eye(3) * L


def _numpy_mpnn(p, pos, ty, L):
    """Independent restatement: explicit image loop, dense tensors."""
    N = len(pos)
    u = pos / L - np.floor(pos / L)
    E_core = 0.0
    S1 = np.zeros((N, p.NB, p.S))
    edges = []
    nmax = int(np.ceil(p.rc / L)) + 1
    shifts = np.array([[a, b, c] for a in range(-nmax, nmax + 1)
                       for b in range(-nmax, nmax + 1) for c in range(-nmax, nmax + 1)])
    for i in range(N):
        d = (u[None, :, :] - u[i][None, None, :] + shifts[:, None, :]) * L
        r = np.linalg.norm(d, axis=-1)
        for si, j in zip(*np.nonzero(r < p.rc)):
            if i == j and not shifts[si].any():
                continue
            edges.append((i, j, r[si, j]))
    n = np.arange(1, p.NB + 1)

    def basis(r):
        x = r / p.rc
        env = 1 - 28 * x**6 + 48 * x**7 - 21 * x**8
        return np.sqrt(2 / p.rc) * np.sin(n * np.pi * x) / r * env

    emb, W1, U1, c1 = (a.astype(np.float64) for a in (p.emb, p.W1, p.U1, p.c1))
    W2, U2, c2, wo = (a.astype(np.float64) for a in (p.W2, p.U2, p.c2, p.wo))
    for i, j, r in edges:
        S1[i, :, ty[j]] += basis(r)
        if r < p.core_rc:
            E_core += 0.5 * p.core_eps * ((p.core_sig / r) ** 12 - (p.core_sig / p.core_rc) ** 12)
    m1 = np.einsum("nc,sc,ins->ic", W1, emb, S1)
    silu = lambda z: z / (1 + np.exp(-z))  # noqa: E731
    h1 = emb[ty] + silu(m1 @ U1 + c1)
    A2 = np.zeros((N, p.NB, p.C))
    for i, j, r in edges:
        A2[i] += basis(r)[:, None] * h1[j][None, :]
    m2 = np.einsum("nc,inc->ic", W2, A2)
    h2 = h1 + silu(m2 @ U2 + c2)
    return float((h2 @ wo).sum() + E_core), len(edges)


@pytest.fixture(scope="module")
def params():
    return init_mpnn(seed=3)


def test_numpy_restatement_matches(params):
    rng = np.random.default_rng(0)
    pos, ty, cell = _liquid(rng, 20, rho=0.12)  # L = 5.5 > rc: images matter
    E, _, ne = OL.mpnn_energy_forces(params, pos, ty, cell, forces=False)
    En, nen = _numpy_mpnn(params, pos, ty, cell[0, 0])
    assert ne == nen
    assert E == pytest.approx(En, rel=1e-11, abs=1e-11)


def test_forces_central_differences(params):
    rng = np.random.default_rng(1)
    pos, ty, cell = _liquid(rng, 16, rho=0.1)
    _, F, _ = OL.mpnn_energy_forces(params, pos, ty, cell)
    h = 1e-5
    fd = np.zeros_like(F)
    for a in range(len(pos)):
        for q in range(3):
            for sgn in (1, -1):
                pp = pos.copy()
                pp[a, q] += sgn * h
                e, _, _ = OL.mpnn_energy_forces(params, pp, ty, cell, forces=False)
                fd[a, q] -= sgn * e / (2 * h)
    np.testing.assert_allclose(F, fd, rtol=1e-6, atol=1e-6 * np.abs(F).max())
    np.testing.assert_allclose(F.sum(0), 0.0, atol=1e-9 * np.abs(F).max())


def test_invariances(params):
    rng = np.random.default_rng(2)
    pos, ty, cell = _liquid(rng, 40, rho=0.3)
    L = cell[0, 0]
    E0, F0, _ = OL.mpnn_energy_forces(params, pos, ty, cell)
    # lattice translation of single atoms and a global shift
    p1 = pos + rng.integers(-2, 3, size=pos.shape) * L + 0.37
    E1, F1, _ = OL.mpnn_energy_forces(params, p1, ty, cell)
    assert E1 == pytest.approx(E0, rel=1e-12)
    np.testing.assert_allclose(F1, F0, rtol=1e-8, atol=1e-9)
    # atom permutation
    perm = rng.permutation(len(pos))
    E2, F2, _ = OL.mpnn_energy_forces(params, pos[perm], ty[perm], cell)
    assert E2 == pytest.approx(E0, rel=1e-12)
    np.testing.assert_allclose(F2, F0[perm], rtol=1e-8, atol=1e-9)
    # axis permutation (a rotation symmetry of the cubic cell)
    E3, F3, _ = OL.mpnn_energy_forces(params, pos[:, [2, 0, 1]], ty, cell)
    assert E3 == pytest.approx(E0, rel=1e-12)
    np.testing.assert_allclose(F3, F0[:, [2, 0, 1]], rtol=1e-8, atol=1e-9)


@pytest.mark.parametrize("r0", [5.0, 1.2])
def test_continuity_at_cutoffs(params, r0):
    """Two atoms in a large box: the energy is continuous where the pair
    leaves the message cutoff (5.0) and the core cutoff (1.2)."""
    L = 12.0
    cell = np.eye(3) * L
    ty = np.array([0, 1])
    e = []
    for r in (r0 - 1e-7, r0 + 1e-7):
        pos = np.array([[1.0, 1.0, 1.0], [1.0 + r, 1.0, 1.0]])
        e.append(OL.mpnn_energy_forces(params, pos, ty, cell, forces=False)[0])
    assert abs(e[0] - e[1]) < 1e-5

"""Learned part on the GPU: the FP32 message-passing judge against its FP64
oracle.  Stated accuracy of the judge (FP32 arithmetic, hardware sin/cos/rsq
and reciprocal): |dE| <= 1e-4 (|E| + N) and |dF| <= 1e-3 |F| (Frobenius)."""

import numpy as np
import pytest

from jaxrens_amd import _lib
from jaxrens_amd.mpnn import MPNNBackend, init_mpnn
from oracle import learned as OL

pytestmark = pytest.mark.gpu


def _liquid(rng, N, rho, S=2):
    L = (N / rho) ** (1 / 3)
    n = int(np.ceil(N ** (1 / 3)))
    g = (np.arange(n) + 0.5) / n * L
    lat = np.array([[x, y, z] for x in g for y in g for z in g])[:N]
    pos = lat + rng.normal(scale=0.06 * L / n, size=(N, 3))
    pos += rng.integers(-2, 3, size=(N, 3)) * L  # unwrapped walkers
    ty = rng.permutation(np.arange(N) % S)
    return pos, ty, np.eye(3) * L


@pytest.mark.parametrize("N,rho,B", [(256, 0.9, 3), (100, 0.5, 2), (1, 0.004, 1)])
def test_mpnn_matches_oracle(N, rho, B):
    print(f"mpnn N={N} rho={rho} B={B}", flush=True)
    rng = np.random.default_rng(N)
    p = init_mpnn(seed=11)
    be = MPNNBackend(p)
    cfgs = [_liquid(rng, N, rho) for _ in range(B)]
    pos = np.stack([c[0] for c in cfgs])
    ty = np.stack([c[1] for c in cfgs])
    cell = np.stack([c[2] for c in cfgs])
    U, H, F = be._eval(pos, ty, cell, pressure=np.full(B, 0.5))
    for b in range(B):
        e, f, ne = OL.mpnn_energy_forces(p, pos[b], ty[b], cell[b])
        print(f"  cfg {b}: edges {ne}, E gpu {U[b]:.6f} oracle {e:.6f}, "
              f"|dF|/|F| {np.linalg.norm(F[b] - f) / max(np.linalg.norm(f), 1e-30):.2e}",
              flush=True)
        assert abs(U[b] - e) <= 1e-4 * (abs(e) + N)
        assert np.linalg.norm(F[b] - f) <= 1e-3 * np.linalg.norm(f) + 1e-6
        V = np.linalg.det(cell[b])
        assert H[b] - U[b] == pytest.approx(0.5 * V, rel=1e-12)


def test_mpnn_rejects_unsupported_cells():
    be = MPNNBackend(init_mpnn(seed=1))
    pos = np.zeros((1, 4, 3))
    ty = np.zeros((1, 4), int)
    with pytest.raises(_lib.JrError):
        be._eval(pos, ty, np.eye(3)[None] * 4.0)  # edge < rc
    tri = np.eye(3) * 8.0
    tri[1, 0] = 0.5
    with pytest.raises(_lib.JrError):
        be._eval(pos, ty, tri[None])

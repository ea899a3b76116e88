"""FP32-pair variant: float32 pair terms with float64 minimum image and
float64 accumulation.  Its error is measured against the FP64 oracle on
LJ-64 configurations; the bounds below are the stated accuracy."""

import numpy as np
import pytest

from jaxrens_amd.backends import LJBackend

pytestmark = pytest.mark.gpu


def _liquid(rng, B, N, rho=0.8):
    L = (N / rho) ** (1 / 3)
    # jittered lattice: no overlaps, liquid-like pair distances
    n = int(np.ceil(N ** (1 / 3)))
    g = (np.arange(n) + 0.5) / n * L
    lat = np.array([[x, y, z] for x in g for y in g for z in g])[:N]
    pos = lat[None] + rng.normal(scale=0.08 * L / n, size=(B, N, 3))
    pos += rng.integers(-3, 4, size=(B, N, 3)) * L  # unwrapped walkers
    return pos, np.tile(np.eye(3) * L, (B, 1, 1))


@pytest.mark.parametrize("N", [64, 128])
def test_fp32_pairs_error_vs_fp64(oracle, N):
    print(f"fp32 pairs N={N}", flush=True)
    rng = np.random.default_rng(N)
    pos, cell = _liquid(rng, 16, N)
    ty = np.zeros((16, N), int)
    f32 = LJBackend(1.0, 1.0, 2.5, precision="fp32_pairs")
    U, H, F = f32._eval(pos, ty, cell, pressure=0.5, forces=True)
    pot = oracle.Pot.from_table(LJBackend(1.0, 1.0, 2.5).table)
    du, df = [], []
    for b in range(16):
        u, h, f = oracle.energy_forces(pot, pos[b], ty[b], cell[b], pressure=0.5)
        du.append(abs(U[b] - u) / abs(u))
        df.append(np.linalg.norm(F[b] - f) / np.linalg.norm(f))
        assert H[b] - U[b] == pytest.approx(h - u, rel=1e-15)  # P V in FP64
    du, df = max(du), max(df)
    print(f"N={N}: max rel energy error {du:.2e}, max rel force error {df:.2e}")
    assert du < 2e-6
    assert df < 2e-6

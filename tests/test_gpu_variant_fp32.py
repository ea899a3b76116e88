"""FP32-pair variant: float32 pair terms with float64 minimum image and
float64 accumulation.  Its error is measured against the FP64 oracle on
LJ-64 and LJ-128 liquid configurations.

Stated accuracy: |U32 - U64| <= 2e-6 * sum_pairs |e_pair| and
|F32 - F64| <= 2e-6 * |F64| (Frobenius over the configuration).  The energy
bound is normalised by the sum of absolute pair energies (the condition number
of the sum), not by |U|: attraction and repulsion can cancel to a total near
zero (one configuration below has U = 0.0116 with sum |e_pair| = 564), where a
relative error in U says nothing about the pair arithmetic.  A float32
reciprocal biased one ulp high on every pair reaches 6.6e-7 on these inputs.
"""

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


def _sum_abs_pair_energy(x, L, rc2=6.25):
    """sum over pairs within the cutoff of |4 (r^-12 - r^-6)| (cubic box)."""
    d = x[None] - x[:, None]
    d -= L * np.round(d / L)
    r2 = (d ** 2).sum(-1)[np.triu_indices(len(x), 1)]
    r2 = r2[r2 < rc2]
    s6 = r2 ** -3
    return float(np.abs(4 * (s6 * s6 - s6)).sum())


@pytest.mark.parametrize("N", [64, 128])
def test_fp32_pairs_error_vs_fp64(oracle, N):
    print(f"fp32 pairs N={N}", flush=True)
    rng = np.random.default_rng(N)
    pos, cell = _liquid(rng, 16, N)
    ty = np.zeros((16, N), int)
    f32 = LJBackend(1.0, 1.0, 2.5, precision="fp32_pairs")
    U, H, F = f32._eval(pos, ty, cell, pressure=0.5, forces=True)
    pot = oracle.Pot.from_table(LJBackend(1.0, 1.0, 2.5).table)
    du, df, dn = [], [], []
    for b in range(16):
        u, h, f = oracle.energy_forces(pot, pos[b], ty[b], cell[b], pressure=0.5)
        cond = _sum_abs_pair_energy(pos[b], cell[b, 0, 0])
        du.append(abs(U[b] - u) / cond)
        dn.append(abs(U[b] - u) / N)
        df.append(np.linalg.norm(F[b] - f) / np.linalg.norm(f))
        assert H[b] - U[b] == pytest.approx(h - u, rel=1e-15)  # P V in FP64
        # the walk must see the same pair set: no energy far from the oracle
        assert abs(U[b] - u) < 1e-4 * N
    du, df, dn = max(du), max(df), max(dn)
    print(f"N={N}: max |dU|/sum|e_pair| {du:.2e}, max |dU|/N {dn:.2e}, "
          f"max rel force error {df:.2e}")
    assert du < 2e-6
    assert df < 2e-6

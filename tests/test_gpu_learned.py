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


# ------------------------------------------------------------------ surrogate
from jaxrens_amd import surrogate as SG  # noqa: E402


def _bf16_bits(x):
    """float64 -> bfloat16 bit patterns, round to nearest even."""
    f = np.ascontiguousarray(np.asarray(x, np.float32)).view(np.uint32).astype(np.uint64)
    f = f + 0x7FFF + ((f >> 16) & 1)
    return (f >> 16).astype(np.uint16)


def _bits_to_f(b):
    return (np.asarray(b, np.uint32) << 16).view(np.float32)


@pytest.mark.parametrize("N,rho,base", [(256, 0.9, None), (100, 0.7, None),
                                        (128, 1.1, (1.0, 0.85, 1.2))])
def test_surrogate_eval_matches_oracle(N, rho, base):
    """BF16 matrix-core surrogate vs the oracle that rounds to BF16 at the same
    points.  Stated accuracy: descriptors within one BF16 ulp, |dE| <= 2e-3
    (sum_i |e_i| + |E_baseline|) and |dF| <= 2e-2 |F|."""
    print(f"surrogate eval N={N} baseline={base}", flush=True)
    rng = np.random.default_rng(N + 1)
    p = SG.init_surrogate(7)
    cfgs = [_liquid(rng, N, rho) for _ in range(2)]
    pos = np.stack([c[0] for c in cfgs])
    ty = np.stack([c[1] for c in cfgs])
    cell = np.stack([c[2] for c in cfgs])
    U, F, X = SG.sur_eval(p, pos, ty, cell, descriptors=True, base=base)
    for b in range(2):
        xo = OL.sur_descriptors(pos[b], ty[b], cell[b])
        xg = _bits_to_f(X[b]).astype(np.float64)
        xr = _bits_to_f(_bf16_bits(xo)).astype(np.float64)
        same = np.mean(xg == xr)
        ulp = np.abs(xg - xr) / np.maximum(np.abs(xr), 1e-30)
        e, f = OL.sur_energy_forces(p, pos[b], ty[b], cell[b], bf16=True, base=base)
        scale = np.abs(SG_row_energies(p, xo)).sum()
        if base is not None:
            eb = OL.sur_energy_forces(np.zeros(SG.N_PARAMS, np.float32), pos[b], ty[b],
                                      cell[b], forces=False, base=base)[0]
            scale += abs(eb)
        rel_f = np.linalg.norm(F[b] - f) / np.linalg.norm(f)
        print(f"  cfg {b}: descriptors equal {same:.4f}, max rel {ulp.max():.2e}; "
              f"E gpu {U[b]:.6f} oracle {e:.6f} (sum|e| {scale:.3f}); |dF|/|F| {rel_f:.2e}",
              flush=True)
        assert ulp.max() <= 2.0 ** -7 + 1e-12
        assert abs(U[b] - e) <= 2e-3 * scale
        assert rel_f <= 2e-2


def SG_row_energies(p, x):
    """Per-row energies of the oracle (unrounded), for error scales."""
    e = []
    for r in x:
        h1 = r @ p[SG.OFF_W1:SG.OFF_B1].reshape(SG.D_IN, SG.HID) + p[SG.OFF_B1:SG.OFF_W2]
        h1 = h1 / (1 + np.exp(-h1))
        z2 = h1 @ p[SG.OFF_W2:SG.OFF_B2].reshape(SG.HID, SG.HID) + p[SG.OFF_B2:SG.OFF_W3]
        e.append((z2 / (1 + np.exp(-z2))) @ p[SG.OFF_W3:SG.OFF_B3] + p[SG.OFF_B3])
    return np.array(e)


def test_surrogate_gradient_matches_oracle():
    """Training gradient of the loss 1/n sum_b ((E_hat_b - E_b)/N_b)^2 on the
    matrix cores vs the oracle (both round the same operands to BF16).
    Stated accuracy: |dg| <= 1e-2 |g| (vector norm) and the loss within
    1e-3 relative."""
    print("surrogate gradient", flush=True)
    rng = np.random.default_rng(5)
    p = SG.init_surrogate(9)
    sizes = [128, 100, 37]
    cfgs = [_liquid(rng, n, 0.9) if n >= 100 else _liquid(rng, n, 0.3) for n in sizes]
    x = np.concatenate([OL.sur_descriptors(*c) for c in cfgs])
    bits = _bf16_bits(x)
    xr = _bits_to_f(bits).astype(np.float64)
    E_t = rng.normal(size=3) * 5.0
    e_rows, _ = SG.sur_grad(p, bits)
    Eh, o = [], 0
    for n in sizes:
        Eh.append(e_rows[o:o + n].sum())
        o += n
    Eh = np.array(Eh)
    c = np.concatenate([np.full(n, 2.0 * (Eh[b] - E_t[b]) / (n * n * len(sizes)))
                        for b, n in enumerate(sizes)])
    _, g = SG.sur_grad(p, bits, c)
    loss_o, g_o = OL.sur_loss_grad(p, sizes, xr, E_t, bf16=True)
    loss = np.mean(((Eh - E_t) / np.array(sizes)) ** 2)
    rel = np.linalg.norm(g - g_o) / np.linalg.norm(g_o)
    print(f"  loss gpu {loss:.6e} oracle {loss_o:.6e}; |dg|/|g| {rel:.2e}", flush=True)
    assert loss == pytest.approx(loss_o, rel=1e-3)
    assert rel <= 1e-2


def test_surrogate_adam_matches_oracle():
    """Adam on the device vs the oracle: equal up to one FP32 rounding of
    each update (the device fuses multiply-adds, the oracle rounds twice)."""
    rng = np.random.default_rng(6)
    n = SG.N_PARAMS
    p = rng.standard_normal(n).astype(np.float32)
    m = (0.1 * rng.standard_normal(n)).astype(np.float32)
    v = (0.01 * rng.random(n)).astype(np.float32)
    g = rng.standard_normal(n).astype(np.float32)
    pg, mg, vg = SG.sur_adam(p, m, v, g, 1e-3, 4)
    po, mo, vo = OL.adam_step(p, m, v, g, 1e-3, 4)
    np.testing.assert_allclose(mg, mo, rtol=1e-6, atol=2e-7)
    np.testing.assert_allclose(vg, vo, rtol=1e-6, atol=1e-9)
    np.testing.assert_allclose(pg, po, rtol=1e-6, atol=5e-7)


def test_surrogate_training_fits_the_judge():
    """Online-style training on the GPU (matrix-core forward/backward + Adam)
    lowers the per-atom energy error against the judge on held-out
    configurations."""
    print("surrogate training", flush=True)
    rng = np.random.default_rng(8)
    judge = MPNNBackend(init_mpnn(seed=11))
    N = 128
    cfgs = [_liquid(rng, N, float(rng.uniform(0.6, 0.95))) for _ in range(48)]
    pos = np.stack([c[0] for c in cfgs])
    ty = np.stack([c[1] for c in cfgs])
    cell = np.stack([c[2] for c in cfgs])
    Ej, _, _ = judge._eval(pos, ty, cell, forces=False)
    p = SG.init_surrogate(2)
    _, _, X = SG.sur_eval(p, pos, ty, cell, forces=False, descriptors=True)
    train, held = np.arange(32), np.arange(32, 48)

    def rmse(params, idx):
        e, _ = SG.sur_grad(params, X[idx].reshape(-1, SG.D_IN))
        Eh = e.reshape(len(idx), N).sum(1)
        return np.sqrt(np.mean(((Eh - Ej[idx]) / N) ** 2))

    r0 = rmse(p, held)
    m = np.zeros_like(p)
    v = np.zeros_like(p)
    xt = X[train].reshape(-1, SG.D_IN)
    for t in range(1, 301):
        e, _ = SG.sur_grad(p, xt)
        Eh = e.reshape(len(train), N).sum(1)
        c = np.repeat(2.0 * (Eh - Ej[train]) / (N * N * len(train)), N)
        _, g = SG.sur_grad(p, xt, c)
        p, m, v = SG.sur_adam(p, m, v, g, 3e-3, t)
    r1 = rmse(p, held)
    print(f"  held-out per-atom RMSE {r0:.4f} -> {r1:.4f}", flush=True)
    assert r1 < 0.25 * r0

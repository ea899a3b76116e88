"""Pins of the FP64 oracle of the radial-descriptor surrogate (CPU only):
descriptor closed form for two atoms, forces and the training-loss gradient
against central differences, invariances, and Adam's first step
(-lr * sign(g) by construction of the bias correction)."""

import numpy as np
import pytest

from jaxrens_amd import surrogate as SG
from oracle import learned as OL


def _cfg(rng, N, rho, S=2):
    L = max((N / rho) ** (1 / 3), 6.5)
    n = int(np.ceil(N ** (1 / 3)))
    g = (np.arange(n) + 0.5) / n * L
    lat = np.array([[x, y, z] for x in g for y in g for z in g])[:N]
    pos = lat + rng.normal(scale=0.08 * L / n, size=(N, 3))
    return pos, rng.integers(0, S, N), np.eye(3) * L


def test_descriptor_two_atoms_closed_form():
    L, r = 10.0, 1.37
    pos = np.array([[1.0, 2.0, 3.0], [1.0 + r, 2.0, 3.0]])
    x = OL.sur_descriptors(pos, [0, 1], np.eye(3) * L)
    mu = 0.8 + 1.7 * np.arange(32) / 31
    t = r / 3.0
    env = 1 - 28 * t**6 + 48 * t**7 - 21 * t**8
    g = 0.25 * np.exp(-(r - mu) ** 2 / (2 * 0.01)) * env
    np.testing.assert_allclose(x[0, 32:64], g, rtol=1e-13, atol=1e-15)  # atom 0 sees species 1
    np.testing.assert_allclose(x[1, 0:32], g, rtol=1e-13, atol=1e-15)
    assert not x[0, 0:32].any() and not x[1, 32:64].any()
    assert x[0, 64] == 1.0 and x[1, 65] == 1.0 and not x[:, 66:].any()


def test_forces_central_differences():
    rng = np.random.default_rng(0)
    p = SG.init_surrogate(3)
    pos, ty, cell = _cfg(rng, 24, 0.9)
    E, F = OL.sur_energy_forces(p, pos, ty, cell, bf16=False)
    h = 1e-5
    fd = np.zeros_like(F)
    for a in range(len(pos)):
        for q in range(3):
            for s in (1, -1):
                pp = pos.copy()
                pp[a, q] += s * h
                fd[a, q] -= s * OL.sur_energy_forces(p, pp, ty, cell, bf16=False, forces=False)[0] / (2 * h)
    np.testing.assert_allclose(F, fd, rtol=1e-6, atol=1e-7 * np.abs(F).max())
    np.testing.assert_allclose(F.sum(0), 0, atol=1e-10 * np.abs(F).max())


def test_invariances():
    rng = np.random.default_rng(1)
    p = SG.init_surrogate(4)
    pos, ty, cell = _cfg(rng, 30, 0.9)
    E0, F0 = OL.sur_energy_forces(p, pos, ty, cell)
    perm = rng.permutation(30)
    E1, F1 = OL.sur_energy_forces(p, pos[perm] + cell[0, 0] * rng.integers(-2, 3, (30, 3)),
                                  ty[perm], cell)
    assert E1 == pytest.approx(E0, rel=1e-12)
    np.testing.assert_allclose(F1, F0[perm], rtol=1e-9, atol=1e-10)


def test_loss_gradient_central_differences():
    rng = np.random.default_rng(2)
    p = SG.init_surrogate(5)
    cfgs = [_cfg(rng, n, 0.9) for n in (12, 20)]
    x = np.concatenate([OL.sur_descriptors(*c) for c in cfgs])
    n_atoms = [12, 20]
    E = np.array([0.3, -1.1])
    loss, g = OL.sur_loss_grad(p, n_atoms, x, E, bf16=False)
    idx = np.concatenate([rng.choice(np.arange(SG.OFF_W1, SG.OFF_B1), 12),
                          rng.choice(np.arange(SG.OFF_W2, SG.OFF_B2), 12),
                          [SG.OFF_B1 + 3, SG.OFF_B2 + 7, SG.OFF_W3 + 5, SG.OFF_B3]])
    for k in idx:
        h = 1e-3 * max(1.0, abs(float(p[k])))
        pp = p.astype(np.float64).copy()
        pp[k] += h
        lp = OL.sur_loss_grad(pp.astype(np.float32), n_atoms, x, E, bf16=False)[0]
        pp[k] -= 2 * h
        lm = OL.sur_loss_grad(pp.astype(np.float32), n_atoms, x, E, bf16=False)[0]
        # parameters are FP32: use the representable step
        hp = (np.float32(p[k] + h) - np.float32(p[k] - h)) / 2
        assert g[k] == pytest.approx((lp - lm) / (2 * hp), rel=2e-3, abs=1e-7), k


def test_adam_first_step_is_signed_lr():
    rng = np.random.default_rng(3)
    n = 1000
    p = rng.standard_normal(n).astype(np.float32)
    g = rng.standard_normal(n).astype(np.float32)
    p1, m1, v1 = OL.adam_step(p, np.zeros(n), np.zeros(n), g, 1e-3, 1)
    np.testing.assert_allclose(p1 - p, -1e-3 * np.sign(g), rtol=1e-4, atol=1e-7)
    np.testing.assert_allclose(m1, 0.1 * g, rtol=1e-6)
    np.testing.assert_allclose(v1, 0.001 * g * g, rtol=3e-5)  # 1 - 0.999f in FP32


def test_pair_baseline_closed_form_and_forces():
    """With the network's output zeroed (every parameter 0), the surrogate is
    its pair baseline alone: eps ((sig/r)^12 - (sig/rc)^12) for r < rc,
    counted once per pair over all images, with forces -d/dx of it."""
    base = (1.3, 0.85, 1.2)
    zero = np.zeros(SG.N_PARAMS, np.float32)
    L = 6.0
    for r in (0.7, 1.0, 1.19, 1.25):
        pos = np.array([[1.0, 1.0, 1.0], [1.0 + r, 1.0, 1.0]])
        E, F = OL.sur_energy_forces(zero, pos, [0, 1], np.eye(3) * L, base=base)
        want = base[0] * ((base[1] / r) ** 12 - (base[1] / base[2]) ** 12) if r < base[2] else 0.0
        dphi = -12 * base[0] * base[1] ** 12 / r ** 13 if r < base[2] else 0.0
        assert E == pytest.approx(want, rel=1e-13, abs=1e-15)
        np.testing.assert_allclose(F[1], [-dphi, 0, 0], rtol=1e-12, atol=1e-15)
        np.testing.assert_allclose(F[0], -F[1], rtol=1e-12, atol=1e-15)
    # the baseline adds to the network's energy; forces match differences
    rng = np.random.default_rng(5)
    p = SG.init_surrogate(3)
    cell = np.eye(3) * 6.5
    pos = rng.random((60, 3)) * 6.5
    ty = rng.integers(0, 2, 60)
    E0, _ = OL.sur_energy_forces(p, pos, ty, cell, bf16=False)
    Eb, Fb = OL.sur_energy_forces(p, pos, ty, cell, bf16=False, base=base)
    Ez, _ = OL.sur_energy_forces(zero, pos, ty, cell, bf16=False, base=base)
    assert Ez > 1.0  # close pairs present
    assert Eb == pytest.approx(E0 + Ez, rel=1e-12)
    h = 1e-6
    for a in range(3):
        for q in range(3):
            pp, pm = pos.copy(), pos.copy()
            pp[a, q] += h
            pm[a, q] -= h
            fd = -(OL.sur_energy_forces(p, pp, ty, cell, bf16=False, forces=False, base=base)[0]
                   - OL.sur_energy_forces(p, pm, ty, cell, bf16=False, forces=False, base=base)[0]) / (2 * h)
            assert Fb[a, q] == pytest.approx(fd, rel=1e-5, abs=1e-6 * np.abs(Fb).max())

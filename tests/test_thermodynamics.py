"""Pins of the thermodynamics restatement (``postprocess/thermodynamics.py``)
by closed forms: a d-dimensional harmonic well U = k |x|^2 / 2 in a large
box V has prior mass X(E) = (2 pi E / k)^(d/2) / (Gamma(d/2 + 1) V); dead
points placed on the expected ladder X_n of the pymatnest weights give
Z(beta) -> (2 pi / (beta k))^(d/2) / V and C_v -> d / 2 (equipartition)."""

import math

import numpy as np
import pytest

from jaxrens_amd import thermodynamics as T
from jaxrens_amd.prior_mass import log_prior_mass


def harmonic_run(d, k, V, K, n_dead):
    log_x = log_prior_mass(n_dead, K)  # expected prior mass after each death
    c = (2 * math.pi / k) ** (d / 2) / (math.gamma(d / 2 + 1) * V)
    dead = np.exp((log_x - math.log(c)) * (2.0 / d))  # X(E) = c E^(d/2)
    # live walkers: uniform in X below the last contour
    u = (np.arange(K) + 0.5) / K
    live = (np.exp(log_x[-1]) * u / c) ** (2.0 / d)
    return dead, live


@pytest.mark.parametrize("d", [3, 9])
def test_harmonic_partition_function_and_equipartition(d):
    # the box holds the well at every energy that matters (X(30) << 1) and the
    # ladder reaches X ~ 1e-20, below the energies that matter at beta = 4
    k, V, K = 2.0, 1.0e12, 4000
    dead, live = harmonic_run(d, k, V, K, n_dead=K * 46)
    for beta in (0.5, 1.0, 4.0):
        lz = T.partition_function(beta, dead, live, K)
        exact = (d / 2) * math.log(2 * math.pi / (beta * k)) - math.log(V)
        assert lz == pytest.approx(exact, abs=0.02 * d / 2 + 0.01)
        assert T.heat_capacity(beta, dead, live, K) == pytest.approx(d / 2, rel=0.03)
        # <E> = d / (2 beta) and F = -log Z / beta
        assert T.expectation(np.concatenate([dead, live]), beta, dead, live, K) == \
            pytest.approx(d / (2 * beta), rel=0.03)
        assert T.free_energy(beta, lz) == pytest.approx(-lz / beta)
    assert T.log_evidence(dead, live, K) == T.partition_function(1.0, dead, live, K)


def test_deviation_form_survives_a_collapsed_distribution():
    """At low T the weight sits on the lowest energy; the deviation form
    gives C_v >= 0 with no cancellation noise (beta E ~ 1e5 here: the
    reference's formula, kept as is, needs beta E well below 1/eps)."""
    dead = np.linspace(1.0e3 + 100.0, 1.0e3 + 1.0, 500)  # offset >> spread
    live = np.full(8, 1.0e3)  # one unit below every dead point
    cv = T.heat_capacity_curve([1e-2, 1e-1, 1.0, 10.0], dead, live, 8)
    assert np.all(cv >= 0.0) and np.all(np.isfinite(cv))
    assert cv[0] < 1e-6


def test_batch_axes_and_live_broadcast():
    rng = np.random.default_rng(0)
    dead = np.sort(rng.random((3, 200)))[:, ::-1] * 10
    live = rng.random(16)
    lz = T.partition_function(0.7, dead, live, 16)
    assert lz.shape == (3,)
    for r in range(3):
        assert lz[r] == pytest.approx(T.partition_function(0.7, dead[r], live, 16))

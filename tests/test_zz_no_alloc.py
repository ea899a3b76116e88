"""jr_eval_energy_forces_noalloc (the allocation-free entry point) against
jr_eval_energy_forces through the C ABI: bit-identical energies, enthalpies
and forces on two calls in a row, the second smaller than the first (the
shared scratch keeps the first call's contents), and a clear error above the
capacity."""

import ctypes as C

import numpy as np
import pytest

from jaxrens_amd import _lib
from jaxrens_amd.backends import LJBackend

pytestmark = pytest.mark.gpu


def _call(fn, pot, pos, ty, na, cell, pressure, mu, forces=True):
    B, Nm = pos.shape[:2]
    U, H = np.full(B, np.nan), np.full(B, np.nan)
    F = np.full((B, Nm, 3), np.nan) if forces else None
    rc = fn(C.byref(pot), 0, B, Nm, _lib.dptr(_lib.f64(pos)), _lib.iptr(_lib.i32(ty)),
            _lib.iptr(_lib.i32(na)), _lib.dptr(_lib.f64(cell)), _lib.dptr(pressure),
            _lib.dptr(mu), _lib.dptr(U), _lib.dptr(H), _lib.dptr(F))
    return rc, U, H, F


def _cfgs(rng, B, Nm, S):
    na = rng.integers(1, Nm + 1, size=B).astype(np.int32)
    L = (na / 0.7) ** (1 / 3)
    pos = rng.random((B, Nm, 3)) * L[:, None, None]
    cell = np.eye(3)[None] * L[:, None, None] + rng.uniform(-0.1, 0.1, (B, 3, 3))
    ty = rng.integers(0, S, size=(B, Nm)).astype(np.int32)
    return pos, ty, na, cell


@pytest.mark.parametrize("species,chunks", [(1, 1), (2, 1), (1, 4)])
def test_noalloc_matches_allocating_entry_point(species, chunks):
    print(f"noalloc eval vs jr_eval_energy_forces, S={species} eval_chunks={chunks}", flush=True)
    lib = _lib.load_library()
    if species == 1:
        lj = LJBackend(1.0, 1.0, 2.5, eval_chunks=chunks)
    else:
        lj = LJBackend(pair_epsilon=np.array([[1.0, 1.5], [1.5, 0.5]]),
                       pair_sigma=np.array([[1.0, 0.8], [0.8, 0.88]]), cutoff=2.5,
                       eval_chunks=chunks)
    pot, keep = lj.table.as_struct(lj._prec_code)
    rng = np.random.default_rng(40 + species)
    # the first call fills most of the scratch, the second is smaller
    for B, Nm in ((400, 128), (37, 60)):
        print(f"  B={B} Nmax={Nm}", flush=True)
        pos, ty, na, cell = _cfgs(rng, B, Nm, species)
        pr = _lib.f64(rng.uniform(0.1, 2.0, B))
        mu = _lib.f64(rng.normal(size=(B, species)))
        rc0, U0, H0, F0 = _call(lib.jr_eval_energy_forces, pot, pos, ty, na, cell, pr, mu)
        rc1, U1, H1, F1 = _call(lib.jr_eval_energy_forces_noalloc, pot, pos, ty, na, cell, pr,
                                mu)
        assert rc0 == 0 and rc1 == 0, lib.jr_global_error()
        assert np.array_equal(U0, U1) and np.array_equal(H0, H1)
        assert np.array_equal(F0, F1)
        assert np.isfinite(U1).all()
        # energy only, no ensemble arrays
        rc2, U2, H2, _ = _call(lib.jr_eval_energy_forces_noalloc, pot, pos, ty, na, cell, None,
                               None, forces=False)
        rc3, U3, H3, _ = _call(lib.jr_eval_energy_forces, pot, pos, ty, na, cell, None, None,
                               forces=False)
        assert rc2 == 0 and rc3 == 0
        assert np.array_equal(U2, U3) and np.array_equal(H2, H3) and np.array_equal(U2, U0)
    del keep


def test_noalloc_rejects_calls_above_capacity():
    lib = _lib.load_library()
    lj = LJBackend(1.0, 1.0, 2.5)
    pot, keep = lj.table.as_struct(lj._prec_code)
    B, Nm = 300, 256  # 76800 slots > 65536
    pos, ty, na, cell = _cfgs(np.random.default_rng(3), B, Nm, 1)
    rc, *_ = _call(lib.jr_eval_energy_forces_noalloc, pot, pos, ty, na, cell, None, None)
    assert rc == -1  # JR_E_ARG
    assert b"scratch capacity" in lib.jr_global_error()
    del keep

"""Host-side run files: the ``.energies`` log (``io/energy_log.py:24-104``)
and its cut-back on resume (``io/restart_truncate.py:28-50``)."""

import numpy as np

from jaxrens_amd.run_loop import EnergyLogger


def test_energy_log_roundtrip_and_header(tmp_path):
    p = tmp_path / "x.energies"
    lg = EnergyLogger(p, n_walkers=200, n_cull=1, n_dof=0, n_atoms=8)
    for i in range(5):
        lg.write_entry(i, -1.0 - i, 10.0 + i)
    lg.close()
    assert p.read_text().splitlines()[0] == "200 1 0 0.0 8"
    d = EnergyLogger.read(p)
    assert d["n_walkers"] == 200 and d["n_atoms"] == 8
    assert np.array_equal(d["iterations"], np.arange(5))
    assert np.array_equal(d["energies"], -1.0 - np.arange(5))


def test_energy_log_resume_cuts_rows_past_the_checkpoint(tmp_path):
    p = tmp_path / "x.energies"
    lg = EnergyLogger(p, n_walkers=16, n_atoms=8)
    for i in range(9):  # rows 0..8 written before a crash, checkpoint at 6
        lg.write_entry(i, float(i), 1.0)
    lg.close()
    lg = EnergyLogger(p, n_walkers=16, n_atoms=8, restart_iteration=6)
    for i in range(6, 8):
        lg.write_entry(i, 100.0 + i, 2.0)
    lg.close()
    lines = p.read_text().splitlines()
    assert lines[0] == "16 1 0 0.0 8" and sum(ln == lines[0] for ln in lines) == 1
    d = EnergyLogger.read(p)
    assert np.array_equal(d["iterations"], np.arange(8))
    assert np.array_equal(d["energies"][:6], np.arange(6.0))
    assert np.array_equal(d["energies"][6:], [106.0, 107.0])


def test_energy_log_fresh_run_overwrites(tmp_path):
    p = tmp_path / "x.energies"
    p.write_text("stale\n1 2 3\n")
    EnergyLogger(p, n_walkers=4).close()
    assert p.read_text() == "4 1 0 0.0 0\n"

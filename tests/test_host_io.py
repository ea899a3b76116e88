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


GOLD = __import__("pathlib").Path(__file__).parent / "golden"


def test_extxyz_matches_the_reference_run_output_byte_for_byte():
    """The four dead frames of the reference's LJ-8 NPT run written again
    through ours; that run kept absolute Cartesians (``wrap_atoms: false`` in
    its config snapshot), so the writer runs with ``wrap=False`` here."""
    from jaxrens_amd.trajectory import ExtxyzTrajectoryWriter, read_extxyz
    src = GOLD / "lj8_npt.traj.extxyz"
    frames = read_extxyz(src)
    assert [f["iter"] for f in frames] == [0, 50, 100, 150]
    out = src.parent.parent / "__traj_roundtrip.extxyz"
    try:
        w = ExtxyzTrajectoryWriter(out, {0: "Ar"}, wrap=False)
        for f in frames:
            w.write_dead_point(f["iter"], f["positions"], [0] * len(f["positions"]),
                               f["cell"], f["ns_energy"])
        w.close()
        assert out.read_text() == src.read_text()
    finally:
        out.unlink(missing_ok=True)


def test_wrap_positions_convention():
    from jaxrens_amd.trajectory import WRAP_EPS, wrap_positions
    rng = np.random.default_rng(0)
    cell = np.array([[3.0, 0.2, -0.1], [0.4, 2.5, 0.3], [-0.2, 0.1, 2.8]])
    x = rng.normal(size=(50, 3)) * 10
    w = wrap_positions(x, cell)
    f = np.linalg.solve(cell.T, w.T).T
    assert np.all(f >= WRAP_EPS - 1e-12) and np.all(f < 1 + WRAP_EPS + 1e-12)
    shift = rng.integers(-4, 5, (50, 3)) @ cell
    np.testing.assert_allclose(wrap_positions(x + shift, cell), w, atol=1e-11)
    d = np.linalg.solve(cell.T, (w - x).T).T  # whole lattice vectors moved
    np.testing.assert_allclose(d, np.round(d), atol=1e-9)
    np.testing.assert_array_equal(wrap_positions(x, np.zeros((3, 3))), x)


def test_extxyz_resume_drops_frames_past_the_checkpoint(tmp_path):
    from jaxrens_amd.trajectory import ExtxyzTrajectoryWriter, read_extxyz
    p = tmp_path / "t.traj.extxyz"
    cell = np.eye(3) * 4.0
    w = ExtxyzTrajectoryWriter(p, {0: "Ar", 1: "Kr"})
    for i in range(6):
        w.write_dead_point(i, np.full((2, 3), 0.5 + i * 0.1), [0, 1], cell, -float(i))
    w.close()
    w = ExtxyzTrajectoryWriter(p, {0: "Ar", 1: "Kr"}, restart_iteration=4)
    w.write_dead_point(4, np.full((2, 3), 1.5), [1, 1], cell, 40.0)
    w.close()
    fr = read_extxyz(p)
    assert [f["iter"] for f in fr] == [0, 1, 2, 3, 4]
    assert fr[4]["ns_energy"] == 40.0 and fr[4]["symbols"] == ["Kr", "Kr"]
    assert fr[1]["symbols"] == ["Ar", "Kr"] and fr[3]["ns_energy"] == -3.0

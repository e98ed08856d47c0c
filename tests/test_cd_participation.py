"""Circular-dichroism set-up and the participation ratio on the CPU: the
oracle's restatements pinned by closed forms and library routines, the
engine's host helpers (no GPU needed) against the oracle."""
import os

import numpy as np
import pytest

from tests.helpers import ring_tube


def test_oracle_participation_closed_forms(oracle):
    """1 / sum |c|^4: k equal weights give k whatever the phases, a single
    site gives 1, two weights p, 1 - p give 1 / (p^2 + (1 - p)^2)"""
    rng = np.random.default_rng(3)
    for k in (1, 2, 7, 600):
        psi = np.zeros(900, complex)
        idx = rng.choice(900, k, replace=False)
        psi[idx] = np.exp(2j * np.pi * rng.random(k)) / np.sqrt(k)
        assert abs(oracle.participation(psi) - k) <= 1e-12 * k
    for p in (0.1, 0.3, 0.5):
        psi = np.array([np.sqrt(p), 1j * np.sqrt(1 - p), 0.0])
        assert abs(oracle.participation(psi) - 1 / (p * p + (1 - p) ** 2)) <= 1e-14


def test_oracle_cd_vectors_against_cross_product(oracle, engine):
    sysd = ring_tube(engine)
    pos, mu = sysd["pos"], sysd["dip"]
    rd, mt, rt = oracle.cd_vectors(pos, mu)
    assert np.allclose(rd, np.cross(pos, mu), rtol=0, atol=1e-12 * np.abs(rd).max())
    # R x mu is orthogonal to both R and mu
    assert np.abs((rd * pos).sum(1)).max() <= 1e-10 * np.abs(rd).max() * np.abs(pos).max()
    assert np.abs((rd * mu).sum(1)).max() <= 1e-10 * np.abs(rd).max() * np.abs(mu).max()
    assert np.allclose(mt, np.linalg.norm(mu, axis=0), rtol=1e-14)
    assert np.allclose(rt, np.linalg.norm(rd, axis=0), rtol=1e-14)
    # a transposed operand flips the sign of every component
    assert not np.allclose(rd, np.cross(mu, pos))


def test_engine_cd_vectors_match_oracle(oracle, engine):
    sysd = ring_tube(engine)
    rng = np.random.default_rng(11)
    for pos, mu in ((sysd["pos"], sysd["dip"]), (rng.normal(size=(37, 3)), rng.normal(size=(37, 3)))):
        rd, mt, rt = engine.cd_vectors(pos, mu)
        ro, mo, to = oracle.cd_vectors(pos, mu)
        assert np.array_equal(rd, ro)
        assert np.array_equal(mt, mo) and np.array_equal(rt, to)
    with pytest.raises(ValueError):
        engine.cd_vectors(np.zeros((4, 3)), np.zeros((5, 3)))


def test_spectrum_runs_kets_bras_and_scales(oracle, engine):
    sysd = ring_tube(engine)
    pos, mu = sysd["pos"], sysd["dip"]
    rd, mt, rt = oracle.cd_vectors(pos, mu)
    cd = engine.spectrum_runs(pos, mu, "cd")
    ab = engine.spectrum_runs(pos, mu, "absorption")
    for d in range(3):
        ket, bra, scale = cd[d]
        assert np.allclose(ket, rd[:, d] / rt[d], rtol=0, atol=1e-15)
        assert np.allclose(bra, mu[:, d] / mt[d], rtol=0, atol=1e-15)
        assert abs(np.linalg.norm(ket) - 1) < 1e-14 and abs(np.linalg.norm(bra) - 1) < 1e-14
        assert scale == pytest.approx(oracle.CD_PREFACTOR * mt[d] * rt[d], rel=1e-15)
        ket, bra, scale = ab[d]
        assert np.array_equal(ket, bra)
        assert scale == pytest.approx(mt[d] ** 2, rel=1e-15)
    assert oracle.CD_PREFACTOR == pytest.approx(1.5707963267948966e-7, rel=1e-16)
    with pytest.raises(ValueError):
        engine.spectrum_runs(pos, mu, "linear")


def test_position_file_and_posfile_key(engine, tmp_path):
    rng = np.random.default_rng(5)
    pos = rng.normal(size=(6, 3)) * 10
    f = tmp_path / "pos.txt"
    f.write_text("".join("%.17g %.17g %.17g\n" % tuple(r) for r in pos))
    assert np.array_equal(engine.read_positions(str(f), 6), pos)
    with pytest.raises(engine.DQError):
        engine.read_positions(str(f), 7)
    inf = tmp_path / "infile.in"
    inf.write_text("cdspec = 1\nposfile = /data/positions.txt\ndipolefile = /data/dip.txt\n"
                   "print_partic = 1\n")
    r = engine.parse_infile(str(inf))
    assert r["cd_spec"] == 1 and r["print_partic"] == 1
    assert r["pos_file"] == "/data/positions.txt" and r["dipole_file"] == "/data/dip.txt"


def test_oracle_run_records_participation(oracle, engine):
    """the per-step participation of orc_run is that of the states it
    propagates: step 0 is psi(0), the last is psi(T) (N = 600, 5 steps)"""
    from tests.helpers import make_oracle
    sysd = ring_tube(engine)
    p, sites = make_oracle(oracle, sysd, 12, 2.0)
    psi0 = np.zeros(600, complex)
    psi0[:3] = np.sqrt([0.5, 0.3, 0.2])
    r = oracle.run(p, sites, sysd["rp"], sysd["col"], sysd["val"], psi0, 4, 5, partic=True)
    assert r["partic"][0] == pytest.approx(1 / (0.25 + 0.09 + 0.04), rel=1e-14)
    assert r["partic"][-1] == pytest.approx(1 / np.sum(np.abs(r["psi"]) ** 4), rel=1e-12)
    assert np.all(r["partic"] >= 1.0) and np.all(r["partic"] <= 600.0)

"""Readers and writers of the reference's on-disk workflow (host code of the
engine library; no GPU call).  Fixtures under tests/golden/reference_inputs are
the reference's own example input files."""
import os

import numpy as np
import pytest

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "reference_inputs")


def test_parse_reference_infile(engine):
    r = engine.parse_infile(os.path.join(GOLD, "infile.in"))
    assert r["dt"] == 1.0 and r["total_steps"] == 400 and r["m"] == 12 and r["ka_substeps"] == 5
    assert r["ensemble_mode"] == 1 and r["ensemble_runs"] == 5
    assert r["noise_model"] == engine.NOISE_ZZ and r["temperature"] == 77.0
    assert r["static_dis"] == 1 and r["sigma_static"] == 300.0
    assert r["print_pops"][:4] == [300, 301, 359, -1]
    assert r["init_state"] == 300 and r["init_from_file"] == 0
    assert r["state_looping"] == 1 and r["steps_per_loop"] == 200
    assert r["write_pops"] == 1 and r["print_all_pops"] == 1 and r["write_binary_pops"] == 0
    assert r["gauss_width"] == 1000.0 and r["absorb_spec"] == 0
    assert r["ham_file"] == "data-example/hamInner.dat"
    assert r["zz_specd_file"] == "data-example/specd_tresp.dat"
    assert r["init_state_file"] == "data-example/sample_multisite_init.dat"
    assert r["out_folder"] == "outDataFolder" and r["prop_mode"] == 0
    assert r["tk_p"] == 10 and r["tk_m_aux"] == 24 and r["tk_mat_file"] == ""    # systemClass.cu:70-71


def test_infile_taylor_krylov_keys(engine, tmp_path):
    f = tmp_path / "tk.in"
    f.write_text("propmode = PROP_MODE_TAYLOR_LANCZOS\ntkmatfile = data/hamB.dat\ntk_p = 6\ntk_m_aux = 16\n")
    r = engine.parse_infile(str(f))
    assert r["prop_mode"] == 1 and r["tk_mat_file"] == "data/hamB.dat"
    assert r["tk_p"] == 6 and r["tk_m_aux"] == 16
    f.write_text("propmode = PROP_MODE_TAYLOR_KRYLOV\n")     # not a mode the reference knows
    with pytest.raises(engine.DQError):
        engine.parse_infile(str(f))


def test_infile_forms_and_errors(engine, tmp_path):
    f = tmp_path / "a.in"
    f.write_text("!comment = 5\nm=20\n dt\t=   2.5   ! fs\nsimtype = SIM_TYPE_HSR\nhamfile = h.dat ! note\n")
    r = engine.parse_infile(str(f))
    assert r["m"] == 20 and r["dt"] == 2.5 and r["noise_model"] == engine.NOISE_HSR and r["ham_file"] == "h.dat"
    f.write_text("staticdis = 2\n")
    with pytest.raises(engine.DQError):
        engine.parse_infile(str(f))
    f.write_text("absorbspec = 1\ncdspec = 1\n")
    with pytest.raises(engine.DQError):
        engine.parse_infile(str(f))


def test_reference_spectral_density_file(engine, oracle):
    dw, om, J = engine.read_spectral_density(os.path.join(GOLD, "specd_tresp.dat"))
    assert dw == 8.135709639 and len(om) == 253
    assert om[0] == 8.135709639 and J[0] == 2.082740586 and om[1] == 16.27141928 and J[1] == 7.325068785
    c = engine.zz_oscillators(om, J, dw, 77.0)
    assert np.array_equal(c, oracle.zz_oscillators(om, J, dw, 77.0)[1])


def test_reference_state_file(engine):
    psi = engine.read_state(os.path.join(GOLD, "sample_multisite_init.dat"), 100)
    nz = np.nonzero(psi)[0]
    assert list(nz) == list(range(39, 69))
    assert np.allclose(psi[nz], 1 / np.sqrt(30), rtol=0, atol=1e-15)
    assert np.linalg.norm(psi) == pytest.approx(1.0, abs=1e-15)


def write_mm(path, n, rp, col, val, diag, symm, base):
    rows = []
    for i in range(n):
        rows.append((i, i, diag[i]))
        for p in range(rp[i], rp[i + 1]):
            j = col[p]
            if symm != "general" and j > i:
                continue
            rows.append((i, j, val[p]))
    with open(path, "w") as f:
        f.write("%%%%MatrixMarket matrix coordinate real %s\n" % symm)
        f.write("BASE_ONE\n" if base else "BASE_ZERO\n")
        f.write("%d %d %d\n" % (n, n, len(rows)))
        for i, j, v in rows:
            f.write("%d %d %.17g\n" % (i + base, j + base, v))


@pytest.mark.parametrize("symm,base", [("symmetric", 1), ("general", 0), ("hermitian", 0)])
def test_hamiltonian_file_round_trip(engine, tmp_path, symm, base):
    pos, dip, wall, rp, col, val = engine.tube([4], [3.0])
    n = len(wall)
    diag = np.linspace(-100, 100, n)
    path = str(tmp_path / "ham.dat")
    write_mm(path, n, rp, col, val, diag, symm, base)
    n2, rp2, col2, val2, es = engine.read_hamiltonian(path)
    assert n2 == n and np.array_equal(rp2, rp) and np.array_equal(col2, col)
    assert np.array_equal(val2, val) and np.array_equal(es, diag)


def test_hamiltonian_file_errors(engine, tmp_path):
    p = tmp_path / "bad.dat"
    p.write_text("%%MatrixMarket matrix coordinate real general\nBASE_ZERO\n2 2 2\n0 1 1.0\n1 0 2.0\n")
    with pytest.raises(engine.DQError):                       # not symmetric
        engine.read_hamiltonian(str(p))
    p.write_text("%%MatrixMarket matrix coordinate real symmetric\nBASE_ZERO\n2 2 3\n0 0 1.0\n1 0 2.0\n")
    with pytest.raises(engine.DQError):                       # fewer entries than nnz
        engine.read_hamiltonian(str(p))
    p.write_text("%%MatrixMarket matrix coordinate real symmetric\nBASE_TWO\n2 2 1\n0 0 1.0\n")
    with pytest.raises(engine.DQError):
        engine.read_hamiltonian(str(p))


def test_dipole_and_site_parameter_files(engine, tmp_path):
    pos, dip, wall, rp, col, val = engine.tube([2], [3.0])
    n = len(wall)
    f = tmp_path / "dips.dat"
    f.write_text("".join("%10.4f %10.4f %10.4f\n" % tuple(d) for d in dip))
    mu = engine.read_dipoles(str(f), n)
    assert np.allclose(mu, np.round(dip, 4), rtol=0, atol=1e-12)
    with pytest.raises(engine.DQError):
        engine.read_dipoles(str(f), n + 1)
    g = tmp_path / "ka.dat"
    g.write_text("".join("%g %g\n" % (50.0 + i, 300.0) for i in range(n)))
    ka = engine.read_site_columns(str(g), n, 2)
    assert ka[7, 0] == 57.0 and np.all(ka[:, 1] == 300.0)


@pytest.mark.parametrize("binary", [False, True])
def test_population_file_continuation(engine, tmp_path, binary):
    rng = np.random.default_rng(1)
    path = str(tmp_path / ("pops.bin" if binary else "pops.dat"))
    a = rng.random((5, 3))
    b = rng.random((5, 3))
    engine.write_pops(path, a, 1.0, nsims=1, binary=binary)
    engine.write_pops(path, b, 1.0, nsims=2, binary=binary)
    ns, dt, p = engine.read_pops(path, binary=binary)
    assert ns == 3 and dt == 1.0
    assert np.array_equal(p, a + b) if binary else np.allclose(p, a + b, rtol=1e-16, atol=0)
    with pytest.raises(engine.DQError):                       # totalsteps differ
        engine.write_pops(path, b[:4], 1.0, binary=binary)
    with pytest.raises(engine.DQError):                       # dt differs
        engine.write_pops(path, b, 2.0, binary=binary)
    assert not os.path.exists(path + ".tmp")
    if not binary:
        lines = open(path).read().splitlines()
        assert lines[:4] == ["numsims 3", "dt 1", "totalsteps 5", "numprintedsites 3"]
        assert lines[4] == "STEP 1" and lines[5] == "%25.17g (0)" % (a + b)[0, 0]
        assert lines[8] == "STEP 2"

"""GPU trajectories against the FP64 oracle (same Krylov depth, Pade order,
Philox stream and reorthogonalisation)."""
import numpy as np
import pytest

from tests.helpers import double_tube, make_engine, make_oracle, outer_wall_state, ring_tube

pytestmark = pytest.mark.gpu


def site_groups(n):
    return np.arange(n, dtype=np.int32)


def test_plain_fp64_n600_1000_steps(engine, oracle):
    """N=600 ring tube, m=16, dt=1 fs, 1000 steps, site pops every 10 fs"""
    sysd = ring_tube(engine)
    n, steps = 600, 1000
    psi0 = np.zeros(n, complex); psi0[0] = 1
    s = make_engine(engine, sysd, 16, 1.0)
    s.set_state(psi0)
    s.set_observables(site_groups(n), n, 10, steps)
    s.run(steps)
    pops = s.populations()[0]
    p, sites = make_oracle(oracle, sysd, 16, 1.0)
    r = oracle.run(p, sites, sysd["rp"], sysd["col"], sysd["val"], psi0, 0, steps,
                   obs_every=10, group=site_groups(n), ngroups=n)
    assert pops.shape == r["pops"].shape
    err = np.abs(pops - r["pops"]).max()
    print("max |P_gpu - P_oracle| =", err)
    assert err <= 1e-10
    assert s.error_sum()[0] == pytest.approx(r["err"], rel=1e-6)
    assert np.abs(s.get_state()[0] - r["psi"]).max() < 1e-9


def test_batched_fp64_matches_oracle_per_trajectory(engine, oracle):
    sysd = ring_tube(engine)
    n, steps, B = 600, 200, 5
    psi0 = np.zeros(n, complex); psi0[0] = 1
    ids = [3, 1, 4, 1000, 5]
    s = make_engine(engine, sysd, 16, 1.0, B=B, force_batched=True, traj_ids=ids)
    s.set_state(psi0)
    grp = (np.arange(n) // 150).astype(np.int32)
    s.set_observables(grp, 4, 10, steps)
    s.run(steps)
    pops = s.populations()
    p, sites = make_oracle(oracle, sysd, 16, 1.0)
    rb = oracle.run_batch(p, sites, sysd["rp"], sysd["col"], sysd["val"], psi0, ids, steps, 10, grp, 4)
    err = np.abs(pops - rb["pops"]).max()
    print("batched FP64 max err", err)
    assert err <= 1e-10


def test_graph_replay_is_bitwise_equal(engine):
    sysd = ring_tube(engine)
    n = 600
    psi0 = np.zeros(n, complex); psi0[0] = 1
    out = []
    for graph in (False, True):
        s = make_engine(engine, sysd, 16, 1.0, B=64, precision=engine.MIXED)
        s.set_state(psi0)
        s.set_observables((np.arange(n) // 300).astype(np.int32), 2, 5, 40)
        s.run(20, graph=graph)
        s.run(20, graph=graph)
        out.append((s.populations(), s.get_state()))
        s.close()
    assert np.array_equal(out[0][0], out[1][0])
    assert np.array_equal(out[0][1], out[1][1])


def test_mixed_precision_error_vs_oracle(engine, oracle):
    sysd = ring_tube(engine)
    n, steps, B = 600, 300, 64
    psi0 = np.zeros(n, complex); psi0[0] = 1
    s = make_engine(engine, sysd, 16, 1.0, B=B, precision=engine.MIXED)
    s.set_state(psi0)
    grp = (np.arange(n) // 300).astype(np.int32)
    s.set_observables(grp, 2, 10, steps)
    s.run(steps)
    pops = s.populations()
    p, sites = make_oracle(oracle, sysd, 16, 1.0)
    ids = [0, 17, 63]
    rb = oracle.run_batch(p, sites, sysd["rp"], sysd["col"], sysd["val"], psi0, ids, steps, 10, grp, 2)
    err = np.abs(pops[ids] - rb["pops"]).max()
    print("mixed vs FP64 oracle, max wall-population error:", err)
    assert err < 1e-6


def test_double_wall_n12000_1000_steps(engine, oracle):
    """N=12,000 double-walled tube, m=24, dt=2 fs, the plain FP64 path for
    1000 steps: populations of 2 walls x 4 axial segments every 10 steps
    (the walls are uncoupled at 2.5 nm, so whole-wall totals alone are
    constant) and the final state against the oracle"""
    sysd = double_tube(engine)
    n, steps = sysd["n"], 1000
    psi0 = outer_wall_state(sysd)
    seg = np.minimum((sysd["pos"][:, 2] / 80.0 * 4).astype(np.int32), 3)
    grp = (sysd["wall"] * 4 + seg).astype(np.int32)
    kw = dict(sigma_static=250.0, ou_sigma=400.0)
    s = make_engine(engine, sysd, 24, 2.0, **kw)
    s.set_state(psi0)
    s.set_observables(grp, 8, 10, steps)
    s.run(steps, graph=True)
    p, sites = make_oracle(oracle, sysd, 24, 2.0, **kw)
    r = oracle.run(p, sites, sysd["rp"], sysd["col"], sysd["val"], psi0, 0, steps,
                   obs_every=10, group=grp, ngroups=8)
    err = np.abs(s.populations()[0] - r["pops"]).max()
    perr = np.abs(s.get_state()[0] - r["psi"]).max()
    print("N=12000 plain FP64, 1000 steps: group pop err", err, "psi err", perr)
    assert err <= 1e-10 and perr <= 1e-9
    assert s.error_sum()[0] == pytest.approx(r["err"], rel=1e-6)


def test_ragged_batch_matches_oracle(engine, oracle):
    sysd = double_tube(engine, rings=8, order=1)
    n = sysd["n"]
    na = [960, 480, 240, 720]
    psi0 = np.zeros((4, n), complex)
    for b in range(4):
        psi0[b, 60] = 1.0
    s = make_engine(engine, sysd, 12, 1.0, B=4, force_batched=True, n_active=na)
    s.set_state(psi0)
    s.set_observables(sysd["wall"], 2, 5, 30)
    s.run(30)
    p, sites = make_oracle(oracle, sysd, 12, 1.0)
    rb = oracle.run_batch(p, sites, sysd["rp"], sysd["col"], sysd["val"], psi0, [0, 1, 2, 3], 30, 5,
                          sysd["wall"], 2, n_active=na)
    err = np.abs(s.populations() - rb["pops"]).max()
    print("ragged FP64 err", err)
    assert err <= 1e-10
    st = s.get_state()
    for b in range(4):
        assert np.all(st[b, na[b]:] == 0)


def test_spectra_correlation_and_spectrum(engine, oracle):
    sysd = ring_tube(engine)
    n, steps, B = 600, 128, 3
    mu = sysd["dip"][:, 0]
    mu2 = float(mu @ mu)
    bra = mu / np.sqrt(mu2)
    s = make_engine(engine, sysd, 16, 1.0, B=B, force_batched=True)
    s.set_state(bra.astype(complex))
    s.set_observables(None, 0, 0, steps, bra=bra.astype(complex))
    s.run(steps)
    C = s.corr()
    p, sites = make_oracle(oracle, sysd, 16, 1.0)
    rb = oracle.run_batch(p, sites, sysd["rp"], sysd["col"], sysd["val"], bra.astype(complex),
                          [0, 1, 2], steps, steps, np.zeros(n, np.int32), 1, bra=bra.astype(complex))
    assert np.abs(C - rb["corr"]).max() < 1e-10
    om, S = s.spectrum(50.0, mu2, 256)
    omo, So = oracle.spectrum(rb["corr"].mean(0), 1.0, 50.0, mu2, 256)
    assert np.allclose(om, omo)
    assert np.abs(S - So).max() < 1e-10 * np.abs(So).max()


def test_errors_are_reported(engine):
    sysd = ring_tube(engine)
    s = engine.System(600, 16, 1.0)
    with pytest.raises(engine.DQError) as ei:
        s.run(1)
    assert ei.value.code == 8
    s.set_hamiltonian(sysd["rp"], sysd["col"], sysd["val"])
    s.set_ou(50.0, 300.0)
    s.reset_disorder()
    with pytest.raises(engine.DQError) as ei:
        s.set_state(np.full(600, np.nan))
    assert ei.value.code == 5
    # huge dt: Pade scaling exponent >= 50 (padeClass.cu:618-623)
    big = engine.System(600, 16, 1e16)
    big.set_hamiltonian(sysd["rp"], sysd["col"], sysd["val"])
    big.set_ou(50.0, 0.0)
    big.reset_disorder()
    big.set_state(np.eye(1, 600, 0)[0])
    with pytest.raises(engine.DQError) as ei:
        big.run(1)
    assert ei.value.code == 6


def test_deepest_krylov_space_batched(engine, oracle):
    """m = 32, the deepest Krylov space (34 x 34 augmented Pade matrix; the
    batched update's largest vector-group counts, KW = 7 and 8): FP64 batched
    against the oracle per trajectory, and the mixed path against FP64 with
    the matrix-core SpMM and with the VALU SpMM"""
    sysd = ring_tube(engine)
    n, steps, B, m, dt = 600, 60, 4, 32, 2.0
    psi0 = np.zeros(n, complex); psi0[[0, 37]] = np.sqrt([0.6, 0.4])
    grp = (np.arange(n) // 150).astype(np.int32)
    ids = [0, 7, 2, 11]
    s = make_engine(engine, sysd, m, dt, B=B, force_batched=True, traj_ids=ids)
    s.set_state(psi0)
    s.set_observables(grp, 4, 5, steps)
    s.run(steps, graph=True)
    pops64, err64 = s.populations(), s.error_sum()
    s.close()
    p, sites = make_oracle(oracle, sysd, m, dt)
    rb = oracle.run_batch(p, sites, sysd["rp"], sysd["col"], sysd["val"], psi0, ids, steps, 5, grp, 4)
    err = np.abs(pops64 - rb["pops"]).max()
    erel = np.abs(err64 - rb["err"]).max() / np.abs(rb["err"]).max()
    print("m=32 batched FP64 vs oracle: max |P - P_oracle| %.3e, error-sum rel %.3e" % (err, erel))
    assert err <= 1e-10 and erel <= 1e-6
    for tun in (dict(spmm_rows_per_wave=16, spmm_traj_per_lane=2, spmm_phases=1, spmm_pipeline=0),
                dict(spmm_rows_per_wave=4, spmm_traj_per_lane=1, spmm_phases=1, spmm_pipeline=0)):
        s = make_engine(engine, sysd, m, dt, B=64, precision=engine.MIXED, traj_ids=ids + list(range(100, 160)))
        s.set_tuning(**tun)
        s.set_state(psi0)
        s.set_observables(grp, 4, 5, steps)
        s.run(steps, graph=True)
        d = np.abs(s.populations()[:B] - pops64).max()
        s.close()
        print("m=32 mixed (rows per wave %d) vs FP64: %.3e" % (tun["spmm_rows_per_wave"], d))
        assert d <= 1e-6

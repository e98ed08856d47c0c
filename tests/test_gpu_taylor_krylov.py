"""Taylor-Krylov propagation on the GPU against the oracle's Taylor-Krylov
cycle (FP64 bar 1e-10), its launch variants bitwise, the mixed path against
FP64, and the full-size workload with the reference's defaults (p = 10,
m_aux = 24)."""
import numpy as np
import pytest

from tests.helpers import double_tube, make_engine, make_oracle, ring_tube

pytestmark = pytest.mark.gpu


def aux_matrix(sysd):
    """auxiliary matrix: the noise-free couplings with a site-dependent diagonal"""
    n = sysd["n"]
    return sysd["rp"], sysd["col"], sysd["val"], np.linspace(-40.0, 40.0, n)


@pytest.mark.parametrize("B,batched,K", [(1, False, 1), (4, True, 1), (2, True, 2)])
def test_taylor_krylov_matches_oracle(engine, oracle, B, batched, K):
    """N = 600 tube with OU noise and static disorder, p = 6, m_aux = 16,
    200 steps of K Taylor-Krylov cycles; site populations every 10 steps,
    C(n) and the error sums (ragged n_active in the batched cases)"""
    sysd = ring_tube(engine)
    n, steps = 600, 200
    na = [600, 360, 480, 240][:B] if batched else None
    rp, col, val, adiag = aux_matrix(sysd)
    psi0 = np.zeros((B, n), complex)
    for b in range(B):
        psi0[b, [5, 66]] = np.sqrt([0.7, 0.3])
    s = make_engine(engine, sysd, 24, 1.0, B=B, K=K, force_batched=batched, n_active=na)
    s.set_taylor_krylov(6, 16, rp, col, val, adiag)
    s.set_state(psi0)
    grp = np.arange(n, dtype=np.int32)
    s.set_observables(grp, n, 10, steps, bra=psi0[0])
    s.run(steps, graph=True)
    pops, C, err = s.populations(), s.corr(), s.error_sum()
    p, sites = make_oracle(oracle, sysd, 24, 1.0, K=K)
    keep = oracle.set_taylor_krylov(p, 6, 16, rp, col, val, adiag)
    rb = oracle.run_batch(p, sites, sysd["rp"], sysd["col"], sysd["val"], psi0, list(range(B)), steps, 10,
                          grp, n, n_active=na, bra=psi0[0])
    perr = np.abs(pops - rb["pops"]).max()
    cerr = np.abs(C - rb["corr"]).max()
    eerr = np.abs(err - rb["err"]).max() / np.abs(rb["err"]).max()
    print("Taylor-Krylov B=%d K=%d: max |P - P_oracle| %.3e, max |C - C_oracle| %.3e, error-sum rel %.3e (sum %.3e)"
          % (B, K, perr, cerr, eerr, rb["err"].max()))
    assert perr <= 1e-10 and cerr <= 1e-10
    assert eerr <= 1e-6
    if batched:
        st = s.get_state()
        for b in range(B):
            assert np.all(st[b, na[b]:] == 0)
    s.close()
    del keep


def test_taylor_krylov_launch_variants_bitwise(engine):
    """graph vs eager, and 1 vs 2 concurrent trajectory slices, B = 128 mixed"""
    sysd = ring_tube(engine)
    n, steps = 600, 40
    rp, col, val, adiag = aux_matrix(sysd)
    psi0 = np.zeros(n, complex)
    psi0[100] = 1.0

    def run(graph, slices):
        s = make_engine(engine, sysd, 24, 2.0, B=128, precision=engine.MIXED)
        s.set_taylor_krylov(10, 24, rp, col, val, adiag)
        s.set_tuning(concurrent_slices=slices)
        s.set_state(psi0)
        s.run(steps, graph=graph)
        out = s.get_state()
        e = s.error_sum()
        s.close()
        return out, e

    a, ea = run(False, 1)
    b, eb = run(True, 1)
    c, ec = run(True, 2)
    assert np.array_equal(a, b) and np.array_equal(a, c)
    assert np.array_equal(ea, eb) and np.array_equal(ea, ec)


def test_taylor_krylov_with_matrix_core_spmm(engine):
    """the auxiliary Lanczos on the matrix-core SpMM and the 16-row Taylor
    powers: mixed B = 128 against the FP64 path, like the 4-row kernels.
    dt = 0.5 fs: at 2 fs (tau |A| / hbar ~ 19 on this tube) the Taylor part's
    cancellations amplify FP32 rounding to ~1e-4 for every SpMM variant"""
    sysd = ring_tube(engine)
    n, steps = 600, 80
    rp, col, val, adiag = aux_matrix(sysd)
    psi0 = np.zeros(n, complex)
    psi0[100] = 1.0
    out = {}
    for key, prec, tun in (("fp64", engine.FP64, dict(spmm_rows_per_wave=4)),
                           ("rows4", engine.MIXED, dict(spmm_rows_per_wave=4)),
                           ("rows16", engine.MIXED, dict(spmm_rows_per_wave=16, spmm_traj_per_lane=2))):
        s = make_engine(engine, sysd, 24, 0.5, B=128, precision=prec, force_batched=True)
        s.set_tuning(**tun)
        s.set_taylor_krylov(10, 24, rp, col, val, adiag)
        s.set_state(psi0)
        s.run(steps, graph=True)
        out[key] = np.abs(s.get_state()) ** 2
        s.close()
    for key in ("rows4", "rows16"):
        d = np.abs(out[key] - out["fp64"]).max()
        print("Taylor-Krylov mixed %s vs FP64: max site-population error %.3e" % (key, d))
        assert d < 1e-6


def test_taylor_krylov_switch_and_errors(engine):
    sysd = ring_tube(engine)
    n = 600
    rp, col, val, adiag = aux_matrix(sysd)
    psi0 = np.zeros(n, complex)
    psi0[0] = 1.0
    s = make_engine(engine, sysd, 16, 1.0, B=2, force_batched=True)
    for bad in ((1, 8), (4, 17), (20, 16), (4, 1)):
        with pytest.raises(engine.DQError):
            s.set_taylor_krylov(bad[0], bad[1], rp, col, val, adiag)
    bad_col = col.copy()
    bad_col[0] = 0        # row 0 holding its own diagonal
    with pytest.raises(engine.DQError):
        s.set_taylor_krylov(4, 8, rp, bad_col, val, adiag)
    s.set_taylor_krylov(4, 8, rp, col, val, adiag)
    s.set_observables(None, 0, 0, 10)
    s.set_state(psi0)
    with pytest.raises(engine.DQError):
        s.adjoint(10, ckpt=5, w_pop=None)
    # p = 0: back to the Lanczos propagator, bitwise the run of a fresh system
    s.set_taylor_krylov(0)
    s.reset_disorder()
    s.set_state(psi0)
    s.run(10)
    a = s.get_state()
    s.close()
    f = make_engine(engine, sysd, 16, 1.0, B=2, force_batched=True)
    f.set_state(psi0)
    f.run(10)
    assert np.array_equal(a, f.get_state())
    f.close()


def test_taylor_krylov_full_size_mixed_vs_fp64(engine):
    """N = 12,000 double-walled tube, B = 256 mixed, the reference's defaults
    p = 10, m_aux = 24, dt = 2 fs, 500 steps; FP64 on 64 trajectories"""
    sysd = double_tube(engine)
    n, steps = sysd["n"], 500
    rp, col, val = sysd["rp"], sysd["col"], sysd["val"]
    # groups: wall x quarter of the tube length; start on the outer wall's first quarter
    z = sysd["pos"][:, 2]
    seg = np.minimum((4 * (z - z.min()) / (z.max() - z.min() + 1e-9)).astype(np.int32), 3)
    groups = (sysd["wall"] * 4 + seg).astype(np.int32)
    psi0 = np.where(groups == 4, sysd["dip"][:, 0], 0.0).astype(complex)
    psi0 /= np.linalg.norm(psi0)

    def run(B, prec):
        s = engine.System(n, 24, 2.0, n_traj=B, precision=prec, sigma_static=250.0, seed=3)
        s.set_hamiltonian(rp, col, val, np.zeros(n))
        s.set_ou(50.0, 400.0)
        s.set_taylor_krylov(10, 24, rp, col, val, np.zeros(n))
        s.reset_disorder()
        s.set_state(psi0)
        s.set_observables(groups, 8, 5, steps)
        s.run(steps, graph=True)
        ms = s.last_run_ms()
        out = s.populations(), s.error_sum()
        s.close()
        return out, ms

    (pm, em), ms = run(256, engine.MIXED)
    (pf, ef), msf = run(64, engine.FP64)
    err = np.abs(pm[:64] - pf).max()
    print("Taylor-Krylov N=12000 B=256 mixed p=10 m_aux=24: %.3f ms/step, %.0f traj-steps/s; "
          "FP64 B=64 %.3f ms/step" % (ms / steps, 256 * steps / (ms / 1e3), msf / steps))
    print("mixed vs FP64 max group-population error %.3e; mean error sums mixed %.3e FP64 %.3e; "
          "P(T) of trajectory 0 %s" % (err, em.mean(), ef.mean(), np.round(pf[0, -1], 4)))
    assert err <= 1e-6
    assert np.all(np.isfinite(pm))

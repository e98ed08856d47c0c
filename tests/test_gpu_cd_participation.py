"""Circular-dichroism runs and the participation ratio on the GPU against the
oracle (FP64 bar 1e-10), and the mixed-precision path against FP64."""
import numpy as np
import pytest

from tests.helpers import double_tube, make_engine, make_oracle, ring_tube

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("B,batched", [(1, False), (4, True)])
def test_participation_matches_oracle(engine, oracle, B, batched):
    """N = 600 ring tube (ragged n_active in the batched case), 200 steps,
    PR(n) every step next to populations and C(n)"""
    sysd = ring_tube(engine)
    n, steps = 600, 200
    na = [600, 360, 480, 240][:B] if batched else None
    psi0 = np.zeros((B, n), complex)
    for b in range(B):
        psi0[b, [0, 61, 122]] = np.sqrt([0.5, 0.3, 0.2])
    s = make_engine(engine, sysd, 16, 1.0, B=B, force_batched=batched, n_active=na)
    s.set_state(psi0)
    bra = np.zeros(n, complex)
    bra[0] = 1.0
    s.set_observables(sysd["wall"].astype(np.int32), 1, 20, steps, bra=bra)
    s.set_participation()
    s.run(steps, graph=True)
    pr = s.participation()
    p, sites = make_oracle(oracle, sysd, 16, 1.0)
    rb = oracle.run_batch(p, sites, sysd["rp"], sysd["col"], sysd["val"], psi0, list(range(B)), steps, 20,
                          sysd["wall"].astype(np.int32), 1, n_active=na, bra=bra, partic=True)
    rel = np.abs(pr - rb["partic"]).max() / np.abs(rb["partic"]).max()
    print("participation B=%d: max rel err %.3e, PR(T) %s" % (B, rel, pr[:, -1]))
    assert rel <= 1e-10
    assert pr[0, 0] == pytest.approx(1 / (0.25 + 0.09 + 0.04), rel=1e-14)
    assert np.abs(s.corr() - rb["corr"]).max() <= 1e-10
    # switching it off leaves the other observables as they were
    s.set_observables(None, 0, 0, steps, bra=bra)
    with pytest.raises(engine.DQError):
        s.participation()
    s.close()


def test_participation_mixed_vs_fp64(engine):
    """B = 64 trajectories of the N = 600 tube, 400 steps"""
    sysd = ring_tube(engine)
    n, steps = 600, 400
    psi0 = np.zeros(n, complex)
    psi0[300] = 1.0

    def run(prec):
        s = make_engine(engine, sysd, 16, 2.0, B=64, precision=prec, force_batched=True)
        s.set_state(psi0)
        s.set_observables(None, 0, 0, steps)
        s.set_participation()
        s.run(steps, graph=True)
        out = s.participation()
        s.close()
        return out

    a, b = run(engine.MIXED), run(engine.FP64)
    rel = np.abs(a - b).max() / np.abs(b).max()
    print("participation mixed vs FP64: max rel err %.3e, mean PR(T) %.2f" % (rel, b[:, -1].mean()))
    # PR = 1 / sum p_i^2 with sum p_i^2 ~ 1 / 300 here: site-population
    # errors of ~1e-7 (FP32 basis vectors) move it by a few 1e-5 relative
    assert rel <= 1e-4
    assert np.all(b[:, 0] == 1.0)


def test_cd_direction_matches_oracle(engine, oracle):
    """one CD direction (ket R x mu_d, bra mu_d) for each d: C(n) and the
    scaled spectrum against the oracle, N = 600, B = 3, 128 steps"""
    sysd = ring_tube(engine)
    steps, nfft = 128, 256
    runs = engine.spectrum_runs(sysd["pos"], sysd["dip"], "cd")
    p, sites = make_oracle(oracle, sysd, 16, 1.0)
    for d, (ket, bra, scale) in enumerate(runs):
        s = make_engine(engine, sysd, 16, 1.0, B=3, force_batched=True)
        s.set_state(ket)
        s.set_observables(None, 0, 0, steps, bra=bra)
        s.run(steps)
        C = s.corr()
        om, S = s.spectrum(40.0, scale, nfft)
        s.close()
        rb = oracle.run_batch(p, sites, sysd["rp"], sysd["col"], sysd["val"], ket, [0, 1, 2], steps, steps,
                              np.zeros(600, np.int32), 1, bra=bra)
        assert np.abs(C - rb["corr"]).max() <= 1e-10, d
        omo, So = oracle.spectrum(rb["corr"].mean(0), 1.0, 40.0, scale, nfft)
        assert np.allclose(om, omo)
        assert np.abs(S - So).max() <= 1e-10 * np.abs(So).max(), d


def test_isotropic_cd_full_size_mixed_vs_fp64(engine):
    """isotropic CD of the N = 12,000 double-walled tube: three directions,
    B = 256 realisations each in mixed precision, 2048 steps of 1 fs;
    the FP64 path on the first 64 realisations is the reference"""
    sysd = double_tube(engine)
    n = sysd["n"]
    steps, nfft, sigma = 2048, 4096, 500.0
    runs = engine.spectrum_runs(sysd["pos"], sysd["dip"], "cd")

    def iso(B, prec):
        specs, corrs, ms = [], [], 0.0
        for d, (ket, bra, scale) in enumerate(runs):
            s = engine.System(n, 24, 1.0, n_traj=B, precision=prec, sigma_static=250.0, seed=7)
            s.set_hamiltonian(sysd["rp"], sysd["col"], sysd["val"], np.zeros(n))
            s.set_ou(50.0, 400.0)
            s.set_trajectories(np.arange(B, dtype=np.uint64) + 1000 * d)
            s.reset_disorder()
            s.set_state(ket)
            s.set_observables(None, 0, 0, steps, bra=bra)
            s.run(steps, graph=True)
            ms += s.last_run_ms()
            om, sp = s.spectrum(sigma, scale, nfft)
            specs.append(sp)
            corrs.append(s.corr()[:64])
            s.close()
        return om, sum(specs) / 3, corrs, ms

    def spectrum_of(c, scale):
        """the engine's spectrum formula in numpy, for a subset average"""
        k = np.arange(steps + 1)
        w = np.where(k < steps, np.exp(-0.5 * ((k + 1) * 1.0 / sigma) ** 2), 1.0)
        x = np.zeros(nfft, complex)
        x[:steps + 1] = c * w
        return np.fft.ifft(x) * nfft * scale / steps

    om, sm, cm, msm = iso(256, engine.MIXED)
    om64, sf, cf, _ = iso(64, engine.FP64)
    assert np.array_equal(om, om64)
    # |C(n)| <= 1 (unit bra and ket, norm-preserving propagation)
    cerr = max(np.abs(a - b).max() for a, b in zip(cm, cf))
    a = sum(spectrum_of(c.mean(0), r[2]) for c, r in zip(cm, runs)) / 3
    b = sum(spectrum_of(c.mean(0), r[2]) for c, r in zip(cf, runs)) / 3
    rel = np.linalg.norm(a - b) / np.linalg.norm(b)
    print("isotropic CD N=12000 B=256 x 3 directions, %d steps: %.3f ms/step, %.0f traj-steps/s"
          % (steps, msm / (3 * steps), 3 * 256 * steps / (msm / 1e3)))
    print("mixed vs FP64 on 64 realisations: max |C_mixed - C_fp64| %.3e (max |C| %s), "
          "isotropic CD spectrum relative L2 error %.3e"
          % (cerr, ["%.3e" % np.abs(c).max() for c in cf], rel))
    print("B=256 CD spectrum: max |Re S| %.3e at %.1f cm^-1"
          % (np.abs(sm.real).max(), om[np.argmax(np.abs(sm.real))]))
    # C(n) carries the phase, which drifts by ~1e-5 rad over 2048 steps with
    # FP32 basis vectors (populations, phase-free, stay within 1e-6); the
    # averaged spectrum is held to the bar of the absorption spectrum
    assert cerr <= 1e-4
    assert rel <= 1e-4
    assert np.all(np.isfinite(sm))

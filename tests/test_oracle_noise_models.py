"""HSR white noise and Zhong-Zhao real noise of the oracle (the other two
site-energy models of systemClass.cu:1881-2013 besides Kubo-Anderson)."""
import numpy as np
import pytest

KB = 0.695035            # cm^-1 / K (includeFiles.cu:13)


def params(oracle, n, dt=1.0, seed=3):
    return oracle.make_params(n, 8, dt, sigma_static=0.0, seed=seed)


def test_hsr_noise_is_white_gaussian(oracle):
    n, T = 4000, 40
    p = params(oracle, n)
    sig = np.linspace(50.0, 150.0, n)
    s = oracle.Sites(n, noise=oracle.NOISE_HSR, hsr_sigma=sig)
    de = np.zeros(n)
    X = np.empty((T, n))
    for k in range(T):
        de = oracle.noise_step(p, s, 7, k, de)
        X[k] = de
    z = X / sig
    assert abs(z.mean()) < 5 / np.sqrt(z.size)
    assert abs(z.var() - 1) < 5 * np.sqrt(2 / z.size)
    lag1 = (z[1:] * z[:-1]).mean()
    assert abs(lag1) < 5 / np.sqrt(z[1:].size)
    # each step is a fresh draw of the same stream: replaying step k gives the same values
    assert np.array_equal(oracle.noise_step(p, s, 7, 5, np.zeros(n)), X[5])


def test_zz_phase_is_uniform_and_cos_matches_numpy(oracle):
    u = np.array([oracle.zz_phase(3, 1, i, k) for i in range(200) for k in range(50)])
    assert np.all((u >= 0) & (u < 1))
    assert abs(u.mean() - 0.5) < 5 * np.sqrt(1 / 12 / u.size)
    assert abs(u.var() - 1 / 12) < 0.01
    rng = np.random.default_rng(0)
    for _ in range(200):
        uu, w, t = rng.random(), rng.random() * 0.06, rng.random() * 4000
        assert oracle.zz_cos(uu, w, t) == pytest.approx(np.cos(2 * np.pi * (uu + w * t)), abs=2e-13)


def spectral_density(nosc=120, domega=8.0):
    om = domega * np.arange(1, nosc + 1)
    lam, wc = 35.0, 160.0                                   # Drude-Lorentz, cm^-1
    J = 2 * lam * om * wc / (om ** 2 + wc ** 2)
    return om, J, domega


def test_zz_amplitudes_classical_and_quantum_limits(oracle):
    """sum_k c_k^2/2 is the site-energy variance (1/pi) sum J (n+1) domega with
    the reference's G = J / (1 - exp(-omega/kT)) = J (n + 1): kB T lambda_d in
    the classical limit (lambda_d = (1/pi) sum J/omega domega; the symmetrised
    coth form would give twice that, the reference keeps n + 1), and
    (1/pi) sum J domega at zero temperature"""
    om, J, dw = spectral_density()
    lam_d = np.sum(J / om) * dw / np.pi
    w, c = oracle.zz_oscillators(om, J, dw, 1e6)
    assert np.sum(c ** 2) / 2 == pytest.approx(KB * 1e6 * lam_d, rel=1e-3)
    w, c = oracle.zz_oscillators(om, J, dw, 1e-3)
    assert np.sum(c ** 2) / 2 == pytest.approx(np.sum(J) * dw / np.pi, rel=1e-12)
    # omega / 2pi in 1/fs with the reference's 5308.84 cm^-1 fs
    assert w == pytest.approx(om / 5308.84 / (2 * np.pi), rel=1e-15)


def test_zz_noise_autocorrelation(oracle):
    """<de(0) de(t)> over sites = sum_k c_k^2/2 cos(2 pi w_k t) (random phases)"""
    n = 6000
    om, J, dw = spectral_density(nosc=40, domega=20.0)
    w, c = oracle.zz_oscillators(om, J, dw, 300.0)
    p = params(oracle, n, dt=2.0)
    s = oracle.Sites(n, noise=oracle.NOISE_ZZ, zz_w2pi=w, zz_coeff=c)
    d0 = oracle.noise_step(p, s, 2, 0, np.zeros(n))
    var = np.sum(c ** 2) / 2
    for step in (0, 3, 10, 25):
        dt_ = oracle.noise_step(p, s, 2, step, np.zeros(n))
        theory = np.sum(c ** 2 / 2 * np.cos(2 * np.pi * w * step * 2.0))
        assert (d0 * dt_).mean() == pytest.approx(theory, abs=6 * var / np.sqrt(n))
    # deterministic given (seed, trajectory): another trajectory has other phases
    d1 = oracle.noise_step(p, s, 3, 0, np.zeros(n))
    assert not np.array_equal(d0, d1)


def test_noise_models_propagate(oracle):
    """a short run under each model: the noise enters the diagonal only, so the
    norm is kept and the none model equals OU with sigma = 0"""
    rp, col, val = None, None, None
    pos, dip, wall = oracle.tube([3], [3.0])
    rp, col, val = oracle.couplings_csr(pos, dip)
    n = len(wall)
    p = oracle.make_params(n, 8, 1.0, sigma_static=100.0, seed=1, reorth=2)
    psi0 = np.zeros(n, complex); psi0[0] = 1
    om, J, dw = spectral_density(nosc=30, domega=25.0)
    w, c = oracle.zz_oscillators(om, J, dw, 77.0)
    for s in (oracle.Sites(n, noise=oracle.NOISE_HSR, hsr_sigma=120.0),
              oracle.Sites(n, noise=oracle.NOISE_ZZ, zz_w2pi=w, zz_coeff=c)):
        r = oracle.run(p, s, rp, col, val, psi0, 0, 20)
        assert abs(np.linalg.norm(r["psi"]) - 1) < 1e-12
    a = oracle.run(p, oracle.Sites(n, noise=oracle.NOISE_NONE), rp, col, val, psi0, 0, 20)
    b = oracle.run(p, oracle.Sites(n, tau=50.0, sigma=0.0), rp, col, val, psi0, 0, 20)
    assert np.array_equal(a["psi"], b["psi"])

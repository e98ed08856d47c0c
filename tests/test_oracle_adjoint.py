"""Oracle adjoint against central finite differences of the forward map."""
import numpy as np
import pytest


@pytest.fixture(scope="module")
def small(oracle):
    pos, dip, wall = oracle.tube([3], [3.0])          # 180 sites, 3 rings
    rp, col, val = oracle.couplings_csr(pos, dip)
    n = len(wall)
    grp = (np.arange(n) >= 90).astype(np.int32)       # "far half" of the tube
    psi0 = np.zeros(n, complex); psi0[0] = 1
    bra = dip[:, 0] / np.linalg.norm(dip[:, 0])
    return dict(n=n, rp=rp, col=col, val=val, grp=grp, psi0=psi0, bra=bra)


@pytest.mark.parametrize("K,m,ckpt", [(1, 8, 7), (2, 12, 50)])
def test_gradient_matches_central_differences(oracle, small, K, m, ckpt):
    s = small
    n, T = s["n"], 30
    p = oracle.make_params(n, m, 1.0, K=K, sigma_static=200.0, seed=1, reorth=1)
    kw = dict(tau=50.0, sigma=300.0, e_site=0.0)
    ct = 0.3 * np.exp(-0.01j * np.arange(T + 1))
    args = (s["rp"], s["col"], s["val"], s["psi0"], 0, T, s["grp"], 2, [0.0, 1.0])
    loss, ge, gg = oracle.adjoint(p, oracle.Sites(n, **kw), *args, bra=s["bra"], w_spec=0.5,
                                  corr_target=ct, ckpt=ckpt)
    L0 = oracle.loss_fd(p, kw, *args, bra=s["bra"], w_spec=0.5, corr_target=ct)
    assert loss == pytest.approx(L0, rel=1e-13)
    h = 1e-3                                   # cm^-1
    for i in (0, 5, 100, 170):
        e = np.zeros(n); e[i] = h
        Lp = oracle.loss_fd(p, kw, *args, bra=s["bra"], w_spec=0.5, corr_target=ct, e_shift=e)
        Lm = oracle.loss_fd(p, kw, *args, bra=s["bra"], w_spec=0.5, corr_target=ct, e_shift=-e)
        fd = (Lp - Lm) / (2 * h)
        assert abs(ge[i] - fd) <= 1e-5 * abs(fd) + 1e-12
    hg = 1e-6
    pp = oracle.make_params(n, m, 1.0, K=K, sigma_static=200.0, seed=1, reorth=1, g=1 + hg)
    pm = oracle.make_params(n, m, 1.0, K=K, sigma_static=200.0, seed=1, reorth=1, g=1 - hg)
    fd = (oracle.loss_fd(pp, kw, *args, bra=s["bra"], w_spec=0.5, corr_target=ct)
          - oracle.loss_fd(pm, kw, *args, bra=s["bra"], w_spec=0.5, corr_target=ct)) / (2 * hg)
    assert abs(gg - fd) <= 1e-6 * abs(fd)


def test_checkpoint_interval_does_not_change_gradient(oracle, small):
    s = small
    n, T = s["n"], 20
    p = oracle.make_params(n, 8, 1.0, sigma_static=200.0, seed=2, reorth=1)
    sites = oracle.Sites(n, tau=50.0, sigma=300.0, e_site=0.0)
    args = (s["rp"], s["col"], s["val"], s["psi0"], 3, T, s["grp"], 2, [0.0, 1.0])
    a = oracle.adjoint(p, sites, *args, ckpt=1)
    b = oracle.adjoint(p, sites, *args, ckpt=6)
    c = oracle.adjoint(p, sites, *args, ckpt=100)
    assert np.array_equal(a[1], b[1]) and np.array_equal(a[1], c[1]) and a[2] == c[2]


def test_taylor_remainder_is_second_order(oracle, small):
    s = small
    n, T = s["n"], 15
    p = oracle.make_params(n, 8, 1.0, sigma_static=200.0, seed=4, reorth=1)
    kw = dict(tau=50.0, sigma=300.0, e_site=0.0)
    args = (s["rp"], s["col"], s["val"], s["psi0"], 0, T, s["grp"], 2, [0.0, 1.0])
    L0, ge, gg = oracle.adjoint(p, oracle.Sites(n, **kw), *args)
    v = np.random.default_rng(0).normal(size=n)
    rem = []
    for t in (1.0, 0.5, 0.25):
        L = oracle.loss_fd(p, kw, *args, e_shift=t * v)
        rem.append(abs(L - L0 - t * ge @ v))
    assert rem[1] / rem[0] == pytest.approx(0.25, abs=0.05)
    assert rem[2] / rem[1] == pytest.approx(0.25, abs=0.05)


def test_three_term_reverse_equals_full_projection_reverse(oracle, small):
    """the engine's O(m) reverse (three-term recurrence on the stored basis)
    against the reverse of the full-projection map as computed"""
    s = small
    n, T = s["n"], 25
    p = oracle.make_params(n, 12, 2.0, sigma_static=200.0, seed=5, reorth=1)
    sites = oracle.Sites(n, tau=50.0, sigma=300.0, e_site=0.0)
    args = (s["rp"], s["col"], s["val"], s["psi0"], 1, T, s["grp"], 2, [0.3, 1.0])
    try:
        oracle.set_adjoint_mode(0)
        a = oracle.adjoint(p, sites, *args, bra=s["bra"], w_spec=0.2)
    finally:
        oracle.set_adjoint_mode(1)
    b = oracle.adjoint(p, sites, *args, bra=s["bra"], w_spec=0.2)
    assert np.abs(a[1] - b[1]).max() < 1e-11 * np.abs(a[1]).max()
    assert abs(a[2] - b[2]) < 1e-11 * abs(a[2])


def spectrum_weights(T, dt, sigma_fs, mu2, nfft):
    """u(n) such that sum_n u(n)|C(n) - C_t(n)|^2 is the L2 distance of the
    zero-padded windowed spectra S = (mu2/T) FFT(win * C): Parseval for an
    unnormalised DFT of length nfft gives sum_k |S_k|^2 = nfft sum_n |s_n|^2"""
    n = np.arange(T + 1)
    # window at t = (n+1) dt for n < T, none on the last sample (spectraClass.cu:591-598)
    win = np.where(n < T, np.exp(-0.5 * ((n + 1) * dt / sigma_fs) ** 2), 1.0)
    return nfft * (mu2 / T) ** 2 * win ** 2


def test_weighted_correlation_loss_is_the_spectrum_distance(oracle, small):
    """the time-domain weights turn the C(t) term into the L2 mismatch of the
    windowed, zero-padded spectra computed by the oracle's own spectrum"""
    s = small
    n, T, dt, nfft = s["n"], 40, 1.0, 128
    p = oracle.make_params(n, 8, dt, sigma_static=200.0, seed=1, reorth=1)
    run = oracle.run(p, oracle.Sites(n, tau=50.0, sigma=300.0, e_site=0.0), s["rp"], s["col"], s["val"],
                     s["psi0"], 0, T, obs_every=T, bra=s["bra"])
    shifted = oracle.run(p, oracle.Sites(n, tau=50.0, sigma=300.0, e_site=40.0), s["rp"], s["col"], s["val"],
                         s["psi0"], 0, T, obs_every=T, bra=s["bra"])
    u = spectrum_weights(T, dt, 20.0, 1.0, nfft)
    direct = float(np.sum(u * np.abs(run["corr"] - shifted["corr"]) ** 2))
    _, sa = oracle.spectrum(run["corr"], dt, 20.0, 1.0, nfft)
    _, sb = oracle.spectrum(shifted["corr"], dt, 20.0, 1.0, nfft)
    spec = float(np.sum(np.abs(sa - sb) ** 2))
    assert direct == pytest.approx(spec, rel=1e-12)


def test_weighted_spectrum_loss_gradient_matches_central_differences(oracle, small):
    s = small
    n, T, dt = s["n"], 30, 1.0
    p = oracle.make_params(n, 8, dt, sigma_static=200.0, seed=1, reorth=1)
    kw = dict(tau=50.0, sigma=300.0, e_site=0.0)
    target = oracle.run(p, oracle.Sites(n, tau=50.0, sigma=300.0, e_site=25.0), s["rp"], s["col"],
                        s["val"], s["psi0"], 0, T, obs_every=T, bra=s["bra"])["corr"]
    u = spectrum_weights(T, dt, 15.0, 1.0, 64)
    args = (s["rp"], s["col"], s["val"], s["psi0"], 0, T, s["grp"], 2, [0.0, 1.0])
    opt = dict(bra=s["bra"], w_spec=1e-3, corr_target=target, corr_weight=u)
    loss, ge, gg = oracle.adjoint(p, oracle.Sites(n, **kw), *args, ckpt=7, **opt)
    assert loss == pytest.approx(oracle.loss_fd(p, kw, *args, **opt), rel=1e-13)
    h = 1e-3
    for i in (0, 50, 170):
        e = np.zeros(n); e[i] = h
        fd = (oracle.loss_fd(p, kw, *args, e_shift=e, **opt)
              - oracle.loss_fd(p, kw, *args, e_shift=-e, **opt)) / (2 * h)
        assert abs(ge[i] - fd) <= 1e-5 * abs(fd) + 1e-12

"""Oracle trajectories against dense propagation and closed forms."""
import numpy as np
import pytest


@pytest.fixture(scope="module")
def ring(oracle):
    pos, dip, wall = oracle.tube([10], [3.0])
    rp, col, val = oracle.couplings_csr(pos, dip)
    n = 600
    H = np.zeros((n, n))
    for i in range(n):
        H[i, col[rp[i]:rp[i + 1]]] = val[rp[i]:rp[i + 1]]
    return rp, col, val, H


def dense_prop(H, psi, t, hbar):
    w, V = np.linalg.eigh(H)
    return V @ (np.exp(-1j * w * t / hbar) * (V.T @ psi))


@pytest.mark.parametrize("K,tol", [(1, 1e-3), (2, 1e-7), (4, 5e-12)])
def test_static_run_matches_dense_expm(oracle, ring, K, tol):
    rp, col, val, H = ring
    n, steps, dt = 600, 20, 1.0
    p = oracle.make_params(n, 16, dt, K=K, sigma_static=200.0, seed=1, reorth=2)
    sites = oracle.Sites(n, tau=50.0, sigma=0.0)
    psi0 = np.zeros(n, complex); psi0[0] = 1
    r = oracle.run(p, sites, rp, col, val, psi0, 0, steps)
    e0, _ = oracle.init_disorder(p, sites, 0)
    ref = dense_prop(H + np.diag(e0), psi0, steps * dt, oracle.hbar_ref())
    assert np.abs(r["psi"] - ref).max() < tol
    assert abs(np.linalg.norm(r["psi"]) - 1) < 1e-14


def test_rabi_dimer_closed_form(oracle):
    J, delta, dt, steps = 80.0, 300.0, 1.0, 200
    rp = np.array([0, 1, 2], np.int32)
    col = np.array([1, 0], np.int32)
    val = np.array([J, J])
    hb = oracle.hbar_ref()
    p = oracle.make_params(2, 2, dt, K=1, reorth=1)
    sites = oracle.Sites(2, tau=50.0, sigma=0.0, e_site=[0.0, delta])
    psi0 = np.array([1, 0], complex)
    r = oracle.run(p, sites, rp, col, val, psi0, 0, steps, obs_every=1,
                   group=np.array([0, 1], np.int32), ngroups=2)
    t = np.arange(steps + 1) * dt
    om = np.sqrt(delta ** 2 + 4 * J ** 2)
    P2 = (2 * J / om) ** 2 * np.sin(om * t / (2 * hb)) ** 2
    assert np.abs(r["pops"][:, 1] - P2).max() < 1e-12
    assert np.abs(r["pops"].sum(1) - 1).max() < 1e-13


def test_noisy_run_conserves_norm_and_is_reproducible(oracle, ring):
    rp, col, val, H = ring
    n = 600
    p = oracle.make_params(n, 16, 1.0, K=1, sigma_static=200.0, seed=1, reorth=2)
    sites = oracle.Sites(n, tau=50.0, sigma=300.0)
    psi0 = np.zeros(n, complex); psi0[0] = 1
    grp = np.arange(n, dtype=np.int32)
    a = oracle.run(p, sites, rp, col, val, psi0, 0, 50, obs_every=10, group=grp, ngroups=n)
    b = oracle.run(p, sites, rp, col, val, psi0, 0, 50, obs_every=10, group=grp, ngroups=n)
    c = oracle.run(p, sites, rp, col, val, psi0, 1, 50, obs_every=10, group=grp, ngroups=n)
    assert np.array_equal(a["pops"], b["pops"])
    assert not np.allclose(a["pops"], c["pops"])
    assert np.abs(a["pops"].sum(1) - 1).max() < 1e-13
    # Expokit estimate of the unsplit dt=1 fs, m=16 step: order 1e-4..1e-2 per step
    assert 1e-6 < a["err"] / 50 < 1e-1


def test_batch_equals_single_runs(oracle, ring):
    rp, col, val, H = ring
    n = 600
    p = oracle.make_params(n, 8, 1.0, sigma_static=100.0, seed=3, reorth=2)
    sites = oracle.Sites(n, tau=50.0, sigma=200.0)
    psi0 = np.zeros(n, complex); psi0[5] = 1
    grp = (np.arange(n) // 300).astype(np.int32)
    rb = oracle.run_batch(p, sites, rp, col, val, psi0, [0, 7], 20, 5, grp, 2, nthreads=2)
    for k, t in enumerate([0, 7]):
        r = oracle.run(p, sites, rp, col, val, psi0, t, 20, obs_every=5, group=grp, ngroups=2)
        assert np.array_equal(rb["pops"][k], r["pops"])


def test_ragged_batch_equals_truncated_system(oracle):
    pos, dip, wall = oracle.tube([4, 4], [3.0, 6.0], order=1)
    rp, col, val = oracle.couplings_csr(pos, dip)
    n = len(wall)
    na = 240          # two rings of both walls
    p = oracle.make_params(n, 8, 1.0, sigma_static=100.0, seed=3, reorth=2)
    sites = oracle.Sites(n, tau=50.0, sigma=200.0)
    psi0 = np.zeros(n, complex); psi0[60] = 1
    rb = oracle.run_batch(p, sites, rp, col, val, psi0, [4], 10, 5, wall, 2, n_active=[na])
    pos2, dip2, wall2 = oracle.tube([2, 2], [3.0, 6.0], order=1)
    rp2, col2, val2 = oracle.couplings_csr(pos2, dip2)
    p2 = oracle.make_params(na, 8, 1.0, sigma_static=100.0, seed=3, reorth=2)
    r = oracle.run(p2, oracle.Sites(na, tau=50.0, sigma=200.0), rp2, col2, val2, psi0[:na], 4, 10,
                   obs_every=5, group=wall2, ngroups=2)
    assert np.abs(rb["pops"][0] - r["pops"]).max() < 1e-13

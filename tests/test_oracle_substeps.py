"""Substep control from the Expokit estimate (oracle): the controlled run
meets its per-step tolerance against exact propagation of the same noisy
Hamiltonian, where the reference's single Krylov cycle per dt does not."""
import numpy as np
import pytest

from oracle import pyoracle as O


@pytest.fixture(scope="module")
def two_rings():
    pos, dip, wall = O.tube([2], [3.0])
    rp, col, val = O.couplings_csr(pos, dip)
    n = len(wall)
    J = np.zeros((n, n))
    for i in range(n):
        J[i, col[rp[i]:rp[i + 1]]] = val[rp[i]:rp[i + 1]]
    return n, rp, col, val, J


def exact_run(p, sites, J, psi0, traj, nsteps, dt):
    """exp(-i dt H(t_k) / hbar) per step, H(t_k) = diag(e0 + de_k) + J, by
    eigendecomposition (the noise of step k from the oracle's Philox stream)"""
    e0, de = O.init_disorder(p, sites, traj)
    psi = psi0 / np.linalg.norm(psi0)
    hb = O.hbar_ref()
    for k in range(nsteps):
        de = O.noise_step(p, sites, traj, k, de)
        w, V = np.linalg.eigh(J + np.diag(de + e0))
        psi = V @ (np.exp(-1j * w * dt / hb) * (V.T @ psi))
    return psi


def test_controlled_run_meets_its_tolerance(two_rings):
    n, rp, col, val, J = two_rings
    dt, m, steps, tol = 2.0, 12, 40, 1e-10
    p = O.make_params(n, m, dt, K=1, sigma_static=250.0, seed=1, reorth=O.REORTH_GRAM)
    sites = O.Sites(n, tau=50.0, sigma=400.0)
    psi0 = np.zeros(n, complex); psi0[3] = 1
    ids = [0, 5]
    r = O.run_controlled(p, sites, rp, col, val, psi0, ids, steps, tol, 64, 10)
    ref = [exact_run(p, sites, J, psi0, t, steps, dt) for t in ids]
    err_c = max(np.abs(r["psi"][k] - ref[k]).max() for k in range(2))
    one = [O.run(p, sites, rp, col, val, psi0, t, steps)["psi"] for t in ids]
    err_1 = max(np.abs(one[k] - ref[k]).max() for k in range(2))
    print("K per chunk", r["k_of_step"][::10], "redone", r["rejects"], "error", err_c, "one cycle per dt", err_1)
    assert r["k_of_step"][-1] > 1 and r["rejects"] >= 1
    assert err_1 > 1e3 * tol                    # the reference's single cycle is far off
    assert err_c <= steps * tol                 # the controlled run meets tol per step
    # every kept chunk met the criterion: rerunning with its final K from the start
    K = int(r["k_of_step"][-1])
    pk = O.make_params(n, m, dt, K=K, sigma_static=250.0, seed=1, reorth=O.REORTH_GRAM)
    if np.all(r["k_of_step"] == K):
        fixed = O.run(pk, sites, rp, col, val, psi0, ids[0], steps)
        assert np.abs(fixed["psi"] - r["psi"][0]).max() < 1e-12


def test_control_with_a_loose_tolerance_keeps_one_cycle(two_rings):
    n, rp, col, val, J = two_rings
    p = O.make_params(n, 12, 1.0, K=1, sigma_static=250.0, seed=1, reorth=O.REORTH_GRAM)
    sites = O.Sites(n, tau=50.0, sigma=400.0)
    psi0 = np.zeros(n, complex); psi0[3] = 1
    r = O.run_controlled(p, sites, rp, col, val, psi0, [2], 20, 1.0, 64, 10)
    one = O.run(p, sites, rp, col, val, psi0, 2, 20)
    assert r["rejects"] == 0 and np.all(r["k_of_step"] == 1)
    assert np.array_equal(r["psi"][0], one["psi"])
    assert r["err"][0] == one["err"]

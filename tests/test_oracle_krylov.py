"""Oracle Lanczos, Pade and Krylov step against closed forms and scipy."""
import json
import os

import numpy as np
import pytest
import scipy.linalg as sl

GOLD = os.path.join(os.path.dirname(__file__), "golden")


@pytest.fixture(scope="module")
def ring(oracle):
    pos, dip, wall = oracle.tube([10], [3.0])
    rp, col, val = oracle.couplings_csr(pos, dip)
    n = 600
    H = np.zeros((n, n))
    for i in range(n):
        H[i, col[rp[i]:rp[i + 1]]] = val[rp[i]:rp[i + 1]]
    return rp, col, val, H


def test_constants_match_reference_run_log(oracle):
    # out:76-79 and out:83 of the reference's example run
    assert round(oracle.hbar_ref(), 2) == 5308.84
    assert float("%.6g" % (-1.0 / oracle.hbar_ref())) == -0.000188365
    c = oracle.pade_coeffs(6)
    logged = [1, 0.5, 0.113636, 0.0151515, 0.00126263, 6.31313e-05, 1.50313e-06]
    assert [float("%.6g" % x) for x in c] == logged


def test_spmv_matches_dense(oracle, ring):
    rp, col, val, H = ring
    rng = np.random.default_rng(0)
    d = rng.normal(0, 300, 600)
    x = rng.normal(size=600) + 1j * rng.normal(size=600)
    for g in (1.0, 0.7):
        ref = (g * H + np.diag(d)) @ x
        assert np.abs(oracle.spmv(rp, col, val, d, x, g=g) - ref).max() < 1e-12 * np.abs(ref).max()


@pytest.mark.parametrize("mode", [0, 1, 2])
def test_lanczos_invariants(oracle, ring, mode):
    rp, col, val, H = ring
    rng = np.random.default_rng(1)
    d = rng.normal(0, 200, 600)
    Hd = H + np.diag(d)
    psi = rng.normal(size=600) + 1j * rng.normal(size=600)
    psi /= np.linalg.norm(psi)
    m = 16
    Q, a, b, nh, broke = oracle.lanczos(rp, col, val, d, psi, m, reorth=mode)
    assert broke == 0
    Qm = Q[1:m + 1].T
    T = np.diag(a[1:m + 1]) + np.diag(b[1:m], 1) + np.diag(b[1:m], -1)
    assert np.abs(Qm.conj().T @ Qm - np.eye(m)).max() < 1e-12
    R = Hd @ Qm - Qm @ T
    R[:, m - 1] -= b[m] * Q[m + 1]
    assert np.abs(R).max() < 1e-11 * np.abs(Hd).max()
    assert nh == pytest.approx(np.linalg.norm(Hd @ Q[m + 1]), rel=1e-12)
    # polynomial exactness: p(H) psi = Q_m p(T) e1 for deg < m
    e1 = np.zeros(m); e1[0] = 1
    v = psi.copy(); w = e1.copy()
    for _ in range(m - 1):
        v = Hd @ v / 4e4
        w = T @ w / 4e4
    assert np.abs(v - Qm @ w).max() < 1e-10 * np.abs(v).max()


def test_reorth_modes_agree(oracle, ring):
    rp, col, val, H = ring
    rng = np.random.default_rng(2)
    d = rng.normal(0, 200, 600)
    psi = np.zeros(600, complex); psi[0] = 1
    out = [oracle.lanczos(rp, col, val, d, psi, 24, reorth=m) for m in (1, 2)]
    assert np.abs(out[0][1] - out[1][1]).max() < 1e-9
    assert np.abs(out[0][2] - out[1][2]).max() < 1e-9


def test_lanczos_exhausts_small_space(oracle):
    """N <= m: the Krylov space is exhausted, the guard stops the recurrence"""
    n = 4
    rp = np.array([0, 1, 3, 5, 6], np.int32)
    col = np.array([1, 0, 2, 1, 3, 2], np.int32)
    val = np.array([100., 100., -50., -50., 80., 80.])
    d = np.array([0., 200., -100., 50.])
    psi = np.array([1, 0.5, 0.2, 0.1], complex); psi /= np.linalg.norm(psi)
    Q, a, b, nh, broke = oracle.lanczos(rp, col, val, d, psi, 8, reorth=1)
    assert broke == 1
    assert b[4] == 0.0 and np.all(a[5:9] == 0.0) and nh == 0.0
    H = np.diag(d); H[0, 1] = H[1, 0] = 100; H[1, 2] = H[2, 1] = -50; H[2, 3] = H[3, 2] = 80
    T = np.diag(a[1:5]) + np.diag(b[1:4], 1) + np.diag(b[1:4], -1)
    assert np.allclose(np.sort(np.linalg.eigvalsh(T)), np.linalg.eigvalsh(H), atol=1e-9)


def test_pade_matches_scipy(oracle):
    for c in json.load(open(os.path.join(GOLD, "expm_kat.json"))):
        n = c["n"]
        A = (np.array(c["A"][0]) + 1j * np.array(c["A"][1])).reshape(n, n)
        E = (np.array(c["E"][0]) + 1j * np.array(c["E"][1])).reshape(n, n)
        got, s = oracle.pade_expm(A)
        assert np.abs(got - E).max() < 5e-13 * max(1.0, np.abs(E).max())


def test_pade_scalar_closed_form(oracle):
    c = oracle.pade_coeffs(6)
    for z in (0.3 + 0.1j, -0.2j, 3.0 - 4.0j, -40j):
        got, s = oracle.pade_expm(np.array([[z]]))
        x = z / 2 ** s
        num = sum(c[k] * x ** k for k in range(7))
        den = sum(c[k] * (-x) ** k for k in range(7))
        r = (num / den) ** (2 ** s)
        assert abs(got[0, 0] - r) < 1e-13 * abs(r)
        assert abs(got[0, 0] - np.exp(z)) < 1e-13 * abs(np.exp(z))
        # s = halvings until |z| 2^-s <= 1/2 (true norm)
        assert abs(z) / 2 ** s <= 0.5 and (s == 0 or abs(z) / 2 ** (s - 1) > 0.5)


def test_pade_rotation_generator(oracle):
    for th in (0.1, 1.3, 17.0):
        A = -1j * th * np.array([[0, 1], [1, 0]])
        got, s = oracle.pade_expm(A)
        ref = np.cos(th) * np.eye(2) - 1j * np.sin(th) * np.array([[0, 1], [1, 0]])
        assert np.abs(got - ref).max() < 1e-13


def test_tbar_structure_and_phi_blocks(oracle):
    """exp of the augmented Tbar carries phi_1, phi_2 in the extra columns"""
    rng = np.random.default_rng(3)
    m = 5
    a = np.zeros(m + 2); b = np.zeros(m + 2)
    a[1:m + 1] = rng.normal(size=m)
    b[1:m] = np.abs(rng.normal(size=m - 1))
    Tb = oracle.build_tbar(m, a, b)
    T = np.diag(a[1:m + 1]) + np.diag(b[1:m], 1) + np.diag(b[1:m], -1)
    assert np.array_equal(Tb[:m, :m], T)
    assert Tb[0, m] == 1 and Tb[m, m + 1] == 1 and Tb.sum() == pytest.approx(T.sum() + 2)
    tau = 0.7j
    E, s = oracle.pade_expm(tau * Tb)
    w, V = np.linalg.eigh(T)
    e1 = np.zeros(m); e1[0] = 1
    z = tau * w
    phi1 = V @ (((np.exp(z) - 1) / z) * (V.T @ e1))
    phi2 = V @ (((np.exp(z) - 1 - z) / z ** 2) * (V.T @ e1))
    assert np.abs(E[:m, m] - tau * phi1).max() < 1e-13
    assert np.abs(E[:m, m + 1] - tau ** 2 * phi2).max() < 1e-13
    assert E[m, m + 1] == pytest.approx(tau)


def test_expokit_estimate_recomputed(oracle):
    m = 4
    E = np.zeros((6, 6), complex)
    E[m - 1, m] = 3e-5 + 4e-5j
    E[m - 1, m + 1] = 1e-9
    est = oracle.expokit_err(m, 200.0, 1e4, E)
    err1 = 200 * 5e-5
    err2 = 200 * 1e4 * 1e-9
    assert est == pytest.approx(err2 / (1 - err2 / err1))
    E[m - 1, m + 1] = 1e-7
    assert oracle.expokit_err(m, 200.0, 1e4, E) == pytest.approx(200 * 5e-5)

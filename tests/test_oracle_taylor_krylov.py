"""The oracle's Taylor-Krylov propagator (taylorKrylovClass.cu:133-434)
pinned by dense linear algebra: with an exact auxiliary Krylov space the
cycle is sum_{k<p} (cA)^k psi / k! + c^p phi_p(cB) A^p psi, which is
expm(cA) psi when B = A (scipy.linalg.expm), and the phi_p part is checked
against an eigendecomposition of B with the phi series."""
import math

import numpy as np
import pytest
import scipy.linalg as sla

from tests.helpers import make_oracle, ring_tube


def dense_to_split(H):
    """symmetric dense -> (rowptr, col, val) of the off-diagonal part, diag"""
    n = H.shape[0]
    rp, col, val = [0], [], []
    for i in range(n):
        for j in range(n):
            if i != j and H[i, j] != 0.0:
                col.append(j)
                val.append(H[i, j])
        rp.append(len(col))
    return (np.array(rp, np.int32), np.array(col, np.int32), np.array(val), np.diag(H).copy())


def random_sym(n, rng, scale=300.0, density=0.5):
    M = rng.normal(size=(n, n)) * scale
    M = np.where(rng.random((n, n)) < density, M, 0.0)
    M = np.triu(M, 1)
    return M + M.T + np.diag(rng.normal(size=n) * scale)


def phi_times(Bm, c, p, v):
    """phi_p(c B) v = sum_j (c B)^j v / (j + p)! by eigendecomposition"""
    lam, U = np.linalg.eigh(Bm)
    z = c * lam
    f = np.zeros_like(z, dtype=complex)
    term = np.ones_like(z, dtype=complex) / math.factorial(p)
    for j in range(60):
        f += term
        term = term * z / (j + p + 1)
    return U @ (f * (U.T @ v))


def test_tk_coefficients(oracle):
    tau, hbar = 2.0, oracle.hbar_ref()
    c = -1j * tau / hbar
    co = oracle.tk_coeffs(12, tau)
    for k in range(12):
        assert co[k] == pytest.approx(c ** k / math.factorial(k), rel=1e-14, abs=0)
    # truncated exponential series: |exp(c x) - sum_{k<p} co_k x^k| ~ |c x|^p / p!
    x = 3000.0
    s = sum(co[k] * x ** k for k in range(12))
    assert abs(s - np.exp(c * x)) <= 2 * abs(c * x) ** 12 / math.factorial(12)


def test_augmented_tbar_gives_phi_blocks(oracle):
    """exp(c Tbar_aug) holds c^j phi_j(c T) e1 in column m + j - 1"""
    rng = np.random.default_rng(1)
    m, aug = 6, 5
    alpha = np.concatenate([[0.0], rng.normal(size=m) * 200])
    beta = np.concatenate([[0.0], np.abs(rng.normal(size=m)) * 100])
    tb = oracle.build_tbar_aug(m, aug, alpha, beta)
    assert tb.shape == (m + aug, m + aug)
    T = np.diag(alpha[1:m + 1]) + np.diag(beta[1:m], 1) + np.diag(beta[1:m], -1)
    assert np.array_equal(tb[:m, :m], T)
    c = -1j * 2e-3
    E = sla.expm(c * tb)
    e1 = np.eye(m)[0]
    for j in range(1, aug + 1):
        assert np.allclose(E[:m, m + j - 1], c ** j * phi_times(T, c, j, e1), rtol=0, atol=1e-14)
    # aug = 2 (the Lanczos propagator's matrix) is the leading block
    assert np.array_equal(oracle.build_tbar_aug(m, 2, alpha, beta), tb[:m + 2, :m + 2])


@pytest.mark.parametrize("p", [2, 5, 10])
def test_tk_cycle_equals_expm_when_aux_is_the_hamiltonian(oracle, p):
    rng = np.random.default_rng(p)
    n = 14
    H = random_sym(n, rng)
    rp, col, val, diag = dense_to_split(H)
    prm = oracle.make_params(n, 4, 1.0, reorth=oracle.REORTH_GRAM)
    keep = oracle.set_taylor_krylov(prm, p, n, rp, col, val, diag)
    psi = rng.normal(size=n) + 1j * rng.normal(size=n)
    psi /= np.linalg.norm(psi)
    tau = 1.0
    out, err = oracle.tk_cycle(prm, rp, col, val, diag, psi, tau)
    ref = sla.expm(-1j * tau / oracle.hbar_ref() * H) @ psi
    assert np.abs(out - ref).max() <= 1e-12
    assert 0 <= err <= 1e-10
    del keep


def test_tk_cycle_with_a_different_aux_matrix(oracle):
    """B != A: the Taylor part follows A, the remainder phi_p(cB) A^p psi"""
    rng = np.random.default_rng(7)
    n, p, tau = 12, 4, 2.0
    A = random_sym(n, rng)
    Bm = A + np.diag(rng.normal(size=n) * 80.0)
    ra, ca, va, da = dense_to_split(A)
    rb, cb, vb, db = dense_to_split(Bm)
    prm = oracle.make_params(n, 4, tau, reorth=oracle.REORTH_GRAM)
    keep = oracle.set_taylor_krylov(prm, p, n, rb, cb, vb, db)
    psi = rng.normal(size=n) + 1j * rng.normal(size=n)
    psi /= np.linalg.norm(psi)
    out, _ = oracle.tk_cycle(prm, ra, ca, va, da, psi, tau)
    c = -1j * tau / oracle.hbar_ref()
    taylor = sum(np.linalg.matrix_power(c * A, k) @ psi / math.factorial(k) for k in range(p))
    ref = taylor + c ** p * phi_times(Bm, c, p, np.linalg.matrix_power(A, p) @ psi)
    assert np.abs(out - ref).max() <= 1e-12
    # and it is not expm(cA) psi: the auxiliary matrix matters
    assert np.abs(out - sla.expm(c * A) @ psi).max() > 1e-6
    del keep


def test_tk_run_matches_the_exact_propagator_on_the_tube(oracle, engine):
    """N = 600 tube without noise or disorder, aux matrix = H: 20 steps of
    Taylor-Krylov (p = 6, m_aux = 24) against expm(-i H dt / hbar)^20 psi;
    a too-short auxiliary space (m_aux = 16) is visibly worse, and the
    summed error estimate bounds the actual error"""
    sysd = ring_tube(engine)
    n = 600
    H = np.zeros((n, n))
    for i in range(n):
        H[i, sysd["col"][sysd["rp"][i]:sysd["rp"][i + 1]]] = sysd["val"][sysd["rp"][i]:sysd["rp"][i + 1]]
    psi0 = np.zeros(n, complex)
    psi0[0] = 1.0
    U = sla.expm(-1j / oracle.hbar_ref() * H)
    ex = psi0
    for _ in range(20):
        ex = U @ ex
    sites = oracle.Sites(n, e_site=0.0, noise=oracle.NOISE_NONE)
    errs = {}
    for m_aux in (24, 16):
        pt, _ = make_oracle(oracle, sysd, 24, 1.0, sigma_static=0.0)
        keep = oracle.set_taylor_krylov(pt, 6, m_aux, sysd["rp"], sysd["col"], sysd["val"], np.zeros(n))
        rt = oracle.run(pt, sites, sysd["rp"], sysd["col"], sysd["val"], psi0, 0, 20)
        errs[m_aux] = np.abs(rt["psi"] - ex).max()
        assert errs[m_aux] <= rt["err"]
        del keep
    assert errs[24] <= 1e-12
    assert errs[16] > 1e3 * errs[24]


def test_tk_ragged_batch_cuts_both_matrices(oracle, engine):
    """a trajectory with n_active = 360 equals a run of the 360-site leading
    block of both the Hamiltonian and the auxiliary matrix"""
    sysd = ring_tube(engine)
    n, na = 600, 360
    psi0 = np.zeros((2, n), complex)
    psi0[:, 10] = 1.0
    p, sites = make_oracle(oracle, sysd, 12, 1.0)
    aux_diag = np.linspace(-50.0, 50.0, n)
    keep = oracle.set_taylor_krylov(p, 5, 12, sysd["rp"], sysd["col"], sysd["val"], aux_diag)
    rb = oracle.run_batch(p, sites, sysd["rp"], sysd["col"], sysd["val"], psi0, [3, 4], 10, 5,
                          np.zeros(n, np.int32), 1, n_active=[n, na], bra=psi0[0], partic=True)
    # the cut system by hand
    rp, col, val = sysd["rp"], sysd["col"], sysd["val"]
    crp, ccol, cval = [0], [], []
    for i in range(na):
        for q in range(rp[i], rp[i + 1]):
            if col[q] < na:
                ccol.append(col[q]); cval.append(val[q])
        crp.append(len(ccol))
    crp, ccol, cval = np.array(crp, np.int32), np.array(ccol, np.int32), np.array(cval)
    pc, _ = make_oracle(oracle, dict(n=na), 12, 1.0)
    sc = oracle.Sites(na, tau=50.0, sigma=300.0, e_site=0.0)
    keep2 = oracle.set_taylor_krylov(pc, 5, 12, crp, ccol, cval, aux_diag[:na])
    r = oracle.run(pc, sc, crp, ccol, cval, psi0[1, :na], 4, 10, bra=psi0[0, :na], partic=True)
    assert np.allclose(rb["corr"][1], r["corr"], rtol=0, atol=1e-14)
    assert np.allclose(rb["partic"][1], r["partic"], rtol=1e-13)
    del keep, keep2

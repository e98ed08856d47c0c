"""GPU unit seams against the oracle: noise, SpMV, Pade, FFT."""
import json
import os

import numpy as np
import pytest

from tests.helpers import make_engine, make_oracle, ring_tube

pytestmark = pytest.mark.gpu
GOLD = os.path.join(os.path.dirname(__file__), "golden")


@pytest.mark.parametrize("B,batched", [(1, False), (3, True)])
def test_noise_bit_exact_fp64(engine, oracle, B, batched):
    sysd = ring_tube(engine)
    s = make_engine(engine, sysd, 16, 1.0, B=B, S=3, force_batched=batched, traj_ids=[5, 0, 9][:B])
    s.set_state(np.eye(1, 600, 0)[0])
    p, sites = make_oracle(oracle, sysd, 16, 1.0, S=3)
    d0 = s.diagonal()
    for k, t in enumerate([5, 0, 9][:B]):
        e0, de = oracle.init_disorder(p, sites, t)
        assert np.array_equal(d0[k], de + e0)
    for step in range(3):
        s.update_noise()
    d3 = s.diagonal()
    for k, t in enumerate([5, 0, 9][:B]):
        e0, de = oracle.init_disorder(p, sites, t)
        for step in range(3):
            de = oracle.ou_step(p, sites, t, step, de)
        assert np.array_equal(d3[k], de + e0), "noise stream differs from the oracle"


def test_noise_mixed_is_rounded_fp64(engine, oracle):
    sysd = ring_tube(engine)
    s = make_engine(engine, sysd, 16, 1.0, B=2, precision=engine.MIXED)
    s.set_state(np.eye(1, 600, 0)[0])
    s.update_noise()
    p, sites = make_oracle(oracle, sysd, 16, 1.0)
    for t in range(2):
        e0, de = oracle.init_disorder(p, sites, t)
        de = oracle.ou_step(p, sites, t, 0, de)
        assert np.array_equal(s.diagonal()[t], (de + e0).astype(np.float32).astype(np.float64))


@pytest.mark.parametrize("prec,batched,tol", [(0, False, 1e-14), (0, True, 1e-14), (1, False, 2e-6)])
def test_spmv_matches_oracle(engine, oracle, prec, batched, tol):
    sysd = ring_tube(engine)
    B = 1 if not batched else 2
    s = make_engine(engine, sysd, 16, 1.0, B=B, precision=prec, force_batched=batched)
    rng = np.random.default_rng(0)
    x = rng.normal(size=600) + 1j * rng.normal(size=600)
    for t in range(B if prec == 0 else 1):
        d = s.diagonal()[t]
        ref = oracle.spmv(sysd["rp"], sysd["col"], sysd["val"], d, x)
        got = s.spmv(t, x)
        assert np.abs(got - ref).max() < tol * np.abs(ref).max()


def test_pade_kernel_matches_oracle(engine, oracle):
    rng = np.random.default_rng(4)
    for m in (4, 16, 24, 32):
        a = rng.normal(0, 1e4, (3, m))
        b = np.abs(rng.normal(0, 5e3, (3, m)))
        theta = 2.0 / oracle.hbar_ref()
        E, s = engine.pade_tbar(m, a, b, theta)
        for c in range(3):
            al = np.zeros(m + 2); be = np.zeros(m + 2)
            al[1:m + 1] = a[c]; be[1:m + 1] = b[c]
            Tb = oracle.build_tbar(m, al, be)
            Eo, so = oracle.pade_expm(-1j * theta * Tb)
            assert s[c] == so
            assert np.abs(E[c] - Eo).max() < 1e-12 * np.abs(Eo).max()


def test_pade_kernel_matches_scipy_fixture(engine):
    c = json.load(open(os.path.join(GOLD, "expm_kat.json")))[-1]
    n = c["n"]
    E_ref = (np.array(c["E"][0]) + 1j * np.array(c["E"][1])).reshape(n, n)
    E, s = engine.pade_tbar(n - 2, np.array(c["alpha"]), np.array(c["beta"] + [0.0]), c["theta"])
    assert np.abs(E[0] - E_ref).max() < 1e-12


def test_fft_kernel_matches_numpy(engine):
    for c in json.load(open(os.path.join(GOLD, "fft_kat.json"))):
        x = np.array(c["x"][0]) + 1j * np.array(c["x"][1])
        y = np.array(c["y"][0]) + 1j * np.array(c["y"][1])
        assert np.abs(engine.fft(x, +1)[0] - y).max() < 1e-12 * np.abs(y).max()
    x = np.random.default_rng(1).normal(size=(3, 8192)) + 0j
    assert np.abs(engine.fft(x, -1) - np.fft.fft(x, axis=1)).max() < 1e-10

"""Oracle noise stream: Philox known answers, Box-Muller pieces, OU statistics."""
import json
import math
import os

import numpy as np
import pytest

GOLD = os.path.join(os.path.dirname(__file__), "golden")


def test_philox_matches_numpy_known_answers(oracle):
    kat = json.load(open(os.path.join(GOLD, "philox_kat.json")))
    for c in kat:
        ctr = [int(x) for x in c["ctr"]]
        key = [int(x) for x in c["key"]]
        got = oracle.philox(ctr, key)
        assert [int(x) for x in got] == [int(x) for x in c["out"]]


def test_log_sqrt_sincos_accuracy(oracle):
    rng = np.random.default_rng(5)
    xs = np.concatenate([rng.random(3000), 2.0 ** -np.arange(1, 53), [1.0, 0.5, 0.999999]])
    for x in xs:
        ref = math.log(x)
        got = oracle.log(x)
        assert abs(got - ref) <= 4e-16 * max(abs(ref), 1e-300) + 1e-300
        assert abs(oracle.sqrt(x) - math.sqrt(x)) <= 2.3e-16 * math.sqrt(x)
    for u in np.concatenate([rng.random(3000), np.arange(16) / 16.0]):
        s, c = oracle.sincos_2pi(u)
        assert abs(s - math.sin(2 * math.pi * u)) < 1e-15
        assert abs(c - math.cos(2 * math.pi * u)) < 1e-15
    assert oracle.sqrt(0.0) == 0.0
    assert oracle.log(1.0) == 0.0


def test_normals_moments_and_streams(oracle):
    z = np.array([oracle.normal(3, 1, i, 0, 3, 0) for i in range(100000)])
    assert abs(z.mean()) < 0.015
    assert abs(z.var() - 1.0) < 0.015
    assert abs((z ** 4).mean() - 3.0) < 0.08
    # different streams / substeps / trajectories give different numbers
    a = oracle.normal(1, 0, 5, 7, 3, 0)
    assert a != oracle.normal(1, 0, 5, 7, 3, 1)
    assert a != oracle.normal(1, 0, 5, 7, 2, 0)
    assert a != oracle.normal(1, 1, 5, 7, 3, 0)
    # substep s uses normal s%4 of block s//4
    z4 = oracle.normal4(1, 0, 5, 7, 3, 1)
    assert oracle.normal(1, 0, 5, 7, 3, 6) == z4[2]


def test_ou_stationary_variance_and_autocorrelation(oracle):
    n = 4000
    sig, tau, dt = 300.0, 50.0, 2.0
    p = oracle.make_params(n, 4, dt, S=1, seed=9)
    sites = oracle.Sites(n, tau=tau, sigma=sig)
    e0, de = oracle.init_disorder(p, sites, 0)
    assert abs(de.std() / sig - 1) < 0.05
    lag = 10
    x0 = de.copy()
    x = de.copy()
    for s in range(lag):
        x = oracle.ou_step(p, sites, 0, s, x)
    # Euler-Maruyama: x_{k+1} = (1 - dt/tau) x_k + sig sqrt(2 dt/tau) xi
    a = 1 - dt / tau
    corr = np.mean(x0 * x) / sig ** 2
    assert abs(corr - a ** lag) < 0.05
    var_inf = 2 * dt / tau / (1 - a * a) * sig ** 2
    assert abs(x.var() / var_inf - 1) < 0.08


def test_ou_update_matches_reference_formula(oracle):
    """one step written out as gpuSysHelper.cu:112 with the oracle's normals"""
    n = 16
    p = oracle.make_params(n, 4, 2.0, S=3, seed=4)
    tau = np.linspace(30, 80, n)
    sig = np.linspace(100, 400, n)
    sites = oracle.Sites(n, tau=tau, sigma=sig)
    e0, de = oracle.init_disorder(p, sites, 2)
    got = oracle.ou_step(p, sites, 2, 5, de)
    subdt = 2.0 / 3
    for i in range(n):
        cur = de[i]
        for s in range(3):
            xi = oracle.normal(4, 2, i, 5, 3, s)
            cur = cur - cur * (1.0 / tau[i]) * subdt + sig[i] * math.sqrt(2.0 / tau[i]) * xi * math.sqrt(subdt)
        assert got[i] == pytest.approx(cur, rel=0, abs=1e-12 * abs(cur) + 1e-13)


def test_static_disorder(oracle):
    n = 20000
    p = oracle.make_params(n, 4, 1.0, sigma_static=250.0, seed=2)
    sites = oracle.Sites(n, tau=50.0, sigma=0.0, e_site=100.0)
    e0, de = oracle.init_disorder(p, sites, 3)
    assert abs(e0.mean() - 100.0) < 5.0
    assert abs(e0.std() / 250.0 - 1) < 0.02
    assert np.all(de == 0.0)
    e0b, _ = oracle.init_disorder(p, sites, 4)
    assert not np.allclose(e0, e0b)

"""HSR, Zhong-Zhao and constant-Hamiltonian noise on the GPU against the
oracle: bit-exact diagonals and FP64 trajectories."""
import numpy as np
import pytest

from tests.helpers import make_oracle, ring_tube

pytestmark = pytest.mark.gpu


def zz_density(nosc=253, domega=8.135709639):
    om = domega * np.arange(1, nosc + 1)
    J = 2 * 35.0 * om * 160.0 / (om ** 2 + 160.0 ** 2)
    return om, J, domega


def systems(engine, oracle, sysd, model, B=1, batched=False, precision=0, dt=1.0):
    n = sysd["n"]
    s = engine.System(n, 16, dt, n_traj=B, precision=precision, sigma_static=200.0, seed=1,
                      force_batched=batched, noise_model=model)
    s.set_hamiltonian(sysd["rp"], sysd["col"], sysd["val"], np.zeros(n))
    sig = np.linspace(80.0, 160.0, n)
    om, J, dw = zz_density()
    coeff = engine.zz_oscillators(om, J, dw, 77.0)
    if model == engine.NOISE_HSR:
        s.set_hsr(sig)
    elif model == engine.NOISE_ZZ:
        s.set_zz(om, coeff)
    s.reset_disorder()
    p, _ = make_oracle(oracle, sysd, 16, dt)
    w, c = oracle.zz_oscillators(om, J, dw, 77.0)
    sites = oracle.Sites(n, e_site=0.0, noise=model, hsr_sigma=sig, zz_w2pi=w, zz_coeff=c)
    return s, p, sites


@pytest.mark.parametrize("model", [1, 2, 3])
@pytest.mark.parametrize("B,batched", [(1, False), (3, True)])
def test_noise_model_diagonal_bit_exact(engine, oracle, model, B, batched):
    sysd = ring_tube(engine)
    s, p, sites = systems(engine, oracle, sysd, model, B=B, batched=batched)
    s.set_state(np.eye(1, 600, 0)[0])
    for step in range(3):
        s.update_noise()
    d3 = s.diagonal()
    for t in range(B):
        e0, de = oracle.init_disorder(p, sites, t)
        for step in range(3):
            de = oracle.noise_step(p, sites, t, step, de)
        assert np.array_equal(d3[t], de + e0), (model, t)
    s.close()


def test_zz_diagonal_mixed_is_rounded_fp64(engine, oracle):
    sysd = ring_tube(engine)
    s, p, sites = systems(engine, oracle, sysd, engine.NOISE_ZZ, B=2, precision=engine.MIXED)
    s.set_state(np.eye(1, 600, 0)[0])
    s.update_noise()
    s.update_noise()
    for t in range(2):
        e0, de = oracle.init_disorder(p, sites, t)
        de = oracle.noise_step(p, sites, t, 1, de)
        assert np.array_equal(s.diagonal()[t], (de + e0).astype(np.float32).astype(np.float64))
    s.close()


@pytest.mark.parametrize("model", [1, 2, 3])
def test_noise_model_trajectory_matches_oracle(engine, oracle, model):
    """plain FP64, N = 600, 300 steps, site populations every 10 fs"""
    sysd = ring_tube(engine)
    n, steps = 600, 300
    s, p, sites = systems(engine, oracle, sysd, model)
    psi0 = np.zeros(n, complex); psi0[0] = 1
    s.set_state(psi0)
    grp = np.arange(n, dtype=np.int32)
    s.set_observables(grp, n, 10, steps)
    s.run(steps, graph=True)
    r = oracle.run(p, sites, sysd["rp"], sysd["col"], sysd["val"], psi0, 0, steps, obs_every=10,
                   group=grp, ngroups=n)
    err = np.abs(s.populations()[0] - r["pops"]).max()
    print("noise model", model, "max |P_gpu - P_oracle|", err)
    assert err <= 1e-10
    s.close()

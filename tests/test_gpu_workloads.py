"""Full-size batched workloads: the mixed-precision path against the FP64 path
on the same trajectories (the FP64 batched path is itself held to 1e-10 of
the FP64 oracle by test_gpu_run.py and test_gpu_tuning.py).

* ragged batch: B = 256 double-walled tubes with ring counts uniform in
  [20, 200] per wall (N_b = 120 r_b, 2,400 .. 24,000 sites), m = 24,
  dt = 2 fs, 2000 steps, autotuned launch choices;
* spectra (the three-direction average, B = 3 x 512): test_gpu_isotropic.py."""
import numpy as np
import pytest

from tests.helpers import double_tube

pytestmark = pytest.mark.gpu


def ragged_states(sysd, na):
    """psi_b proportional to mu_x on the outer-wall sites of tube b"""
    B, n = len(na), sysd["n"]
    psi = np.zeros((B, n), complex)
    outer = np.where(sysd["wall"] == 1, sysd["dip"][:, 0], 0.0)
    for b in range(B):
        v = np.zeros(n)
        v[:na[b]] = outer[:na[b]]
        psi[b] = v / np.linalg.norm(v)
    return psi


def test_ragged_batch_mixed_vs_fp64(engine):
    sysd = double_tube(engine, rings=200, order=1)
    n = sysd["n"]
    rng = np.random.default_rng(2024)
    rings = rng.integers(20, 201, size=256)
    na = (120 * rings).astype(np.int32)
    ids = np.arange(256, dtype=np.uint64)
    # groups: wall x 50-ring segment (ring-major order: site // 120 is the ring)
    ring = np.arange(n) // 120
    groups = (sysd["wall"] * 4 + np.minimum(ring // 50, 3)).astype(np.int32)
    psi0 = ragged_states(sysd, na)
    steps, every = 2000, 10

    def mk(B, prec):
        s = engine.System(n, 24, 2.0, n_traj=B, precision=prec, sigma_static=250.0, seed=1)
        s.set_hamiltonian(sysd["rp"], sysd["col"], sysd["val"], np.zeros(n))
        s.set_ou(50.0, 400.0)
        s.set_trajectories(ids[:B], na[:B])
        s.reset_disorder()
        s.set_state(psi0[:B])
        s.set_observables(groups, 8, every, steps)
        return s

    s = mk(256, engine.MIXED)
    tuning = s.autotune(reps=1)
    s.run(steps, graph=True)
    ms = s.last_run_ms()
    pm = s.populations()
    sm = s.get_state()
    s.close()
    f = mk(64, engine.FP64)
    f.run(steps, graph=True)
    pf = f.populations()
    sf = f.get_state()
    f.close()
    err = np.abs(pm[:64] - pf).max()
    serr = np.abs(np.abs(sm[:64]) ** 2 - np.abs(sf) ** 2).max()
    print("ragged B=256 (sum N_b = %d, N_max = %d): %.3f ms/step, %.0f traj-steps/s, tuning %s"
          % (na.sum(), n, ms / steps, 256 * steps / (ms / 1e3), tuning))
    print("mixed vs FP64 (64 trajectories, 2000 steps): max group population error %.3e, "
          "final site population error %.3e" % (err, serr))
    assert err <= 1e-6
    for b in range(256):
        assert np.all(sm[b, na[b]:] == 0)

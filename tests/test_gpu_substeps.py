"""Substep control on the GPU (padeClass.cu:749-782 supplies the estimate):
the engine's controlled run against the oracle's controlled run."""
import numpy as np
import pytest

from tests.helpers import make_engine, make_oracle, ring_tube

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("graph", [False, True])
def test_controlled_run_against_the_oracle(engine, oracle, graph):
    """600-site ring tube, dt = 2 fs, m = 16, 60 steps in chunks of 10 at
    1e-10 per dt, FP64: the same K sequence and redo count as the oracle,
    populations and final states within 1e-10"""
    sysd = ring_tube(engine)
    n, dt, m, steps, tol = sysd["n"], 2.0, 16, 60, 1e-10
    ids = [0, 5, 11]
    grp = (np.arange(n) // 75).astype(np.int32)
    psi0 = np.zeros(n, complex); psi0[3] = 1
    s = make_engine(engine, sysd, m, dt, B=3, force_batched=True, traj_ids=ids)
    s.set_state(psi0)
    s.set_observables(grp, 8, 5, steps)
    s.set_krylov_control(tol, k_max=64, chunk=10)
    s.run(steps, graph=graph)
    K, rej = s.krylov_control()
    pops, psi, err = s.populations(), s.get_state(), s.error_sum()
    assert s.get_step() == steps
    s.close()
    p, sites = make_oracle(oracle, sysd, m, dt)
    r = oracle.run_controlled(p, sites, sysd["rp"], sysd["col"], sysd["val"], psi0, ids, steps, tol, 64, 10,
                              obs_every=5, group=grp, ngroups=8)
    print("K", K, "redone", rej, "oracle K", r["k_of_step"][-1], "redone", r["rejects"])
    assert K == r["k_of_step"][-1] and rej == r["rejects"] and K > 1
    assert np.abs(pops - r["pops"]).max() <= 1e-10
    assert np.abs(psi - r["psi"]).max() <= 1e-10
    assert np.abs(err - r["err"]).max() <= 1e-9 * max(1.0, np.abs(r["err"]).max())


def test_control_off_restores_the_configured_substeps(engine):
    sysd = ring_tube(engine)
    s = make_engine(engine, sysd, 16, 1.0, B=1, K=2)
    s.set_state(np.eye(sysd["n"])[0])
    s.set_krylov_control(1e-30, k_max=8, chunk=5)
    s.run(10)
    assert s.krylov_control()[0] == 8          # K stops at k_max when tol cannot be met
    s.set_krylov_control(0.0)
    assert s.krylov_control()[0] == 2
    s.close()

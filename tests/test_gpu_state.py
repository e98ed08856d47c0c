"""State handling on the GPU: the initial state is normalised like the
reference's (systemClass.cu:1383-1481), a run resumes bit for bit from a
snapshot taken in another system, and the adjoint leaves the system at the
forward run's final state."""
import os
import tempfile

import numpy as np
import pytest

from tests.helpers import make_engine, make_oracle, ring_tube

pytestmark = pytest.mark.gpu


def _groups(n):
    return (np.arange(n) // 75).astype(np.int32)


def test_set_state_normalises_like_the_oracle(engine, oracle):
    """a scaled state and a state with weight beyond n_active propagate as
    their normalised active parts, and match the oracle (which normalises too)"""
    sysd = ring_tube(engine)
    n, steps = sysd["n"], 40
    rng = np.random.default_rng(11)
    psi = rng.normal(size=n) + 1j * rng.normal(size=n)
    na = np.array([600, 437], np.int32)
    grp = _groups(n)

    def run(states, n_active):
        s = make_engine(engine, sysd, 16, 1.0, B=2, force_batched=True, n_active=n_active)
        s.set_state(states)
        s.set_observables(grp, 8, 5, steps)
        s.run(steps)
        out = s.populations(), s.get_state()
        s.close()
        return out

    pa, sa = run(np.stack([3.7 * psi, 0.01 * psi]), na)
    cut = psi.copy(); cut[437:] = 0
    pb, sb = run(np.stack([psi / np.linalg.norm(psi), cut / np.linalg.norm(cut)]), na)
    assert np.abs(pa - pb).max() <= 1e-13
    assert abs(pa[0, 0].sum() - 1) < 1e-13 and abs(pa[1, 0].sum() - 1) < 1e-13
    p, sites = make_oracle(oracle, sysd, 16, 1.0)
    r = oracle.run_batch(p, sites, sysd["rp"], sysd["col"], sysd["val"], np.stack([3.7 * psi, 0.01 * cut]),
                         [0, 1], steps, 5, grp, 8, n_active=na)
    assert np.abs(pa - r["pops"]).max() <= 1e-10
    s = make_engine(engine, sysd, 16, 1.0, B=1)
    with pytest.raises(engine.DQError):
        s.set_state(np.zeros(n))
    s.close()


@pytest.mark.parametrize("precision,batched", [(0, False), (0, True), (1, True)])
def test_snapshot_resume_is_bitwise(engine, precision, batched):
    """run(30), snapshot to a file, new system, restore, run(30) == run(60)"""
    sysd = ring_tube(engine)
    n = sysd["n"]
    B = 1 if not batched else 3
    psi0 = np.zeros(n, complex); psi0[0] = 1
    grp = _groups(n)

    def mk():
        s = make_engine(engine, sysd, 16, 1.0, B=B, S=3, precision=precision, force_batched=batched,
                        traj_ids=None if B == 1 else [7, 2, 9])
        s.set_observables(grp, 8, 5, 60)
        return s

    u = mk()
    u.set_state(psi0)
    u.run(60, graph=True)
    ref = u.populations(), u.get_state(), u.error_sum(), u.diagonal()
    u.close()

    a = mk()
    a.set_state(psi0)
    a.run(30, graph=True)
    snap = a.snapshot()
    a.close()
    path = os.path.join(tempfile.mkdtemp(), "snap.npz")
    np.savez(path, **snap)
    loaded = dict(np.load(path))
    loaded["step"] = int(loaded["step"])
    assert loaded["step"] == 30

    b = mk()
    b.restore(loaded)
    assert np.array_equal(b.snapshot()["noise"], snap["noise"])
    b.run(30, graph=False)
    pops, psi, err, diag = b.populations(), b.get_state(), b.error_sum(), b.diagonal()
    b.close()
    # records from the restored step on, the final state, error sums, diagonal
    assert np.array_equal(pops[:, 6:], ref[0][:, 6:])
    assert np.array_equal(psi, ref[1])
    assert np.array_equal(err, ref[2])
    assert np.array_equal(diag, ref[3])


def test_adjoint_leaves_the_forward_final_state(engine):
    """after adjoint(T) with checkpoints inside the run: the error sums and the
    diagonal are the forward run's, and run(k) continues it as an unsplit
    run(T + k) would"""
    pos, dip, wall, rp, col, val = engine.tube([3], [3.0])
    sysd = dict(n=len(wall), pos=pos, dip=dip, wall=wall, rp=rp, col=col, val=val)
    n, T, k = sysd["n"], 30, 10
    grp = (np.arange(n) >= 90).astype(np.int32)
    psi0 = np.zeros(n, complex); psi0[0] = 1

    def mk():
        s = make_engine(engine, sysd, 8, 1.0, B=2, force_batched=True, traj_ids=[0, 3])
        s.set_state(psi0)
        s.set_observables(grp, 2, 10, T + k)
        return s

    u = mk()
    u.run(T)
    err_T, diag_T = u.error_sum(), u.diagonal()
    u.run(k)
    ref = u.populations(), u.get_state(), u.error_sum()
    u.close()

    a = mk()
    a.adjoint(T, ckpt=7, w_pop=[0.0, 1.0])
    assert a.get_step() == T
    assert np.array_equal(a.error_sum(), err_T)
    assert np.array_equal(a.diagonal(), diag_T)
    a.run(k)
    assert np.array_equal(a.get_state(), ref[1])
    assert np.array_equal(a.error_sum(), ref[2])
    assert np.array_equal(a.populations()[:, T // 10:], ref[0][:, T // 10:])
    a.close()

"""GPU reverse sweep against the oracle adjoint and central differences."""
import numpy as np
import pytest

from tests.helpers import make_engine, make_oracle

pytestmark = pytest.mark.gpu


def small_tube(engine):
    pos, dip, wall, rp, col, val = engine.tube([3], [3.0])
    n = len(wall)
    return dict(n=n, pos=pos, dip=dip, wall=wall, rp=rp, col=col, val=val)


@pytest.mark.parametrize("K,m,ckpt", [(1, 8, 7), (2, 12, 50)])
def test_gpu_adjoint_matches_oracle(engine, oracle, K, m, ckpt):
    sysd = small_tube(engine)
    n, T = sysd["n"], 30
    grp = (np.arange(n) >= 90).astype(np.int32)
    psi0 = np.zeros(n, complex); psi0[0] = 1
    bra = sysd["dip"][:, 0] / np.linalg.norm(sysd["dip"][:, 0])
    ct = 0.3 * np.exp(-0.01j * np.arange(T + 1))
    ids = [0, 3]
    s = make_engine(engine, sysd, m, 1.0, B=2, K=K, force_batched=True, traj_ids=ids)
    s.set_state(psi0)
    s.set_observables(grp, 2, T, T, bra=bra.astype(complex))
    loss, ge, gg = s.adjoint(T, ckpt=ckpt, w_pop=[0.0, 1.0], w_spec=0.5, corr_target=ct)
    p, sites = make_oracle(oracle, sysd, m, 1.0, K=K)
    for k, t in enumerate(ids):
        lo, geo, ggo = oracle.adjoint(p, sites, sysd["rp"], sysd["col"], sysd["val"], psi0, t, T, grp, 2,
                                      [0.0, 1.0], bra=bra, w_spec=0.5, corr_target=ct, ckpt=ckpt)
        print("traj", t, "loss", loss[k], lo, "grad err", np.abs(ge[k] - geo).max() / np.abs(geo).max(),
              "g", gg[k], ggo)
        assert loss[k] == pytest.approx(lo, rel=1e-10)
        assert np.abs(ge[k] - geo).max() <= 1e-8 * np.abs(geo).max()
        assert gg[k] == pytest.approx(ggo, rel=1e-8)


def test_gpu_adjoint_matches_central_differences(engine, oracle):
    sysd = small_tube(engine)
    n, T, m = sysd["n"], 20, 8
    grp = (np.arange(n) >= 90).astype(np.int32)
    psi0 = np.zeros(n, complex); psi0[0] = 1
    s = make_engine(engine, sysd, m, 1.0, B=1, force_batched=True)
    s.set_state(psi0)
    s.set_observables(grp, 2, T, T)
    loss, ge, gg = s.adjoint(T, ckpt=5, w_pop=[0.0, 1.0])
    p, _ = make_oracle(oracle, sysd, m, 1.0)
    kw = dict(tau=50.0, sigma=300.0, e_site=0.0)
    h = 1e-3
    for i in (0, 7, 120):
        e = np.zeros(n); e[i] = h
        args = (sysd["rp"], sysd["col"], sysd["val"], psi0, 0, T, grp, 2, [0.0, 1.0])
        fd = (oracle.loss_fd(p, kw, *args, e_shift=e) - oracle.loss_fd(p, kw, *args, e_shift=-e)) / (2 * h)
        assert abs(ge[0, i] - fd) <= 1e-5 * abs(fd) + 1e-13

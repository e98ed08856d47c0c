"""Adjoint of the transfer-yield + spectrum-mismatch loss: a 600-site
double-walled tube against central differences (h = 1e-3 cm^-1) and the FP64
oracle's adjoint over T = 2 ps with checkpoints every 50 steps, the
torch.autograd wrapper, and the cost of the reverse sweep relative to the
forward run on the 12,000-site tube with B = 64."""
import time

import numpy as np
import pytest

from tests.helpers import double_tube, outer_wall_state

pytestmark = pytest.mark.gpu

M, DT, T, CKPT = 16, 1.0, 2000, 50
SIG_STATIC, SIG_OU, TAU = 200.0, 300.0, 50.0
W_SPEC = 1e-4


def coupled_tube(engine):
    """5 + 5 rings x 60 sites, inner R = 3 nm, outer R = 5 nm: the walls are
    2 nm apart, inside the 2.5 nm cutoff, so excitation moves between them
    (at the literal 3 / 6 nm radii the walls do not couple and the yield is 0)"""
    pos, dip, wall, rp, col, val = engine.tube([5, 5], [3.0, 5.0])
    return dict(n=len(wall), pos=pos, dip=dip, wall=wall, rp=rp, col=col, val=val)


def spectrum_weights(nsteps, dt, sigma_fs, mu2, nfft):
    n = np.arange(nsteps + 1)
    win = np.where(n < nsteps, np.exp(-0.5 * ((n + 1) * dt / sigma_fs) ** 2), 1.0)
    return nfft * (mu2 / nsteps) ** 2 * win ** 2


def build(engine, sysd, e_site, g=1.0, B=1, m=M, dt=DT, precision=0):
    s = engine.System(sysd["n"], m, dt, n_traj=B, precision=precision, coupling_scale=g,
                      sigma_static=SIG_STATIC, seed=1, force_batched=True)
    s.set_hamiltonian(sysd["rp"], sysd["col"], sysd["val"], e_site)
    s.set_ou(TAU, SIG_OU)
    s.reset_disorder()
    psi0 = outer_wall_state(sysd)
    s.set_state(psi0)
    grp = (sysd["wall"] == 0).astype(np.int32)          # group 1 = inner wall
    s.set_observables(grp, 2, T, T, bra=psi0)
    return s


@pytest.fixture(scope="module")
def problem(engine):
    sysd = coupled_tube(engine)
    n = sysd["n"]
    # synthetic target: the same run with the site energies shifted by 30 cm^-1
    s = build(engine, sysd, np.full(n, 30.0))
    s.run(T)
    target = s.corr()[0]
    s.close()
    u = spectrum_weights(T, DT, 1000.0, 1.0, 4096)
    kw = dict(w_pop=[0.0, 1.0], w_spec=W_SPEC, corr_target=target, corr_weight=u)
    return sysd, kw


def loss_of(engine, sysd, kw, e_site, g=1.0):
    s = build(engine, sysd, e_site, g)
    s.run(T)
    yield_ = s.populations()[0, -1, 1]
    c = s.corr()[0]
    s.close()
    return yield_ + W_SPEC * float(np.sum(kw["corr_weight"] * np.abs(c - kw["corr_target"]) ** 2))


def test_transfer_loss_gradient_vs_central_differences(engine, problem):
    sysd, kw = problem
    n = sysd["n"]
    s = build(engine, sysd, np.zeros(n))
    loss, ge, gg = s.adjoint(T, ckpt=CKPT, **kw)
    s.close()
    L0 = loss_of(engine, sysd, kw, np.zeros(n))
    print("loss", loss[0], "forward", L0)
    assert loss[0] == pytest.approx(L0, rel=1e-10)
    h = 1e-3
    for i in (0, 299, 300, 599):
        e = np.zeros(n); e[i] = h
        fd = (loss_of(engine, sysd, kw, e) - loss_of(engine, sysd, kw, -e)) / (2 * h)
        print("site", i, "adjoint", ge[0, i], "fd", fd, "rel", abs(ge[0, i] - fd) / abs(fd))
        assert abs(ge[0, i] - fd) <= 1e-5 * abs(fd) + 1e-13
    # the loss is strongly curved in g (it scales every coupling over 2 ps): the
    # central difference error is ~3.5e-5 relative at h = 1e-6 and ~3.6e-7 at
    # h = 1e-7 (measured with the oracle), so the step here is 1e-8
    hg = 1e-8
    fd = (loss_of(engine, sysd, kw, np.zeros(n), 1 + hg) - loss_of(engine, sysd, kw, np.zeros(n), 1 - hg)) / (2 * hg)
    print("coupling scale: adjoint", gg[0], "fd", fd)
    assert abs(gg[0] - fd) <= 1e-5 * abs(fd)


def test_transfer_loss_gradient_vs_oracle_adjoint(engine, oracle, problem):
    sysd, kw = problem
    n = sysd["n"]
    s = build(engine, sysd, np.zeros(n))
    loss, ge, gg = s.adjoint(T, ckpt=CKPT, **kw)
    s.close()
    p = oracle.make_params(n, M, DT, reorth=oracle.REORTH_GRAM, sigma_static=SIG_STATIC, seed=1)
    sites = oracle.Sites(n, tau=TAU, sigma=SIG_OU, e_site=0.0)
    grp = (sysd["wall"] == 0).astype(np.int32)
    lo, geo, ggo = oracle.adjoint(p, sites, sysd["rp"], sysd["col"], sysd["val"], outer_wall_state(sysd), 0, T,
                                  grp, 2, kw["w_pop"], bra=outer_wall_state(sysd), w_spec=W_SPEC,
                                  corr_target=kw["corr_target"], corr_weight=kw["corr_weight"], ckpt=CKPT)
    err = np.abs(ge[0] - geo).max() / np.abs(geo).max()
    print("loss", loss[0], lo, "grad_e rel err", err, "g", gg[0], ggo)
    assert loss[0] == pytest.approx(lo, rel=1e-9)
    assert err <= 1e-8
    assert gg[0] == pytest.approx(ggo, rel=1e-8)


def test_torch_autograd_wrapper(engine, problem):
    torch = pytest.importorskip("torch")
    from tests.torch_autograd import engine_loss
    sysd, kw = problem
    n = sysd["n"]
    e = torch.zeros(n, dtype=torch.float64, requires_grad=True)
    g = torch.tensor(1.0, dtype=torch.float64, requires_grad=True)
    mk = lambda es, gs: build(engine, sysd, es, gs, B=2)
    L = engine_loss(e, g, mk, T, CKPT, **kw)
    assert L.shape == (2,)
    (L[0] + 2.0 * L[1]).backward()
    s = build(engine, sysd, np.zeros(n), B=2)
    loss, ge, gg = s.adjoint(T, ckpt=CKPT, **kw)
    s.close()
    assert np.allclose(L.detach().numpy(), loss, rtol=0, atol=0)
    assert np.allclose(e.grad.numpy(), ge[0] + 2.0 * ge[1], rtol=1e-14, atol=0)
    assert g.grad.item() == pytest.approx(gg[0] + 2.0 * gg[1], rel=1e-14)
    # directional derivative along a random direction against central differences
    v = np.random.default_rng(0).standard_normal(n)
    h = 1e-3
    def f(eps):
        with torch.no_grad():
            return engine_loss(torch.from_numpy(eps * v), torch.tensor(1.0, dtype=torch.float64), mk, T, CKPT,
                               **kw).numpy()
    fd = (f(h) - f(-h)) / (2 * h)
    an = ge @ v
    print("directional derivative", an, fd)
    assert np.all(np.abs(an - fd) <= 1e-5 * np.abs(fd))


def test_adjoint_cost_n12000_b64(engine):
    """cost of the reverse sweep (with its recomputation) relative to the
    forward run: N = 12,000 literal tube, B = 64, m = 24, dt = 2 fs, T = 2 ps"""
    sysd = double_tube(engine)
    n, B, steps = sysd["n"], 64, 1000
    psi0 = outer_wall_state(sysd)
    grp = (sysd["wall"] == 0).astype(np.int32)

    def mk():
        s = engine.System(n, 24, 2.0, n_traj=B, sigma_static=250.0, seed=1)
        s.set_hamiltonian(sysd["rp"], sysd["col"], sysd["val"], np.zeros(n))
        s.set_ou(50.0, 400.0)
        s.reset_disorder()
        s.set_state(psi0)
        s.set_observables(grp, 2, 10, steps, bra=psi0)
        return s

    s = mk()
    s.run(20, graph=True)
    s.close()
    s = mk()
    t0 = time.perf_counter()
    s.run(steps, graph=True)
    fwd = time.perf_counter() - t0
    s.close()
    s = mk()
    ct = 0.5 * np.exp(-0.001j * np.arange(steps + 1))
    t0 = time.perf_counter()
    loss, ge, gg = s.adjoint(steps, ckpt=50, w_pop=[0.0, 1.0], w_spec=1e-3, corr_target=ct)
    adj = time.perf_counter() - t0
    s.close()
    print("N=12000 B=64 T=%d: forward %.3f s, adjoint %.3f s, ratio %.2f" % (steps, fwd, adj, adj / fwd))
    assert np.all(np.isfinite(ge)) and np.all(np.isfinite(gg))
    assert adj / fwd <= 3.0

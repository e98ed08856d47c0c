"""Shared set-up of the benchmark systems for engine and oracle."""
import numpy as np


def ring_tube(engine):
    """10 rings x 60 sites, R = 3 nm (the N = 600 single-wall system)"""
    pos, dip, wall, rp, col, val = engine.tube([10], [3.0])
    return dict(n=600, pos=pos, dip=dip, wall=wall, rp=rp, col=col, val=val)


def double_tube(engine, rings=100, order=0):
    pos, dip, wall, rp, col, val = engine.tube([rings, rings], [3.0, 6.0], order=order)
    return dict(n=len(wall), pos=pos, dip=dip, wall=wall, rp=rp, col=col, val=val)


def outer_wall_state(sysd):
    """normalised dipole-weighted superposition over the outer wall:
    psi_i ∝ mu_i . x for sites of wall 1, 0 elsewhere"""
    psi = np.where(sysd["wall"] == 1, sysd["dip"][:, 0], 0.0).astype(np.complex128)
    return psi / np.linalg.norm(psi)


def make_engine(engine, sysd, m, dt, B=1, K=1, S=1, precision=0, sigma_static=200.0, ou_sigma=300.0,
                tau=50.0, seed=1, force_batched=False, traj_ids=None, n_active=None):
    s = engine.System(sysd["n"], m, dt, n_traj=B, krylov_substeps=K, ou_substeps=S,
                      precision=precision, sigma_static=sigma_static, seed=seed,
                      force_batched=force_batched)
    s.set_hamiltonian(sysd["rp"], sysd["col"], sysd["val"], np.zeros(sysd["n"]))
    s.set_ou(tau, ou_sigma)
    if traj_ids is not None or n_active is not None:
        s.set_trajectories(traj_ids, n_active)
    s.reset_disorder()
    return s


def make_oracle(oracle, sysd, m, dt, K=1, S=1, sigma_static=200.0, ou_sigma=300.0, tau=50.0, seed=1):
    p = oracle.make_params(sysd["n"], m, dt, K=K, S=S, reorth=oracle.REORTH_GRAM,
                           sigma_static=sigma_static, seed=seed)
    sites = oracle.Sites(sysd["n"], tau=tau, sigma=ou_sigma, e_site=0.0)
    return p, sites

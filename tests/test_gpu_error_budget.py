"""Mixed-precision error over a full-length run: the N = 12,000 double-walled
tube at m = 24, dt = 2 fs for 5000 steps (the plain-FP64 workload's length),
32 trajectories, mixed precision against the FP64 batched path (itself within
1e-10 of the oracle), populations of 2 walls x 4 axial segments every 10
steps and the final site populations, for the SpMM choices: the row-blocked
SpMM (FP64 couplings and sums), the matrix-core SpMM with FP64 sums after
every K-step, and the matrix-core SpMM with FP32 sums over 32 union slots
(the fastest, which the bound of 1e-6 holds for over about 3000 steps)."""
import numpy as np
import pytest

from tests.helpers import double_tube, outer_wall_state

pytestmark = pytest.mark.gpu

CHOICES = [("row-blocked", dict(spmm_rows_per_wave=4), 1e-6),
           ("matrix cores, FP64 sums per K-step", dict(spmm_rows_per_wave=16, spmm_traj_per_lane=2,
                                                        spmm_exact_sums=1), 1e-6),
           ("matrix cores, FP32 sums of 32 slots", dict(spmm_rows_per_wave=16, spmm_traj_per_lane=2,
                                                         spmm_exact_sums=0), 2.5e-6)]


def test_mixed_error_budget_5000_steps(engine):
    sysd = double_tube(engine)
    n, steps, B, every = sysd["n"], 5000, 32, 10
    seg = np.minimum((sysd["pos"][:, 2] / 80.0 * 4).astype(np.int32), 3)
    grp = (sysd["wall"] * 4 + seg).astype(np.int32)
    psi0 = outer_wall_state(sysd)

    def run(prec, tuning):
        s = engine.System(n, 24, 2.0, n_traj=B, precision=prec, sigma_static=250.0, seed=1, force_batched=True)
        s.set_hamiltonian(sysd["rp"], sysd["col"], sysd["val"], np.zeros(n))
        s.set_ou(50.0, 400.0)
        s.set_trajectories(np.arange(B, dtype=np.uint64))
        s.reset_disorder()
        if tuning:
            s.set_tuning(**tuning)
        s.set_state(psi0)
        s.set_observables(grp, 8, every, steps)
        s.run(steps, graph=True)
        out = s.populations(), np.abs(s.get_state()) ** 2
        s.close()
        return out

    ref = run(engine.FP64, None)
    for name, tuning, bound in CHOICES:
        pops, site = run(engine.MIXED, tuning)
        dp = np.abs(pops - ref[0])                                 # [B, records, 8]
        print(name + ": max group-population error " +
              ", ".join("%.3e (<= %d steps)" % (dp[:, : k // every + 1].max(), k) for k in (1000, 2000, 3000, 5000)) +
              "; final site populations %.3e" % np.abs(site - ref[1]).max())
        assert dp.max() <= bound, name
        assert np.abs(site - ref[1]).max() <= bound, name

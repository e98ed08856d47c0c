"""Batched mixed-precision SpMM, every kernel choice, against the oracle's
SpMV on the 12,000-site double tube with ragged trajectory sizes; the
LDS-window SpMM (16-row blocks on the matrix cores, x rows staged in LDS)
bitwise against the 16-row gather SpMM and through whole propagations."""
import numpy as np
import pytest

from tests.helpers import double_tube, make_engine, outer_wall_state, ring_tube

pytestmark = pytest.mark.gpu

# row-blocked gathers (R = 2, 4, 8; 64 or 128 trajectories per wave), the
# 16-row matrix-core SpMM (16, 32, 64 trajectories per wave) and the
# LDS-window SpMM (tiles of 120 / 240 rows, 16 / 32 trajectories per wave)
GATHER = [dict(spmm_rows_per_wave=r, spmm_traj_per_lane=t, spmm_phases=1, spmm_pipeline=0,
               spmm_window_rows=0) for r in (2, 4, 8) for t in (1, 2)]
GATHER += [dict(spmm_rows_per_wave=4, spmm_traj_per_lane=1, spmm_phases=2, spmm_pipeline=1, spmm_window_rows=0)]
MFMA16 = [dict(spmm_rows_per_wave=16, spmm_traj_per_lane=t, spmm_phases=1, spmm_pipeline=0, spmm_window_rows=0)
          for t in (1, 2, 4)]
# 16-row blocks on 16 x 16 x 4 MFMAs reading the LDS window
WINDOW = [dict(spmm_rows_per_wave=16, spmm_traj_per_lane=t, spmm_phases=1, spmm_pipeline=0, spmm_window_rows=T,
               spmm_window_waves=w) for T in (240, 128) for t in (1, 2) for w in (4, 16)]

# 128 trajectories: sizes not multiples of 16, sorted internally by
# n_active, so the slices see ragged ends
N_ACTIVE = [12000, 11993, 7777, 6001, 5999, 250, 12000, 9001] * 16


def _spmv_system(engine, sysd, B, n_active):
    return make_engine(engine, sysd, 24, 2.0, B=B, precision=engine.MIXED, sigma_static=250.0, ou_sigma=400.0,
                       n_active=n_active)


def _oracle_spmv(oracle, sysd, d, x, na):
    """y = H x on the leading na x na block (x and d as the engine holds them)"""
    xa = np.where(np.arange(len(x)) < na, x, 0)
    y = oracle.spmv(sysd["rp"], sysd["col"], sysd["val"], d, xa)
    y[na:] = 0
    return y


def test_every_spmm_choice_against_oracle(engine, oracle):
    sysd = double_tube(engine)
    n, B = sysd["n"], len(N_ACTIVE)
    s = _spmv_system(engine, sysd, B, np.array(N_ACTIVE, np.int32))
    rng = np.random.default_rng(7)
    x = rng.normal(size=n) + 1j * rng.normal(size=n)
    x32 = x.astype(np.complex64).astype(np.complex128)        # the engine rounds x to FP32
    d = s.diagonal()
    trajs = [0, 1, 2, 3, 4, 5, 9, 58]
    refs = {t: _oracle_spmv(oracle, sysd, d[t], x32, N_ACTIVE[t]) for t in trajs}
    got = {}
    for c in GATHER + MFMA16 + WINDOW:
        s.set_tuning(**c)
        for t in trajs:
            y = s.spmv(t, x)
            ref = refs[t]
            err = np.abs(y - ref).max() / np.abs(ref).max()
            assert err < 2e-6, (c, t, err)
            assert np.all(y[N_ACTIVE[t]:] == 0), (c, t)
            got[(tuple(sorted(c.items())), t)] = y
    # the 16-row window SpMM runs the MFMA sequence of the 16-row gather SpMM
    for c in WINDOW:
        m16 = [m for m in MFMA16 if m["spmm_traj_per_lane"] == c["spmm_traj_per_lane"]][0]
        for t in trajs:
            a = got[(tuple(sorted(c.items())), t)]
            b = got[(tuple(sorted(m16.items())), t)]
            assert np.array_equal(a, b), (c, t, np.abs(a - b).max())
    s.close()


def test_window_spmm_is_the_same_for_every_slice_shape(engine):
    """the 16- and 32-trajectory slices and the wave counts change no bit of H x"""
    sysd = double_tube(engine)
    s = _spmv_system(engine, sysd, 128, None)
    x = np.random.default_rng(3).normal(size=sysd["n"]) * (1 + 0.5j)
    for group in (WINDOW,):
        ref = None
        for c in group:
            s.set_tuning(**c)
            y = s.spmv(5, x)
            if ref is None:
                ref = y
            assert np.array_equal(y, ref), c
    s.close()


def _propagate(engine, sysd, tuning, precision, B=64, steps=200, graph=True, n_active=None, slices=1):
    n = sysd["n"]
    s = make_engine(engine, sysd, 16, 1.0, B=B, precision=precision, force_batched=True, n_active=n_active)
    if tuning:
        s.set_tuning(concurrent_slices=slices, **tuning)
    s.set_state(np.eye(1, n, 0)[0])
    grp = (np.arange(n) // 75).astype(np.int32)
    s.set_observables(grp, 8, 5, steps)
    s.run(steps, graph=graph)
    out = s.populations(), s.get_state(), s.error_sum()
    s.close()
    return out


def test_window_propagation_against_fp64(engine):
    """600-site ring tube, 200 steps: the window SpMM in the mixed path stays
    within the FP32 budget of the FP64 path and of the VALU SpMM's mixed run"""
    sysd = ring_tube(engine)
    p64, _, _ = _propagate(engine, sysd, None, engine.FP64)
    pv, _, _ = _propagate(engine, sysd, GATHER[2], engine.MIXED)
    for c in (WINDOW[1], WINDOW[6]):
        pw, psiw, errw = _propagate(engine, sysd, c, engine.MIXED)
        e64 = np.abs(pw - p64).max()
        ev = np.abs(pw - pv).max()
        print(c, "vs FP64", e64, "vs VALU mixed", ev)
        assert e64 < 1e-6
        assert ev < 1e-6


def test_window_ragged_graph_and_slices_bitwise(engine):
    """ragged sizes, four concurrent slices in the captured step: bitwise the
    eager one-stream run"""
    sysd = ring_tube(engine)
    na = np.array([600, 599, 421, 300, 61, 600, 517, 8] * 32, np.int32)
    for c in (WINDOW[2], WINDOW[5]):
        a = _propagate(engine, sysd, c, engine.MIXED, B=256, steps=60, graph=False, n_active=na)
        b = _propagate(engine, sysd, c, engine.MIXED, B=256, steps=60, graph=True, n_active=na, slices=4)
        for u, v in zip(a, b):
            assert np.array_equal(u, v)
        pops = a[0]
        for t in range(256):
            assert abs(pops[t, 0].sum() - 1.0) < 1e-5


def test_window_choice_rejections_leave_the_system_working(engine):
    sysd = ring_tube(engine)
    s = make_engine(engine, sysd, 16, 1.0, B=64, precision=engine.FP64, force_batched=True)
    with pytest.raises(engine.DQError):
        s.set_tuning(**WINDOW[0])                       # FP64
    s.close()
    s = make_engine(engine, sysd, 16, 1.0, B=64, precision=engine.MIXED)
    with pytest.raises(engine.DQError):
        s.set_tuning(spmm_rows_per_wave=16, spmm_traj_per_lane=2, spmm_window_rows=120, spmm_window_waves=8)
    with pytest.raises(engine.DQError):
        s.set_tuning(spmm_rows_per_wave=4, spmm_traj_per_lane=2, spmm_window_rows=128, spmm_window_waves=8)
    s.set_tuning(**WINDOW[1])
    before = s.tuning()
    # a coupling between the first and the last site of the 12,000-site tube
    # widens the windows past the LDS of a workgroup
    big = double_tube(engine)
    n2 = big["n"]
    rows = [list(zip(big["col"][big["rp"][i]:big["rp"][i + 1]], big["val"][big["rp"][i]:big["rp"][i + 1]]))
            for i in range(n2)]
    rows[0].append((n2 - 1, 3.0))
    rows[n2 - 1].append((0, 3.0))
    rp2 = np.zeros(n2 + 1, np.int32)
    c2, v2 = [], []
    for i in range(n2):
        r = sorted(rows[i])
        c2 += [int(a) for a, _ in r]
        v2 += [float(b) for _, b in r]
        rp2[i + 1] = len(c2)
    wide = engine.System(n2, 16, 1.0, n_traj=64, precision=engine.MIXED)
    wide.set_tuning(spmm_rows_per_wave=16, spmm_traj_per_lane=2, spmm_window_rows=2048, spmm_window_waves=8)
    with pytest.raises(engine.DQError):
        wide.set_hamiltonian(rp2, np.array(c2, np.int32), np.array(v2), np.zeros(n2))
    wide.close()
    n = sysd["n"]
    # the narrow system is unaffected and still runs
    assert s.tuning() == before
    s.set_state(np.eye(1, n, 0)[0])
    s.run(5)
    s.close()

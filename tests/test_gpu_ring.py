"""Batched mixed-precision SpMM, every kernel choice, against the oracle's
SpMV on the 12,000-site double tube with ragged trajectory sizes; the ring
SpMM (x rows staged in an LDS ring, 4-row blocks on the matrix cores) the
same bit for bit for every row range and wave count, and through whole
propagations against the FP64 path."""
import numpy as np
import pytest

from tests.helpers import double_tube, make_engine, outer_wall_state, ring_tube

pytestmark = pytest.mark.gpu

# row-blocked gathers (R = 2, 4, 8; 64 or 128 trajectories per wave), the
# 16-row matrix-core SpMM (16, 32, 64 trajectories per wave) and the ring
# SpMM (ranges of 128 .. 1536 rows per workgroup, 4 .. 16 waves)
GATHER = [dict(spmm_rows_per_wave=r, spmm_traj_per_lane=t, spmm_phases=1, spmm_pipeline=0,
               spmm_window_rows=0) for r in (2, 4, 8) for t in (1, 2)]
GATHER += [dict(spmm_rows_per_wave=4, spmm_traj_per_lane=1, spmm_phases=2, spmm_pipeline=1, spmm_window_rows=0)]
MFMA16 = [dict(spmm_rows_per_wave=16, spmm_traj_per_lane=t, spmm_phases=1, spmm_pipeline=0, spmm_window_rows=0,
               spmm_exact_sums=e) for e in (0, 1) for t in (1, 2, 4)]
# 4-row blocks on 4 x 4 x 1 MFMAs reading the LDS ring
RING = [dict(spmm_rows_per_wave=4, spmm_traj_per_lane=1, spmm_phases=1, spmm_pipeline=0, spmm_window_rows=T,
             spmm_window_waves=w, spmm_exact_sums=e) for e in (0, 1) for T, w in ((768, 16), (384, 8), (128, 4),
                                                                                   (1536, 16))]

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
    for c in GATHER + MFMA16 + RING:
        s.set_tuning(**c)
        for t in trajs:
            y = s.spmv(t, x)
            ref = refs[t]
            err = np.abs(y - ref).max() / np.abs(ref).max()
            assert err < 2e-6, (c, t, err)
            assert np.all(y[N_ACTIVE[t]:] == 0), (c, t)
            got[(tuple(sorted(c.items())), t)] = y
    # a block's products do not depend on the row range or the wave count
    for e in (0, 1):
        group = [c for c in RING if c["spmm_exact_sums"] == e]
        for t in trajs:
            a = got[(tuple(sorted(group[0].items())), t)]
            for c in group[1:]:
                b = got[(tuple(sorted(c.items())), t)]
                assert np.array_equal(a, b), (c, t, np.abs(a - b).max())
    s.close()


def test_ring_spmm_is_the_same_for_every_range(engine):
    """row ranges and wave counts change no bit of H x (every trajectory slot
    position of a 128-slot batch)"""
    sysd = double_tube(engine)
    s = _spmv_system(engine, sysd, 128, None)
    x = np.random.default_rng(3).normal(size=sysd["n"]) * (1 + 0.5j)
    for group in ([c for c in RING if not c["spmm_exact_sums"]], [c for c in RING if c["spmm_exact_sums"]]):
        ref = None
        for c in group:
            s.set_tuning(**c)
            y = np.stack([s.spmv(t, x) for t in (5, 16, 127)])
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


def test_ring_propagation_against_fp64(engine):
    """600-site ring tube, 200 steps: the ring SpMM in the mixed path stays
    within the FP32 budget of the FP64 path and of the VALU SpMM's mixed run"""
    sysd = ring_tube(engine)
    p64, _, _ = _propagate(engine, sysd, None, engine.FP64)
    pv, _, _ = _propagate(engine, sysd, GATHER[2], engine.MIXED)
    for c in (RING[0], RING[2], RING[5]):
        pw, psiw, errw = _propagate(engine, sysd, c, engine.MIXED)
        e64 = np.abs(pw - p64).max()
        ev = np.abs(pw - pv).max()
        print(c, "vs FP64", e64, "vs VALU mixed", ev)
        assert e64 < 1e-6
        assert ev < 1e-6


def test_ring_ragged_graph_and_slices_bitwise(engine):
    """ragged sizes, four concurrent slices in the captured step: bitwise the
    eager one-stream run"""
    sysd = ring_tube(engine)
    na = np.array([600, 599, 421, 300, 61, 600, 517, 8] * 32, np.int32)
    for c in (RING[2], RING[4]):
        a = _propagate(engine, sysd, c, engine.MIXED, B=256, steps=60, graph=False, n_active=na)
        b = _propagate(engine, sysd, c, engine.MIXED, B=256, steps=60, graph=True, n_active=na, slices=4)
        for u, v in zip(a, b):
            assert np.array_equal(u, v)
        pops = a[0]
        for t in range(256):
            assert abs(pops[t, 0].sum() - 1.0) < 1e-5


def test_ring_choice_rejections_leave_the_system_working(engine):
    sysd = ring_tube(engine)
    s = make_engine(engine, sysd, 16, 1.0, B=64, precision=engine.FP64, force_batched=True)
    with pytest.raises(engine.DQError):
        s.set_tuning(**RING[0])                         # FP64
    s.close()
    s = make_engine(engine, sysd, 16, 1.0, B=64, precision=engine.MIXED)
    with pytest.raises(engine.DQError):
        s.set_tuning(spmm_rows_per_wave=4, spmm_window_rows=120, spmm_window_waves=8)
    with pytest.raises(engine.DQError):
        s.set_tuning(spmm_rows_per_wave=4, spmm_window_rows=128, spmm_window_waves=2)
    s.set_tuning(**RING[1])
    before = s.tuning()
    # a coupling between the first and the last site of the 12,000-site tube
    # widens the band past the LDS ring of 1024 rows
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
    wide.set_tuning(spmm_rows_per_wave=4, spmm_window_rows=768, spmm_window_waves=16)
    with pytest.raises(engine.DQError):
        wide.set_hamiltonian(rp2, np.array(c2, np.int32), np.array(v2), np.zeros(n2))
    wide.close()
    n = sysd["n"]
    # the narrow system is unaffected and still runs
    assert s.tuning() == before
    s.set_state(np.eye(1, n, 0)[0])
    s.run(5)
    s.close()


def _exact_coupling_matrix(nt=32, seed=5):
    """n = 3 nt sites: site t < nt couples to nt + 2t with 2^24 + k_t (k_t
    odd, so its FP32 value is 2^24 + k_t -/+ 1) and to nt + 2t + 1 with -2^24
    (exact in FP32); the pair sits in one group of 16 union slots"""
    rng = np.random.default_rng(seed)
    n = 3 * nt
    k = 2 * rng.integers(1, 60, nt) + 1
    big = float(2 ** 24)
    rows = [[] for _ in range(n)]
    for t in range(nt):
        rows[t] += [(nt + 2 * t, big + k[t]), (nt + 2 * t + 1, -big)]
        rows[nt + 2 * t].append((t, big + k[t]))
        rows[nt + 2 * t + 1].append((t, -big))
    rp = np.zeros(n + 1, np.int32)
    col, val = [], []
    for i in range(n):
        for c, v in sorted(rows[i]):
            col.append(c)
            val.append(v)
        rp[i + 1] = len(col)
    x = np.zeros(n, complex)
    x[nt:] = 1.0 + 0.5j                                  # x_t = 0 for t < nt
    return n, rp, np.array(col, np.int32), np.array(val), x, k


@pytest.mark.parametrize("choice", GATHER + MFMA16 + RING[:1] + RING[-1:])
def test_mixed_spmm_uses_exact_couplings(engine, choice):
    """couplings 2^24 + k_t and -2^24 on x = 1 + i/2: w_t = k_t (1 + i/2)
    exactly, which FP32 couplings alone (2^24 + k_t -/+ 1) would miss: the
    row-blocked SpMMs use the FP64 couplings, the matrix-core ones add the
    low parts of the FP32 couplings (every lane position of their operands:
    32 rows, unions of two 16-slot groups; the ring SpMM multiplies the two
    columns of a pair in different lane halves, which sum apart)"""
    n, rp, col, val, x, k = _exact_coupling_matrix()
    nt = n // 3
    s = engine.System(n, 8, 1.0, n_traj=128, precision=engine.MIXED, sigma_static=0.0, force_batched=True)
    s.set_hamiltonian(rp, col, val, np.zeros(n))
    s.set_ou(50.0, 0.0)
    s.reset_disorder()
    try:
        s.set_tuning(**choice)
    except engine.DQError as e:
        s.close()
        pytest.skip("choice not available at n = %d: %s" % (n, e))
    for traj in (0, 101):
        y = s.spmv(traj, x)
        assert np.array_equal(y[:nt], k * (1.0 + 0.5j)), (choice, traj, y[:nt].real - k)
        assert np.all(y[nt:] == 0)
    s.close()

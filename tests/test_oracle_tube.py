"""Oracle and engine tube generators against the dipole formula."""
import json
import os

import numpy as np
import pytest

GOLD = os.path.join(os.path.dirname(__file__), "golden")


def dense(rowptr, col, val, n):
    H = np.zeros((n, n))
    for i in range(n):
        H[i, col[rowptr[i]:rowptr[i + 1]]] = val[rowptr[i]:rowptr[i + 1]]
    return H


def test_oracle_couplings_match_numpy_pairs(oracle):
    pos, dip, wall = oracle.tube([10], [3.0])
    rp, col, val = oracle.couplings_csr(pos, dip)
    H = dense(rp, col, val, 600)
    for c in json.load(open(os.path.join(GOLD, "tube_kat.json"))):
        if c["r"] < 2.5:
            assert H[c["i"], c["j"]] == pytest.approx(c["J"], rel=1e-12)
        else:
            assert H[c["i"], c["j"]] == 0.0


def test_nnz_counts_single_and_double_wall(oracle):
    pos, dip, wall = oracle.tube([10], [3.0])
    rp, col, val = oracle.couplings_csr(pos, dip)
    assert len(col) == 42720          # 71.2 per row
    pos, dip, wall = oracle.tube([100, 100], [3.0, 6.0])
    rp, col, val = oracle.couplings_csr(pos, dip)
    assert len(col) == 734160         # 61.18 per row at N = 12000
    # no inter-wall coupling at the 2.5 nm cutoff (radial gap 3 nm)
    rows = np.repeat(np.arange(12000), np.diff(rp))
    assert np.all(wall[rows] == wall[col])


def test_engine_generator_matches_oracle(oracle, engine):
    for rings, radius in [([10], [3.0]), ([6, 6], [3.0, 6.0])]:
        for order in (0, 1):
            if order == 1 and len(rings) == 1:
                continue
            pos, dip, wall, rp, col, val = engine.tube(rings, radius, order=order)
            opos, odip, owall = oracle.tube(rings, radius, order=order)
            assert np.allclose(pos, opos, atol=1e-14) and np.allclose(dip, odip, atol=1e-13)
            assert np.array_equal(wall, owall)
            orp, ocol, oval = oracle.couplings_csr(opos, odip)
            n = len(wall)
            assert np.array_equal(rp, orp) and np.array_equal(col, ocol)
            assert np.allclose(val, oval, rtol=1e-12, atol=1e-9)
            H = dense(rp, col, val, n)
            assert np.array_equal(H, H.T)       # engine mirrors J_ij exactly


def test_ring_major_order_gives_leading_blocks(engine):
    """tube of r rings = leading 120 r sites of the longer ring-major tube"""
    _, _, wall_l, rp_l, col_l, val_l = engine.tube([12, 12], [3.0, 6.0], order=1)
    _, _, wall_s, rp_s, col_s, val_s = engine.tube([5, 5], [3.0, 6.0], order=1)
    n = 600
    Hl = dense(rp_l, col_l, val_l, 1440)[:n, :n]
    Hs = dense(rp_s, col_s, val_s, n)
    assert np.array_equal(Hl, Hs)
    assert np.array_equal(wall_l[:n], wall_s)

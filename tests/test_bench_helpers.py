"""bench.py's bookkeeping on the CPU: algorithmic byte counts, the kernel
names a tuning launches and the lookup of the committed traffic summary."""
import os
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)

import bench  # noqa: E402


def test_algorithmic_bytes():
    n, nnz, B, m = 12000, 734160, 256, 24
    # SpMM: 8 B per nonzero, rowptr, x read + w write (8 B each) + diagonal (4 B) per element
    assert bench.spmm_bytes(n, nnz, B, True) == nnz * 8 + 4 * (n + 1) + B * n * 20
    assert bench.spmm_bytes(n, nnz, B, False) == nnz * 12 + 4 * (n + 1) + B * n * 40
    # update j reads j+1 basis vectors and w, writes v_{j+1}
    assert bench.update_bytes(n, B, m, True) == sum((j + 3) * n * B * 8 for j in range(m))
    assert abs(bench.update_bytes(n, B, m, True) / m - 356.352e6) < 1.0


def test_kernel_names_and_traffic_lookup():
    t = dict(spmm_rows_per_wave=16, spmm_traj_per_lane=2, spmm_pipeline=0, tile_rows=128)
    assert bench.spmm_kernel_name(t) == "dq::k_spmm_mfma<4>"
    assert bench.spmm_kernel_name(dict(t, spmm_rows_per_wave=4)) == "dq::k_spmm_b<float,2,4,false>"
    up = bench.measured_traffic(bench.update_kernel_name(t))
    sp = bench.measured_traffic(bench.spmm_kernel_name(t))
    assert up is not None and sp is not None
    # traffic of a launch is at least its algorithmic bytes and within 30 % of them
    # (the matrix-core SpMM launch also reads the previous update's Gram tiles,
    # 4.6 MB on average: within 40 %)
    assert 1.0 <= up / (bench.update_bytes(12000, 256, 24, True) / 24) < 1.3
    assert 1.0 <= sp / bench.spmm_bytes(12000, 734160, 256, True) < 1.4
    ring = bench.measured_traffic(bench.spmm_kernel_name(dict(t, spmm_window_rows=768)))
    assert ring is not None and 1.0 <= ring / bench.spmm_bytes(12000, 734160, 256, True) < 1.5
    # no line for a choice that was not measured
    assert bench.measured_traffic(bench.update_kernel_name(dict(t, tile_rows=256))) is None

"""The C-ABI library loads on the CPU and exports every declared entry point."""
import ctypes as C
import os
import re

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    text = open(os.path.join(ROOT, "include", "dynaquant.h")).read()
    return sorted(set(re.findall(r"\b(dq_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol(engine):
    lib = engine.lib()
    names = declared_symbols()
    assert len(names) >= 25
    for name in names:
        assert hasattr(lib, name), name
    assert set(names) == set(engine.EXPORTS)


def test_create_without_gpu_reports_no_device_or_works(engine):
    try:
        s = engine.System(10, 4, 1.0)
    except engine.DQError as e:
        assert e.code == 7          # NO_DEVICE on a CPU-only host
    else:
        s.close()


def test_invalid_config_rejected_before_device(engine):
    for kw in (dict(m=1), dict(m=33), dict(dt_fs=-1.0)):
        args = dict(n=10, m=4, dt_fs=1.0)
        args.update(kw)
        with pytest.raises(engine.DQError) as ei:
            engine.System(**args)
        assert ei.value.code == 1


def test_tube_generator_runs_on_host(engine):
    pos, dip, wall, rp, col, val = engine.tube([10], [3.0])
    assert len(wall) == 600 and len(col) == 42720
    assert np.allclose(np.linalg.norm(dip, axis=1), 10.0)


def test_zz_oscillators_host_helper_matches_oracle(engine, oracle):
    """the engine's spectral-density amplitudes are the oracle's (host code only)"""
    om = 8.0 * np.arange(1, 101)
    J = 2 * 35.0 * om * 160.0 / (om ** 2 + 160.0 ** 2)
    for T in (77.0, 300.0):
        c = engine.zz_oscillators(om, J, 8.0, T)
        w, co = oracle.zz_oscillators(om, J, 8.0, T)
        assert np.array_equal(c, co)
    with pytest.raises(engine.DQError):
        engine.zz_oscillators(om, J, 8.0, 0.0)

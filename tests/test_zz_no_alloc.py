"""dq_spectrum_noalloc (the spectrum in the library's fixed per-GPU
workspace) against dq_spectrum, through the C ABI: bit for bit, on two calls
in a row whose second input is smaller, and its size limit."""
import numpy as np
import pytest

from tests.helpers import make_engine, ring_tube

pytestmark = pytest.mark.gpu


def _system(engine, sysd, B, steps):
    s = make_engine(engine, sysd, 16, 1.0, B=B, force_batched=B > 1,
                    traj_ids=None if B == 1 else np.arange(B, dtype=np.uint64) + 3)
    psi0 = np.zeros(sysd["n"], complex)
    psi0[:60] = 1 / np.sqrt(60)
    s.set_state(psi0)
    s.set_observables(None, 0, 0, steps, bra=psi0)
    s.run(steps, graph=True)
    return s


def test_noalloc_spectrum_equals_dq_spectrum(engine):
    sysd = ring_tube(engine)
    big = _system(engine, sysd, 4, 300)
    small = _system(engine, sysd, 1, 100)
    cases = [(big, 1024, 120.0, 2.5), (small, 256, 40.0, 1.0), (big, 512, 0.0, 1.0)]
    for s, nfft, sigma, mu2 in cases:
        om, sp = s.spectrum(sigma, mu2, nfft)
        om0, sp0 = s.spectrum_noalloc(sigma, mu2, nfft)
        print("nfft", nfft, "steps", s.get_step(), "max |S|", np.abs(sp).max())
        assert np.abs(sp).max() > 0
        assert np.array_equal(om0, om)
        assert np.array_equal(sp0, sp)
    with pytest.raises(engine.DQError):
        big.spectrum_noalloc(100.0, 1.0, 1 << 18)       # beyond the fixed workspace
    om0, sp0 = small.spectrum_noalloc(40.0, 1.0, 256)   # still working after the refusal
    assert np.array_equal(sp0, small.spectrum(40.0, 1.0, 256)[1])
    big.close()
    small.close()

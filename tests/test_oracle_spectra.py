"""Oracle FFT and spectrum against numpy and the single-site closed form."""
import json
import os

import numpy as np
import pytest

GOLD = os.path.join(os.path.dirname(__file__), "golden")


def test_fft_matches_numpy_fixture(oracle):
    for c in json.load(open(os.path.join(GOLD, "fft_kat.json"))):
        x = np.array(c["x"][0]) + 1j * np.array(c["x"][1])
        y = np.array(c["y"][0]) + 1j * np.array(c["y"][1])
        assert np.abs(oracle.fft(x, +1) - y).max() < 1e-12 * np.abs(y).max()
        if c["n"] <= 64:
            assert np.abs(oracle.dft(x, +1) - y).max() < 1e-12 * np.abs(y).max()
        assert np.abs(oracle.fft(x, -1) - np.fft.fft(x)).max() < 1e-12 * np.abs(y).max()


def test_single_site_spectrum_is_gaussian_line(oracle):
    """C(t) = exp(-i e t/hbar): the windowed transform peaks at omega = -e"""
    hb = oracle.hbar_ref()
    eps, dt, T, nfft, sig = 1500.0, 1.0, 4096, 8192, 300.0
    n = np.arange(T + 1)
    corr = np.exp(-1j * eps * n * dt / hb)
    om, S = oracle.spectrum(corr, dt, sig, 1.0, nfft)
    k = np.argmax(np.abs(S))
    dw = 33356.40951981521 / (dt * nfft)
    # inverse-direction transform of exp(-i eps t/hbar) peaks at +eps (in the
    # spectra's 1/(2 pi c) units, which use the true lightspeed)
    line = eps / (2 * np.pi * hb) * 33356.40951981521
    assert abs(om[k] - line) <= dw
    # window and scaling as written: brute force at one frequency
    w = np.where(n < T, np.exp(-0.5 * ((n + 1) * dt / sig) ** 2), 1.0)
    ref = np.sum(corr * w * np.exp(2j * np.pi * k * n / nfft)) / T
    assert abs(S[k] - ref) < 1e-12 * abs(ref)


def test_spectrum_axis_wraps(oracle):
    om, S = oracle.spectrum(np.ones(9, complex), 2.0, 0.0, 2.0, 16)
    assert om[0] == 0.0 and om[8] > 0 and om[9] < 0
    assert om[1] == pytest.approx(1.0 / (299792458. * 100. * 1e-15) / (2.0 * 16))
    assert S[0] == pytest.approx(2.0 * 9 / 8)

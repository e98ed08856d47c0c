"""Orientationally averaged absorption (spectraClass.cu:437-461, 699-744): the
three dipole directions run as one batch with one bra per direction, and the
weighted spectrum is the mean of the three direction spectra."""
import numpy as np
import pytest

from tests.helpers import make_engine, make_oracle, ring_tube

pytestmark = pytest.mark.gpu


def test_isotropic_absorption_against_the_oracle(engine, oracle):
    """600-site ring tube, three directions x 3 realisations, FP64: each
    trajectory's C(n) against the oracle with that direction's ket and bra,
    and the weighted spectrum against the mean of the oracle's three
    direction spectra"""
    sysd = ring_tube(engine)
    n, T, per, sigma, nfft = sysd["n"], 300, 3, 100.0, 512
    mu = sysd["dip"]
    psi0, bras, idx, w = engine.direction_batch(None, mu, per)
    B = 3 * per
    s = make_engine(engine, sysd, 16, 1.0, B=B, force_batched=True, traj_ids=np.arange(B, dtype=np.uint64))
    s.set_state(psi0)
    s.set_observables(None, 0, 0, T, bra=bras[0])
    s.set_bras(bras, idx)
    s.run(T, graph=True)
    C = s.corr()
    om, spec = s.spectrum_weighted(sigma, w, nfft)
    s.close()
    p, sites = make_oracle(oracle, sysd, 16, 1.0)
    mt2 = (mu * mu).sum(0)
    ref = np.zeros(nfft, complex)
    for d in range(3):
        ids = np.arange(d * per, (d + 1) * per)
        r = oracle.run_batch(p, sites, sysd["rp"], sysd["col"], sysd["val"], psi0[ids[0]], ids, T, T,
                             np.zeros(n, np.int32), 1, bra=bras[d])
        err = np.abs(C[ids] - r["corr"]).max()
        print("direction", d, "max |C_gpu - C_oracle| =", err)
        assert err <= 1e-10
        omo, sd = oracle.spectrum(r["corr"].mean(0), 1.0, sigma, mt2[d], nfft)
        ref += sd / 3.0
    assert np.allclose(om, omo, rtol=0, atol=0)
    rel = np.abs(spec - ref).max() / np.abs(ref).max()
    print("isotropic spectrum vs the mean of the oracle's direction spectra:", rel)
    assert rel <= 1e-10


def test_isotropic_absorption_cfg4_mixed_vs_fp64(engine):
    """N = 12,000 double tube, 3 directions x 512 realisations, dt = 1 fs,
    4096 steps, 1 ps Gaussian window, nfft 8192, mixed precision against the
    FP64 path on 64 realisations per direction"""
    from tests.helpers import double_tube
    sysd = double_tube(engine)
    n, T, nfft, sigma = sysd["n"], 4096, 8192, 1000.0
    mu = sysd["dip"]

    def mk(per, prec):
        psi0, bras, idx, w = engine.direction_batch(None, mu, per)
        B = 3 * per
        # trajectory ids: direction d, realisation k -> 1000 d + k
        ids = (1000 * idx + np.tile(np.arange(per), 3)).astype(np.uint64)
        s = engine.System(n, 24, 1.0, n_traj=B, precision=prec, sigma_static=250.0, seed=7)
        s.set_hamiltonian(sysd["rp"], sysd["col"], sysd["val"], np.zeros(n))
        s.set_ou(50.0, 400.0)
        s.set_trajectories(ids)
        s.reset_disorder()
        s.set_state(psi0)
        s.set_observables(None, 0, 0, T, bra=bras[0])
        s.set_bras(bras, idx)
        return s, w, idx

    s, w, idx = mk(512, engine.MIXED)
    s.run(T, graph=True)
    ms = s.last_run_ms()
    om, spec = s.spectrum_weighted(sigma, w, nfft)
    # the 3 x 64 realisations the FP64 run repeats
    sub = np.zeros_like(w)
    for d in range(3):
        sub[d * 512:d * 512 + 64] = w[d * 512] * 512 / 64
    _, spec_sub = s.spectrum_weighted(sigma, sub, nfft)
    cm = s.corr()
    s.close()
    f, wf, _ = mk(64, engine.FP64)
    f.run(T, graph=True)
    _, spec64 = f.spectrum_weighted(sigma, wf, nfft)
    cf = f.corr()
    f.close()
    # the GPU FFT path at full size against numpy, on the FP64 weighted C(n)
    k = np.arange(T + 1)
    win = np.where(k < T, np.exp(-0.5 * ((k + 1) * 1.0 / sigma) ** 2), 1.0)
    xw = np.zeros(nfft, complex)
    xw[:T + 1] = (wf[:, None] * cf).sum(0) * win
    ref64 = np.fft.ifft(xw) * nfft / T
    assert np.abs(spec64 - ref64).max() <= 1e-12 * np.abs(spec64).max()
    rel = np.linalg.norm(spec_sub - spec64) / np.linalg.norm(spec64)
    cerr = max(np.abs(cm[d * 512:d * 512 + 64] - cf[d * 64:(d + 1) * 64]).max() for d in range(3))
    print("isotropic spectra 3 x 512 N=12000 4096 steps: %.3f ms/step, %.0f traj-steps/s"
          % (ms / T, 1536 * T / (ms / 1e3)))
    print("mixed vs FP64 on 3 x 64 realisations: spectrum relative L2 error %.3e, max |dC| %.3e" % (rel, cerr))
    print("isotropic B=1536 spectrum peak at %.1f cm^-1" % om[np.argmax(spec.real)])
    assert rel <= 1e-4

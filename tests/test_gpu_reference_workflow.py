"""The reference's on-disk workflow end to end on the GPU: its own infile.in
(Zhong-Zhao noise at 77 K from its spectral-density file, static disorder
300 cm^-1, m = 12, dt = 1 fs, 5 Krylov-OU substeps, 400 steps, initial state
site 300, ENS_AVG_MODE_FIRST with 5 runs, all populations printed), a
Hamiltonian file written in its format for the 600-site ring tube (its own
hamInner.dat is not in the reference tree), the population file written and
read back, everything against the FP64 oracle fed the same inputs."""
import os

import numpy as np
import pytest

from tests.helpers import ring_tube
from tests.test_io_files import GOLD, write_mm

pytestmark = pytest.mark.gpu


def test_reference_infile_to_population_file(engine, oracle, tmp_path):
    inf = engine.parse_infile(os.path.join(GOLD, "infile.in"))
    assert inf["noise_model"] == engine.NOISE_ZZ and inf["ensemble_mode"] == 1
    sysd = ring_tube(engine)
    ham = tmp_path / "hamInner.dat"
    write_mm(str(ham), sysd["n"], sysd["rp"], sysd["col"], sysd["val"], np.zeros(sysd["n"]), "symmetric", 0)
    n, rp, col, val, e_site = engine.read_hamiltonian(str(ham))
    domega, om, J = engine.read_spectral_density(os.path.join(GOLD, "specd_tresp.dat"))
    coeff = engine.zz_oscillators(om, J, domega, inf["temperature"])
    runs, steps = inf["ensemble_runs"], inf["total_steps"]
    sigma = inf["sigma_static"] if inf["static_dis"] else 0.0

    s = engine.System(n, inf["m"], inf["dt"], n_traj=runs, ou_substeps=inf["ka_substeps"],
                      precision=engine.FP64, sigma_static=sigma, seed=11, force_batched=True,
                      noise_model=inf["noise_model"])
    s.set_hamiltonian(rp, col, val, e_site)
    s.set_zz(om, coeff)
    s.reset_disorder()
    psi0 = np.zeros(n, complex)
    psi0[inf["init_state"]] = 1.0
    s.set_state(psi0)
    sites = np.arange(n, dtype=np.int32)
    s.set_observables(sites, n, 1, steps)
    s.run(steps, graph=True)
    pops = s.populations()                       # [run][record 0..T][site]
    s.close()

    # ENS_AVG_MODE_FIRST: the file holds the sum over runs for steps 1..T
    summed = pops[:, 1:, :].sum(0)
    out = tmp_path / "pops.dat"
    engine.write_pops(str(out), summed, inf["dt"], nsims=runs)
    nsims, dt, back = engine.read_pops(str(out))
    assert nsims == runs and dt == inf["dt"] and back.shape == (steps, n)
    assert np.abs(back - summed).max() <= 1e-15 * runs     # %25.17g text round trip

    p = oracle.make_params(n, inf["m"], inf["dt"], S=inf["ka_substeps"], reorth=oracle.REORTH_GRAM,
                           sigma_static=sigma, seed=11)
    w, c = oracle.zz_oscillators(om, J, domega, inf["temperature"])
    osites = oracle.Sites(n, e_site=e_site, noise=oracle.NOISE_ZZ, zz_w2pi=w, zz_coeff=c)
    rb = oracle.run_batch(p, osites, rp, col, val, psi0, np.arange(runs), steps, 1, sites, n)
    err = np.abs(pops - rb["pops"]).max()
    print("reference infile workflow: %d runs x %d steps of the 600-site tube, ZZ noise at %g K: "
          "max |P - P_oracle| %.3e, summed P(T) of sites 300, 301, 359: %s"
          % (runs, steps, inf["temperature"], err, np.round(summed[-1, [300, 301, 359]], 5)))
    assert err <= 1e-10

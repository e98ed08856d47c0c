"""Writes the golden fixtures under tests/golden/ with numpy and scipy.

The reference (nicolassawaya/DynaQuant) ships no tests or golden vectors and
cannot run here (CUDA + CULA).  These fixtures pin the oracle's building
blocks against independent implementations:
  philox_kat.json   numpy.random.Philox (Philox4x64-10) output blocks
  expm_kat.json     scipy.linalg.expm of complex test matrices
  fft_kat.json      numpy.fft.ifft * n of test sequences
  tube_kat.json     dipole couplings of a few site pairs from numpy
Run: python tests/golden/make_golden.py
"""
import json
import os

import numpy as np
import scipy.linalg as sl

HERE = os.path.dirname(os.path.abspath(__file__))


def philox():
    out = []
    rng = np.random.default_rng(123)
    cases = [([1, 0, 0, 0], [0, 0]), ([1, 2, 3, 4], [5, 6]),
             ([2**64 - 1, 7, 0, 1], [2**63, 11])]
    for _ in range(5):
        ctr = [int(x) for x in rng.integers(1, 2**63, 4)]
        key = [int(x) for x in rng.integers(0, 2**63, 2)]
        cases.append((ctr, key))
    for ctr, key in cases:
        # numpy increments the counter before each block: start one below
        c = list(ctr)
        c0 = (c[0] - 1) % 2**64
        borrow = c[0] == 0
        start = [c0, c[1] - 1 if borrow else c[1], c[2], c[3]]
        bg = np.random.Philox(counter=np.array(start, dtype=np.uint64),
                              key=np.array(key, dtype=np.uint64))
        block = [int(x) for x in bg.random_raw(4)]
        out.append({"ctr": [str(x) for x in ctr], "key": [str(x) for x in key],
                    "out": [str(x) for x in block]})
    return out


def expm():
    rng = np.random.default_rng(7)
    out = []
    for n, scale in [(1, 0.3), (2, 5.0), (5, 2.0), (12, 8.0), (26, 15.0)]:
        A = (rng.normal(size=(n, n)) + 1j * rng.normal(size=(n, n))) * scale / np.sqrt(n)
        E = sl.expm(A)
        out.append({"n": n, "A": [A.real.ravel().tolist(), A.imag.ravel().tolist()],
                    "E": [E.real.ravel().tolist(), E.imag.ravel().tolist()]})
    # purely imaginary times a real symmetric tridiagonal-with-augmentation
    m = 6
    T = np.zeros((m + 2, m + 2))
    a = rng.normal(size=m) * 1e4
    b = np.abs(rng.normal(size=m - 1)) * 5e3
    T[:m, :m] = np.diag(a) + np.diag(b, 1) + np.diag(b, -1)
    T[0, m] = 1.0
    T[m, m + 1] = 1.0
    A = -1j * (2.0 / 5308.835865122816) * T
    E = sl.expm(A)
    out.append({"n": m + 2, "A": [A.real.ravel().tolist(), A.imag.ravel().tolist()],
                "E": [E.real.ravel().tolist(), E.imag.ravel().tolist()],
                "alpha": a.tolist(), "beta": b.tolist(), "theta": 2.0 / 5308.835865122816})
    return out


def fft():
    rng = np.random.default_rng(11)
    out = []
    for n in (8, 64, 1024):
        x = rng.normal(size=n) + 1j * rng.normal(size=n)
        y = np.fft.ifft(x) * n          # e^{+2 pi i k j / n}, unnormalised
        out.append({"n": n, "x": [x.real.tolist(), x.imag.tolist()],
                    "y": [y.real.tolist(), y.imag.tolist()]})
    return out


def tube_pairs():
    """J between a few sites of the 10 x 60 ring tube, straight from the formula"""
    R, dz, n_ring, mu, tilt = 3.0, 0.8, 60, 10.0, np.deg2rad(30.0)

    def site(r, k):
        phi = 2 * np.pi * k / n_ring + (np.pi / n_ring if r % 2 else 0.0)
        p = np.array([R * np.cos(phi), R * np.sin(phi), dz * r])
        t = np.array([-np.sin(phi), np.cos(phi), 0.0])
        u = mu * (np.cos(tilt) * t + np.sin(tilt) * np.array([0, 0, 1.0]))
        return p, u

    out = []
    for (r1, k1), (r2, k2) in [((0, 0), (0, 1)), ((0, 0), (1, 0)), ((0, 0), (1, 59)),
                               ((3, 10), (5, 12)), ((2, 7), (2, 9)), ((0, 0), (0, 30))]:
        p1, u1 = site(r1, k1)
        p2, u2 = site(r2, k2)
        d = p2 - p1
        r = np.linalg.norm(d)
        rh = d / r
        J = 5.04 * (u1 @ u2 - 3 * (u1 @ rh) * (u2 @ rh)) / r**3
        out.append({"i": r1 * n_ring + k1, "j": r2 * n_ring + k2, "r": r, "J": J})
    return out


if __name__ == "__main__":
    for name, fn in [("philox_kat", philox), ("expm_kat", expm), ("fft_kat", fft),
                     ("tube_kat", tube_pairs)]:
        with open(os.path.join(HERE, name + ".json"), "w") as f:
            json.dump(fn(), f)
        print("wrote", name)

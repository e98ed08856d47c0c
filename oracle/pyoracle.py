"""ctypes binding of the FP64 CPU oracle (oracle/_build/liboracle.so).

TEST INFRASTRUCTURE: imported only by tests/, the smoke check and the
CPU-baseline leg of bench.py, always as the checker.  The engine in
dynaquant_amd/ never imports it.
"""
import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "_build", "liboracle.so")
_lib = None

REORTH_NONE, REORTH_CGS, REORTH_GRAM = 0, 1, 2

_dp = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
_ip = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")
_up = np.ctypeslib.ndpointer(dtype=np.uint64, flags="C_CONTIGUOUS")


class Params(C.Structure):
    _fields_ = [
        ("n", C.c_int), ("m", C.c_int), ("krylov_substeps", C.c_int),
        ("ou_substeps", C.c_int), ("reorth", C.c_int),
        ("breakdown_tol", C.c_double), ("dt_fs", C.c_double),
        ("hbar", C.c_double), ("g", C.c_double),
        ("sigma_static", C.c_double), ("seed", C.c_uint64),
        ("tk_p", C.c_int), ("tk_m", C.c_int), ("tk_rowptr", C.c_void_p), ("tk_col", C.c_void_p),
        ("tk_val", C.c_void_p), ("tk_diag", C.c_void_p),
    ]


NOISE_OU, NOISE_HSR, NOISE_ZZ, NOISE_NONE = 0, 1, 2, 3


class SiteParams(C.Structure):
    _fields_ = [
        ("ou_tau", C.c_void_p), ("ou_sigma", C.c_void_p),
        ("e_site", C.c_void_p), ("e_shift", C.c_void_p),
        ("noise_model", C.c_int), ("zz_nosc", C.c_int),
        ("hsr_sigma", C.c_void_p), ("zz_w2pi", C.c_void_p), ("zz_coeff", C.c_void_p),
    ]


class LanczosOpts(C.Structure):
    _fields_ = [("m", C.c_int), ("reorth", C.c_int), ("breakdown_tol", C.c_double)]


def build():
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = C.CDLL(_LIB_PATH)
        L.orc_hbar_ref.restype = C.c_double
        L.orc_pade_coeffs.argtypes = [C.c_int, _dp]
        L.orc_philox4x64_10.argtypes = [_up, _up, _up]
        L.orc_log.restype = C.c_double
        L.orc_log.argtypes = [C.c_double]
        L.orc_sqrt.restype = C.c_double
        L.orc_sqrt.argtypes = [C.c_double]
        L.orc_sincos_2pi.argtypes = [C.c_double, C.POINTER(C.c_double), C.POINTER(C.c_double)]
        L.orc_normal4.argtypes = [C.c_uint64] * 6 + [_dp]
        L.orc_normal.restype = C.c_double
        L.orc_normal.argtypes = [C.c_uint64] * 6
        L.orc_tube_sites.restype = C.c_int
        L.orc_tube_sites.argtypes = [C.c_int, _ip, _dp, C.c_int, C.c_double, C.c_double,
                                     C.c_double, C.c_int, _dp, _dp, _ip]
        L.orc_couplings_csr.restype = C.c_int64
        L.orc_couplings_csr.argtypes = [C.c_int, _dp, _dp, C.c_double, C.c_double,
                                        C.c_void_p, C.c_void_p, C.c_void_p]
        L.orc_spmv.argtypes = [C.c_int, _ip, _ip, _dp, C.c_double, _dp, _dp, _dp]
        L.orc_lanczos.restype = C.c_int
        L.orc_lanczos.argtypes = [C.c_int, _ip, _ip, _dp, C.c_double, _dp, _dp,
                                  C.POINTER(LanczosOpts), _dp, _dp, _dp,
                                  C.POINTER(C.c_double)]
        L.orc_build_tbar.argtypes = [C.c_int, _dp, _dp, _dp]
        L.orc_pade_expm.restype = C.c_int
        L.orc_pade_expm.argtypes = [C.c_int, _dp, _dp, C.POINTER(C.c_int)]
        L.orc_expokit_err.restype = C.c_double
        L.orc_expokit_err.argtypes = [C.c_int, C.c_double, C.c_double, _dp]
        L.orc_init_disorder.argtypes = [C.POINTER(Params), C.POINTER(SiteParams), C.c_uint64, _dp, _dp]
        L.orc_ou_step.argtypes = [C.POINTER(Params), C.POINTER(SiteParams), C.c_uint64, C.c_int64, _dp]
        L.orc_run.restype = C.c_int
        L.orc_run.argtypes = [C.POINTER(Params), C.POINTER(SiteParams), _ip, _ip, _dp, _dp,
                              C.c_uint64, C.c_int64, C.c_int, C.c_void_p, C.c_int,
                              C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]
        L.orc_participation.restype = C.c_double
        L.orc_participation.argtypes = [C.c_int, _dp]
        L.orc_tk_coeffs.argtypes = [C.c_int, C.c_double, C.c_double, _dp]
        L.orc_build_tbar_aug.argtypes = [C.c_int, C.c_int, _dp, _dp, _dp]
        L.orc_tk_cycle.restype = C.c_double
        L.orc_tk_cycle.argtypes = [C.c_int, _ip, _ip, _dp, C.c_double, _dp, C.POINTER(Params),
                                   C.c_double, _dp]
        L.orc_run_batch.restype = C.c_int
        L.orc_run_batch.argtypes = [C.POINTER(Params), C.POINTER(SiteParams), _ip, _ip, _dp, _dp,
                                    _up, C.c_int, C.c_void_p, C.c_int64, C.c_int, C.c_void_p,
                                    C.c_int, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p,
                                    C.c_void_p, C.c_int]
        L.orc_run_controlled.restype = C.c_int
        L.orc_run_controlled.argtypes = [C.POINTER(Params), C.POINTER(SiteParams), _ip, _ip, _dp, _dp,
                                         _up, C.c_int, C.c_int64, C.c_double, C.c_int, C.c_int, C.c_int,
                                         C.c_void_p, C.c_int, C.c_void_p, _dp, _dp, C.c_void_p,
                                         C.POINTER(C.c_int64)]
        L.orc_fft.restype = C.c_int
        L.orc_fft.argtypes = [C.c_int, _dp, C.c_int]
        L.orc_dft.argtypes = [C.c_int, _dp, _dp, C.c_int]
        L.orc_spectrum.restype = C.c_int
        L.orc_spectrum.argtypes = [C.c_int, _dp, C.c_double, C.c_double, C.c_double, C.c_int,
                                   _dp, _dp]
        L.orc_adjoint.restype = C.c_double
        L.orc_adjoint.argtypes = [C.POINTER(Params), C.POINTER(SiteParams), _ip, _ip, _dp, _dp,
                                  C.c_uint64, C.c_int64, C.c_void_p, C.c_int, C.c_void_p,
                                  C.c_void_p, C.c_double, C.c_void_p, C.c_void_p, C.c_int, _dp,
                                  C.POINTER(C.c_double)]
        L.orc_set_adjoint_mode.argtypes = [C.c_int]
        L.orc_noise_step.argtypes = [C.POINTER(Params), C.POINTER(SiteParams), C.c_uint64, C.c_int64, _dp]
        L.orc_zz_phase.restype = C.c_double
        L.orc_zz_phase.argtypes = [C.c_uint64, C.c_uint64, C.c_uint64, C.c_uint64]
        L.orc_zz_cos.restype = C.c_double
        L.orc_zz_cos.argtypes = [C.c_double, C.c_double, C.c_double]
        L.orc_zz_oscillators.argtypes = [C.c_int, _dp, _dp, C.c_double, C.c_double, _dp, _dp]
        _lib = L
    return _lib


def set_adjoint_mode(mode):
    lib().orc_set_adjoint_mode(int(mode))


def _ptr(a):
    return None if a is None else a.ctypes.data


def _c(a):
    return np.ascontiguousarray(a)


def hbar_ref():
    return lib().orc_hbar_ref()


def pade_coeffs(p=6):
    c = np.zeros(p + 1)
    lib().orc_pade_coeffs(p, c)
    return c


def philox(ctr, key):
    out = np.zeros(4, np.uint64)
    lib().orc_philox4x64_10(np.asarray(ctr, np.uint64), np.asarray(key, np.uint64), out)
    return out


def normal4(seed, traj, c0, c1, c2, c3):
    z = np.zeros(4)
    lib().orc_normal4(seed, traj, c0, c1, c2, c3, z)
    return z


def normal(seed, traj, site, step, stream, sub):
    return lib().orc_normal(seed, traj, site, step, stream, sub)


def log(x):
    return lib().orc_log(float(x))


def sqrt(x):
    return lib().orc_sqrt(float(x))


def sincos_2pi(u):
    s, c = C.c_double(), C.c_double()
    lib().orc_sincos_2pi(float(u), C.byref(s), C.byref(c))
    return s.value, c.value


def tube(rings, radius, per_ring=60, ring_dz=0.8, mu_debye=10.0, tilt_deg=30.0, order=0):
    rings = np.asarray(rings, np.int32)
    radius = np.asarray(radius, np.float64)
    n = int(rings.sum()) * per_ring
    pos = np.zeros(3 * n)
    dip = np.zeros(3 * n)
    wall = np.zeros(n, np.int32)
    r = lib().orc_tube_sites(len(rings), rings, radius, per_ring, ring_dz, mu_debye,
                             tilt_deg, order, pos, dip, wall)
    if r < 0:
        raise ValueError("bad tube spec")
    return pos.reshape(n, 3), dip.reshape(n, 3), wall


def couplings_csr(pos, dip, cutoff=2.5, kappa=5.04):
    pos = _c(np.asarray(pos, np.float64).ravel())
    dip = _c(np.asarray(dip, np.float64).ravel())
    n = pos.size // 3
    rowptr = np.zeros(n + 1, np.int32)
    nnz = lib().orc_couplings_csr(n, pos, dip, cutoff, kappa, rowptr.ctypes.data, None, None)
    col = np.zeros(max(nnz, 1), np.int32)
    val = np.zeros(max(nnz, 1))
    lib().orc_couplings_csr(n, pos, dip, cutoff, kappa, rowptr.ctypes.data,
                            col.ctypes.data, val.ctypes.data)
    return rowptr, col[:nnz].copy(), val[:nnz].copy()


def c2i(z):
    """complex array -> interleaved float64"""
    z = np.asarray(z, np.complex128)
    return _c(z.view(np.float64).reshape(-1)).copy()


def i2c(a):
    return np.asarray(a, np.float64).view(np.complex128).copy()


def spmv(rowptr, col, val, diag, x, g=1.0):
    n = rowptr.size - 1
    y = np.zeros(2 * n)
    lib().orc_spmv(n, _c(rowptr), _c(col), _c(val), g, _c(np.asarray(diag, np.float64)), c2i(x), y)
    return i2c(y)


def lanczos(rowptr, col, val, diag, psi, m, reorth=REORTH_CGS, tol=1e-10, g=1.0):
    n = rowptr.size - 1
    Q = np.zeros(2 * n * (m + 2))
    alpha = np.zeros(m + 2)
    beta = np.zeros(m + 2)
    nh = C.c_double()
    o = LanczosOpts(m, reorth, tol)
    broke = lib().orc_lanczos(n, _c(rowptr), _c(col), _c(val), g, _c(np.asarray(diag, np.float64)),
                              c2i(psi), C.byref(o), Q, alpha, beta, C.byref(nh))
    return i2c(Q).reshape(m + 2, n), alpha, beta, nh.value, broke


def build_tbar(m, alpha, beta):
    t = np.zeros((m + 2) * (m + 2))
    lib().orc_build_tbar(m, _c(np.asarray(alpha, np.float64)), _c(np.asarray(beta, np.float64)), t)
    return t.reshape(m + 2, m + 2).T  # column-major -> numpy [row, col]


def pade_expm(A):
    A = np.asarray(A, np.complex128)
    n = A.shape[0]
    a = c2i(A.T.ravel())  # column-major
    e = np.zeros(2 * n * n)
    s = C.c_int()
    r = lib().orc_pade_expm(n, a, e, C.byref(s))
    if r != 0:
        raise RuntimeError("pade failed %d" % r)
    return i2c(e).reshape(n, n).T, s.value


def expokit_err(m, beta_m, nrm_hq, E):
    e = c2i(np.asarray(E, np.complex128).T.ravel())
    return lib().orc_expokit_err(m, beta_m, nrm_hq, e)


def make_params(n, m, dt, K=1, S=1, reorth=REORTH_CGS, tol=1e-10, hbar=None, g=1.0,
                sigma_static=0.0, seed=1):
    return Params(n, m, K, S, reorth, tol, dt, hbar_ref() if hbar is None else hbar, g,
                  sigma_static, seed)


def set_taylor_krylov(p, tk_p, tk_m, rowptr, col, val, diag):
    """switch p to Taylor-Krylov propagation with the auxiliary matrix
    (rowptr, col, val) + diag; returns the arrays, which must stay alive as
    long as p is used"""
    keep = (_c(np.asarray(rowptr, np.int32)), _c(np.asarray(col, np.int32)),
            _c(np.asarray(val, np.float64)), _c(np.asarray(diag, np.float64)))
    p.tk_p, p.tk_m = tk_p, tk_m
    p.tk_rowptr, p.tk_col, p.tk_val, p.tk_diag = (a.ctypes.data for a in keep)
    return keep


def tk_coeffs(p, tau, hbar=None):
    out = np.zeros(2 * p)
    lib().orc_tk_coeffs(p, tau, hbar_ref() if hbar is None else hbar, out)
    return i2c(out)


def build_tbar_aug(m, aug, alpha, beta):
    """(m + aug)^2 augmented tridiagonal from alpha[1..m], beta[1..m-1]
    (index 0 unused, as orc_lanczos returns them), row-major numpy"""
    n = m + aug
    tb = np.zeros(n * n)
    lib().orc_build_tbar_aug(m, aug, _c(np.asarray(alpha, np.float64)), _c(np.asarray(beta, np.float64)), tb)
    return tb.reshape(n, n).T


def tk_cycle(p, rowptr, col, val, diag, psi, tau):
    """one Taylor-Krylov cycle: (psi', unclamped Krylov error estimate)"""
    a = c2i(np.asarray(psi, np.complex128))
    err = lib().orc_tk_cycle(p.n, _c(rowptr), _c(col), _c(val), p.g, _c(np.asarray(diag, np.float64)),
                             C.byref(p), tau, a)
    return i2c(a), err


class Sites:
    """keeps the arrays alive for a SiteParams struct; noise = NOISE_OU (tau,
    sigma), NOISE_HSR (hsr_sigma), NOISE_ZZ (zz_w2pi, zz_coeff) or NOISE_NONE"""

    def __init__(self, n, tau=None, sigma=None, e_site=None, e_shift=None, noise=NOISE_OU,
                 hsr_sigma=None, zz_w2pi=None, zz_coeff=None):
        f = lambda v: None if v is None else _c(np.broadcast_to(np.asarray(v, np.float64), (n,)).copy())
        self.tau, self.sigma, self.e_site, self.e_shift = f(tau), f(sigma), f(e_site), f(e_shift)
        self.hsr = f(hsr_sigma)
        self.zw = None if zz_w2pi is None else _c(np.asarray(zz_w2pi, np.float64))
        self.zc = None if zz_coeff is None else _c(np.asarray(zz_coeff, np.float64))
        nosc = 0 if self.zw is None else len(self.zw)
        self.sp = SiteParams(_ptr(self.tau), _ptr(self.sigma), _ptr(self.e_site), _ptr(self.e_shift),
                             noise, nosc, _ptr(self.hsr), _ptr(self.zw), _ptr(self.zc))


def init_disorder(p, sites, traj):
    e0 = np.zeros(p.n)
    de = np.zeros(p.n)
    lib().orc_init_disorder(C.byref(p), C.byref(sites.sp), traj, e0, de)
    return e0, de


def ou_step(p, sites, traj, step, de):
    de = _c(np.array(de, np.float64))
    lib().orc_ou_step(C.byref(p), C.byref(sites.sp), traj, step, de)
    return de


def noise_step(p, sites, traj, step, de):
    """the noise offset of step `step` for the Sites' noise model"""
    de = _c(np.array(de, np.float64))
    lib().orc_noise_step(C.byref(p), C.byref(sites.sp), traj, step, de)
    return de


def zz_phase(seed, traj, site, k):
    return lib().orc_zz_phase(seed, traj, site, k)


def zz_cos(u, w2pi, t):
    return lib().orc_zz_cos(u, w2pi, t)


def zz_oscillators(omega_cm, J, domega, temperature):
    """(omega/2pi in 1/fs, amplitudes in cm^-1) of a sampled spectral density"""
    om = _c(np.asarray(omega_cm, np.float64))
    jj = _c(np.asarray(J, np.float64))
    w = np.zeros(len(om))
    c = np.zeros(len(om))
    lib().orc_zz_oscillators(len(om), om, jj, domega, temperature, w, c)
    return w, c


def run(p, sites, rowptr, col, val, psi0, traj, nsteps, obs_every=0, group=None,
        ngroups=0, bra=None, partic=False):
    n = p.n
    nobs = nsteps // obs_every + 1 if obs_every > 0 else 0
    pops = np.zeros(max(nobs * ngroups, 1))
    corr = np.zeros(2 * (nsteps + 1)) if bra is not None else None
    pr = np.zeros(nsteps + 1) if partic else None
    psi = np.zeros(2 * n)
    err = np.zeros(1)
    grp = None if group is None else _c(np.asarray(group, np.int32))
    braa = None if bra is None else c2i(bra)
    rc = lib().orc_run(C.byref(p), C.byref(sites.sp), _c(rowptr), _c(col), _c(val), c2i(psi0),
                       traj, nsteps, obs_every, _ptr(grp), ngroups, _ptr(braa),
                       pops.ctypes.data, _ptr(corr), _ptr(pr), psi.ctypes.data, err.ctypes.data)
    if rc != 0:
        raise RuntimeError("orc_run failed %d" % rc)
    return dict(pops=pops[: nobs * ngroups].reshape(nobs, ngroups) if nobs else None,
                corr=None if corr is None else i2c(corr), psi=i2c(psi), err=err[0], partic=pr)


def run_batch(p, sites, rowptr, col, val, psi0, traj_ids, nsteps, obs_every, group, ngroups,
              n_active=None, bra=None, nthreads=0, partic=False):
    n = p.n
    traj_ids = np.asarray(traj_ids, np.uint64)
    nb = traj_ids.size
    psi0 = np.asarray(psi0, np.complex128)
    if psi0.ndim == 1:
        psi0 = np.broadcast_to(psi0, (nb, n))
    p0 = c2i(psi0.ravel())
    nobs = nsteps // obs_every + 1
    pops = np.zeros(nb * nobs * ngroups)
    corr = np.zeros(nb * 2 * (nsteps + 1)) if bra is not None else None
    pr = np.zeros(nb * (nsteps + 1)) if partic else None
    err = np.zeros(nb)
    grp = _c(np.asarray(group, np.int32))
    na = None if n_active is None else _c(np.asarray(n_active, np.int32))
    braa = None if bra is None else c2i(bra)
    rc = lib().orc_run_batch(C.byref(p), C.byref(sites.sp), _c(rowptr), _c(col), _c(val), p0,
                             traj_ids, nb, _ptr(na), nsteps, obs_every, grp.ctypes.data, ngroups,
                             _ptr(braa), pops.ctypes.data, _ptr(corr), _ptr(pr), err.ctypes.data, nthreads)
    if rc != 0:
        raise RuntimeError("orc_run_batch failed")
    return dict(pops=pops.reshape(nb, nobs, ngroups),
                corr=None if corr is None else i2c(corr).reshape(nb, nsteps + 1), err=err,
                partic=None if pr is None else pr.reshape(nb, nsteps + 1))


def run_controlled(p, sites, rowptr, col, val, psi0, traj_ids, nsteps, tol, kmax, chunk, obs_every=0,
                   group=None, ngroups=0):
    """substep control restated: the batch advanced chunk by chunk, K doubled
    while K x the largest cycle estimate exceeds tol"""
    n = p.n
    traj_ids = np.asarray(traj_ids, np.uint64)
    nb = traj_ids.size
    psi0 = np.asarray(psi0, np.complex128)
    if psi0.ndim == 1:
        psi0 = np.broadcast_to(psi0, (nb, n))
    nobs = nsteps // obs_every + 1 if obs_every > 0 else 0
    pops = np.zeros(max(nb * nobs * ngroups, 1))
    psi = np.zeros(2 * n * nb)
    err = np.zeros(nb)
    ks = np.zeros(nsteps, np.int32)
    rej = C.c_int64()
    grp = None if group is None else _c(np.asarray(group, np.int32))
    rc = lib().orc_run_controlled(C.byref(p), C.byref(sites.sp), _c(rowptr), _c(col), _c(val),
                                  c2i(psi0.ravel()), traj_ids, nb, nsteps, tol, kmax, chunk, obs_every,
                                  _ptr(grp), ngroups, pops.ctypes.data if nobs else None, psi, err,
                                  ks.ctypes.data, C.byref(rej))
    if rc != 0:
        raise RuntimeError("orc_run_controlled failed")
    return dict(pops=pops[:nb * nobs * ngroups].reshape(nb, nobs, ngroups) if nobs else None,
                psi=i2c(psi).reshape(nb, n), err=err, k_of_step=ks, rejects=int(rej.value))


def participation(psi):
    """1 / sum_i |psi_i|^4 (systemClass.cu:2209-2257)"""
    a = c2i(np.asarray(psi, np.complex128))
    return lib().orc_participation(len(a) // 2, a)


def cd_vectors(pos, mu):
    """R x mu per site and the component norms (spectraClass.cu:346-411):
    rd[i] = (y mu_z - z mu_y, z mu_x - x mu_z, x mu_y - y mu_x), mu_tot[d] =
    sqrt(sum_i mu_d,i^2), rd_tot[d] = sqrt(sum_i rd_d,i^2), summed in site
    order."""
    pos = np.asarray(pos, np.float64)
    mu = np.asarray(mu, np.float64)
    n = len(pos)
    rd = np.zeros((n, 3))
    m2 = [0.0, 0.0, 0.0]
    r2 = [0.0, 0.0, 0.0]
    for i in range(n):
        x, y, z = (float(v) for v in pos[i])
        ux, uy, uz = (float(v) for v in mu[i])
        c = (y * uz - z * uy, z * ux - x * uz, x * uy - y * ux)
        rd[i] = c
        for d in range(3):
            m2[d] += (ux, uy, uz)[d] * (ux, uy, uz)[d]
            r2[d] += c[d] * c[d]
    return rd, np.sqrt(np.array(m2)), np.sqrt(np.array(r2))


CD_PREFACTOR = np.pi / 2 * 1e-7     # spectraClass.cu:645-647


def fft(x, sign=+1):
    a = c2i(x)
    if lib().orc_fft(len(x), a, sign) != 0:
        raise ValueError("length must be a power of two")
    return i2c(a)


def dft(x, sign=+1):
    y = np.zeros(2 * len(x))
    lib().orc_dft(len(x), c2i(x), y, sign)
    return i2c(y)


def spectrum(corr, dt, sigma, mu2, nfft):
    T = len(corr) - 1
    om = np.zeros(nfft)
    sp = np.zeros(2 * nfft)
    if lib().orc_spectrum(T, c2i(corr), dt, sigma, mu2, nfft, om, sp) != 0:
        raise ValueError("bad spectrum args")
    return om, i2c(sp)


def adjoint(p, sites, rowptr, col, val, psi0, traj, nsteps, group=None, ngroups=0, w_pop=None,
            bra=None, w_spec=0.0, corr_target=None, ckpt=50, corr_weight=None):
    """(loss, dL/d e_site[n], dL/dg) by the exact reverse sweep"""
    n = p.n
    grp = None if group is None else _c(np.asarray(group, np.int32))
    wp = None if w_pop is None else _c(np.asarray(w_pop, np.float64))
    braa = None if bra is None else c2i(bra)
    ct = None if corr_target is None else c2i(corr_target)
    cw = None if corr_weight is None else _c(np.asarray(corr_weight, np.float64))
    ge = np.zeros(n)
    gg = C.c_double()
    loss = lib().orc_adjoint(C.byref(p), C.byref(sites.sp), _c(rowptr), _c(col), _c(val), c2i(psi0),
                             traj, nsteps, _ptr(grp), ngroups, _ptr(wp), _ptr(braa), w_spec, _ptr(ct),
                             _ptr(cw), ckpt, ge, C.byref(gg))
    return loss, ge, gg.value


def loss_fd(p, sites_kw, rowptr, col, val, psi0, traj, nsteps, group, ngroups, w_pop,
            bra=None, w_spec=0.0, corr_target=None, e_shift=None, corr_weight=None):
    """the same loss evaluated by orc_run (for finite differences)"""
    n = p.n
    s = Sites(n, e_shift=e_shift, **sites_kw)
    r = run(p, s, rowptr, col, val, psi0, traj, nsteps, obs_every=nsteps, group=group,
            ngroups=ngroups, bra=bra)
    L = float(np.dot(w_pop, r["pops"][-1])) if w_pop is not None else 0.0
    if bra is not None and w_spec != 0.0:
        ct = 0.0 if corr_target is None else np.asarray(corr_target)
        u = 1.0 if corr_weight is None else np.asarray(corr_weight)
        L += w_spec * float(np.sum(u * np.abs(r["corr"] - ct) ** 2))
    return L

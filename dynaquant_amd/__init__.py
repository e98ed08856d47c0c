"""dynaquant_amd -- MI355X (gfx950) stochastic exciton-dynamics engine.

Python mirror of the reference's system / propagate / spectra seams over
the C ABI in include/dynaquant.h (libdynaquant_amd.so, built in-tree by
dynaquant_amd.build).  Loading fails loudly when the library is missing:
there is no CPU fallback.
"""
import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libdynaquant_amd.so")

FP64, MIXED = 0, 1
NOISE_OU, NOISE_HSR, NOISE_ZZ, NOISE_NONE = 0, 1, 2, 3

STATUS = {0: "OK", 1: "INVALID_ARG", 2: "OOM", 3: "HIP_ERROR", 4: "BREAKDOWN",
          5: "NONFINITE", 6: "SCALING_OVERFLOW", 7: "NO_DEVICE", 8: "STATE"}

EXPORTS = [
    "dq_system_create", "dq_system_destroy", "dq_last_error", "dq_sync",
    "dq_system_set_hamiltonian_csr", "dq_system_set_ou_params", "dq_system_set_trajectories",
    "dq_system_reset_disorder", "dq_system_set_state", "dq_system_get_state",
    "dq_system_set_observables", "dq_tube_sites", "dq_tube_couplings", "dq_update_noise",
    "dq_propagate_step", "dq_run", "dq_get_step", "dq_get_populations", "dq_get_corr",
    "dq_get_error_sum", "dq_get_diagonal", "dq_spmv", "dq_pade_tbar", "dq_spectrum", "dq_fft",
    "dq_last_run_ms", "dq_set_profiling", "dq_get_kernel_times", "dq_adjoint",
    "dq_get_tuning", "dq_set_tuning", "dq_autotune",
    "dq_system_set_hsr_params", "dq_system_set_zz_params", "dq_zz_oscillators",
    "dq_read_hamiltonian_file", "dq_read_state_file", "dq_read_dipole_file", "dq_read_site_columns",
    "dq_read_spectral_density", "dq_parse_infile", "dq_write_pops_file", "dq_read_pops_file",
    "dq_read_position_file", "dq_cd_vectors", "dq_system_set_participation", "dq_get_participation",
    "dq_system_set_taylor_krylov", "dq_system_snapshot", "dq_system_restore", "dq_system_set_bras",
    "dq_spectrum_weighted", "dq_system_set_krylov_control", "dq_get_krylov_control",
    "dq_spectrum_noalloc",
]

KERNEL_CLASSES = ["prologue", "spmm", "scalars", "update", "reduce", "spmm_last", "pade",
                  "combine", "norm"]


class DQError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__("%s (%d): %s" % (STATUS.get(code, "?"), code, msg))
        self.code = code


class Config(C.Structure):
    _fields_ = [
        ("n_sites", C.c_int32), ("n_traj", C.c_int32), ("krylov_m", C.c_int32),
        ("krylov_substeps", C.c_int32), ("ou_substeps", C.c_int32), ("precision", C.c_int32),
        ("force_batched", C.c_int32), ("noise_model", C.c_int32),
        ("dt_fs", C.c_double), ("hbar_cm_fs", C.c_double), ("coupling_scale", C.c_double),
        ("sigma_static", C.c_double), ("breakdown_tol", C.c_double), ("seed", C.c_uint64),
    ]


class Tuning(C.Structure):
    _fields_ = [
        ("spmm_rows_per_wave", C.c_int32), ("spmm_traj_per_lane", C.c_int32),
        ("spmm_phases", C.c_int32), ("tile_rows", C.c_int32), ("spmm_pipeline", C.c_int32),
        ("concurrent_slices", C.c_int32), ("spmm_window_rows", C.c_int32),
        ("spmm_window_waves", C.c_int32), ("spmm_exact_sums", C.c_int32),
    ]


TUNING_KEYS = ["spmm_rows_per_wave", "spmm_traj_per_lane", "spmm_phases", "tile_rows", "spmm_pipeline",
               "concurrent_slices", "spmm_window_rows", "spmm_window_waves", "spmm_exact_sums"]


class TubeSpec(C.Structure):
    _fields_ = [
        ("n_walls", C.c_int32), ("per_ring", C.c_int32), ("order", C.c_int32),
        ("rings", C.c_int32 * 4), ("radius_nm", C.c_double * 4), ("ring_dz_nm", C.c_double),
        ("mu_debye", C.c_double), ("tilt_deg", C.c_double), ("cutoff_nm", C.c_double),
        ("kappa", C.c_double),
    ]


_lib = None
_dp = C.POINTER(C.c_double)


def lib():
    """Load libdynaquant_amd.so; raises if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError("libdynaquant_amd.so not built: run python -m dynaquant_amd.build")
        L = C.CDLL(LIB_PATH)
        L.dq_last_error.restype = C.c_char_p
        for name in EXPORTS:
            f = getattr(L, name)
            if name != "dq_last_error":
                f.restype = C.c_int
        _lib = L
    return _lib


def _check(code):
    if code != 0:
        raise DQError(code, lib().dq_last_error().decode())


def _p(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


def _c(a, dt):
    return np.ascontiguousarray(a, dtype=dt)


def c2i(z):
    return np.ascontiguousarray(np.asarray(z, np.complex128)).view(np.float64).copy()


# --------------------------------------------------------------------------
# Hamiltonian generator
# --------------------------------------------------------------------------
def tube_spec(rings, radius, per_ring=60, order=0, ring_dz=0.8, mu_debye=10.0, tilt_deg=30.0,
              cutoff=2.5, kappa=5.04):
    s = TubeSpec()
    s.n_walls = len(rings)
    s.per_ring = per_ring
    s.order = order
    for w in range(len(rings)):
        s.rings[w] = int(rings[w])
        s.radius_nm[w] = float(radius[w])
    s.ring_dz_nm, s.mu_debye, s.tilt_deg, s.cutoff_nm, s.kappa = ring_dz, mu_debye, tilt_deg, cutoff, kappa
    return s


def tube(rings, radius, **kw):
    """-> (pos[n,3], dip[n,3], wall[n], rowptr, col, val) of a dipole tube."""
    s = tube_spec(rings, radius, **kw)
    n = C.c_int32()
    _check(lib().dq_tube_sites(C.byref(s), None, None, None, C.byref(n)))
    n = n.value
    pos = np.zeros(3 * n)
    dip = np.zeros(3 * n)
    wall = np.zeros(n, np.int32)
    _check(lib().dq_tube_sites(C.byref(s), _p(pos), _p(dip), _p(wall), None))
    rowptr = np.zeros(n + 1, np.int32)
    nnz = C.c_int64()
    _check(lib().dq_tube_couplings(C.byref(s), _p(rowptr), None, None, C.byref(nnz)))
    col = np.zeros(max(nnz.value, 1), np.int32)
    val = np.zeros(max(nnz.value, 1))
    _check(lib().dq_tube_couplings(C.byref(s), _p(rowptr), _p(col), _p(val), None))
    return pos.reshape(n, 3), dip.reshape(n, 3), wall, rowptr, col[: nnz.value], val[: nnz.value]


# --------------------------------------------------------------------------
# System (systemClass / propagate / spectraClass seams)
# --------------------------------------------------------------------------
class System:
    def __init__(self, n, m, dt_fs, n_traj=1, krylov_substeps=1, ou_substeps=1, precision=FP64,
                 hbar=0.0, coupling_scale=1.0, sigma_static=0.0, breakdown_tol=0.0, seed=1,
                 force_batched=False, noise_model=NOISE_OU):
        cfg = Config(n, n_traj, m, krylov_substeps, ou_substeps, precision, int(force_batched),
                     int(noise_model), dt_fs, hbar, coupling_scale, sigma_static, breakdown_tol, seed)
        self.cfg = cfg
        self.n, self.B, self.m = n, n_traj, m
        h = C.c_void_p()
        _check(lib().dq_system_create(C.byref(cfg), C.byref(h)))
        self._h = h
        self.every = 0
        self.ngroups = 0

    def close(self):
        if self._h:
            lib().dq_system_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def set_hamiltonian(self, rowptr, col, val, e_site=None):
        rowptr, col, val = _c(rowptr, np.int32), _c(col, np.int32), _c(val, np.float64)
        es = None if e_site is None else _c(e_site, np.float64)
        _check(lib().dq_system_set_hamiltonian_csr(self._h, self.n, _p(rowptr), _p(col), _p(val), _p(es)))

    def set_ou(self, tau_fs, sigma):
        tau = _c(np.broadcast_to(tau_fs, (self.n,)), np.float64)
        sig = _c(np.broadcast_to(sigma, (self.n,)), np.float64)
        _check(lib().dq_system_set_ou_params(self._h, _p(tau), _p(sig)))

    def set_hsr(self, sigma):
        sig = _c(np.broadcast_to(sigma, (self.n,)), np.float64)
        _check(lib().dq_system_set_hsr_params(self._h, _p(sig)))

    def set_zz(self, omega_cm, coeff):
        om = _c(omega_cm, np.float64)
        co = _c(coeff, np.float64)
        if om.shape != co.shape:
            raise ValueError("omega and coeff differ in length")
        _check(lib().dq_system_set_zz_params(self._h, len(om), _p(om), _p(co)))

    def set_taylor_krylov(self, p, m_aux=None, rowptr=None, col=None, val=None, diag=None):
        """Taylor-Krylov propagation of order p with an m_aux-step Krylov part
        on the auxiliary matrix (rowptr, col, val) + diag; p = 0: Lanczos"""
        if p == 0:
            _check(lib().dq_system_set_taylor_krylov(self._h, 0, 0, None, None, None, None))
            return
        rp = _c(rowptr, np.int32)
        cl = _c(col, np.int32)
        vl = _c(val, np.float64)
        dg = _c(np.broadcast_to(np.asarray(diag, np.float64), (self.n,)), np.float64)
        _check(lib().dq_system_set_taylor_krylov(self._h, int(p), int(m_aux), _p(rp), _p(cl), _p(vl), _p(dg)))

    def set_trajectories(self, traj_ids=None, n_active=None):
        t = None if traj_ids is None else _c(traj_ids, np.uint64)
        a = None if n_active is None else _c(n_active, np.int32)
        _check(lib().dq_system_set_trajectories(self._h, _p(t), _p(a)))

    def reset_disorder(self):
        _check(lib().dq_system_reset_disorder(self._h))

    def set_state(self, psi):
        psi = np.asarray(psi, np.complex128)
        if psi.ndim == 1:
            psi = np.broadcast_to(psi, (self.B, self.n))
        a = c2i(psi.reshape(-1))
        _check(lib().dq_system_set_state(self._h, _p(a)))

    def get_state(self):
        a = np.zeros(2 * self.n * self.B)
        _check(lib().dq_system_get_state(self._h, _p(a)))
        return a.view(np.complex128).reshape(self.B, self.n)

    def set_krylov_control(self, tol, k_max=16, chunk=10):
        """substeps from the Expokit estimate: tol per dt (<= 0: off)"""
        _check(lib().dq_system_set_krylov_control(self._h, C.c_double(tol), int(k_max), int(chunk)))

    def krylov_control(self):
        """(current K, chunks redone)"""
        k = C.c_int32()
        r = C.c_int64()
        _check(lib().dq_get_krylov_control(self._h, C.byref(k), C.byref(r)))
        return int(k.value), int(r.value)

    def snapshot(self):
        """resume point: dict(psi [B, n], noise [B, n], err_sum [B], step)"""
        psi = np.zeros(2 * self.B * self.n)
        noise = np.zeros(self.B * self.n)
        err = np.zeros(self.B)
        step = C.c_int64()
        _check(lib().dq_system_snapshot(self._h, _p(psi), _p(noise), _p(err), C.byref(step)))
        return dict(psi=psi.view(np.complex128).reshape(self.B, self.n), noise=noise.reshape(self.B, self.n),
                    err_sum=err, step=int(step.value))

    def restore(self, snap):
        """continue from a snapshot (after reset_disorder)"""
        psi = c2i(np.asarray(snap["psi"]).reshape(-1))
        noise = _c(np.asarray(snap["noise"]).reshape(-1), np.float64)
        err = None if snap.get("err_sum") is None else _c(snap["err_sum"], np.float64)
        _check(lib().dq_system_restore(self._h, _p(psi), _p(noise), _p(err), C.c_int64(int(snap["step"]))))

    def set_bras(self, bras, bra_of_traj):
        """C_b(n) = <bras[bra_of_traj[b]]|psi_b(n)> (after set_observables)"""
        bras = np.atleast_2d(np.asarray(bras, np.complex128))
        idx = _c(np.asarray(bra_of_traj, np.int32), np.int32)
        if idx.size != self.B or bras.shape[1] != self.n:
            raise ValueError("bra_of_traj needs B entries and bras n columns")
        _check(lib().dq_system_set_bras(self._h, int(bras.shape[0]), _p(c2i(bras.ravel())), _p(idx)))

    def spectrum_weighted(self, sigma_fs, weight, nfft):
        """(omega [nfft], spectrum [nfft]) of sum_b weight[b] C_b(n), scaled 1/T"""
        om = np.zeros(nfft)
        sp = np.zeros(2 * nfft)
        w = _c(weight, np.float64)
        if w.size != self.B:
            raise ValueError("one weight per trajectory")
        _check(lib().dq_spectrum_weighted(self._h, C.c_double(sigma_fs), _p(w), int(nfft), _p(om), _p(sp)))
        return om, sp.view(np.complex128)

    def set_observables(self, group=None, ngroups=0, every=0, max_steps=0, bra=None):
        g = None if group is None else _c(group, np.int32)
        br = None if bra is None else c2i(bra)
        _check(lib().dq_system_set_observables(self._h, _p(g), ngroups, every, max_steps, _p(br)))
        self.every, self.ngroups = every, ngroups

    def set_participation(self, on=True):
        """record 1 / sum_i |psi_i|^4 every step (after set_observables)"""
        _check(lib().dq_system_set_participation(self._h, int(on)))

    def update_noise(self):
        _check(lib().dq_update_noise(self._h))

    def step(self):
        _check(lib().dq_propagate_step(self._h))

    def run(self, nsteps, graph=False):
        _check(lib().dq_run(self._h, C.c_int64(nsteps), int(graph)))

    def last_run_ms(self):
        ms = C.c_double()
        _check(lib().dq_last_run_ms(self._h, C.byref(ms)))
        return ms.value

    def set_profiling(self, on=True):
        _check(lib().dq_set_profiling(self._h, int(on)))

    def tuning(self):
        """current launch choices of the batched kernels"""
        t = Tuning()
        _check(lib().dq_get_tuning(self._h, C.byref(t)))
        return {k: int(getattr(t, k)) for k in TUNING_KEYS}

    def set_tuning(self, **kw):
        t = Tuning()
        _check(lib().dq_get_tuning(self._h, C.byref(t)))
        for k, v in kw.items():
            if k not in TUNING_KEYS:
                raise KeyError(k)
            setattr(t, k, int(v))
        _check(lib().dq_set_tuning(self._h, C.byref(t)))

    def autotune(self, reps=3, cache=None):
        """time every launch choice on this system, apply the fastest, return it"""
        t = Tuning()
        path = None if cache is None else os.fsencode(cache)
        _check(lib().dq_autotune(self._h, int(reps), path, C.byref(t)))
        return {k: int(getattr(t, k)) for k in TUNING_KEYS}

    def kernel_times(self):
        """{class: (total ms, launches)} accumulated in profiling mode"""
        ms = np.zeros(len(KERNEL_CLASSES))
        cnt = np.zeros(len(KERNEL_CLASSES), np.int64)
        _check(lib().dq_get_kernel_times(self._h, _p(ms), _p(cnt)))
        return {k: (float(ms[i]), int(cnt[i])) for i, k in enumerate(KERNEL_CLASSES)}

    def adjoint(self, nsteps, ckpt=50, w_pop=None, w_spec=0.0, corr_target=None, corr_weight=None):
        """(loss[B], dL/d e_site [B, n], dL/dg [B]) by the reverse sweep"""
        wp = None if w_pop is None else _c(w_pop, np.float64)
        ct = None if corr_target is None else c2i(corr_target)
        cw = None if corr_weight is None else _c(corr_weight, np.float64)
        if cw is not None and len(cw) != nsteps + 1:
            raise ValueError("corr_weight needs nsteps + 1 entries")
        loss = np.zeros(self.B)
        ge = np.zeros(self.B * self.n)
        gg = np.zeros(self.B)
        _check(lib().dq_adjoint(self._h, C.c_int64(nsteps), int(ckpt), _p(wp), C.c_double(w_spec),
                                _p(ct), _p(cw), _p(loss), _p(ge), _p(gg)))
        return loss, ge.reshape(self.B, self.n), gg

    def get_step(self):
        s = C.c_int64()
        _check(lib().dq_get_step(self._h, C.byref(s)))
        return s.value

    def populations(self):
        nrec = C.c_int64()
        _check(lib().dq_get_populations(self._h, None, C.byref(nrec)))
        G = self.ngroups
        out = np.zeros(self.B * nrec.value * G)
        _check(lib().dq_get_populations(self._h, _p(out), C.byref(nrec)))
        return out.reshape(self.B, nrec.value, G)

    def corr(self):
        T = self.get_step()
        out = np.zeros(2 * self.B * (T + 1))
        _check(lib().dq_get_corr(self._h, _p(out)))
        return out.view(np.complex128).reshape(self.B, T + 1)

    def participation(self):
        """pr[b][n] = 1 / sum_i |psi_b,i(n)|^4, n = 0..step"""
        T = self.get_step()
        out = np.zeros(self.B * (T + 1))
        _check(lib().dq_get_participation(self._h, _p(out)))
        return out.reshape(self.B, T + 1)

    def error_sum(self):
        out = np.zeros(self.B)
        _check(lib().dq_get_error_sum(self._h, _p(out)))
        return out

    def diagonal(self):
        out = np.zeros(self.B * self.n)
        _check(lib().dq_get_diagonal(self._h, _p(out)))
        return out.reshape(self.B, self.n)

    def spmv(self, traj, x):
        xi = c2i(x)
        y = np.zeros(2 * self.n)
        _check(lib().dq_spmv(self._h, traj, _p(xi), _p(y)))
        return y.view(np.complex128)

    def spectrum(self, sigma_fs, mu2, nfft):
        om = np.zeros(nfft)
        sp = np.zeros(2 * nfft)
        _check(lib().dq_spectrum(self._h, C.c_double(sigma_fs), C.c_double(mu2), nfft, _p(om), _p(sp)))
        return om, sp.view(np.complex128)

    def spectrum_noalloc(self, sigma_fs, mu2, nfft):
        """spectrum() in the library's fixed device workspace (nfft <= 131072)"""
        om = np.zeros(nfft)
        sp = np.zeros(2 * nfft)
        _check(lib().dq_spectrum_noalloc(self._h, C.c_double(sigma_fs), C.c_double(mu2), nfft, _p(om),
                                         _p(sp)))
        return om, sp.view(np.complex128)


def zz_oscillators(omega_cm, J, domega, temperature):
    """amplitudes (cm^-1) of a spectral density J(omega_k) sampled every domega"""
    om = _c(omega_cm, np.float64)
    jj = _c(J, np.float64)
    out = np.zeros(len(om))
    _check(lib().dq_zz_oscillators(len(om), _p(om), _p(jj), C.c_double(domega), C.c_double(temperature),
                                   _p(out)))
    return out


# ---------------------------------------------------------------------------
# files of the reference's workflow
# ---------------------------------------------------------------------------
class Infile(C.Structure):
    _fields_ = [
        ("dt", C.c_double), ("temperature", C.c_double), ("sigma_static", C.c_double),
        ("gauss_width", C.c_double), ("total_steps", C.c_int32), ("ensemble_runs", C.c_int32),
        ("ensemble_mode", C.c_int32), ("m", C.c_int32), ("ka_substeps", C.c_int32),
        ("noise_model", C.c_int32), ("prop_mode", C.c_int32), ("static_dis", C.c_int32),
        ("absorb_spec", C.c_int32), ("cd_spec", C.c_int32), ("lineshape_dress", C.c_int32),
        ("state_looping", C.c_int32), ("steps_per_loop", C.c_int32), ("write_pops", C.c_int32),
        ("print_all_pops", C.c_int32), ("write_binary_pops", C.c_int32), ("write_states", C.c_int32),
        ("print_partic", C.c_int32), ("debug_diags", C.c_int32), ("print_pops", C.c_int32 * 10),
        ("init_state", C.c_int32), ("init_from_file", C.c_int32),
        ("ham_file", C.c_char * 512), ("dipole_file", C.c_char * 512), ("init_state_file", C.c_char * 512),
        ("ka_params_file", C.c_char * 512), ("hsr_file", C.c_char * 512), ("zz_specd_file", C.c_char * 512),
        ("out_folder", C.c_char * 512), ("pos_file", C.c_char * 512), ("tk_mat_file", C.c_char * 512),
        ("tk_p", C.c_int32), ("tk_m_aux", C.c_int32),
    ]


def parse_infile(path):
    r = Infile()
    _check(lib().dq_parse_infile(os.fsencode(path), C.byref(r)))
    out = {}
    for name, _ in Infile._fields_:
        v = getattr(r, name)
        if isinstance(v, bytes):
            v = v.decode()
        elif name == "print_pops":
            v = list(v)
        out[name] = v
    return out


def read_hamiltonian(path):
    """(n, rowptr, col, val, e_site) of a Hamiltonian file"""
    n = C.c_int32()
    nnz = C.c_int64()
    _check(lib().dq_read_hamiltonian_file(os.fsencode(path), C.byref(n), C.byref(nnz), None, None, None, None))
    rp = np.zeros(n.value + 1, np.int32)
    col = np.zeros(max(nnz.value, 1), np.int32)
    val = np.zeros(max(nnz.value, 1))
    es = np.zeros(n.value)
    _check(lib().dq_read_hamiltonian_file(os.fsencode(path), C.byref(n), C.byref(nnz), _p(rp), _p(col), _p(val),
                                          _p(es)))
    return n.value, rp, col[:nnz.value], val[:nnz.value], es


def read_state(path, n):
    a = np.zeros(2 * n)
    _check(lib().dq_read_state_file(os.fsencode(path), int(n), _p(a)))
    return a.view(np.complex128)


def read_dipoles(path, n):
    a = np.zeros(3 * n)
    _check(lib().dq_read_dipole_file(os.fsencode(path), int(n), _p(a)))
    return a.reshape(n, 3)


def read_positions(path, n):
    a = np.zeros(3 * n)
    _check(lib().dq_read_position_file(os.fsencode(path), int(n), _p(a)))
    return a.reshape(n, 3)


CD_PREFACTOR = np.pi / 2 * 1e-7


def cd_vectors(pos, mu):
    """(rd[n][3] = pos x mu per site, mu_tot[3], rd_tot[3]) on the host"""
    pos = _c(pos, np.float64)
    mu = _c(mu, np.float64)
    n = pos.shape[0]
    if pos.shape != (n, 3) or mu.shape != (n, 3):
        raise ValueError("pos and mu must both be [n][3]")
    rd = np.zeros((n, 3))
    mt = np.zeros(3)
    rt = np.zeros(3)
    _check(lib().dq_cd_vectors(n, _p(pos), _p(mu), _p(rd), _p(mt), _p(rt)))
    return rd, mt, rt


def spectrum_runs(pos, mu, kind="absorption"):
    """the three direction runs of an isotropic spectrum: a list of
    (psi0, bra, scale) for d = x, y, z, where the run propagates psi0, records
    C(n) = <bra|psi(n)> and dq_spectrum takes `scale` (spectraClass.cu:437-496,
    641-648); the isotropic spectrum is the mean of the three."""
    mu = _c(mu, np.float64)
    if kind == "absorption":
        mt = np.sqrt((mu * mu).sum(0))
        return [((mu[:, d] / mt[d]).astype(np.complex128), (mu[:, d] / mt[d]).astype(np.complex128),
                 mt[d] ** 2) for d in range(3)]
    if kind != "cd":
        raise ValueError("kind is 'absorption' or 'cd'")
    rd, mt, rt = cd_vectors(pos, mu)
    return [((rd[:, d] / rt[d]).astype(np.complex128), (mu[:, d] / mt[d]).astype(np.complex128),
             CD_PREFACTOR * mt[d] * rt[d]) for d in range(3)]


def direction_batch(pos, mu, per_direction, kind="absorption"):
    """the three direction ensembles of an orientationally averaged spectrum
    (spectraClass.cu:437-496, 699-744) as one batch of 3 * per_direction
    trajectories, direction d taking trajectories d * per_direction ..:
    returns (psi0 [3 B_d, n], bras [3, n], bra_of_traj [3 B_d], weight [3 B_d])
    for set_state, set_bras and spectrum_weighted; weight = scale_d / (3 B_d),
    so spectrum_weighted is the mean of the three direction spectra."""
    runs = spectrum_runs(pos, mu, kind)
    psi0 = np.repeat(np.stack([r[0] for r in runs]), per_direction, axis=0)
    bras = np.stack([r[1] for r in runs])
    idx = np.repeat(np.arange(3, dtype=np.int32), per_direction)
    weight = np.repeat(np.array([r[2] for r in runs]) / (3.0 * per_direction), per_direction)
    return psi0, bras, idx, weight


def read_site_columns(path, n, ncol):
    a = np.zeros(n * ncol)
    _check(lib().dq_read_site_columns(os.fsencode(path), int(n), int(ncol), _p(a)))
    return a.reshape(n, ncol)


def read_spectral_density(path):
    """(domega, omega[nosc], J[nosc]) in cm^-1"""
    k = C.c_int32()
    dw = C.c_double()
    _check(lib().dq_read_spectral_density(os.fsencode(path), C.byref(k), C.byref(dw), None, None))
    om = np.zeros(k.value)
    jj = np.zeros(k.value)
    _check(lib().dq_read_spectral_density(os.fsencode(path), C.byref(k), C.byref(dw), _p(om), _p(jj)))
    return dw.value, om, jj


def write_pops(path, pops, dt, nsims=1, binary=False):
    """write (or continue) a population file; pops[t-1][site] for steps 1..T"""
    p = _c(pops, np.float64)
    T, P = p.shape
    _check(lib().dq_write_pops_file(os.fsencode(path), int(binary), C.c_double(dt), T, P, _p(p), int(nsims)))


def read_pops(path, binary=False):
    """(nsims, dt, pops[T][P])"""
    ns, T, P = C.c_int32(), C.c_int32(), C.c_int32()
    dt = C.c_double()
    _check(lib().dq_read_pops_file(os.fsencode(path), int(binary), C.byref(ns), C.byref(dt), C.byref(T), C.byref(P),
                                   None))
    out = np.zeros(T.value * P.value)
    _check(lib().dq_read_pops_file(os.fsencode(path), int(binary), C.byref(ns), C.byref(dt), C.byref(T), C.byref(P),
                                   _p(out)))
    return ns.value, dt.value, out.reshape(T.value, P.value)


def pade_tbar(m, alpha, beta, theta):
    """exp(-i theta Tbar) on the GPU for each row of alpha/beta ([count, m])."""
    alpha = _c(np.atleast_2d(alpha), np.float64)
    beta = _c(np.atleast_2d(beta), np.float64)
    count = alpha.shape[0]
    n = m + 2
    E = np.zeros(count * n * n * 2)
    s = np.zeros(count, np.int32)
    _check(lib().dq_pade_tbar(m, count, _p(alpha), _p(beta), C.c_double(theta), _p(E), _p(s)))
    E = E.view(np.complex128).reshape(count, n, n).transpose(0, 2, 1)  # column-major -> [row, col]
    return E, s


def fft(x, sign=+1):
    """unnormalised FFT on the GPU of each row of x (length a power of two)"""
    x = np.atleast_2d(np.asarray(x, np.complex128))
    a = c2i(x.reshape(-1))
    _check(lib().dq_fft(x.shape[1], x.shape[0], _p(a), sign))
    return a.view(np.complex128).reshape(x.shape)


def hbar_ref():
    """hbar in cm^-1 fs as the reference computes it (cppUtility.cpp:10-35)"""
    pi = np.arctan(1.0) * 4
    return (1 / 100.) * (1e15) * 1. / (299792548. * 2 * pi)

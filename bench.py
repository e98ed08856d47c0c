"""Benchmark: batched mixed-precision stochastic exciton propagation.

Workload (one MI355X): double-walled dipole tube, inner wall 100 rings x 60
sites at R = 3 nm, outer wall 100 x 60 at R = 6 nm (N = 12,000, 734,160
off-diagonal couplings at the 2.5 nm cutoff), B = 256 trajectories with
their own static-disorder (sigma 250 cm^-1) and OU-noise (sigma 400 cm^-1,
tau 50 fs) Philox streams, Krylov m = 24, one Krylov cycle per dt = 2 fs,
FP32 complex Krylov vectors and SpMM with FP64 alpha/beta, Gram sums, Pade
and psi.  A step = one dt for all 256 trajectories (noise, Lanczos with full
reorthogonalisation, Pade, combine, renormalise, wall populations every 10
steps), one hipGraph launch.  value = B * steps / seconds.

Also reported: the plain FP64 single-trajectory GPU path on the same tube,
the mixed-vs-FP64 wall-population error over the whole run, the roofline of
the dominant kernel (HIP events handed to each of its launches, which hold
the kernel's own start and end) and of the SpMM (HIP events recorded around
each launch), and the FP64 CPU oracle timed on the host cores.

python bench.py [--gpus 1] [--steps 2000] [--warmup 20]
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
N_RINGS, M, DT, B = 100, 24, 2.0, 256
SIG_STATIC, SIG_OU, TAU = 250.0, 400.0, 50.0
EVERY = 10


def build_tube(dq):
    pos, dip, wall, rp, col, val = dq.tube([N_RINGS, N_RINGS], [3.0, 6.0])
    psi = np.where(wall == 1, dip[:, 0], 0.0).astype(np.complex128)
    psi /= np.linalg.norm(psi)
    # the same superposition on the inner wall, whose 0.31 nm site spacing
    # makes it the stiff part of the tube (one Krylov cycle per dt is not
    # converged there)
    psi_in = np.where(wall == 0, dip[:, 0], 0.0).astype(np.complex128)
    psi_in /= np.linalg.norm(psi_in)
    # population groups: 2 walls x 4 axial segments (the per-wall populations
    # are their sums; the walls are uncoupled at 2.5 nm, so per-wall totals
    # alone are constant and would hide any precision error)
    seg = np.minimum((pos[:, 2] / (0.8 * N_RINGS) * 4).astype(np.int32), 3)
    groups = (wall * 4 + seg).astype(np.int32)
    return dict(n=len(wall), wall=wall, groups=groups, rp=rp, col=col, val=val, psi0=psi, psi_inner=psi_in)


def make_system(dq, t, n_traj, precision, max_steps, traj_ids=None, psi=None):
    s = dq.System(t["n"], M, DT, n_traj=n_traj, precision=precision, sigma_static=SIG_STATIC, seed=1)
    s.set_hamiltonian(t["rp"], t["col"], t["val"], np.zeros(t["n"]))
    s.set_ou(TAU, SIG_OU)
    if traj_ids is not None:
        s.set_trajectories(traj_ids)
    s.reset_disorder()
    s.set_state(t["psi0"] if psi is None else psi)
    s.set_observables(t["groups"], 8, EVERY, max_steps)
    return s


def spmm_bytes(n, nnz, b, mixed):
    """algorithmic bytes of one batched SpMM: matrix once, x read, w write, diagonal"""
    ba, bv, bd = (4, 8, 4) if mixed else (8, 16, 8)
    return nnz * (ba + 4) + 4 * (n + 1) + b * n * (2 * bv + bd)


def update_bytes(n, b, m, mixed):
    """algorithmic bytes of the m update launches of one cycle: iteration j
    reads v_0..v_j and w, writes v_{j+1}"""
    bv = 8 if mixed else 16
    return sum((j + 3) * n * b * bv for j in range(m))


TRAFFIC_FILE = os.path.join(ROOT, "profiles", "traffic_N12000_B256_m24.txt")


def spmm_kernel_name(tuning):
    """the SpMM kernel a tuning launches (mixed precision)"""
    if tuning.get("spmm_window_rows", 0) > 0:
        return "dq::k_spmm_ring<%s>" % ("true" if tuning.get("spmm_exact_sums", 0) else "false")
    if tuning["spmm_rows_per_wave"] == 16:
        return "dq::k_spmm_mfma<%d>" % (2 * tuning["spmm_traj_per_lane"])
    return "dq::k_spmm_b<float,%d,%d,%s>" % (tuning["spmm_traj_per_lane"], tuning["spmm_rows_per_wave"],
                                             "true" if tuning["spmm_pipeline"] else "false")


def update_kernel_name(tuning):
    return "dq::k_update_b<float,KW,NT> tile_rows=%d" % tuning["tile_rows"]


def measured_traffic(kernel):
    """HBM bytes per launch of `kernel` (the names above) from the committed
    rocprofv3 FETCH_SIZE / WRITE_SIZE summary (already scaled by the copy-kernel
    factor recorded there); None when that summary has no line for it"""
    try:
        with open(TRAFFIC_FILE) as f:
            for line in f:
                key, _, val = line.rpartition("=")
                if key.strip() == "traffic_bytes_per_launch " + kernel:
                    return float(val)
    except OSError:
        pass
    return None


def traffic_probe(dq, t, tuning, steps=2):
    """workload for the rocprofv3 --pmc runs: a few batched steps without the
    graph, then one 256 MiB device copy whose counters give the scale factor"""
    import ctypes
    s = make_system(dq, t, B, dq.MIXED, steps + 1)
    if tuning:
        s.set_tuning(**tuning)
    print("# traffic probe launch choices: %s" % s.tuning(), file=sys.stderr, flush=True)
    s.run(steps, graph=False)
    s.close()
    hip = ctypes.CDLL("libamdhip64.so")
    nbytes = 256 << 20
    a, b = ctypes.c_void_p(), ctypes.c_void_p()
    ok = hip.hipMalloc(ctypes.byref(a), ctypes.c_size_t(nbytes)) == 0
    ok = ok and hip.hipMalloc(ctypes.byref(b), ctypes.c_size_t(nbytes)) == 0
    ok = ok and hip.hipMemset(a, 1, ctypes.c_size_t(nbytes)) == 0
    ok = ok and hip.hipDeviceSynchronize() == 0
    ok = ok and hip.hipMemcpy(b, a, ctypes.c_size_t(nbytes), 3) == 0   # device to device
    ok = ok and hip.hipDeviceSynchronize() == 0
    hip.hipFree(a)
    hip.hipFree(b)
    if not ok:
        raise RuntimeError("traffic probe: HIP copy failed")
    print("# traffic probe: %d batched steps, then one device copy of %d bytes" % (steps, nbytes),
          file=sys.stderr, flush=True)


def hbm_ceiling(reps=10):
    """what plain streaming reaches on this GPU beyond the 256 MB last-level
    cache, for context beside the 8 TB/s peak: a read-only reduction over a
    4 GiB buffer and a 2 GiB -> 2 GiB copy (PyTorch's own kernels), best of
    `reps` event-timed runs each"""
    import torch
    nbytes = 4 << 30
    x = torch.ones(nbytes // 4, dtype=torch.float32, device="cuda")
    y = torch.empty(nbytes // 8, dtype=torch.float32, device="cuda")
    out = {}
    for name, fn, moved in (("read_only_sum_4GiB", lambda: x.sum(), nbytes),
                            ("copy_2GiB", lambda: y.copy_(x[: nbytes // 8]), nbytes)):
        fn()
        torch.cuda.synchronize()
        best = float("inf")
        for _ in range(reps):
            a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            a.record()
            fn()
            b.record()
            b.synchronize()
            best = min(best, a.elapsed_time(b))
        out[name] = {"ms": best, "GB_per_s": moved / (best * 1e-3) / 1e9}
    print(json.dumps({"hbm_ceiling": out, "peak_GB_per_s": 8000.0}), flush=True)


def error_check(dq, t, tuning, steps, every=50, ntraj=64):
    """max |P_mixed - P_fp64| over every record of a `steps`-step run of the
    first `ntraj` trajectories: per site, and per group (2 walls x 4 axial
    segments), with the headline's launch choices in the mixed run"""
    n = t["n"]
    sites = np.arange(n, dtype=np.int32)
    out = []
    for prec in (dq.MIXED, dq.FP64):
        s = dq.System(n, M, DT, n_traj=ntraj, precision=prec, sigma_static=SIG_STATIC, seed=1)
        s.set_hamiltonian(t["rp"], t["col"], t["val"], np.zeros(n))
        s.set_ou(TAU, SIG_OU)
        s.set_trajectories(np.arange(ntraj, dtype=np.uint64))
        s.reset_disorder()
        if prec == dq.MIXED:
            c = dict(tuning)
            c["concurrent_slices"] = 1
            try:
                s.set_tuning(**c)
            except dq.DQError:
                pass                      # a choice the smaller batch cannot take: defaults
        s.set_state(t["psi0"])
        s.set_observables(sites, n, every, steps)
        s.run(steps, graph=True)
        out.append(s.populations())
        s.close()
    site_err = float(np.abs(out[0] - out[1]).max())
    g = t["groups"]
    grp = [np.stack([p[..., g == k].sum(-1) for k in range(8)], -1) for p in out]
    pop_err = float(np.abs(grp[0] - grp[1]).max())
    return pop_err, site_err


def cpu_baseline(t, seconds_target=15.0):
    """FP64 CPU oracle on the host cores: one trajectory per OpenMP thread"""
    from oracle import pyoracle as O
    cores = min(os.cpu_count() or 1, 16)
    p = O.make_params(t["n"], M, DT, reorth=O.REORTH_GRAM, sigma_static=SIG_STATIC, seed=1)
    sites = O.Sites(t["n"], tau=TAU, sigma=SIG_OU, e_site=0.0)
    steps = 2
    t0 = time.perf_counter()
    O.run_batch(p, sites, t["rp"], t["col"], t["val"], t["psi0"], np.arange(cores), steps, EVERY,
                t["groups"], 8, nthreads=cores)
    dt = time.perf_counter() - t0
    steps2 = max(2, int(steps * seconds_target / max(dt, 1e-3)))
    steps2 = min(steps2, 2000)
    t0 = time.perf_counter()
    O.run_batch(p, sites, t["rp"], t["col"], t["val"], t["psi0"], np.arange(cores), steps2, EVERY,
                t["groups"], 8, nthreads=cores)
    dt = time.perf_counter() - t0
    # single core: the N = 12,000 workload and the 600-site single-wall tube (m = 16, dt = 1 fs)
    t1 = time.perf_counter()
    O.run_batch(p, sites, t["rp"], t["col"], t["val"], t["psi0"], np.arange(1), 10, EVERY, t["groups"], 8,
                nthreads=1)
    single12k = (time.perf_counter() - t1) / 10 * 1e3
    import dynaquant_amd as dq
    pos, dip, wall, rp6, col6, val6 = dq.tube([10], [3.0])
    p6 = O.make_params(600, 16, 1.0, reorth=O.REORTH_GRAM, sigma_static=200.0, seed=1)
    s6 = O.Sites(600, tau=TAU, sigma=300.0, e_site=0.0)
    psi6 = np.zeros(600, complex)
    psi6[0] = 1.0
    t1 = time.perf_counter()
    O.run_batch(p6, s6, rp6, col6, val6, psi6, np.arange(1), 200, 10, np.arange(600, dtype=np.int32), 600,
                nthreads=1)
    single600 = (time.perf_counter() - t1) / 200 * 1e3
    model = "unknown"
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return {"value": cores * steps2 / dt, "unit": "traj-steps/s", "cores": cores, "kind": "port",
            "sample": "%d trajectories x %d steps of the same N=12000 m=24 dt=2fs workload, FP64 "
                      "oracle, one trajectory per OpenMP thread, %.1f s" % (cores, steps2, dt),
            "cpu_model": model, "host_cpus": os.cpu_count(),
            "single_core_ms_per_traj_step": {"N=12000 m=24 dt=2fs": single12k,
                                             "N=600 m=16 dt=1fs": single600}}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2000)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--plain-steps", type=int, default=300)
    ap.add_argument("--profile-steps", type=int, default=10)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-check", action="store_true")
    ap.add_argument("--err-steps", type=int, default=2000,
                    help="length of the mixed-vs-FP64 error run (independent of --steps)")
    ap.add_argument("--ktol", type=float, default=1e-10,
                    help="error allowed per dt for the substep-controlled run (<= 0: skip it)")
    ap.add_argument("--ktol-steps", type=int, default=200)
    ap.add_argument("--no-autotune", action="store_true",
                    help="keep the default launch choices (for kernel-trace profiles, whose per-kernel "
                         "averages would otherwise include the autotuner's trial launches)")
    ap.add_argument("--traffic-probe", action="store_true",
                    help="run only the short workload profiled for HBM traffic counters")
    ap.add_argument("--tuning", type=str, default=None,
                    help="launch choices as JSON (dq_tuning field names) instead of the autotuner")
    ap.add_argument("--hbm-ceiling", action="store_true",
                    help="only time plain streaming (read-only sum, copy) beyond the last-level cache")
    args = ap.parse_args()
    if args.hbm_ceiling:
        hbm_ceiling()
        return
    if args.gpus != 1:
        print("bench.py: this engine runs on one GPU; trajectories shard across GPUs with no "
              "exchange, but only the single-GPU figure is measured here", file=sys.stderr)
        sys.exit(2)
    import dynaquant_amd as dq

    t = build_tube(dq)
    n, nnz = t["n"], len(t["col"])
    K, W = args.steps, args.warmup
    print("# workload N=%d nnz_off=%d B=%d m=%d dt=%g" % (n, nnz, B, M, DT), file=sys.stderr, flush=True)
    fixed = json.loads(args.tuning) if args.tuning else None
    if args.traffic_probe:
        traffic_probe(dq, t, fixed)
        return

    # ---- headline: batched mixed precision, hipGraph per step
    s = make_system(dq, t, B, dq.MIXED, W + K + args.profile_steps)
    if fixed:
        s.set_tuning(**fixed)
        tuning, how = s.tuning(), "given"
    elif args.no_autotune:
        tuning, how = s.tuning(), "default"
    else:
        tuning, how = s.autotune(reps=2), "autotuned"
    print("# launch choices (%s): %s" % (how, tuning), file=sys.stderr, flush=True)
    s.run(W, graph=True)
    s.run(K, graph=True)
    ms = s.last_run_ms()
    value = B * K / (ms / 1e3)
    print("# mixed: %.3f ms/step, %.0f traj-steps/s" % (ms / K, value), file=sys.stderr, flush=True)

    # ---- the same batch with K from the Expokit estimate (chunks of 50 steps):
    # from the headline state (outer wall) and from the same superposition on
    # the stiff inner wall
    ctl = None
    if args.ktol > 0:
        ctl = {}
        for name, psi in (("outer_wall_state", t["psi0"]), ("inner_wall_state", t["psi_inner"])):
            c = make_system(dq, t, B, dq.MIXED, args.ktol_steps, psi=psi)
            c.set_tuning(**tuning)
            c.set_krylov_control(args.ktol, k_max=16, chunk=50)
            c.run(args.ktol_steps, graph=True)
            kc, redone = c.krylov_control()
            cms = c.last_run_ms()
            c.close()
            ctl[name] = {"converged_K": kc, "chunks_redone": redone,
                         "traj_steps_per_s": B * args.ktol_steps / (cms / 1e3)}
            print("# substep control at %.1e per dt, %s: K=%d (%d chunks redone), %.0f traj-steps/s"
                  % (args.ktol, name, kc, redone, ctl[name]["traj_steps_per_s"]), file=sys.stderr, flush=True)
        ctl.update({"tol_per_dt": args.ktol, "steps": args.ktol_steps,
                    "note": "mixed batch with substep control: chunks of 50 steps, K doubled while K x the "
                            "largest Krylov-cycle estimate exceeds tol; time includes redone chunks"})

    # ---- per-kernel times from HIP events on the system's stream
    s.set_profiling(True)
    s.run(args.profile_steps)
    kt = s.kernel_times()
    s.close()
    cyc = args.profile_steps
    spmm_ms = kt["spmm"][0] / kt["spmm"][1]
    upd_ms = kt["update"][0] / kt["update"][1]
    b_spmm = spmm_bytes(n, nnz, B, True)
    b_upd = update_bytes(n, B, M, True) / M
    spmm_gbs = b_spmm / (spmm_ms * 1e-3) / 1e9
    upd_gbs = b_upd / (upd_ms * 1e-3) / 1e9
    totals = {k: v[0] for k, v in kt.items()}
    dominant = max(totals, key=totals.get)
    print("# kernel ms per step: " + ", ".join("%s %.3f" % (k, v / cyc) for k, v in totals.items()),
          file=sys.stderr, flush=True)
    upd_name, spmm_name = update_kernel_name(tuning), spmm_kernel_name(tuning)
    upd_roof = {"bound": "hbm", "achieved": upd_gbs, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": upd_gbs / HBM_PEAK_GBS, "traffic": measured_traffic(upd_name),
                "kernel": upd_name + " (KW = ceil((j+1)/4) = 1..6 for j = 0..23; the NT = max(0, KW - 2)"
                          " oldest groups of 4 vectors read with the nontemporal hint)",
                "bytes_per_launch": b_upd, "avg_launch_ms": upd_ms,
                "timing": "HIP events passed to every update launch of the profiled steps "
                          "(hipExtLaunchKernelGGL: the kernel's own start and end) on the system's stream"}
    spmm_roof = {"bound": "hbm", "achieved": spmm_gbs, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                 "frac": spmm_gbs / HBM_PEAK_GBS, "traffic": measured_traffic(spmm_name),
                 "kernel": spmm_name, "bytes_per_launch": b_spmm, "avg_launch_ms": spmm_ms,
                 "timing": "HIP events recorded around every SpMM launch of the profiled steps"}
    roof = spmm_roof if dominant == "spmm" else upd_roof

    # ---- plain FP64 single-trajectory path, same tube
    p = make_system(dq, t, 1, dq.FP64, 20 + args.plain_steps)
    p.run(20, graph=True)
    p.run(args.plain_steps, graph=True)
    plain_ms = p.last_run_ms()
    p.close()
    plain_value = args.plain_steps / (plain_ms / 1e3)
    print("# plain FP64: %.3f ms/step" % (plain_ms / args.plain_steps), file=sys.stderr, flush=True)

    # ---- mixed vs FP64 on a fixed-length run (independent of --steps): the
    # first 64 trajectories of the batch, site populations every 50 steps
    pop_err = site_err = None
    if not args.no_check:
        pop_err, site_err = error_check(dq, t, tuning, args.err_steps)
        print("# mixed vs FP64 over %d steps (64 trajectories, every 50 steps): max group-population "
              "error %.3e, max site-population error %.3e" % (args.err_steps, pop_err, site_err),
              file=sys.stderr, flush=True)

    cpu = None if args.no_cpu else cpu_baseline(t)
    line = {
        "metric": "traj-steps/s (N=12k, B=256), %HBM peak; mixed vs FP64 at 1e-6 pop error",
        "value": value, "unit": "traj-steps/s", "n_gpus": 1, "steps": K, "warmup": W,
        "ms_per_step": ms / K, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "fp32-vectors/fp64-reductions", "data": "synthetic",
        "config": {"workload": "double-walled dipole tube N=12000 (2x100 rings x 60), B=256 "
                               "trajectories, m=24, dt=2 fs, 1 Krylov cycle per dt, mixed "
                               "precision, hipGraph per step",
                   "n_sites": n, "nnz_offdiag": nnz, "batch": B, "krylov_m": M, "dt_fs": DT,
                   "krylov_substeps": 1, "precision": "mixed", "tuning": tuning},
        "roofline": roof,
        "spmm_roofline": spmm_roof,
        "plain_fp64_traj_steps_per_s": plain_value,
        "speedup_vs_plain_fp64": value / plain_value,
        "pop_err_vs_fp64": pop_err,
        "site_pop_err_vs_fp64": site_err,
        "pop_err_note": "%d-step runs of the first 64 trajectories, every 50 steps: max over records of "
                        "the 8 group populations (2 walls x 4 axial segments) and of the 12,000 site "
                        "populations, against the FP64 batched GPU path (itself <=1e-10 from the FP64 "
                        "oracle)" % args.err_steps,
        "substep_control": ctl,
        "kernel_ms_per_step": {k: v / cyc for k, v in totals.items()},
        "cpu_baseline": cpu,
    }
    print(json.dumps(line))


if __name__ == "__main__":
    main()

#!/usr/bin/env python3
"""Benchmark: fTWA RK4 trajectory-steps/s on the 8x8 periodic Hubbard lattice, 65,536 trajectories.

One "step" is one RK4 "3/8" step (4 stages, both spins) of the whole 65,536-trajectory batch.
The timed region is one fpsftwa_run of K steps of the FP32 batched engine (sampling, K steps with an
observable record at every time point, and the FP64 ensemble reduction); it prints one JSON line.
The default K = 4000 is the whole run at dt = 0.005, t_max = 20 (4001 time points).

Beside it, on the same workload:
  * the plain FP64 GPU path (dense Delta, one launch per time point), timed over a few steps: the
    speedup denominator of the headline metric;
  * the FP64 batched engine over the same K steps: the FP64 reference for the FP32 error check, and
    itself timed.  The error report gives, per observable, the largest difference and the time up
    to which every observable stays within 1e-5 (the trajectories are chaotic, so rounding
    differences grow exponentially in t);
  * the roofline of the dominant kernel (the upper-triangle integrator), from HIP events on its stream.
    The state stays on chip within a launch, so the bound is the vector ALU: (64 z + 104) L^2 flop per
    trajectory-step against the FP32 vector peak, with the streamed-state HBM figure beside it;
  * untimed, the accuracy over the whole configured run (t_max = 20): FP32 against FP64 batched, and
    FP64 batched against the plain FP64 engine on 1,024 trajectories, with the path that meets 1e-5;
  * the CPU baseline: the FP64 oracle (a C/OpenMP restatement of the reference algorithm) on a bounded
    sample of the same workload on the host cores.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "fTWA RK4 traj-steps/s, 8x8 Hubbard 64k traj; speedup vs plain FP64 GPU at 1e-5 err"
HBM_PEAK_GBS = 8000.0


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def workload(seed=202):
    from fermiphasespace_amd import lattices as lat
    pos, box = lat.grid(8, 8)
    U = 2.0 + 6.0 * lat.uniform_draw(seed, 0)
    tM, Um = lat.system_arrays(pos, box, True, t=-1.0, U=U)
    occ = np.array(lat.checkerboard(8, 8)["IC_NuDiag"], dtype=float)
    return {"tM": tM, "U": Um, "occ": occ, "Uval": U, "L": 64}


def traffic_from_profile(path):
    """HBM bytes per launch of the stencil kernel on this workload, measured with rocprofv3 counters."""
    try:
        with open(path) as f:
            d = json.load(f)
        return d.get("hbm_bytes_per_launch"), d
    except (OSError, ValueError):
        return None, None


def error_report(r32, r64, dt, tol=1e-5):
    """Per-observable FP32-vs-FP64 ensemble differences, and the time up to which all stay within tol.

    Both runs start from the same samples.  The per-trajectory dynamics is chaotic, so any rounding
    difference grows roughly exponentially in t; the horizon is where that growth reaches tol.
    """
    diffs = {k: np.abs(r32[k] - r64[k]) for k in ("Niiu", "Niid", "NuNd")}
    per_t = np.max(np.stack([d.max(axis=0) for d in diffs.values()]), axis=0)  # [T]
    above = np.nonzero(per_t > tol)[0]
    t_bad = int(above[0]) if above.size else per_t.shape[0]
    e_rel = np.abs(r32["E"] - r64["E"]).max(axis=0) / np.maximum(np.abs(r64["E"]).max(axis=0), 1e-30)
    rep = {"tolerance": tol,
           "max_abs": {k: float(d.max()) for k, d in diffs.items()},
           "E_max_rel": float(e_rel.max()),
           "stat_err_max": {"Niiu": float(np.max(r32["Niiu_err"])), "NuNd": float(np.max(r32["NuNd_err"]))},
           "within_tolerance": bool(above.size == 0),
           "t_within_tolerance": round(t_bad * dt if above.size else (per_t.shape[0] - 1) * dt, 4),
           "max_abs_by_t": [[round(i * dt, 3), float(per_t[i])]
                            for i in np.linspace(0, per_t.shape[0] - 1, 11).astype(int)]}
    if t_bad > 0:
        rep["max_abs_within_horizon"] = {k: float(d[:, :t_bad].max()) for k, d in diffs.items()}
        rep["E_max_rel_within_horizon"] = float(e_rel[:t_bad].max())
    return rep


def cpu_baseline(w, dt, target_s=15.0):
    from oracle import oracle as O
    nthreads = os.cpu_count() or 1
    sysO = O.System(tM=w["tM"], coulomb_type=1, U=w["U"])
    # calibrate on a tiny sample, then size the sample to about target_s seconds
    B0, K0 = nthreads, 2
    t0 = time.perf_counter()
    O.run(sysO, dt, K0 * dt, 1, w["occ"], w["occ"], n_traj=B0, seed=202, nthreads=nthreads)
    c0 = time.perf_counter() - t0
    rate0 = B0 * K0 / max(c0, 1e-9)
    K = max(4, int(round(target_s * rate0 / (4 * nthreads))))
    B = 4 * nthreads
    t0 = time.perf_counter()
    O.run(sysO, dt, K * dt, 1, w["occ"], w["occ"], n_traj=B, seed=202, nthreads=nthreads)
    el = time.perf_counter() - t0
    return {"value": B * K / el, "unit": "traj-steps/s", "cores": nthreads, "kind": "port",
            "sample": f"FP64 oracle (dense Delta, reference algorithm) on the same 8x8 workload: {B} trajectories x "
                      f"{K} RK4 steps, {el:.1f} s on {nthreads} OpenMP threads"}


def full_run_accuracy(eng, w, dt, seed, B, args, res32=None, res64=None, K_full=4000, tol=1e-5):
    """Untimed: the FP32 and FP64 batched engines over the whole configured run (t_max = 20) from the same
    samples, and the FP64 batched engine against the plain FP64 engine (the reference's operation order)
    on the first `horizon_plain_traj` trajectories over the same run.  Reports the time up to which each
    path stays within tol of its FP64 reference, and which path meets tol over the whole run."""
    from fermiphasespace_amd import _lib
    keys = ("Niiu", "Niid", "NuNd", "E", "Niiu_err", "NuNd_err")
    t_start = time.perf_counter()
    if res32 is None:
        eng.set_tuning(args.block, args.time_chunk, args.kernel)
        eng.set_params(dt, K_full * dt, 1, B, seed=seed, engine="fp32")
        eng.run()
        res32 = {k: eng.result(k) for k in keys}
        eng.set_tuning(0, 0, "auto")
    if res64 is None:
        eng.set_params(dt, K_full * dt, 1, B, seed=seed, engine="fp64")
        eng.run()
        res64 = {k: eng.result(k) for k in keys}
    rep32 = error_report(res32, res64, dt, tol)
    log(f"full run fp32 vs fp64 batched: within {tol} up to t = {rep32['t_within_tolerance']}")
    # FP64 batched against plain FP64 on a trajectory subset (the plain engine runs one launch per time
    # point and about 2e5 trajectory-steps/s, so the whole batch would take 20 minutes)
    Bp = min(B, args.horizon_plain_traj)
    sub = _lib.Engine(w["tM"], coulomb_type=_lib.COULOMB_ONSITE, U=w["U"])
    sub.set_initial_state(w["occ"], w["occ"])
    sub.set_results(_lib.RES_DIAG | _lib.RES_DOUBLONS | _lib.RES_ENERGIES)
    got = {}
    for engine in ("fp64", "fp64_plain"):
        sub.set_params(dt, K_full * dt, 1, Bp, seed=seed, engine=engine)
        sub.run()
        got[engine] = {k: sub.result(k) for k in keys}
    sub.close()
    rep64 = error_report(got["fp64"], got["fp64_plain"], dt, tol)
    log(f"full run fp64 batched vs fp64 plain ({Bp} traj): within {tol} up to t = {rep64['t_within_tolerance']}")
    meets = [name for name, rep in (("fp32_batched", rep32), ("fp64_batched", rep64)) if rep["within_tolerance"]]
    return {"t_max": K_full * dt, "steps": K_full, "tolerance": tol,
            "fp32_vs_fp64_batched": {k: rep32[k] for k in ("within_tolerance", "t_within_tolerance", "max_abs",
                                                           "E_max_rel", "stat_err_max", "max_abs_by_t")},
            "fp64_batched_vs_fp64_plain": dict({k: rep64[k] for k in ("within_tolerance", "t_within_tolerance",
                                                                      "max_abs", "E_max_rel", "max_abs_by_t")},
                                               n_traj=Bp),
            "meets_tolerance_over_full_run": meets,
            "seconds": round(time.perf_counter() - t_start, 1)}


# ---- the other workloads (parity cases of the build; one JSON line each, not the headline) ----------


def measured_peaks(args):
    """fpsftwa_measure_peaks on this device (about a second), or None with --no-peaks."""
    if getattr(args, "no_peaks", False):
        return None
    from fermiphasespace_amd import _lib
    try:
        pk = _lib.measure_peaks()
    except _lib.FtwaError as e:  # e.g. no 2 GiB free for the copy: the bench line goes out without them
        log("measured peaks unavailable:", e)
        return None
    log("measured peaks", {k: round(v, 2) for k, v in pk.items()})
    return pk


def _timed_run(eng, sync):
    sync()
    t0 = time.perf_counter()
    eng.run()
    sync()
    return time.perf_counter() - t0


def side_chain10(args, sync):
    """10-site open chain, t = 1, U = 4, Neel, 1,000 trajectories, dt 0.01, 10 sub-steps, t_max 10 (the
    reference's CPU-sized case): FP64 batched and plain engines, and the CPU oracle on the same run."""
    from fermiphasespace_amd import _lib
    from oracle import oracle as O
    L, B, dt, nsub = 10, 1000, 0.01, 10
    t_max = 10.0 if args.steps is None else args.steps * dt
    tM = -(np.eye(L, k=1) + np.eye(L, k=-1))
    nu = np.array([1.0, 0] * 5)
    eng = _lib.Engine(tM, U=4.0 * np.eye(L))
    eng.set_initial_state(nu, 1 - nu)
    eng.set_results(_lib.RES_DIAG | _lib.RES_DOUBLONS | _lib.RES_ENERGIES)
    steps = int(round(t_max / dt)) * nsub
    out = {"metric": "fTWA RK4 traj-steps/s, 10-site chain 1000 traj", "unit": "traj-steps/s", "n_gpus": 1,
           "higher_is_better": True, "dtype": "fp64", "data": "synthetic",
           "config": {"workload": "chain10_open_neel", "L": L, "n_traj": B, "dt": dt, "n_sub": nsub, "t_max": t_max,
                      "U": 4.0}}
    for name in ("fp64", "fp64_plain", "fp32"):
        eng.set_params(dt, 2 * dt, nsub, B, seed=101, engine=name)
        eng.run()
        eng.set_params(dt, t_max, nsub, B, seed=101, engine=name)
        el = _timed_run(eng, sync)
        out[name] = {"value": round(B * steps / el, 1), "seconds": round(el, 3)}
        log(f"chain10 {name}: {B * steps / el:.4e} traj-steps/s")
    out["value"] = out["fp64"]["value"]
    eng.close()
    if not args.no_cpu:  # the whole configured run fits the CPU budget at L = 10
        nth = os.cpu_count() or 1
        t0 = time.perf_counter()
        O.run(O.System(tM=tM, coulomb_type=1, U=4.0 * np.eye(L)), dt, t_max, nsub, nu, 1 - nu, n_traj=B, seed=101,
              nthreads=nth)
        el = time.perf_counter() - t0
        out["cpu_baseline"] = {"value": round(B * steps / el, 1), "unit": "traj-steps/s", "cores": nth, "kind": "port",
                               "sample": f"FP64 oracle on the whole run: {B} trajectories x {steps} RK4 steps, "
                                         f"{el:.1f} s on {nth} OpenMP threads"}
    print(json.dumps(out), flush=True)


def side_gpsr4x4(args, sync):
    """GPSR (the reference's FHGPSRSolver) on the 4x4 periodic Hubbard lattice (L = 16), t = 1, U = 2,
    half-filled charge-density wave, 65,536 trajectories, dt 0.01, 50 steps (GPSR trajectories start to
    diverge soon after t = 0.5 at this U): the FP64 and FP32 batched GPSR kernels over the whole run, the plain
    FP64 engine over 2 steps, and the CPU oracle on a bounded sample."""
    from fermiphasespace_amd import _lib
    from fermiphasespace_amd import lattices as lat
    from oracle import oracle as O
    B, dt, U = args.n_traj, 0.01, 2.0
    steps = 50 if args.steps is None else args.steps
    pos, box = lat.grid(4, 4)
    tM, Um = lat.system_arrays(pos, box, True, t=-1.0, U=U)
    L = tM.shape[0]
    nu = np.array(lat.checkerboard(4, 4)["IC_NuDiag"], dtype=float)
    eng = _lib.Engine(tM, U=Um)
    eng.set_method("gpsr")
    eng.set_initial_state(nu, nu)
    eng.set_results(_lib.RES_DIAG | _lib.RES_DOUBLONS | _lib.RES_ENERGIES)
    out = {"metric": "GPSR RK4 traj-steps/s, 4x4 periodic Hubbard 65536 traj", "unit": "traj-steps/s", "n_gpus": 1,
           "higher_is_better": True, "dtype": "fp64", "data": "synthetic",
           "config": {"workload": "gpsr4x4_periodic_cdw", "L": L, "n_traj": B, "dt": dt, "steps": steps, "U": U}}
    for name, n in (("fp64", steps), ("fp32", steps), ("fp64_plain", 2)):
        eng.set_params(dt, 2 * dt, 1, B, seed=31, engine=name)
        eng.run()
        eng.set_params(dt, n * dt, 1, B, seed=31, engine=name)
        el = _timed_run(eng, sync)
        kms = eng.timing()["kernel_ms"]
        out[name] = {"value": round(B * n / el, 1), "seconds": round(el, 3), "kernel_ms": round(kms, 3), "steps": n}
        log(f"gpsr4x4 {name}: {B * n / el:.4e} traj-steps/s ({n} steps)")
    out["value"] = out["fp64"]["value"]
    out["ms_per_step"] = round(1e3 * out["fp64"]["seconds"] / steps, 4)
    out["vs_plain_fp64"] = round(out["fp64"]["value"] / out["fp64_plain"]["value"], 1)
    eng.close()
    if not args.no_cpu:  # bounded sample: trajectories of the same run until about cpu-seconds of work
        nth = os.cpu_count() or 1
        s = O.System(tM=tM, coulomb_type=1, U=Um, gpsr=True, gpsr_seed=31)
        t0 = time.perf_counter()
        O.run(s, dt, steps * dt, 1, nu, nu, n_traj=nth, seed=31, nthreads=nth)  # one trajectory per thread
        per = time.perf_counter() - t0
        nb = max(nth, int(args.cpu_seconds * nth / max(per, 1e-9)))
        t0 = time.perf_counter()
        O.run(s, dt, steps * dt, 1, nu, nu, n_traj=nb, seed=31, nthreads=nth)
        el = time.perf_counter() - t0
        out["cpu_baseline"] = {"value": round(nb * steps / el, 1), "unit": "traj-steps/s", "cores": nth, "kind": "port",
                               "sample": f"FP64 GPSR oracle: {nb} trajectories x {steps} steps of the same run, "
                                         f"{el:.1f} s on {nth} OpenMP threads"}
    print(json.dumps(out), flush=True)


def side_mixed(args, sync):
    """Honeycomb 24 / kagome 48 / square 100 / bcc 54 with site potentials, FP32, dt 0.01, t_max 10:
    per-system kernel by the autotuner, all four systems at once (fpsftwa_run_many)."""
    from fermiphasespace_amd import _lib
    from fermiphasespace_amd import lattices as lat
    dt = 0.01
    K = 1000 if args.steps is None else args.steps
    ws = lat.workload_mixed()
    engs, tune = [], {}
    for w in ws:
        e = _lib.Engine(w["tM"], U=w["U"], pot=w["pot"])
        e.set_initial_state(w["n_up"], w["n_dn"])
        e.set_results(_lib.RES_DIAG | _lib.RES_DOUBLONS | _lib.RES_ENERGIES)
        e.set_params(dt, K * dt, 1, w["n_traj"], seed=303, engine="fp32")
        tune[w["name"]] = e.autotune(trial_steps=8)
        engs.append(e)
    tot = sum(w["n_traj"] for w in ws) * K
    sync()
    t0 = time.perf_counter()
    _lib.run_many(engs)
    sync()
    el = time.perf_counter() - t0
    per = {}
    for w, e in zip(ws, engs):
        el1 = _timed_run(e, sync)
        per[w["name"]] = {"L": w["L"], "n_traj": w["n_traj"], "kernel": tune[w["name"]]["kernel"],
                          "alone_traj_steps_per_s": round(w["n_traj"] * K / el1, 1)}
        e.close()
    out = {"metric": "fTWA RK4 traj-steps/s, mixed-lattice batch (4 systems at once)", "value": round(tot / el, 1),
           "unit": "traj-steps/s", "n_gpus": 1, "higher_is_better": True, "dtype": "fp32", "data": "synthetic",
           "seconds": round(el, 3), "config": {"workload": "mixed_honeycomb24_kagome48_square100_bcc54", "dt": dt,
                                               "steps": K}, "systems": per}
    log("mixed", json.dumps(out))
    print(json.dumps(out), flush=True)


def side_moire196(args, sync):
    """Twisted bilayer L = 196, long-range hopping (cut-off 4a) and Coulomb V = U/(1 + r/a), Gaussian
    pulse, 4,096 trajectories, dt 0.005, t_max 10: the dense path (one GEMM per RK stage for the batch).
    The roofline is the GEMM's: 64 L^3 flop per trajectory-step (with the pulse, [Y_re Y_im] times the
    2L x 2L real block matrix per spin and stage) against the FP32 MFMA peak, over the integrator time."""
    from fermiphasespace_amd import _lib
    from fermiphasespace_amd import lattices as lat
    dt, B = 0.005, 4096
    K = 2000 if args.steps is None else args.steps
    w = lat.workload_moire196()
    L = w["L"]
    eng = _lib.Engine(w["tM"], coulomb_type=_lib.COULOMB_GENERAL, U=w["U"], V=w["V"], **w["pulse"])
    eng.set_initial_state(w["n_up"], w["n_dn"])
    eng.set_results(_lib.RES_DIAG | _lib.RES_DOUBLONS | _lib.RES_ENERGIES)
    eng.set_params(dt, 2 * dt, 1, B, seed=404, engine="fp32")
    eng.run()
    eng.set_params(dt, K * dt, 1, B, seed=404, engine="fp32")
    el = _timed_run(eng, sync)
    tm = eng.timing()
    flop = 64.0 * L ** 3 * B * K
    achieved = flop / (tm["kernel_ms"] * 1e-3) / 1e12
    out = {"metric": "fTWA RK4 traj-steps/s, moire L=196 4096 traj (dense path)", "value": round(B * K / el, 1),
           "unit": "traj-steps/s", "n_gpus": 1, "higher_is_better": True, "dtype": "fp32", "data": "synthetic",
           "seconds": round(el, 3),
           "config": {"workload": "moire196_longrange_pulse", "L": L, "n_traj": B, "dt": dt, "steps": K},
           "roofline": {"bound": "mfma", "achieved": round(achieved, 2), "peak": 157.3, "unit": "TFLOP/s",
                        "frac": round(achieved / 157.3, 4), "traffic": None,
                        "note": "GEMM flop over the whole integrator time (GEMM + update + Hartree + record kernels)",
                        "kernel_ms": round(tm["kernel_ms"], 3), "launches": tm["launches"]}}
    pk = measured_peaks(args)
    if pk:
        out["roofline"]["measured"] = {"fp32_mfma_tflops": round(pk["fp32_mfma_tflops"], 2),
                                       "frac_of_mfma": round(achieved / pk["fp32_mfma_tflops"], 4)}
    # the autotuner's pick (kernel family and, on this path, the library's GEMM solution), same run
    t0 = time.perf_counter()
    choice = eng.autotune(trial_steps=4)
    tune_s = time.perf_counter() - t0
    eng.set_params(dt, K * dt, 1, B, seed=404, engine="fp32")
    el2 = _timed_run(eng, sync)
    tm2 = eng.timing()
    out["autotuned"] = {"value": round(B * K / el2, 1), "seconds": round(el2, 3), "choice": choice,
                        "tune_seconds": round(tune_s, 2), "kernel_ms": round(tm2["kernel_ms"], 3),
                        "achieved_tflops": round(flop / (tm2["kernel_ms"] * 1e-3) / 1e12, 2)}
    eng.close()
    log("moire196", json.dumps(out))
    print(json.dumps(out), flush=True)


def side_adjoint6x6(args, sync):
    """6x6 lattice, 200 pulse knots, 16,384 trajectories, dt 0.01, t_max 10, checkpoints every 50 steps:
    gradient of the ensemble double occupancy at t_max (FP32 and FP64).  value = trajectory-steps per second
    of the FP32 gradient (one gradient = forward pass, recompute and reverse of every step).  Beside it,
    central differences of the FP64 engine's J at h = 4e-6, 2e-6, 1e-6, Richardson-extrapolated (the h^2
    and h^4 terms removed: at this horizon J is strongly curved, so a single h is far from its limit), for
    dJ/dU (uniform shift) and dJ/d(knot 50), against the FP64 adjoint."""
    from fermiphasespace_amd import _lib
    from fermiphasespace_amd import lattices as lat
    dt, B = 0.01, 16384
    t_max = 10.0 if args.steps is None else args.steps * dt
    nst = int(round(t_max / dt))
    w = lat.workload_adjoint6x6()
    L = w["L"]

    def engine(U0=4.0, knots=None, name="fp64"):
        e = _lib.Engine(w["tM"], U=U0 * np.eye(L), pulse_type=_lib.PULSE_KNOTS,
                        knots=w["knots"] if knots is None else knots, pol=w["pol"], disp=w["disp"])
        e.set_initial_state(w["n_up"], w["n_dn"])
        e.set_params(dt, t_max, 1, B, seed=505, engine=name)
        return e

    def forward(**kw):
        e = engine(**kw)
        e.set_results(_lib.RES_DIAG | _lib.RES_DOUBLONS)
        e.run()
        J = float(e.result("NuNd")[:, -1].mean())
        e.close()
        return J

    out = {"metric": "adjoint fTWA gradient traj-steps/s, 6x6 lattice 200 knots 16384 traj", "unit": "traj-steps/s",
           "n_gpus": 1, "higher_is_better": True, "data": "synthetic", "dtype": "fp32",
           "config": {"workload": "adjoint6x6_knots200", "L": L, "n_traj": B, "dt": dt, "t_max": t_max,
                      "steps": nst, "checkpoint_every": 50}}
    grads = {}
    for name in ("fp32", "fp64"):
        e = engine(name=name)
        e.adjoint(50)  # warm-up (allocations, code objects)
        sync()
        t0 = time.perf_counter()
        g = e.adjoint(50)
        sync()
        el = time.perf_counter() - t0
        e.close()
        grads[name] = g
        out[name] = {"seconds": round(el, 3), "traj_steps_per_s": round(B * nst / el, 1), "J": g["J"],
                     "dJdU_sum": float(np.sum(g["grad_u"])), "dJdk50": float(g["grad_pulse"][50])}
        log(f"adjoint {name}: {el:.2f} s, {B * nst / el:.4e} traj-steps/s")
    out["value"] = out["fp32"]["traj_steps_per_s"]

    def richardson(f):
        d = {h: (f(h) - f(-h)) / (2 * h) for h in (4e-6, 2e-6, 1e-6)}
        r1a, r1b = (4 * d[2e-6] - d[4e-6]) / 3, (4 * d[1e-6] - d[2e-6]) / 3
        return (16 * r1b - r1a) / 15, d[1e-6]

    def knot50(hh):
        k = w["knots"].copy()
        k[50] += hh
        return forward(knots=k)

    fd_u, fd_u1 = richardson(lambda hh: forward(U0=4.0 + hh))
    fd_k, fd_k1 = richardson(knot50)
    ad = out["fp64"]
    out["finite_differences_fp64"] = {
        "h": [4e-6, 2e-6, 1e-6], "extrapolation": "Richardson, h^2 and h^4 terms removed",
        "dJdU": fd_u, "dJdk50": fd_k, "dJdU_h1e-6": fd_u1, "dJdk50_h1e-6": fd_k1,
        "rel_diff_dJdU": abs(ad["dJdU_sum"] - fd_u) / max(abs(fd_u), 1e-300),
        "rel_diff_dJdk50": abs(ad["dJdk50"] - fd_k) / max(abs(fd_k), 1e-300)}
    gp32, gp64 = grads["fp32"]["grad_pulse"], grads["fp64"]["grad_pulse"]
    out["fp32_vs_fp64_knots_max_rel"] = float(np.max(np.abs(gp32 - gp64)) / np.max(np.abs(gp64)))
    log("adjoint", json.dumps(out))
    print(json.dumps(out), flush=True)


def main():
    ap = argparse.ArgumentParser(description=__doc__.splitlines()[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--workload", default="hubbard8x8",
                    choices=["hubbard8x8", "chain10", "mixed", "moire196", "adjoint6x6", "gpsr4x4"],
                    help="hubbard8x8 is the headline; the others print their own line (default length: the "
                         "whole configured run)")
    ap.add_argument("--steps", type=int, default=None,
                    help="timed RK4 steps (hubbard8x8: 4000 = t_max 20 at dt 0.005)")
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--n-traj", type=int, default=65536)
    ap.add_argument("--block", type=int, default=0, help="threads per trajectory workgroup (0, 512, 1024)")
    ap.add_argument("--time-chunk", type=int, default=0)
    ap.add_argument("--kernel", default="auto", choices=["auto", "full", "hermitian", "pipelined", "sequential"],
                    help="FP32 integrator: automatic (here the spin-sequential upper-triangle kernel in its "
                         "direct-commutator form), the spin-sequential kernel with X = T Y in LDS (sequential), "
                         "upper-triangle with both spins (hermitian), spin-pipelined, or full-matrix")
    ap.add_argument("--plain-steps", type=int, default=10)
    ap.add_argument("--no-check", action="store_true", help="skip the FP64 batched reference run")
    ap.add_argument("--no-horizon", action="store_true",
                    help="skip the untimed full-run (t_max 20) accuracy check of FP32 and FP64 batched")
    ap.add_argument("--horizon-plain-traj", type=int, default=1024,
                    help="trajectories of the full-run plain-vs-batched FP64 comparison")
    ap.add_argument("--no-plain", action="store_true")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-peaks", action="store_true",
                    help="skip the measured device peaks (copy, FMA and MFMA loops) beside the published ones")
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--states-steps", type=int, default=0,
                    help="also time this many FP32 steps recording the tight-binding eigenstate occupations "
                         "(tbStates), to report their cost")
    ap.add_argument("--calibrate-copy", type=int, default=0, metavar="MIB",
                    help="only run a device copy of MIB MiB (torch), the reference kernel for HBM counter scaling")
    args = ap.parse_args()
    if args.calibrate_copy:
        import torch
        n = args.calibrate_copy * (1 << 20) // 4
        x = torch.ones(n, dtype=torch.float32, device="cuda")
        y = torch.empty_like(x)
        for _ in range(3):
            y.copy_(x)
        torch.cuda.synchronize()
        print(json.dumps({"copy_bytes_read": n * 4, "copy_bytes_written": n * 4, "dispatches": 3}))
        return
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for --gpus {args.gpus}); trajectories shard trivially "
              "across GPUs but multi-GPU runs are not part of this build.", file=sys.stderr)
        sys.exit(2)
    if args.workload != "hubbard8x8":
        try:
            import torch

            sync_side = torch.cuda.synchronize if torch.cuda.is_available() else (lambda: None)
        except Exception:
            sync_side = lambda: None  # noqa: E731
        {"chain10": side_chain10, "mixed": side_mixed, "moire196": side_moire196,
         "adjoint6x6": side_adjoint6x6, "gpsr4x4": side_gpsr4x4}[args.workload](args, sync_side)
        return
    if args.steps is None:
        args.steps = 4000

    from fermiphasespace_amd import _lib
    try:
        import torch  # only to bracket the timed region with a device-wide synchronize

        sync = torch.cuda.synchronize if torch.cuda.is_available() else (lambda: None)
    except Exception:  # torch is plumbing here, not required
        sync = lambda: None  # noqa: E731

    dt, seed = 0.005, 202
    B, K, W = args.n_traj, args.steps, args.warmup
    w = workload(seed)
    L = w["L"]
    info = _lib.device_info()
    log("device", info, "U =", w["Uval"])
    eng = _lib.Engine(w["tM"], coulomb_type=_lib.COULOMB_ONSITE, U=w["U"])
    eng.set_initial_state(w["occ"], w["occ"])
    eng.set_tuning(args.block, args.time_chunk, args.kernel)
    eng.set_results(_lib.RES_DIAG | _lib.RES_DOUBLONS | _lib.RES_ENERGIES)

    # ---- warmup ----
    eng.set_params(dt, W * dt, 1, B, seed=seed, engine="fp32")
    eng.run()
    log(f"warmup {W} steps done")

    # ---- timed: K steps of the FP32 batched engine ----
    eng.set_params(dt, K * dt, 1, B, seed=seed, engine="fp32")
    sync()
    t0 = time.perf_counter()
    eng.run()
    sync()
    wall = time.perf_counter() - t0
    tm = eng.timing()
    value = B * K / wall
    log(f"fp32 batched: {K} steps x {B} traj in {wall:.3f} s -> {value:.4e} traj-steps/s; kernel {tm}")
    res32 = {k: eng.result(k) for k in ("Niiu", "Niid", "NuNd", "E", "Niiu_err", "NuNd_err")}

    # roofline of the integrator: algorithmic bytes 32 L^2 per trajectory-step (FP32)
    alg_bytes = 32.0 * L * L * B * K
    kern_s = tm["kernel_ms"] * 1e-3
    achieved = alg_bytes / kern_s / 1e9
    kname = {"full": "stencil_kernel<float,16,8,4,false,64>", "hermitian": "herm_kernel<16,8,5,4,false,false>",
             "pipelined": "herm_pipe_kernel<5,4>", "sequential": "herm_seq_kernel<float,8,4,4,2,1>"}.get(
                 args.kernel, "herm_wide_kernel<4,4,true>")
    pfile = {"hermitian": "herm_fp32_hubbard8x8_traffic.json",
             "sequential": "herm_seq_fp32_hubbard8x8_traffic.json"}.get(args.kernel, "herm_wide_fp32_hubbard8x8_traffic.json")
    traffic, prof = traffic_from_profile(os.path.join(ROOT, "profiles", pfile))
    if prof is None or prof.get("kernel") != kname:
        traffic = None  # no committed counter figures for this kernel
    z = 4
    valu_flop = (64 * z + 104) * L * L  # reference-formula flops per trajectory-step, sparse tM
    valu_achieved = valu_flop * B * K / kern_s / 1e12
    # The state stays on chip for a whole launch (measured HBM traffic is a few per cent of the streamed
    # state's bytes), so HBM does not bind: the line reports the arithmetic figure by the fixed formula
    # (64 z + 104) L^2 flop per trajectory-step against the FP32 vector peak, with the VALU issue, LDS
    # activity and LDS bank-conflict cycles the committed counters measured beside it (issue and LDS are
    # both near their limits), and the streamed-state HBM figure as a secondary field.
    roof = {"bound": "valu", "achieved": round(valu_achieved, 2), "peak": 157.3, "unit": "TFLOP/s",
            "frac": round(valu_achieved / 157.3, 4), "traffic": traffic,
            "kernel": kname, "launches": tm["launches"],
            "avg_launch_ms": round(tm["kernel_ms"] / max(1, tm["launches"]), 4),
            "alg_flop_per_traj_step": valu_flop,
            "lds_active_frac": prof.get("lds_active_frac") if prof else None,
            "valu_issue_frac": prof.get("valu_issue_frac") if prof else None,
            "lds_bank_conflict_frac": prof.get("lds_bank_conflict_frac") if prof else None,
            "hbm": {"achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                    "frac": round(achieved / HBM_PEAK_GBS, 4), "alg_bytes_per_traj_step": 32 * L * L,
                    "note": "streamed-state model: one read and one write of both spins' L x L blocks per step"}}
    pk = measured_peaks(args)
    if pk:
        # the same achieved figures against what a copy kernel and a (non-packed) FMA loop reach here
        roof["measured"] = {"hbm_copy_gbps": round(pk["hbm_copy_gbps"], 1),
                            "hbm_frac_of_copy": round(achieved / pk["hbm_copy_gbps"], 4),
                            "fp32_fma_tflops": round(pk["fp32_fma_tflops"], 2),
                            "valu_frac_of_fma": round(valu_achieved / pk["fp32_fma_tflops"], 4),
                            "fp64_fma_tflops": round(pk["fp64_fma_tflops"], 2),
                            "fp32_mfma_tflops": round(pk["fp32_mfma_tflops"], 2)}

    out = {"metric": METRIC, "value": round(value, 1), "unit": "traj-steps/s", "n_gpus": 1, "steps": K,
           "warmup": W, "ms_per_step": round(wall * 1e3 / K, 4), "higher_is_better": True, "scaling": "weak",
           "vs_baseline": None, "dtype": "fp32", "data": "synthetic",
           "config": {"workload": "hubbard8x8_periodic_cdw", "L": L, "n_traj": B, "dt": dt, "U": round(w["Uval"], 6),
                      "seed": seed, "steps_timed": K, "t_timed": K * dt, "t_max": K * dt, "integrator": "rk4_3_8",
                      "records_every_step": True},
           "roofline": roof}

    # ---- optional: the same FP32 run recording the tight-binding eigenstate occupations ----
    if args.states_steps > 0:
        Ks = args.states_steps
        Vt = np.linalg.eigh(w["tM"])[1]
        eng.set_bases(np.stack([np.stack([Vt, Vt])]))
        eng.set_params(dt, Ks * dt, 1, B, seed=seed, engine="fp32")
        sync()
        t0 = time.perf_counter()
        eng.run()
        sync()
        els = time.perf_counter() - t0
        ts = eng.timing()
        M = eng.result("sum_basis")[0]
        tr_err = float(np.abs(M[0].sum(axis=0) - eng.result("sum_u").sum(axis=0)).max())
        out["tb_states"] = {"value": round(B * Ks / els, 1), "unit": "traj-steps/s", "steps": Ks,
                            "seconds": round(els, 3), "kernel_ms": round(ts["kernel_ms"], 3),
                            "trace_check_max_abs": tr_err}
        log(f"fp32 with tbStates: {Ks} steps in {els:.3f} s -> {B * Ks / els:.4e} traj-steps/s; "
            f"trace check {tr_err:.2e}")
        eng.set_bases(None)

    # ---- FP64 batched reference over the same steps: FP32 error vs FP64 ----
    eng.set_tuning(0, 0, "auto")  # the FP32 kernel choice does not apply to the FP64 engines
    el64 = None
    if not args.no_check:
        eng.set_params(dt, K * dt, 1, B, seed=seed, engine="fp64")
        t0 = time.perf_counter()
        eng.run()
        el64 = time.perf_counter() - t0
        res64 = {k: eng.result(k) for k in ("Niiu", "Niid", "NuNd", "E")}
        out["fp64_batched"] = {"value": round(B * K / el64, 1), "unit": "traj-steps/s", "seconds": round(el64, 3)}
        out["err_vs_fp64"] = error_report(res32, res64, dt)
        log("fp32 vs fp64:", json.dumps(out["err_vs_fp64"]))

    # ---- untimed: accuracy over the whole configured run (t_max = 20, 4000 steps) ----
    if not args.no_horizon:
        out["accuracy_full_run"] = full_run_accuracy(eng, w, dt, seed, B, args, res32 if K == 4000 else None,
                                                     res64 if (K == 4000 and el64 is not None) else None)
        eng.set_tuning(0, 0, "auto")

    # ---- plain FP64 GPU path: the speedup denominator ----
    if not args.no_plain:
        P = args.plain_steps
        eng.set_params(dt, P * dt, 1, B, seed=seed, engine="fp64_plain")
        t0 = time.perf_counter()
        eng.run()
        elp = time.perf_counter() - t0
        tp = eng.timing()
        # both sides on the same footing: wall time of a whole run (sampling, steps, records, reduction),
        # and kernel time of the integrator launches alone
        plain_value = B * P / elp
        plain_kernel_value = B * P / (tp["kernel_ms"] * 1e-3)
        out["fp64_plain"] = {"value": round(plain_value, 1), "unit": "traj-steps/s", "steps": P,
                             "seconds": round(elp, 3), "kernel_ms": round(tp["kernel_ms"], 3)}
        out["speedup_vs_fp64_plain"] = round(value / plain_value, 2)
        out["speedup_vs_fp64_plain_kernel_only"] = round((B * K / kern_s) / plain_kernel_value, 2)
        if el64 is not None:
            out["fp64_batched"]["speedup_vs_fp64_plain"] = round((B * K / el64) / plain_value, 2)
        log(f"fp64 plain: {P} steps in {elp:.3f} s -> {plain_value:.4e} traj-steps/s; speedup {value / plain_value:.1f}x")
    eng.close()

    # ---- CPU baseline (oracle, host cores) ----
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(w, dt, args.cpu_seconds)
        log("cpu baseline", out["cpu_baseline"])
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()

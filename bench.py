#!/usr/bin/env python3
"""Benchmark: single-shot RTM on a 751 x 2301 faulted-layer model.

Workload (one "step" = one complete RTM of the shot):
  grid 751 x 2301 at 4 m, 40-cell absorbing pad, seeded faulted-layer model
  (1500-4500 m/s, 12 layers, dips +/-15 deg, 3 normal faults), migration
  velocity = that model Gaussian-smoothed with sigma 60 m; 20 Hz Ricker at the
  surface mid-x, 2301 surface receivers, dt = 0.4 ms, nt = 7500; snapshots
  every 4 steps (FP32 by default, --snap fp16 for FP16 scaled by max|S| per snapshot tile).
  Observed data are modeled from the true model before timing starts.

Metric: FD point-updates per second (Gpts/s) = sum over both propagations of
nz_FD * nx_FD * nt / wall time, inputs resident in HBM.  Prints one JSON line.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "Gpts/s FD updates + FWI time-to-misfit, learned-precond vs plain L-BFGS"
HBM_PEAK_GBS = 8000.0
TRAFFIC_FILE = os.path.join(ROOT, "profiles", "fd_step_rtm751x2301_traffic.json")

NZ, NX, DZ, DT, NT, NTAPER, SNAPI = 751, 2301, 4.0, 4e-4, 7500, 40, 4


def log(*a):
    print("[bench]", *a, file=sys.stderr, flush=True)


def build_inputs(seed=2, nt=NT):
    from seismic_amd.synthetic import faulted_layers, gaussian_smooth, ricker
    true = faulted_layers(NZ, NX, DZ, DZ, seed=seed)
    mig = gaussian_smooth(true, 60.0 / DZ)
    w = ricker(20.0, nt, DT)
    rx = (np.arange(NX) * DZ).astype(np.float32)
    rz = np.zeros_like(rx)
    return true, mig, w, rz, rx


def cpu_baseline(mig, w, rz, rx, budget_s=15.0):
    """The reference's own Fortran kernels (flang build, OpenMP) composed in
    Propagator.f90 order, on a bounded sample of this workload: the full
    831 x 2381 FD grid for `ns` forward steps (source injection + snapshots)
    and `ns` reverse-time steps (adjoint-style stencil + 2301 receiver
    injections).  Falls back to the single-threaded C oracle port."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_lib as O
    g = O.make_grid(NZ, NX, DZ, DZ, DT, ntaper=NTAPER, snapi=SNAPI)
    shot = O.Shot(0.0, DZ * (NX // 2), w, rz, rx)
    pts_step = (NZ + 2 * NTAPER) * (NX + 2 * NTAPER)
    cores = int(os.environ.get("OMP_NUM_THREADS", "0") or 0) or (os.cpu_count() or 1)
    if O.ref_available():
        vpad, oxw, nxw, b = O.padded_vel(g, mig, shot)
        rng = np.random.default_rng(0)

        def run(ns):
            data = (rng.standard_normal((len(rx), ns)) * 1e-3).astype(np.float32)
            t0 = time.perf_counter()
            O.ref_propagate(g, vpad, oxw, nxw, 0, 1, [shot.sz], [shot.sx], w[None, :ns], -1, [], [], ns, 1,
                            ns // SNAPI)
            O.ref_propagate(g, vpad, oxw, nxw, 1, 1, rz, rx, data, -1, [], [], ns, -1, ns // SNAPI)
            return time.perf_counter() - t0
        t = run(8)
        ns = int(max(16, min(800, 8 * budget_s / max(t, 1e-3))) // SNAPI * SNAPI)
        t = run(ns)
        return {"value": round(2 * ns * pts_step / t / 1e9, 4), "unit": "Gpts/s", "cores": cores,
                "kind": "reference",
                "sample": "reference Fortran kernels (flang -O2 -fopenmp, %d threads): %d forward + %d reverse "
                          "steps of the full %dx%d FD grid, %.1f s" % (cores, ns, ns, NZ + 2 * NTAPER,
                                                                       NX + 2 * NTAPER, t)}
    # port: C oracle, single thread, forward modeling of a truncated record
    ns = 64
    sub = O.Shot(0.0, DZ * (NX // 2), w[:ns], rz, rx)
    t0 = time.perf_counter()
    O.modeling(g, mig, sub, extract_kind=0)
    t = time.perf_counter() - t0
    return {"value": round(ns * pts_step / t / 1e9, 4), "unit": "Gpts/s", "cores": 1, "kind": "port",
            "sample": "C oracle, %d forward steps of the full FD grid, %.1f s" % (ns, t)}


def fwi_leg(niter=20):
    """Config 4 FWI: 201 x 601 at 10 m, 1500->3500 m/s gradient + 4 seeded
    +/-250 m/s Gaussians, initial model = Gaussian sigma 200 m; 32 shots of a
    5 Hz Ricker, 601 surface receivers, dt 1 ms, nt 3000; L-BFGS m = 5 for
    `niter` iterations, plain and with the learned preconditioner.  Time to
    misfit = wall seconds until the preconditioned run's accepted misfit is at
    or below the plain run's final accepted misfit (null if never)."""
    import seismic_amd as sa
    from seismic_amd.synthetic import gradient_anomalies, gaussian_smooth, ricker
    nz, nx, dx, dt, nt = 201, 601, 10.0, 1e-3, 3000
    true = gradient_anomalies(nz, nx, dx, dx, seed=4)
    init = gaussian_smooth(true, 200.0 / dx)
    w = ricker(5.0, nt, dt)
    rx = (np.arange(nx) * dx).astype(np.float32)
    rz = np.zeros_like(rx)
    shots = [sa.Shot(0.0, float(sx), w, rz, rx) for sx in np.linspace(0.0, dx * (nx - 1), 32)]
    e = sa.Engine(sa.make_grid(nz, nx, dx, dx, dt, ntaper=40, snapi=4))
    e.set_model(true)
    dobs = e.forward(shots, extract_kind=sa.SFD_EXTRACT_NEAREST)
    kw = dict(niter=niter, neval=3 * niter, bound=0, illupow=0.0)
    _, rp = e.fwi_lbfgs(shots, dobs, init, **kw)
    log("fwi plain: %d iterations, %.2f s" % (rp["iterations"], rp["seconds"]))
    _, rc = e.fwi_lbfgs(shots, dobs, init, precond=dict(train_after=3, train_steps=300), **kw)
    log("fwi preconditioned: %d iterations, %.2f s" % (rc["iterations"], rc["seconds"]))
    fp = rp["f_eval"][rp["accepted"]]
    fc, tc = rc["f_eval"][rc["accepted"]], rc["t_eval"][rc["accepted"]]
    target = float(fp[-1])
    hit = np.nonzero(fc <= target)[0]
    t_pc = float(tc[hit[0]]) if hit.size else None
    t_plain = float(rp["t_eval"][rp["accepted"]][-1])
    return {"workload": "fwi-201x601-32shots-nt3000-lbfgs%d" % niter,
            "plain": {"seconds": round(rp["seconds"], 3), "iterations": rp["iterations"],
                      "evaluations": rp["evaluations"], "final_misfit_rel": float(fp[-1] / rp["f_eval"][0])},
            "preconditioned": {"seconds": round(rc["seconds"], 3), "iterations": rc["iterations"],
                               "evaluations": rc["evaluations"], "with_network": rc["pc_iterations"],
                               "restarts": rc["pc_fallbacks"], "train_seconds": round(rc["pc_train_seconds"], 3),
                               "final_misfit_rel": float(fc[-1] / rc["f_eval"][0])},
            "time_to_plain_final_misfit_s": {"plain": round(t_plain, 3),
                                             "preconditioned": None if t_pc is None else round(t_pc, 3)},
            "speedup": None if t_pc is None else round(t_plain / t_pc, 3)}


def batched_leg(seed=3):
    """Config 3: 96 shots spread uniformly over the surface of the 751 x 2301
    faulted model (4 m), each with a centred receiver spread of U{200..1200}
    traces (clipped to the model) and nt in U{4000..7500} at dt 0.4 ms (seeded);
    observed data from the true model, then one batched forward + adjoint
    gradient call (compute_fct_gdt) on the smoothed model with FP16 snapshots
    every 4 steps, tile shape chosen by the autotuner.  Gpts/s counts both
    propagations of every shot over its own FD region and record length."""
    import seismic_amd as sa
    true, mig, _, _, _ = build_inputs(nt=16)
    rng = np.random.default_rng(seed)
    shots, pts = [], 0.0
    for sx in np.linspace(0.0, DZ * (NX - 1), 96):
        nr = int(rng.integers(200, 1201))
        nt = int(rng.integers(4000, 7501))
        j0 = int(round(sx / DZ)) - nr // 2
        rx = (np.clip(np.arange(j0, j0 + nr), 0, NX - 1) * DZ).astype(np.float32)
        rx = np.unique(rx)
        from seismic_amd.synthetic import ricker
        shots.append(sa.Shot(0.0, float(sx), ricker(20.0, nt, DT), np.zeros_like(rx), rx))
    e = sa.Engine(sa.make_grid(NZ, NX, DZ, DZ, DT, ntaper=NTAPER, snapi=SNAPI))
    e.set_model(true)
    t0 = time.perf_counter()
    dobs = e.forward(shots, extract_kind=sa.SFD_EXTRACT_NEAREST)
    t_obs = time.perf_counter() - t0
    pts_fwd = e.last_timing()["pts"]
    e.set_model(mig)
    e.set_tile(0, 0)
    t0 = time.perf_counter()
    f, grad, _, _ = e.fwi_gradient(shots, dobs, snap_mode=sa.SFD_SNAP_FP16S)
    wall = time.perf_counter() - t0
    tm, ks, tile = e.last_timing(), e.kernel_stats(), e.get_tile()
    return {"workload": "batched-96-shots-751x2301-ragged", "shots": len(shots),
            "receivers": [int(min(s.ntr for s in shots)), int(max(s.ntr for s in shots))],
            "nt": [int(min(s.nt for s in shots)), int(max(s.nt for s in shots))],
            "gradient_s": round(wall, 3), "Gpts_per_s": round(tm["pts"] / wall / 1e9, 3),
            "pass_ms": [round(tm["fwd_ms"], 1), round(tm["bwd_ms"], 1)],
            "observed_data_s": round(t_obs, 3), "observed_data_Gpts_per_s": round(pts_fwd / t_obs / 1e9, 3),
            "avg_launch_us": [round(ks["fwd_us"], 2), round(ks["bwd_us"], 2)],
            "tile": list(tile["tile"]), "strip": tile["strip"], "time_block": tile["time_block"],
            "autotune": tile["log"],
            "snapshots": "fp16", "misfit": f}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--snap", choices=["fp32", "fp16"], default="fp32")
    ap.add_argument("--tile", default="64x32")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--graphs", type=int, default=1, help="0: eager launches (for profilers)")
    ap.add_argument("--strip", type=int, default=0, help="columns per wavefront (0 auto, 4, 8, 16)")
    ap.add_argument("--time-block", type=int, default=1, choices=[1, 2], help="time steps per launch")
    ap.add_argument("--cpu-budget", type=float, default=15.0)
    ap.add_argument("--nt", type=int, default=NT, help="record length (shorter records for counter passes)")
    ap.add_argument("--no-fwi", action="store_true", help="skip the config-4 FWI time-to-misfit leg")
    ap.add_argument("--no-batched", action="store_true", help="skip the config-3 batched gradient leg")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="run only the 512 MiB copy kernel (counter calibration) and exit")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py runs on one GPU only: this build has no multi-GPU path "
              "(shots would shard across GPUs with one gradient all-reduce); got --gpus %d" % args.gpus,
              file=sys.stderr)
        sys.exit(2)

    import seismic_amd as sa
    import seismic_amd.abi  # noqa: F401
    if args.copy_calibration:
        import ctypes
        ms = ctypes.c_double()
        nbytes = 512 << 20
        rc = sa.abi.load().sfd_copy_calibration(nbytes, 5, ctypes.byref(ms))
        print(json.dumps({"copy_bytes": nbytes, "launches": 6, "avg_ms": ms.value, "rc": rc,
                          "GBps": 2 * nbytes / (ms.value * 1e-3) / 1e9 if ms.value > 0 else None}))
        sys.exit(0 if rc == 0 else 1)
    info = sa.device_info()
    log("device", info["name"], "CUs", info["cus"])
    true, mig, w, rz, rx = build_inputs(nt=args.nt)
    g = sa.make_grid(NZ, NX, DZ, DZ, DT, ntaper=NTAPER, snapi=SNAPI)
    e = sa.Engine(g)
    tz, tx = (int(v) for v in args.tile.split("x"))
    e.set_tile(tz, tx)
    e.set_option(sa.abi.SFD_OPT_GRAPHS, args.graphs)
    e.set_option(sa.abi.SFD_OPT_STRIP_PX, args.strip)
    e.set_option(sa.abi.SFD_OPT_TIME_BLOCK, args.time_block)
    shot = sa.Shot(0.0, DZ * (NX // 2), w, rz, rx)
    log("modeling observed data (untimed)")
    e.set_model(true)
    dobs = e.forward([shot])[0]
    e.set_model(mig)
    snap = sa.SFD_SNAP_FP16S if args.snap == "fp16" else sa.SFD_SNAP_FP32
    # inputs resident in HBM before the timed region: the observed data are
    # uploaded once (sfd_set_traces) and every RTM reads them on the device
    e.set_traces([dobs], [shot])
    for k in range(args.warmup):
        e.rtm([shot], None, snap_mode=snap)
        log("warmup %d done" % k)
    # every sfd_* call returns after the engine's stream has completed, so the
    # wall clock brackets exactly the timed device work
    t0 = time.perf_counter()
    pts = 0.0
    for k in range(args.steps):
        e.rtm([shot], None, snap_mode=snap)
        pts += e.last_timing()["pts"]
        log("step %d done" % k)
    wall = time.perf_counter() - t0
    tm = e.last_timing()
    ks = e.kernel_stats()
    value = pts / wall / 1e9
    # the same RTM with the observed data handed over from host memory each
    # call (PCIe-inclusive; reported beside the value, never as it)
    t1 = time.perf_counter()
    e.rtm([shot], [dobs], snap_mode=snap)
    host_rate = e.last_timing()["pts"] / (time.perf_counter() - t1) / 1e9
    # the same RTM planned afresh (no reuse of the previous call's tables):
    # host planning and table upload inside the measured call
    e.set_option(sa.abi.SFD_OPT_PLAN_REUSE, 0)
    t1 = time.perf_counter()
    e.rtm([shot], None, snap_mode=snap)
    fresh_rate = e.last_timing()["pts"] / (time.perf_counter() - t1) / 1e9
    e.set_option(sa.abi.SFD_OPT_PLAN_REUSE, 1)

    # roofline of the step kernel: algorithmic bytes per launch / measured duration
    nfd = (NZ + 2 * NTAPER) * (NX + 2 * NTAPER)
    npix = NZ * NX
    sb = 2 if args.snap == "fp16" else 4
    bytes_fwd = 24.0 * nfd + sb * npix / SNAPI
    bytes_bwd = 24.0 * nfd + (sb + 16.0) * npix / SNAPI
    ach_fwd = bytes_fwd / (ks["fwd_us"] * 1e-6) / 1e9 if ks["fwd_us"] > 0 else None
    ach_bwd = bytes_bwd / (ks["bwd_us"] * 1e-6) / 1e9 if ks["bwd_us"] > 0 else None
    traffic = None
    if os.path.exists(TRAFFIC_FILE):
        with open(TRAFFIC_FILE) as f:
            traffic = json.load(f).get("bytes_per_launch_fwd")
    roof = {"bound": "hbm", "kernel": "fd_step_kernel<%d,%d,fwd>" % (tz, tx),
            "achieved": round(ach_fwd, 1) if ach_fwd else None, "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(ach_fwd / HBM_PEAK_GBS, 4) if ach_fwd else None, "traffic": traffic,
            "algorithmic_bytes_per_launch": bytes_fwd, "avg_launch_us": round(ks["fwd_us"], 3),
            "samples": ks["fwd_n"],
            "reverse_time_kernel": {"kernel": "fd_step_kernel<%d,%d,adj>" % (tz, tx),
                                    "achieved": round(ach_bwd, 1) if ach_bwd else None,
                                    "algorithmic_bytes_per_launch": bytes_bwd,
                                    "avg_launch_us": round(ks["bwd_us"], 3)}}
    batched = None
    if not args.no_batched:
        log("batched 96-shot gradient leg")
        batched = batched_leg()
    fwi = None
    if not args.no_fwi:
        log("fwi time-to-misfit leg")
        fwi = fwi_leg()
    cpu = None
    if not args.no_cpu:
        log("cpu baseline")
        cpu = cpu_baseline(mig, w, rz, rx, budget_s=args.cpu_budget)
    out = {"metric": METRIC, "value": round(value, 3), "unit": "Gpts/s", "n_gpus": 1, "steps": args.steps,
           "warmup": args.warmup, "ms_per_step": round(wall / args.steps * 1000, 2), "higher_is_better": True,
           "scaling": "weak", "vs_baseline": None, "dtype": "f32", "data": "synthetic",
           "config": {"workload": "rtm-single-shot-751x2301-nt%d" % args.nt, "grid": [NZ, NX], "dx_m": DZ,
                      "dt_s": DT, "nt": args.nt, "ntaper": NTAPER, "receivers": NX, "snapi": SNAPI, "snapshots": args.snap,
                      "tile": args.tile, "strip": args.strip, "time_block": args.time_block, "pass_ms": [round(tm["fwd_ms"], 1), round(tm["bwd_ms"], 1)],
                      "inputs": "observed data resident in HBM (sfd_set_traces) before the timed region",
                      "host_inputs_Gpts_per_s": round(host_rate, 3),
                      "plan": "tables planned by the warm-up call and reused (SFD_OPT_PLAN_REUSE)",
                      "fresh_plan_Gpts_per_s": round(fresh_rate, 3),
                      "device": info["name"]},
           "roofline": roof, "cpu_baseline": cpu, "fwi": fwi,
           "batched": batched}
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()

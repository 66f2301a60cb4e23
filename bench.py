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
    ap.add_argument("--cpu-budget", type=float, default=15.0)
    ap.add_argument("--nt", type=int, default=NT, help="record length (shorter records for counter passes)")
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
    shot = sa.Shot(0.0, DZ * (NX // 2), w, rz, rx)
    log("modeling observed data (untimed)")
    e.set_model(true)
    dobs = e.forward([shot])[0]
    e.set_model(mig)
    snap = sa.SFD_SNAP_FP16S if args.snap == "fp16" else sa.SFD_SNAP_FP32
    for k in range(args.warmup):
        e.rtm([shot], [dobs], snap_mode=snap)
        log("warmup %d done" % k)
    t0 = time.perf_counter()
    pts = 0.0
    for k in range(args.steps):
        e.rtm([shot], [dobs], snap_mode=snap)
        pts += e.last_timing()["pts"]
        log("step %d done" % k)
    wall = time.perf_counter() - t0
    tm = e.last_timing()
    ks = e.kernel_stats()
    value = pts / wall / 1e9

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
    cpu = None
    if not args.no_cpu:
        log("cpu baseline")
        cpu = cpu_baseline(mig, w, rz, rx, budget_s=args.cpu_budget)
    out = {"metric": METRIC, "value": round(value, 3), "unit": "Gpts/s", "n_gpus": 1, "steps": args.steps,
           "warmup": args.warmup, "ms_per_step": round(wall / args.steps * 1000, 2), "higher_is_better": True,
           "scaling": "weak", "vs_baseline": None, "dtype": "f32", "data": "synthetic",
           "config": {"workload": "rtm-single-shot-751x2301-nt%d" % args.nt, "grid": [NZ, NX], "dx_m": DZ,
                      "dt_s": DT, "nt": args.nt, "ntaper": NTAPER, "receivers": NX, "snapi": SNAPI, "snapshots": args.snap,
                      "tile": args.tile, "strip": args.strip, "pass_ms": [round(tm["fwd_ms"], 1), round(tm["bwd_ms"], 1)],
                      "device": info["name"]},
           "roofline": roof, "cpu_baseline": cpu}
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()

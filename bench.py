#!/usr/bin/env python3
"""Benchmark: FWI gradient of a 96-shot ragged batch on a 751 x 2301 model.

Workload (one "step" = one complete misfit + gradient evaluation of the
batch, compute_fct_gdt, progs/afwi/compute_function_gradient.f90:208-350):
  751 x 2301 at 4 m, 40-cell absorbing pad, seeded faulted-layer model
  (1500-4500 m/s, 12 layers, dips +/-15 deg, 3 normal faults); the gradient
  is taken at that model Gaussian-smoothed with sigma 60 m.  96 shots spread
  uniformly over the surface, each with a centred spread of U{200..1200}
  receivers (clipped to the model) and nt in U{4000..7500} at dt 0.4 ms,
  20 Hz Ricker; observed data modeled from the true model before timing and
  kept resident in HBM.  FP32 arithmetic, source snapshots every 4 steps
  stored as FP16 scaled per snapshot tile.  Every timed call plans its batch
  afresh (windows, padding, injection / extraction tables, uploads).

Metric: FD point-updates per second (Gpts/s) = sum over shots of
nz_FD * nx_FD * nt over both propagations / wall time.  Prints one JSON
line; extra keys carry the single-shot 751 x 2301 RTM, the 201 x 601 FWI
time-to-misfit (plain vs learned-preconditioned L-BFGS) and one gradient of
the 1911 x 5395 salt model (64 shots).
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
TRAFFIC_FILE = os.path.join(ROOT, "profiles", "fd_step_batched96_traffic.json")


def log(*a):
    print("[bench]", *a, file=sys.stderr, flush=True)


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def engine_for(sa, w):
    nz, nx, dz, dx, dt = w["grid"]
    return sa.Engine(sa.make_grid(nz, nx, dz, dx, dt, ntaper=w["ntaper"], snapi=w["snapi"]))


def shots_of(sa, w, idx=None, nt_max=0):
    idx = range(len(w["shots"])) if idx is None else idx
    out = []
    for k in idx:
        sz, sx, wav, rz, rx = w["shots"][k]
        if nt_max:
            wav = wav[:nt_max]
        out.append(sa.Shot(sz, sx, wav, rz, rx))
    return out


def pass_roofline(kt, k, kernel):
    """Achieved algorithmic bandwidth of one pass's launches (event-timed)."""
    us, by, n = kt["us"][k], kt["bytes"][k], kt["steps"][k]
    if us <= 0 or n <= 0:
        return None
    ach = by / (us * 1e-6) / 1e9
    return {"kernel": kernel, "achieved": round(ach, 1), "frac": round(ach / HBM_PEAK_GBS, 4),
            "algorithmic_bytes_per_launch": round(by / n), "avg_launch_us": round(us / n, 3), "launches": n,
            "pass_kernel_ms": round(us / 1e3, 2)}


def batched_headline(sa, args):
    """The 96-shot ragged batch: observed data (untimed), autotune in the
    first warm-up call, then `steps` timed gradient evaluations."""
    from seismic_amd import synthetic as syn
    w = syn.batched_96shot_workload()
    shots = shots_of(sa, w, nt_max=args.nt_max)
    e = engine_for(sa, w)
    e.set_option(sa.abi.SFD_OPT_GRAPHS, args.graphs)
    e.set_model(w["true"])
    t0 = time.perf_counter()
    dobs = e.forward(shots, extract_kind=sa.SFD_EXTRACT_NEAREST)
    t_obs = time.perf_counter() - t0
    obs_pts = e.last_timing()["pts"]
    log("observed data: %d shots in %.2f s" % (len(shots), t_obs))
    e.set_model(w["mig"])
    e.set_traces(dobs, shots)              # inputs resident in HBM before the timed region
    snap = sa.SFD_SNAP_FP16S
    if args.warmup > 0:
        e.set_tile(0, 0)                   # the first warm-up call autotunes tile / strip / time block
    # as in sfd_fwi_lbfgs's evaluations, the residual gathers (the adjoint
    # sources, about 1.5 GB here) stay on the device: misfit, gradient and
    # illumination come back
    for k in range(args.warmup):
        e.fwi_gradient(shots, None, snap_mode=snap, residuals=False)
        log("warmup %d done" % k)
    tile = e.get_tile()
    # every sfd_* call returns after the engine's stream has completed, so
    # the wall clock brackets exactly the timed device work
    t0 = time.perf_counter()
    pts = 0.0
    for k in range(args.steps):
        f, _, _, _ = e.fwi_gradient(shots, None, snap_mode=snap, residuals=False)
        pts += e.last_timing()["pts"]
        log("step %d done" % k)
    wall = time.perf_counter() - t0
    tm, kt = e.last_timing(), e.kernel_totals()
    # the same evaluation with the observed data handed over from host memory
    t1 = time.perf_counter()
    e.fwi_gradient(shots, dobs, snap_mode=snap, residuals=False)
    host_rate = e.last_timing()["pts"] / (time.perf_counter() - t1) / 1e9
    fwd = pass_roofline(kt, 0, "fd_step_kernel<4,false,false,0|1> (forward time, snapshot stores)")
    bwd = pass_roofline(kt, 1, "fd_step_kernel<4,false,false,0> + image_kernel<true> (reverse time, FWI imaging)")
    return dict(w=w, shots=shots, wall=wall, pts=pts, misfit=f, tm=tm, kt=kt, fwd=fwd, bwd=bwd, tile=tile,
                host_rate=host_rate, t_obs=t_obs, obs_rate=obs_pts / t_obs / 1e9)


def rtm_leg(sa, steps=3):
    """Single-shot RTM, 751 x 2301 at 4 m, nt 7500, FP32 snapshots: repeated
    migrations of the same shot (plan tables reused), observed data resident."""
    from seismic_amd import synthetic as syn
    w = syn.rtm_751x2301_workload()
    s = shots_of(sa, w)
    e = engine_for(sa, w)
    e.set_model(w["true"])
    d = e.forward(s)
    e.set_model(w["mig"])
    e.set_traces(d, s)
    e.rtm(s, None)
    t0 = time.perf_counter()
    for _ in range(steps):
        e.rtm(s, None)
    wall = time.perf_counter() - t0
    rate = steps * e.last_timing()["pts"] / wall / 1e9
    kt = e.kernel_totals()
    e.set_option(sa.abi.SFD_OPT_PLAN_REUSE, 0)
    t1 = time.perf_counter()
    e.rtm(s, None)
    fresh = e.last_timing()["pts"] / (time.perf_counter() - t1) / 1e9
    return {"workload": "rtm-single-shot-751x2301-nt7500-fp32snap", "Gpts_per_s": round(rate, 3),
            "ms_per_rtm": round(wall / steps * 1e3, 2), "fresh_plan_Gpts_per_s": round(fresh, 3),
            "forward": pass_roofline(kt, 0, "fd_step_kernel<4,false,true,0|1>"),
            "reverse": pass_roofline(kt, 1, "fd_step_kernel<4,true,true,0|3> (imaging blocks)")}


def fwi_leg(sa, niter=20):
    """201 x 601 at 10 m, 1500->3500 m/s gradient + 4 seeded +/-250 m/s
    Gaussians, initial model Gaussian sigma 200 m; 32 shots of a 5 Hz Ricker,
    601 surface receivers, dt 1 ms, nt 3000; L-BFGS m = 5 for `niter`
    iterations without bounds (bound 0) and without illumination weighting
    (illu_pow 0): at the reference's defaults (bound 2, illu_pow 1,
    Inversion_types.f90:57-60) its L-BFGS stops within 1-4 iterations on this
    workload, with or without a near-surface mask.  Plain and with the
    learned preconditioner.  Time to misfit = wall seconds until the
    preconditioned run's accepted misfit is at or below the plain run's final
    accepted misfit (null if never)."""
    from seismic_amd import synthetic as syn
    w = syn.fwi_201x601_workload()
    shots = shots_of(sa, w)
    e = engine_for(sa, w)
    e.set_model(w["true"])
    dobs = e.forward(shots, extract_kind=sa.SFD_EXTRACT_NEAREST)
    kw = dict(niter=niter, neval=3 * niter, bound=0, illupow=0.0)
    _, rp = e.fwi_lbfgs(shots, dobs, w["mig"], **kw)
    log("fwi plain: %d iterations, %.2f s" % (rp["iterations"], rp["seconds"]))
    _, rc = e.fwi_lbfgs(shots, dobs, w["mig"], precond=dict(train_after=3, train_steps=300), **kw)
    log("fwi preconditioned: %d iterations, %.2f s" % (rc["iterations"], rc["seconds"]))
    return time_to_misfit("fwi-201x601-32shots-nt3000-lbfgs%d" % niter, rp, rc)


def time_to_misfit(name, rp, rc):
    fp = rp["f_eval"][rp["accepted"]]
    fc, tc = rc["f_eval"][rc["accepted"]], rc["t_eval"][rc["accepted"]]
    target = float(fp[-1])
    hit = np.nonzero(fc <= target)[0]
    t_pc = float(tc[hit[0]]) if hit.size else None
    t_plain = float(rp["t_eval"][rp["accepted"]][-1])
    return {"workload": name,
            "plain": {"seconds": round(rp["seconds"], 3), "iterations": rp["iterations"],
                      "evaluations": rp["evaluations"], "final_misfit_rel": float(fp[-1] / rp["f_eval"][0])},
            "preconditioned": {"seconds": round(rc["seconds"], 3), "iterations": rc["iterations"],
                               "evaluations": rc["evaluations"], "with_network": rc["pc_iterations"],
                               "restarts": rc["pc_fallbacks"], "train_seconds": round(rc["pc_train_seconds"], 3),
                               "final_misfit_rel": float(fc[-1] / rc["f_eval"][0])},
            "time_to_plain_final_misfit_s": {"plain": round(t_plain, 3),
                                             "preconditioned": None if t_pc is None else round(t_pc, 3)},
            "speedup": None if t_pc is None else round(t_plain / t_pc, 3)}


def salt_leg(sa):
    """One misfit + gradient of the 1911 x 5395 salt model: 64 shots, 1200
    receivers each, nt 8000, 7 Hz, FP16 snapshots every 10 steps (one scale
    per snapshot), observed data resident in HBM."""
    from seismic_amd import synthetic as syn
    w = syn.salt_fwi_workload()
    shots = shots_of(sa, w)
    e = engine_for(sa, w)
    e.set_model(w["true"])
    dobs = e.forward(shots, extract_kind=sa.SFD_EXTRACT_NEAREST)
    e.set_model(w["mig"])
    e.set_traces(dobs, shots)
    t0 = time.perf_counter()
    f, _, _, _ = e.fwi_gradient(shots, None, snap_mode=sa.SFD_SNAP_FP16G, residuals=False)
    wall = time.perf_counter() - t0
    tm, kt = e.last_timing(), e.kernel_totals()
    return {"workload": "salt-fwi-gradient-1911x5395-64shots-nt8000-fp16snap", "gradient_s": round(wall, 3),
            "Gpts_per_s": round(tm["pts"] / wall / 1e9, 3), "waves": e.last_waves(), "misfit": f,
            "forward": pass_roofline(kt, 0, "fd_step_kernel<4,false,false,0|1>"),
            "reverse": pass_roofline(kt, 1, "fd_step_kernel<4,false,false,0> + image_kernel<true>")}


def density_leg(sa, nt=2000):
    """Variable-density forward modeling (FD_2D_derivatives_acoustic_forward_grid)
    of one shot on the 751 x 2301 faulted model at 4 m, nt 2000, 2301 surface
    receivers, density from the velocity by Gardner's relation (310 v^0.25),
    beside the same shot at constant density.  Algorithmic bytes per FD point
    and step: 24 at constant density (u_-1, u0, u1, u2, v^2 read, u3 written);
    28 with a density (the step kernel forms the staggered gradients from u2
    and delp = 1/rho2 itself: u_-1, u0, u1, u2, delp, v^2 rho read, u3
    written)."""
    from seismic_amd import synthetic as syn
    w = syn.rtm_751x2301_workload(nt=nt)
    s = shots_of(sa, w)
    e = engine_for(sa, w)
    e.set_model(w["true"])
    rho = (310.0 * np.power(w["true"].astype(np.float64), 0.25)).astype(np.float32)
    out = {"workload": "forward-751x2301-nt%d-one-shot" % nt}
    for name, r, bpp in (("constant_density", None, 24.0), ("variable_density", rho, 28.0)):
        e.set_density(r)
        e.forward(s)
        t0 = time.perf_counter()
        e.forward(s)
        wall = time.perf_counter() - t0
        kt = e.kernel_totals()
        us, pts, steps = kt["us"][0], kt["pts"][0], kt["steps"][0]
        ach = bpp * pts / (us * 1e-6) / 1e9
        out[name] = {"Gpts_per_s": round(pts / wall / 1e9, 3), "us_per_step": round(us / max(steps, 1), 3),
                     "algorithmic_bytes_per_point": bpp, "achieved_GBps": round(ach, 1),
                     "frac": round(ach / HBM_PEAK_GBS, 4)}
    e.close()
    return out


def forward3d_leg(sa, n=300, nt=400):
    """3D forward modeling (genpar%twoD false) of an n^3 model at 10 m (FD
    region (n+40)^3 with ntaper 20), velocity 1500 -> 3000 m/s with depth,
    15 Hz Ricker at 20 m depth in the middle, 100 receivers along x at 20 m
    depth, dt 1 ms, nt 400; constant density, then a density.  The step's
    algorithmic bytes at constant density: 24 per FD point (the fused step
    kernel reads u_-1, u0, u1, u2, v^2 and writes u3); achieved = those bytes
    over the whole call's wall time divided by nt (bands, extraction and
    set-up included); then one RTM of the same shot with snapshots every 4
    steps (both passes counted)."""
    from seismic_amd.synthetic import ricker
    g = sa.make_grid3d(n, n, n, 10.0, 10.0, 10.0, 1e-3, ntaper=20, aperture_x=10.0 * (n - 1),
                       aperture_y=10.0 * (n - 1), fmax=0.0)
    z = np.arange(n, dtype=np.float32)[:, None, None] / n
    vel = np.ascontiguousarray(np.broadcast_to(1500.0 + 1500.0 * z, (n, n, n)), np.float32)
    e = sa.Engine3D(g)
    e.set_model(vel)
    rx = np.linspace(100.0, 10.0 * (n - 10), 100).astype(np.float32)
    shot = sa.Shot3D(20.0, 5.0 * n, 5.0 * n, ricker(15.0, nt, 1e-3), np.full_like(rx, 20.0), rx,
                     np.full_like(rx, 5.0 * n))
    F = (n + 40) ** 3
    out = {"workload": "forward3d-%d^3-nt%d" % (n, nt), "fd_points": F}
    for name, rho in (("constant_density", None), ("variable_density", 2000.0 + 500.0 * np.broadcast_to(z, (n, n, n)))):
        if rho is not None:
            e.set_density(np.ascontiguousarray(rho, np.float32))
        e.forward([shot])
        t0 = time.perf_counter()
        e.forward([shot])
        wall = time.perf_counter() - t0
        out[name] = {"Gpts_per_s": round(F * nt / wall / 1e9, 3), "us_per_step": round(wall / nt * 1e6, 1)}
    ach = 24.0 * F * nt / (out["constant_density"]["us_per_step"] * nt * 1e-6) / 1e9
    out["constant_density"].update({"algorithmic_bytes_per_point": 24.0, "achieved_GBps": round(ach, 1),
                                    "frac": round(ach / HBM_PEAK_GBS, 4)})
    e.close()
    # 3D RTM of the same shot (constant density, snapshots every 4 steps on
    # the window): forward pass with snapshots, reverse pass with imaging
    g4 = sa.make_grid3d(n, n, n, 10.0, 10.0, 10.0, 1e-3, ntaper=20, aperture_x=10.0 * (n - 1),
                        aperture_y=10.0 * (n - 1), fmax=0.0, snapi=4)
    e = sa.Engine3D(g4)
    e.set_model(vel)
    tr = e.forward([shot])
    e.rtm([shot], tr)
    t0 = time.perf_counter()
    e.rtm([shot], tr)
    wall = time.perf_counter() - t0
    out["rtm_snapi4"] = {"Gpts_per_s": round(2 * F * nt / wall / 1e9, 3), "ms_per_rtm": round(wall * 1e3, 1)}
    e.close()
    return out


def cpu_baseline(w, budget_s=15.0):
    """The reference's own Fortran kernels (flang build, OpenMP over columns)
    composed in Propagator.f90 order on a bounded sample of the batched
    workload: one of its shots (number 48) for `ns` forward steps (sinc
    source injection, nearest extraction, snapshots) and `ns` reverse steps
    (simple injection of its receivers, the AGRAD pass without the imaging
    condition, whose Imaging.f90 needs SEPlib).  Falls back to the C oracle."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_lib as O
    nz, nx, dz, dx, dt = w["grid"]
    g = O.make_grid(nz, nx, dz, dx, dt, ntaper=w["ntaper"], snapi=w["snapi"])
    sz, sx, wav, rz, rx = w["shots"][48]
    shot = O.Shot(sz, sx, wav, rz, rx)
    cores = int(os.environ.get("OMP_NUM_THREADS", "0") or 0) or (os.cpu_count() or 1)
    oxw, nxw = O.shot_window(g, shot)
    pts_step = (nz + 2 * w["ntaper"]) * (nxw + 2 * w["ntaper"])
    snapi = w["snapi"]
    if O.ref_available():
        vpad, oxw, nxw, _ = O.padded_vel(g, w["mig"], shot)
        rng = np.random.default_rng(0)

        def run(ns):
            data = (rng.standard_normal((len(rx), ns)) * 1e-3).astype(np.float32)
            t0 = time.perf_counter()
            O.ref_propagate(g, vpad, oxw, nxw, 0, 1, [sz], [sx], wav[None, :ns], 1, rz, rx, ns, 1, ns // snapi)
            O.ref_propagate(g, vpad, oxw, nxw, 0, 2, rz, rx, data, -1, [], [], ns, -1, 0)
            return time.perf_counter() - t0
        t = run(8)
        ns = int(max(16, min(2000, 8 * budget_s / max(t, 1e-3))) // snapi * snapi)
        t = run(ns)
        return {"value": round(2 * ns * pts_step / t / 1e9, 4), "unit": "Gpts/s", "cores": cores,
                "cpu": cpu_model(), "kind": "reference",
                "sample": "reference Fortran kernels (flang -O2 -fopenmp, %d threads): shot 48 of the 96 "
                          "(%dx%d FD region), %d forward steps with snapshots + %d reverse steps, %.1f s"
                          % (cores, nz + 2 * w["ntaper"], nxw + 2 * w["ntaper"], ns, ns, t)}
    ns = 200
    sub = O.Shot(sz, sx, wav[:ns], rz, rx)
    t0 = time.perf_counter()
    O.modeling(g, w["mig"], sub, extract_kind=1)
    t = time.perf_counter() - t0
    return {"value": round(ns * pts_step / t / 1e9, 4), "unit": "Gpts/s", "cores": cores, "cpu": cpu_model(),
            "kind": "port", "sample": "C oracle (OpenMP over columns), shot 48, %d forward steps, %.1f s" % (ns, t)}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--nt-max", type=int, default=0,
                    help="cap every shot's record (short runs for counter passes; not the benchmark)")
    ap.add_argument("--only-batched", action="store_true", help="the headline only (profiling)")
    ap.add_argument("--graphs", type=int, default=1,
                    help="0: launch every step eagerly instead of replaying HIP graphs (profiler runs)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-fwi", action="store_true", help="skip the 201x601 FWI time-to-misfit leg")
    ap.add_argument("--no-rtm", action="store_true", help="skip the single-shot 751x2301 RTM leg")
    ap.add_argument("--no-salt", action="store_true", help="skip the 1911x5395 salt gradient leg")
    ap.add_argument("--no-density", action="store_true", help="skip the 2D variable-density leg")
    ap.add_argument("--no-3d", action="store_true", help="skip the 3D forward leg")
    ap.add_argument("--leg", choices=["rtm", "fwi", "salt", "density", "3d"],
                    help="run only this leg and print its JSON (profiling)")
    ap.add_argument("--cpu-budget", type=float, default=15.0)
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
    if args.leg:
        legs = {"rtm": rtm_leg, "fwi": fwi_leg, "salt": salt_leg, "density": density_leg, "3d": forward3d_leg}
        print(json.dumps({args.leg: legs[args.leg](sa)}), flush=True)
        return
    b = batched_headline(sa, args)
    value = b["pts"] / b["wall"] / 1e9
    tr = None
    if os.path.exists(TRAFFIC_FILE):
        with open(TRAFFIC_FILE) as f:
            tr = json.load(f)
    roof = None
    if b["fwd"]:
        # HBM traffic per launch: the calibrated FETCH_SIZE + WRITE_SIZE counters
        # of the forward-pass launches (measured with every shot running, records
        # capped at 400 steps) over their algorithmic bytes, applied to this
        # run's algorithmic bytes per launch
        traffic = round(tr["ratio_fwd"] * b["fwd"]["algorithmic_bytes_per_launch"]) if tr else None
        roof = {"bound": "hbm", "kernel": b["fwd"]["kernel"], "achieved": b["fwd"]["achieved"],
                "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": b["fwd"]["frac"], "traffic": traffic,
                "traffic_measured": None if not tr else {
                    "bytes_per_launch": tr["bytes_per_launch_fwd"],
                    "algorithmic_bytes_per_launch": tr["algorithmic_bytes_per_launch_fwd"],
                    "ratio": tr["ratio_fwd"], "source": "profiles/fd_step_batched96_traffic.json"},
                "algorithmic_bytes_per_launch": b["fwd"]["algorithmic_bytes_per_launch"],
                "avg_launch_us": b["fwd"]["avg_launch_us"], "launches": b["fwd"]["launches"],
                "reverse_time_pass": b["bwd"]}
    if roof:
        # the streaming ceiling at this size: a plain 5-read / 1-write kernel
        # over the step kernel's algorithmic bytes per launch
        import ctypes
        npts = int(b["fwd"]["algorithmic_bytes_per_launch"] // 24)
        ms = ctypes.c_double()
        rc = sa.abi.load().sfd_stream_calibration(npts, 20, ctypes.byref(ms))
        if rc == 0 and ms.value > 0:
            ceil = 24.0 * npts / (ms.value * 1e-3) / 1e9
            roof["stream_ceiling"] = {"kernel": "stream5_calib_kernel (5 FP32 reads, 1 write per point)",
                                      "points": npts, "avg_launch_us": round(ms.value * 1e3, 3),
                                      "achieved": round(ceil, 1), "frac": round(ceil / HBM_PEAK_GBS, 4),
                                      "step_kernel_frac_of_ceiling": round(b["fwd"]["achieved"] / ceil, 4)}
        log("stream ceiling:", roof.get("stream_ceiling"))
    w = b["w"]
    shots = b["shots"]
    extra = {}
    if not args.only_batched:
        if not args.no_rtm:
            log("single-shot RTM leg")
            extra["rtm751x2301"] = rtm_leg(sa)
        if not args.no_fwi:
            log("fwi time-to-misfit leg")
            extra["fwi"] = fwi_leg(sa)
        if not args.no_salt:
            log("salt gradient leg")
            extra["salt"] = salt_leg(sa)
        if not args.no_density:
            log("variable-density leg")
            extra["density751x2301"] = density_leg(sa)
        if not args.no_3d:
            log("3D forward leg")
            extra["forward3d"] = forward3d_leg(sa)
    cpu = None
    if not args.no_cpu and not args.only_batched:
        log("cpu baseline")
        cpu = cpu_baseline(w, budget_s=args.cpu_budget)
    steps_total = sum(s.nt for s in shots)
    out = {"metric": METRIC, "value": round(value, 3), "unit": "Gpts/s", "n_gpus": 1, "steps": args.steps,
           "warmup": args.warmup, "ms_per_step": round(b["wall"] / max(args.steps, 1) * 1000, 2),
           "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
           "data": "synthetic",
           "config": {"workload": "fwi-gradient-96-ragged-shots-751x2301-fp16snap", "grid": list(w["grid"][:2]),
                      "dx_m": w["grid"][2], "dt_s": w["grid"][4], "ntaper": w["ntaper"], "snapi": w["snapi"],
                      "shots": len(shots), "receivers": [int(min(s.ntr for s in shots)), int(max(s.ntr for s in shots))],
                      "nt": [int(min(s.nt for s in shots)), int(max(s.nt for s in shots))],
                      "nt_capped_at": args.nt_max or None, "shot_steps": steps_total,
                      "snapshots": "fp16, one max|S| scale per snapshot and 64x4 block",
                      "inputs": "observed data resident in HBM (sfd_set_traces) before the timed region",
                      "plan": "every timed call plans its batch afresh",
                      "outputs": "misfit, gradient and illumination to host; residual gathers stay on the device "
                                 "(as in sfd_fwi_lbfgs's evaluations)",
                      "host_inputs_Gpts_per_s": round(b["host_rate"], 3),
                      "pass_ms": [round(b["tm"]["fwd_ms"], 1), round(b["tm"]["bwd_ms"], 1)],
                      "setup_ms": round(b["tm"]["setup_ms"], 1),
                      "observed_data_Gpts_per_s": round(b["obs_rate"], 3),
                      "tile": list(b["tile"]["tile"]), "strip": b["tile"]["strip"],
                      "time_block": b["tile"]["time_block"], "autotune": b["tile"]["log"],
                      "misfit": b["misfit"], "device": info["name"]},
           "roofline": roof, "cpu_baseline": cpu}
    out.update(extra)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()

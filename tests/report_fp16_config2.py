"""Image error of the two snapshot storage forms on the 751 x 2301 single-shot
RTM (the bench workload) against the FP64 oracle.

  python tests/report_fp16_config2.py gpu OUT.npz   # on the GPU: images with FP32 and FP16 snapshots
  python tests/report_fp16_config2.py cpu OUT.npz   # on the CPU: FP64 oracle image, then the errors

The observed data are the FP32 forward modeling of the true model; the GPU
forward is bitwise equal to the FP32 oracle's, so the CPU side recomputes
them with the oracle instead of moving 69 MB of traces.  The FP64 oracle RTM
holds 1875 FP64 snapshots (26 GB of host memory) and runs single-threaded
(tens of minutes)."""
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "tests"))
import bench  # noqa: E402


def rel(a, b):
    return float(np.linalg.norm((np.asarray(a, np.float64) - b).ravel()) / np.linalg.norm(np.ravel(b)))


def gpu(out):
    import seismic_amd as sa
    true, mig, w, rz, rx = bench.build_inputs()
    e = sa.Engine(sa.make_grid(bench.NZ, bench.NX, bench.DZ, bench.DZ, bench.DT, ntaper=bench.NTAPER,
                               snapi=bench.SNAPI))
    shot = sa.Shot(0.0, bench.DZ * (bench.NX // 2), w, rz, rx)
    e.set_model(true)
    dobs = e.forward([shot])[0]
    e.set_model(mig)
    i32, l32 = e.rtm([shot], [dobs], snap_mode=sa.SFD_SNAP_FP32)
    i16, l16 = e.rtm([shot], [dobs], snap_mode=sa.SFD_SNAP_FP16S)
    np.savez(out, i32=i32, i16=i16, l32=l32, l16=l16, dobs_sum=float(np.sum(dobs.astype(np.float64))))
    print("gpu images written", out, flush=True)


def cpu(out):
    import oracle_lib as O
    z = np.load(out)
    true, mig, w, rz, rx = bench.build_inputs()
    g = O.make_grid(bench.NZ, bench.NX, bench.DZ, bench.DZ, bench.DT, ntaper=bench.NTAPER, snapi=bench.SNAPI)
    shot = O.Shot(0.0, bench.DZ * (bench.NX // 2), w, rz, rx)
    t0 = time.time()
    dobs = O.modeling(g, true, shot, extract_kind=0, prec="f")
    print("oracle FP32 observed data %.0f s (sum matches GPU: %s)" % (
        time.time() - t0, bool(float(np.sum(dobs.astype(np.float64))) == float(z["dobs_sum"]))), flush=True)
    t0 = time.time()
    o64 = O.rtm(g, mig, shot, dobs, prec="d")
    print("oracle FP64 RTM %.0f s" % (time.time() - t0), flush=True)
    e32, e16 = rel(z["i32"], o64["image"]), rel(z["i16"], o64["image"])
    l32, l16 = rel(z["l32"], o64["illum"]), rel(z["l16"], o64["illum"])
    print("image rel L2 vs FP64: FP32 snapshots %.3e, FP16 snapshots %.3e" % (e32, e16))
    print("illumination rel L2 vs FP64: FP32 snapshots %.3e, FP16 snapshots %.3e" % (l32, l16))


if __name__ == "__main__":
    {"gpu": gpu, "cpu": cpu}[sys.argv[1]](sys.argv[2])

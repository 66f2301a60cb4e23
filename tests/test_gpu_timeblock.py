"""Two steps per launch (SFD_OPT_TIME_BLOCK = 2, on its 56 x 32 tiling):
every result must be bitwise equal to the one-step kernel's on the same
tiling (same arithmetic per cell, same order, same FP16 scale blocks), which
is itself bitwise equal to the oracle."""
import numpy as np
import pytest

import oracle_lib as O
from models import ricker, layered_anomaly, smooth_random

sa = pytest.importorskip("seismic_amd")
import seismic_amd.abi  # noqa: E402,F401
pytestmark = pytest.mark.gpu


def engine(nz, nx, ntaper, vel, tb, dz=10.0, dt=1e-3, snapi=4):
    e = sa.Engine(sa.make_grid(nz, nx, dz, dz, dt, ntaper=ntaper, snapi=snapi))
    e.set_model(vel)
    e.set_tile(56, 32)
    e.set_option(sa.abi.SFD_OPT_TIME_BLOCK, tb)
    return e


def test_config1_forward_two_step_bitwise():
    print("case: two-step config1 forward", flush=True)
    nz, nx, nt = 201, 401, 2000
    vel = layered_anomaly(nz, nx, 10.0, 10.0, seed=1)
    w = ricker(15.0, nt, 1e-3)
    rx = np.arange(nx, dtype=np.float32) * 10.0
    rz = np.full_like(rx, 20.0)
    gpu = engine(nz, nx, 40, vel, 2).forward([sa.Shot(20.0, 2000.0, w, rz, rx)])[0]
    ref = O.modeling(O.make_grid(nz, nx, 10.0, 10.0, 1e-3, ntaper=40), vel, O.Shot(20.0, 2000.0, w, rz, rx))
    print("bitwise", np.array_equal(gpu, ref), flush=True)
    assert np.array_equal(gpu, ref)


@pytest.mark.parametrize("ext", [0, 1, 2, 3])
def test_two_step_extractions_bitwise(ext):
    print("case: two-step ext=%d" % ext, flush=True)
    nz, nx, nt = 47, 83, 401   # odd nt: the last step runs one-step
    vel = smooth_random(nz, nx, seed=ext)
    w = ricker(20.0, nt, 1e-3)
    rx = np.linspace(13.3, 10.0 * (nx - 1) - 7.7, 37).astype(np.float32)
    rz = np.full_like(rx, 31.7)
    shot = sa.Shot(44.4, 333.3, w, rz, rx)
    a = engine(nz, nx, 12, vel, 1).forward([shot], extract_kind=ext)[0]
    b = engine(nz, nx, 12, vel, 2).forward([shot], extract_kind=ext)[0]
    assert np.array_equal(a, b)


def test_two_step_ragged_batch_bitwise():
    print("case: two-step ragged batch", flush=True)
    nz, nx = 50, 160
    vel = smooth_random(nz, nx, seed=11)
    rng = np.random.default_rng(4)
    shots = []
    for k in range(5):
        nt = int(rng.integers(150, 400))
        sx = float(rng.uniform(200, 1400))
        ntr = int(rng.integers(10, 60))
        rx = np.clip(sx + (np.arange(ntr) - ntr / 2) * 10.0, 0, 10.0 * (nx - 1)).astype(np.float32)
        shots.append(sa.Shot(20.0, sx, ricker(20.0 + k, nt, 1e-3), np.full_like(rx, 20.0), rx))
    a = engine(nz, nx, 14, vel, 1).forward(shots)
    b = engine(nz, nx, 14, vel, 2).forward(shots)
    for x, y in zip(a, b):
        assert np.array_equal(x, y)


@pytest.mark.parametrize("mode", ["rtm", "lsrtm", "rtm16", "dense"])
def test_two_step_rtm_bitwise(mode):
    print("case: two-step %s" % mode, flush=True)
    nz, nx, nt = 48, 72, 322
    vel = smooth_random(nz, nx, seed=5)
    w = ricker(20.0, nt, 1e-3)
    if mode == "dense":
        rx = np.arange(20.0, 10.0 * (nx - 3), 5.0).astype(np.float32)
    else:
        rx = np.linspace(10.0, 10.0 * (nx - 2), 30).astype(np.float32)
    rz = np.full_like(rx, 12.0)
    data = (np.random.default_rng(5).standard_normal((len(rx), nt)) * 1e-2).astype(np.float32)
    shot = sa.Shot(18.0, 355.0, w, rz, rx)
    out = []
    for tb in (1, 2):
        e = engine(nz, nx, 12, vel, tb)
        if mode == "lsrtm":
            out.append(e.lsrtm([shot], [data]))
        else:
            snap = sa.SFD_SNAP_FP16S if mode == "rtm16" else sa.SFD_SNAP_FP32
            out.append(e.rtm([shot], [data], snap_mode=snap))
    assert np.array_equal(out[0][0], out[1][0]) and np.array_equal(out[0][1], out[1][1])


@pytest.mark.parametrize("snap", ["fp32", "fp16"])
def test_two_step_fwi_gradient_bitwise(snap):
    print("case: two-step fwi gradient %s" % snap, flush=True)
    nz, nx, nt = 40, 96, 300
    true = smooth_random(nz, nx, seed=21, vmin=1600, vmax=2600)
    init = smooth_random(nz, nx, seed=22, vmin=1600, vmax=2500)
    w = ricker(18.0, nt, 1e-3)
    shots = []
    for sx in [200.0, 470.0, 760.0]:
        rx = np.linspace(max(0.0, sx - 300), min(10.0 * (nx - 1), sx + 300), 41).astype(np.float32)
        shots.append(sa.Shot(12.0, sx, w, np.full_like(rx, 15.0), rx))
    e0 = engine(nz, nx, 12, true, 1)
    dobs = e0.forward(shots, extract_kind=sa.SFD_EXTRACT_NEAREST)
    mode = sa.SFD_SNAP_FP16S if snap == "fp16" else sa.SFD_SNAP_FP32
    res = []
    for tb in (1, 2):
        e = engine(nz, nx, 12, init, tb)
        res.append(e.fwi_gradient(shots, dobs, rect_z=2, rect_x=3, do_z=1, do_x=1, snap_mode=mode))
    assert res[0][0] == res[1][0]
    assert np.array_equal(res[0][1], res[1][1]) and np.array_equal(res[0][2], res[1][2])

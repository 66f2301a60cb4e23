"""GPU parity of variable density (withRho) against the oracle, which is
pinned to the reference's compiled kernels in test_oracle_density.py:
forward modeling (every extraction kind, with free surface and ghosts,
ragged batches) and RTM, bitwise; unsupported drivers refuse."""
import numpy as np
import pytest

import oracle_lib as O
from models import ricker, smooth_random

sa = pytest.importorskip("seismic_amd")
pytestmark = pytest.mark.gpu


def rel(a, b):
    return float(np.linalg.norm((np.asarray(a, np.float64) - b).ravel()) /
                 max(np.linalg.norm(np.asarray(b, np.float64).ravel()), 1e-300))


def say(*a):
    print(*a, flush=True)


def setup(nz, nx, types=(0, 0, 0), seed=1):
    r, s, f = types
    gg = sa.make_grid(nz, nx, 10.0, 10.0, 1e-3, ntaper=12, snapi=4, rec_type=r, shot_type=s, surf_type=f)
    og = O.make_grid(nz, nx, 10.0, 10.0, 1e-3, ntaper=12, snapi=4)
    og.rec_type, og.shot_type, og.surf_type = r, s, f
    vel = smooth_random(nz, nx, seed=seed)
    rho = smooth_random(nz, nx, seed=seed + 100, vmin=1.0, vmax=2.6)
    e = sa.Engine(gg)
    e.set_model(vel)
    e.set_density(rho)
    return e, og, vel, rho


def shots(sz, sx, w, rz, rx):
    return sa.Shot(sz, sx, w, rz, rx), O.Shot(sz, sx, w, rz, rx)


@pytest.mark.parametrize("ext", [0, 1, 2, 3])
@pytest.mark.parametrize("types", [(0, 0, 0), (1, 1, 1)])
def test_forward_with_density(ext, types):
    say("case: density forward ext=%d types=%s" % (ext, types))
    nz, nx, nt = 36, 48, 450
    e, og, vel, rho = setup(nz, nx, types)
    w = ricker(25.0, nt, 1e-3)
    rx = np.linspace(13.0, 10.0 * (nx - 1) - 17.0, 21).astype(np.float32)
    gs, os_ = shots(20.0 if types[2] else 37.0, 241.0, w, np.full_like(rx, 23.0), rx)
    gpu = e.forward([gs], extract_kind=ext)[0]
    O.set_density(rho)
    try:
        ref = O.modeling(og, vel, os_, extract_kind=ext)
    finally:
        O.set_density(None)
    say("rel=%.3e bitwise=%s" % (rel(gpu, ref), np.array_equal(gpu, ref)))
    assert np.array_equal(gpu, ref)


def test_ragged_batch_with_density():
    say("case: density ragged batch")
    nz, nx = 40, 120
    e, og, vel, rho = setup(nz, nx, seed=31)
    gs, os_ = [], []
    for sx, nt, n in [(300.0, 420, 31), (700.0, 250, 17), (950.0, 330, 23)]:
        rx = np.linspace(sx - 250, sx + 250, n).astype(np.float32)
        a, b = shots(30.0, sx, ricker(22.0, nt, 1e-3), np.full_like(rx, 23.0), rx)
        gs.append(a)
        os_.append(b)
    gpu = e.forward(gs, extract_kind=0)
    O.set_density(rho)
    try:
        for k in range(3):
            ref = O.modeling(og, vel, os_[k], extract_kind=0)
            assert np.array_equal(gpu[k], ref), (k, rel(gpu[k], ref))
    finally:
        O.set_density(None)


def test_rtm_with_density():
    say("case: density rtm")
    nz, nx, nt = 40, 60, 320
    e, og, vel, rho = setup(nz, nx, seed=7)
    w = ricker(25.0, nt, 1e-3)
    rx = np.linspace(13.0, 10.0 * (nx - 1) - 17.0, 25).astype(np.float32)
    gs, os_ = shots(20.0, 301.0, w, np.full_like(rx, 23.0), rx)
    O.set_density(rho)
    try:
        data = O.modeling(og, vel, os_, extract_kind=0)
        out = O.rtm(og, vel, os_, data)
    finally:
        O.set_density(None)
    img, ill = e.rtm([gs], [data])
    say("image rel=%.3e illum rel=%.3e" % (rel(img, out["image"]), rel(ill, out["illum"])))
    assert np.array_equal(img, out["image"]) and np.array_equal(ill, out["illum"])
    i16, _ = e.rtm([gs], [data], snap_mode=sa.SFD_SNAP_FP16S)
    assert rel(i16, img) < 1e-3
    e.set_density(None)            # back to constant density: the oracle's constant-density image
    img0, _ = e.rtm([gs], [data])
    assert np.array_equal(img0, O.rtm(og, vel, os_, data)["image"])


def test_density_refuses_other_drivers():
    say("case: density refusals")
    nz, nx, nt = 36, 48, 100
    e, og, vel, rho = setup(nz, nx)
    w = ricker(25.0, nt, 1e-3)
    rx = np.linspace(13.0, 10.0 * (nx - 1) - 17.0, 11).astype(np.float32)
    gs = sa.Shot(37.0, 241.0, w, np.full_like(rx, 23.0), rx)
    d = np.zeros((len(rx), nt), np.float32)
    for call in (lambda: e.lsrtm([gs], [d]), lambda: e.born([gs], np.zeros((nz, nx), np.float32)),
                 lambda: e.fwi_gradient([gs], [d])):
        with pytest.raises(sa.SeisFDError) as ei:
            call()
        assert ei.value.code == -4
    e.set_tile(56, 32)
    with pytest.raises(sa.SeisFDError) as ei:
        e.forward([gs])
    assert ei.value.code == -4

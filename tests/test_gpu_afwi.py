"""GPU parity of sfd_afwi_gradient (3DAFWI.x's gradient, AFWI_to_memory) with
the oracle, which test_oracle_afwi.py pins by finite differences: constant
density (FWI imaging after the adjoint-style stencil and sinc injection of
the residual) and variable density (the v-rho imaging condition into two
gradients, vprho 0 and 1), with free surface and ghosts, a ragged batch, the
non-L2 norms with a mute, bitwise in FP32; FP16 snapshots close to FP32."""
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


def setup(nz, nx, types=(0, 0, 0), seed=1, density=False, snapi=2):
    r, s, f = types
    gg = sa.make_grid(nz, nx, 10.0, 10.0, 1e-3, ntaper=12, snapi=snapi, rec_type=r, shot_type=s, surf_type=f)
    og = O.make_grid(nz, nx, 10.0, 10.0, 1e-3, ntaper=12, snapi=snapi)
    og.rec_type, og.shot_type, og.surf_type = r, s, f
    vel = smooth_random(nz, nx, seed=seed)
    rho = smooth_random(nz, nx, seed=seed + 100, vmin=1.0, vmax=2.6) if density else None
    e = sa.Engine(gg)
    e.set_model(vel)
    if density:
        e.set_density(rho)
    return e, og, vel, rho


def make_shots(nx, specs, w):
    gs, os_ = [], []
    for sz, sx, nt, n in specs:
        rx = np.linspace(max(13.0, sx - 200.0), min(10.0 * (nx - 1) - 17.0, sx + 200.0), n).astype(np.float32)
        rz = np.full_like(rx, 23.0)
        gs.append(sa.Shot(sz, sx, w[:nt], rz, rx))
        os_.append(O.Shot(sz, sx, w[:nt], rz, rx))
    return gs, os_


def observed(og, vel, rho, os_):
    vt = vel.copy()
    vt[18:26, :] *= 1.06
    O.set_density(None if rho is None else rho * 1.1)
    try:
        return [O.modeling(og, vt, s, extract_kind=3) for s in os_]
    finally:
        O.set_density(None)


def oracle_afwi(og, vel, rho, os_, dobs, **kw):
    O.set_density(rho)
    try:
        return O.afwi(og, vel, os_, dobs, **kw)
    finally:
        O.set_density(None)


@pytest.mark.parametrize("density,vprho", [(False, 0), (True, 0), (True, 1)])
def test_afwi_bitwise(density, vprho):
    say("case: afwi density=%s vprho=%d" % (density, vprho))
    nz, nx = 40, 72
    e, og, vel, rho = setup(nz, nx, density=density)
    w = ricker(22.0, 400, 1e-3)
    gs, os_ = make_shots(nx, [(30.0, 250.0, 400, 19), (30.0, 480.0, 400, 15)], w)
    dobs = observed(og, vel, rho, os_)
    got = e.afwi_gradient(gs, dobs, vprho_param=vprho)
    want = oracle_afwi(og, vel, rho, os_, dobs, vprho=vprho)
    npar = 2 if density else 1
    say("f %.9e / %.9e, grad rel %s" % (got["f"], want["f"], [rel(got["grad"][k], want["grad"][k]) for k in range(npar)]))
    assert got["f"] == want["f"]
    for k in range(npar):
        assert np.abs(want["grad"][k]).max() > 0
        assert np.array_equal(got["grad"][k], want["grad"][k]), k
    assert np.array_equal(got["illum"], want["illum"])


@pytest.mark.parametrize("types", [(1, 1, 1), (0, 0, -1)])
def test_afwi_density_with_surface(types):
    """The v-rho imaging starts at each column's surface row (Imaging.f90:
    455-458); ghosts in the sinc injections and extractions."""
    say("case: afwi density types=%s" % (types,))
    nz, nx = 40, 64
    e, og, vel, rho = setup(nz, nx, types=types, seed=5, density=True)
    elev = (10.0 + 12.0 * np.sin(np.arange(nx) / 9.0)).astype(np.float32)
    e.set_elevation(elev)
    O.set_elevation(elev)
    try:
        w = ricker(22.0, 360, 1e-3)
        gs, os_ = make_shots(nx, [(20.0, 300.0, 360, 17)], w)
        dobs = observed(og, vel, rho, os_)
        got = e.afwi_gradient(gs, dobs)
        want = oracle_afwi(og, vel, rho, os_, dobs)
    finally:
        O.set_elevation(None)
    for k in range(2):
        assert np.array_equal(got["grad"][k], want["grad"][k]), (k, rel(got["grad"][k], want["grad"][k]))
    assert np.array_equal(got["illum"], want["illum"])


def test_afwi_ragged_batch_norm_and_mute():
    say("case: afwi ragged, Huber, mute")
    nz, nx = 40, 120
    e, og, vel, rho = setup(nz, nx, seed=9, density=True)
    w = ricker(20.0, 420, 1e-3)
    gs, os_ = make_shots(nx, [(30.0, 300.0, 420, 21), (30.0, 700.0, 260, 13), (30.0, 950.0, 330, 17)], w)
    dobs = observed(og, vel, rho, os_)
    rng = np.random.default_rng(4)
    mute = [rng.uniform(0.3, 1.0, d.shape).astype(np.float32) for d in dobs]
    got = e.afwi_gradient(gs, dobs, mute=mute, norm=sa.SFD_NORM_HUBER, norm_thresh=0.0)
    want = oracle_afwi(og, vel, rho, os_, dobs, mute=mute, norm=12, norm_thresh=0.0)
    assert got["f"] == want["f"]
    for k in range(2):
        assert np.array_equal(got["grad"][k], want["grad"][k]), (k, rel(got["grad"][k], want["grad"][k]))


def test_afwi_fp16_snapshots_and_resident_data():
    say("case: afwi fp16")
    nz, nx = 40, 72
    e, og, vel, rho = setup(nz, nx, seed=3, density=True)
    w = ricker(22.0, 400, 1e-3)
    gs, os_ = make_shots(nx, [(30.0, 350.0, 400, 19)], w)
    dobs = observed(og, vel, rho, os_)
    ref = e.afwi_gradient(gs, dobs)
    e.set_traces(dobs, gs)
    res = e.afwi_gradient(gs, None)
    assert np.array_equal(res["grad"], ref["grad"])
    for mode in (sa.SFD_SNAP_FP16S, sa.SFD_SNAP_FP16G):
        h = e.afwi_gradient(gs, dobs, snap_mode=mode)
        errs = [rel(h["grad"][k], ref["grad"][k]) for k in range(2)]
        say("fp16 mode %d rel %s" % (mode, errs))
        assert max(errs) < 1e-2 and h["f"] == ref["f"]


def test_afwi_errors():
    nz, nx = 36, 48
    e, og, vel, rho = setup(nz, nx, density=True)
    w = ricker(25.0, 100, 1e-3)
    rx = np.array([241.0], np.float32)
    one = sa.Shot(37.0, 241.0, w, np.full_like(rx, 23.0), rx)   # a one-column window
    d = np.zeros((1, 100), np.float32)
    with pytest.raises(sa.SeisFDError) as ei:
        e.afwi_gradient([one], [d])
    assert ei.value.code == -1
    rx = np.linspace(100.0, 300.0, 5).astype(np.float32)
    gs = sa.Shot(37.0, 241.0, w, np.full_like(rx, 23.0), rx)
    with pytest.raises(sa.SeisFDError) as ei:
        e.afwi_gradient([gs], [np.zeros((5, 100), np.float32)], vprho_param=2)
    assert ei.value.code == -1

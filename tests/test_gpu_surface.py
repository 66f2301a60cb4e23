"""GPU parity of the free surface and the ghost sources/receivers
(Boundary_set_free_surface_grid, Injection/Extraction ghost branches,
Boundary_3d's surf_type rules) against the oracle, which is pinned to the
reference's compiled kernels in test_oracle_surface.py.  Through the C-ABI;
bar: bitwise (relative L2 <= 1e-5 asserted, bitwise reported and asserted
where the arithmetic is the oracle's)."""
import numpy as np
import pytest

import oracle_lib as O
from models import ricker, smooth_random

sa = pytest.importorskip("seismic_amd")
pytestmark = pytest.mark.gpu

TOL = 1e-5
TYPES = [(0, 0, 1), (0, 0, -1), (0, 1, 0), (1, 0, 0), (1, 1, 0), (1, 0, 1), (0, 1, 1), (1, 1, 1)]


def rel(a, b):
    return float(np.linalg.norm((np.asarray(a, np.float64) - b).ravel()) /
                 max(np.linalg.norm(np.asarray(b, np.float64).ravel()), 1e-300))


def say(*a):
    print(*a, flush=True)


def grids(nz, nx, types, ntaper=12, snapi=4, oz=0.0):
    r, s, f = types
    gg = sa.make_grid(nz, nx, 10.0, 10.0, 1e-3, ntaper=ntaper, snapi=snapi, oz=oz, rec_type=r, shot_type=s,
                      surf_type=f)
    og = O.make_grid(nz, nx, 10.0, 10.0, 1e-3, ntaper=ntaper, snapi=snapi, oz=oz)
    og.rec_type, og.shot_type, og.surf_type = r, s, f
    return gg, og


def shots(sz, sx, w, rz, rx):
    return sa.Shot(sz, sx, w, rz, rx), O.Shot(sz, sx, w, rz, rx)


def _spread(nx, rz, n=21):
    rx = np.linspace(13.0, 10.0 * (nx - 1) - 17.0, n).astype(np.float32)
    return np.full_like(rx, rz), rx


@pytest.mark.parametrize("types", TYPES)
@pytest.mark.parametrize("ext", [0, 1, 2, 3])
def test_forward_with_surface_and_ghosts(types, ext):
    say("case: forward types=%s ext=%d" % (types, ext))
    nz, nx, nt = 36, 48, 450   # two graph blocks and a tail
    vel = smooth_random(nz, nx, seed=1)
    w = ricker(25.0, nt, 1e-3)
    rz, rx = _spread(nx, 23.0)
    sz = 20.0 if types[1] and types[2] > 0 else 37.0
    gs, os_ = shots(sz, 241.0, w, rz, rx)
    gg, og = grids(nz, nx, types)
    O.set_elevation(None)
    e = sa.Engine(gg)
    e.set_model(vel)
    gpu = e.forward([gs], extract_kind=ext)[0]
    ref = O.modeling(og, vel, os_, extract_kind=ext)
    say("rel=%.3e bitwise=%s" % (rel(gpu, ref), np.array_equal(gpu, ref)))
    assert np.array_equal(gpu, ref)


def _elevation(kind, nx):
    if kind == "offgrid":
        return np.full(nx, 3.0, np.float32)
    if kind == "below_top":
        return np.full(nx, 20.0, np.float32)
    x = np.arange(nx, dtype=np.float32)
    return (12.0 + 9.0 * np.sin(x / 5.0)).astype(np.float32)


@pytest.mark.parametrize("kind", ["offgrid", "below_top", "varying"])
def test_free_surface_elevations_forward_and_rtm(kind):
    """Off-grid surface (sinc mirror), a surface below model row 1 (zero
    snapshot rows: RTM image and illumination), and a per-column elevation."""
    say("case: elevation %s" % kind)
    nz, nx, nt = 40, 56, 320
    vel = smooth_random(nz, nx, seed=3)
    w = ricker(25.0, nt, 1e-3)
    rz, rx = _spread(nx, 53.0)
    gs, os_ = shots(57.0, 281.0, w, rz, rx)
    gg, og = grids(nz, nx, (0, 0, 1))
    e_ = _elevation(kind, nx)
    e = sa.Engine(gg)
    e.set_model(vel)
    e.set_elevation(e_)
    O.set_elevation(e_)
    try:
        gpu = e.forward([gs], extract_kind=0)[0]
        ref = O.modeling(og, vel, os_, extract_kind=0)
        data = ref.copy()
        img, ill = e.rtm([gs], [data])
        out = O.rtm(og, vel, os_, data)
    finally:
        O.set_elevation(None)
    say("gather rel=%.3e image rel=%.3e illum rel=%.3e" % (rel(gpu, ref), rel(img, out["image"]),
                                                          rel(ill, out["illum"])))
    assert np.array_equal(gpu, ref)
    assert np.array_equal(img, out["image"]) and np.array_equal(ill, out["illum"])
    if kind == "below_top":
        assert not np.any(ill[:2, :])


@pytest.mark.parametrize("types", [(0, 1, 1), (1, 1, 1), (0, 1, 0), (1, 0, -1)])
@pytest.mark.parametrize("lsrtm", [False, True])
def test_rtm_with_surface_and_ghosts(types, lsrtm):
    say("case: %s types=%s" % ("lsrtm" if lsrtm else "rtm", types))
    nz, nx, nt = 40, 60, 320
    vel = smooth_random(nz, nx, seed=7)
    w = ricker(25.0, nt, 1e-3)
    rz, rx = _spread(nx, 23.0, n=25)
    gs, os_ = shots(20.0, 301.0, w, rz, rx)
    gg, og = grids(nz, nx, types)
    O.set_elevation(None)
    data = O.modeling(og, vel, os_, extract_kind=0)
    e = sa.Engine(gg)
    e.set_model(vel)
    img, ill = (e.lsrtm if lsrtm else e.rtm)([gs], [data])
    out = O.rtm(og, vel, os_, data, lsrtm=lsrtm)
    say("image rel=%.3e illum rel=%.3e" % (rel(img, out["image"]), rel(ill, out["illum"])))
    assert np.array_equal(img, out["image"]) and np.array_equal(ill, out["illum"])


@pytest.mark.parametrize("types", [(1, 1, 1), (0, 0, 1)])
def test_born_with_surface(types):
    say("case: born types=%s" % (types,))
    nz, nx, nt = 44, 70, 300
    vel = smooth_random(nz, nx, seed=13)
    refl = np.zeros((nz, nx), np.float32)
    refl[25, 10:60] = 0.1
    refl[2, 20:40] = -0.05
    w = ricker(20.0, nt, 1e-3)
    rz, rx = _spread(nx, 12.0, n=25)
    gs, os_ = shots(18.0, 345.0, w, rz, rx)
    gg, og = grids(nz, nx, types)
    O.set_elevation(None)
    e = sa.Engine(gg)
    e.set_model(vel)
    gpu = e.born([gs], np.asfortranarray(refl))[0]
    ref = O.born(og, vel, np.asfortranarray(refl), os_)
    say("born rel=%.3e bitwise=%s" % (rel(gpu, ref), np.array_equal(gpu, ref)))
    assert np.array_equal(gpu, ref)


@pytest.mark.parametrize("types", [(1, 0, 1), (1, 1, 1), (1, 0, 0)])
def test_fwi_gradient_with_surface_and_ghosts(types):
    """AMOD with the sinc ghost source and nearest ghost receivers, AGRAD with
    the simple receiver ghost (rec_type), the surface on both passes; a
    ragged 3-shot batch."""
    say("case: fwi gradient types=%s" % (types,))
    nz, nx = 40, 96
    true = smooth_random(nz, nx, seed=21, vmin=1600, vmax=2600)
    init = smooth_random(nz, nx, seed=22, vmin=1600, vmax=2500)
    gg, og = grids(nz, nx, types)
    O.set_elevation(None)
    gsh, osh, dobs = [], [], []
    for k, (sx, nt) in enumerate([(200.0, 300), (470.0, 260), (760.0, 280)]):
        w = ricker(18.0, nt, 1e-3)
        rx = np.linspace(max(0.0, sx - 300), min(10.0 * (nx - 1), sx + 300), 41).astype(np.float32)
        rz = np.full_like(rx, 15.0)
        a, b = shots(12.0, sx, w, rz, rx)
        gsh.append(a)
        osh.append(b)
        dobs.append(O.modeling(og, true, b, extract_kind=1))
    kw = dict(illupow=1.0, rect_z=3, rect_x=4, do_z=1, do_x=1)
    ref = O.fwi_gradient(og, init, osh, dobs, **kw)
    e = sa.Engine(gg)
    e.set_model(init)
    f, grad, illu, res = e.fwi_gradient(gsh, dobs, **kw)
    say("misfit %.9e vs %.9e grad rel=%.3e" % (f, ref["f"], rel(grad, ref["grad"])))
    assert f == ref["f"]
    assert np.array_equal(grad, ref["grad"]) and np.array_equal(illu, ref["illum"])


def test_ragged_batch_with_surface():
    """Batched forward of three ragged shots with a free surface and ghosts:
    the surface kernel covers every running shot, the extraction kernel the
    running tiles."""
    say("case: ragged batch surface")
    nz, nx = 40, 120
    vel = smooth_random(nz, nx, seed=31)
    gg, og = grids(nz, nx, (1, 1, 1))
    O.set_elevation(None)
    e = sa.Engine(gg)
    e.set_model(vel)
    gs, os_ = [], []
    for sx, nt, n in [(300.0, 420, 31), (700.0, 250, 17), (950.0, 330, 23)]:
        w = ricker(22.0, nt, 1e-3)
        rx = np.linspace(sx - 250, sx + 250, n).astype(np.float32)
        a, b = shots(20.0, sx, w, np.full_like(rx, 23.0), rx)
        gs.append(a)
        os_.append(b)
    gpu = e.forward(gs, extract_kind=0)
    for k in range(3):
        ref = O.modeling(og, vel, os_[k], extract_kind=0)
        assert np.array_equal(gpu[k], ref), (k, rel(gpu[k], ref))


def test_fp16_snapshots_with_surface_below_top():
    """FP16 snapshots (per 64x4 block and per snapshot) with rows above the
    surface zeroed: the image within 1e-3 of the FP32-snapshot image."""
    say("case: fp16 surface")
    nz, nx, nt = 40, 56, 320
    vel = smooth_random(nz, nx, seed=3)
    w = ricker(25.0, nt, 1e-3)
    rz, rx = _spread(nx, 53.0)
    gs, os_ = shots(57.0, 281.0, w, rz, rx)
    gg, og = grids(nz, nx, (0, 0, 1))
    e_ = _elevation("below_top", nx)
    O.set_elevation(e_)
    try:
        data = O.modeling(og, vel, os_, extract_kind=0)
    finally:
        O.set_elevation(None)
    e = sa.Engine(gg)
    e.set_model(vel)
    e.set_elevation(e_)
    i32, _ = e.rtm([gs], [data])
    for mode in (sa.SFD_SNAP_FP16S, sa.SFD_SNAP_FP16G):
        i16, l16 = e.rtm([gs], [data], snap_mode=mode)
        say("mode %d rel=%.3e" % (mode, rel(i16, i32)))
        assert rel(i16, i32) < 1e-3
        assert not np.any(l16[:2, :])


def test_two_step_and_exact_adjoint_with_surface():
    """A free surface needs the one-step kernel: asking for two steps per
    launch still gives the oracle's gather; the exact adjoints refuse."""
    say("case: surface two-step / exact adjoint")
    nz, nx, nt = 36, 48, 240
    vel = smooth_random(nz, nx, seed=1)
    w = ricker(25.0, nt, 1e-3)
    rz, rx = _spread(nx, 23.0)
    gs, os_ = shots(37.0, 241.0, w, rz, rx)
    gg, og = grids(nz, nx, (0, 0, 1))
    O.set_elevation(None)
    e = sa.Engine(gg)
    e.set_model(vel)
    e.set_option(sa.abi.SFD_OPT_TIME_BLOCK, 2)
    gpu = e.forward([gs], extract_kind=0)[0]
    assert np.array_equal(gpu, O.modeling(og, vel, os_, extract_kind=0))
    with pytest.raises(sa.SeisFDError) as ei:
        e.adjoint_exact([gs], [gpu])
    assert ei.value.code == -4

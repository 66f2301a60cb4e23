"""GPU parity: the HIP engine (through the C-ABI) against the CPU oracle on
the same seeded inputs.  Bar: relative L2 <= 1e-5 in FP32; the engine is
built to reproduce the reference's FP32 rounding, so most cases are also
reported (and asserted where stated) bitwise."""
import numpy as np
import pytest

import oracle_lib as O
from models import ricker, layered_anomaly, smooth_random

sa = pytest.importorskip("seismic_amd")
import seismic_amd.abi  # noqa: E402,F401
pytestmark = pytest.mark.gpu

TOL = 1e-5


def rel(a, b):
    return float(np.linalg.norm((np.asarray(a, np.float64) - b).ravel()) /
                 max(np.linalg.norm(np.asarray(b, np.float64).ravel()), 1e-300))


def say(*a):
    print(*a, flush=True)


def engine(nz, nx, dz, dx, dt, ntaper, snapi=4, vel=None, vel_bwd=None):
    e = sa.Engine(sa.make_grid(nz, nx, dz, dx, dt, ntaper=ntaper, snapi=snapi))
    if vel is not None:
        e.set_model(vel, vel_bwd)
    return e


def both_shots(sz, sx, w, rz, rx):
    return sa.Shot(sz, sx, w, rz, rx), O.Shot(sz, sx, w, rz, rx)


def test_config1_forward_gather():
    """Config 1: 201x401 at 10 m, three layers + Gaussian anomaly, 15 Hz Ricker
    at 20 m depth mid-x, 401 receivers at 20 m, dt=1 ms, nt=2000, 40-pt pad."""
    say("case: config1 forward 201x401 nt=2000")
    nz, nx, nt = 201, 401, 2000
    vel = layered_anomaly(nz, nx, 10.0, 10.0, seed=1)
    w = ricker(15.0, nt, 1e-3)
    rx = np.arange(nx, dtype=np.float32) * 10.0
    rz = np.full_like(rx, 20.0)
    gs, os_ = both_shots(20.0, 2000.0, w, rz, rx)
    e = engine(nz, nx, 10.0, 10.0, 1e-3, 40, vel=vel)
    gpu = e.forward([gs])[0]
    ref = O.modeling(O.make_grid(nz, nx, 10.0, 10.0, 1e-3, ntaper=40), vel, os_, extract_kind=0)
    err = rel(gpu, ref)
    say("config1 gather rel L2 = %.3e bitwise=%s" % (err, np.array_equal(gpu, ref)))
    assert err <= TOL
    assert np.array_equal(gpu, ref)


@pytest.mark.parametrize("ext", [0, 1, 2, 3])
def test_forward_offgrid_all_extractions(ext):
    say("case: forward off-grid ext=%d" % ext)
    nz, nx, nt = 47, 83, 400
    vel = smooth_random(nz, nx, seed=ext)
    w = ricker(20.0, nt, 1e-3)
    rx = np.linspace(13.3, 10.0 * (nx - 1) - 7.7, 37).astype(np.float32)
    rz = np.full_like(rx, 31.7)
    gs, os_ = both_shots(44.4, 333.3, w, rz, rx)
    e = engine(nz, nx, 10.0, 10.0, 1e-3, 12, vel=vel)
    gpu = e.forward([gs], extract_kind=ext)[0]
    ref = O.modeling(O.make_grid(nz, nx, 10.0, 10.0, 1e-3, ntaper=12), vel, os_, extract_kind=ext)
    err = rel(gpu, ref)
    say("ext=%d rel L2 = %.3e bitwise=%s" % (ext, err, np.array_equal(gpu, ref)))
    assert err <= TOL


@pytest.mark.parametrize("tile", [(64, 32, 0), (64, 32, 4), (64, 64, 0)])
def test_tile_shapes_bitwise(tile):
    """Every compiled tile shape gives the same gather (grid sizes chosen so
    tile-origin shifts and partial tiles occur)."""
    say("case: tile shape %s" % (tile,))
    nz, nx, nt = 61, 129, 250
    vel = smooth_random(nz, nx, seed=7)
    w = ricker(25.0, nt, 1e-3)
    rx = np.linspace(5.0, 10.0 * (nx - 1) - 5.0, 50).astype(np.float32)
    rz = np.full_like(rx, 15.0)
    gs, os_ = both_shots(25.0, 640.0, w, rz, rx)
    e = engine(nz, nx, 10.0, 10.0, 1e-3, 15, vel=vel)
    e.set_tile(*tile[:2])
    e.set_option(sa.abi.SFD_OPT_STRIP_PX, tile[2])
    gpu = e.forward([gs])[0]
    ref = O.modeling(O.make_grid(nz, nx, 10.0, 10.0, 1e-3, ntaper=15), vel, os_, extract_kind=0)
    say("tile %s rel L2 = %.3e bitwise=%s" % (tile, rel(gpu, ref), np.array_equal(gpu, ref)))
    assert rel(gpu, ref) <= TOL


def test_autotuner_picks_a_compiled_shape():
    """sfd_set_tile(h, 0, 0): the next call times every compiled step-kernel
    shape on its batch and keeps the fastest; the gather is unchanged."""
    say("case: autotuner")
    nz, nx, nt = 61, 129, 250
    vel = smooth_random(nz, nx, seed=7)
    w = ricker(25.0, nt, 1e-3)
    rx = np.linspace(5.0, 10.0 * (nx - 1) - 5.0, 50).astype(np.float32)
    rz = np.full_like(rx, 15.0)
    gs, os_ = both_shots(25.0, 640.0, w, rz, rx)
    e = engine(nz, nx, 10.0, 10.0, 1e-3, 15, vel=vel)
    e.set_tile(0, 0)
    gpu = e.forward([gs])[0]
    t = e.get_tile()
    say("autotune: %s -> %s strip %d" % (t["log"], t["tile"], t["strip"]))
    assert t["log"].count("us") == 4 and t["tile"] in [(64, 32), (64, 64)]
    ref = O.modeling(O.make_grid(nz, nx, 10.0, 10.0, 1e-3, ntaper=15), vel, os_, extract_kind=0)
    assert np.array_equal(gpu, ref)


def test_batched_ragged_shots():
    """Several shots with uneven receiver spreads and record lengths in one call."""
    say("case: batched ragged 5 shots")
    nz, nx = 50, 160
    vel = smooth_random(nz, nx, seed=11)
    rng = np.random.default_rng(4)
    og = O.make_grid(nz, nx, 10.0, 10.0, 1e-3, ntaper=14)
    e = engine(nz, nx, 10.0, 10.0, 1e-3, 14, vel=vel)
    gshots, oshots = [], []
    for k in range(5):
        nt = int(rng.integers(150, 400))
        sx = float(rng.uniform(200, 1400))
        ntr = int(rng.integers(10, 60))
        rx = np.clip(sx + (np.arange(ntr) - ntr / 2) * 10.0, 0, 10.0 * (nx - 1)).astype(np.float32)
        rz = np.full_like(rx, 20.0)
        w = ricker(20.0 + k, nt, 1e-3)
        a, b = both_shots(20.0, sx, w, rz, rx)
        gshots.append(a)
        oshots.append(b)
    outs = e.forward(gshots)
    for k, (o, s) in enumerate(zip(outs, oshots)):
        ref = O.modeling(og, vel, s, extract_kind=0)
        say("shot %d nt=%d ntr=%d rel=%.3e" % (k, s.nt, s.ntr, rel(o, ref)))
        assert rel(o, ref) <= TOL


def _rtm_case(seed=5, nz=48, nx=72, nt=320, snapi=4):
    vel = smooth_random(nz, nx, seed=seed)
    w = ricker(20.0, nt, 1e-3)
    rx = np.linspace(10.0, 10.0 * (nx - 2), 30).astype(np.float32)
    rz = np.full_like(rx, 12.0)
    rng = np.random.default_rng(seed)
    data = (rng.standard_normal((len(rx), nt)) * 1e-2).astype(np.float32)
    return vel, w, rx, rz, data


@pytest.mark.parametrize("lsrtm", [False, True])
def test_rtm_image_matches_oracle(lsrtm):
    say("case: %s small" % ("lsrtm" if lsrtm else "rtm"))
    nz, nx, nt = 48, 72, 320
    vel, w, rx, rz, data = _rtm_case(nz=nz, nx=nx, nt=nt)
    gs, os_ = both_shots(18.0, 355.0, w, rz, rx)
    e = engine(nz, nx, 10.0, 10.0, 1e-3, 12, vel=vel)
    if lsrtm:
        img, ill = e.lsrtm([gs], [data])
    else:
        img, ill = e.rtm([gs], [data])
    out = O.rtm(O.make_grid(nz, nx, 10.0, 10.0, 1e-3, ntaper=12), vel, os_, data, lsrtm=lsrtm)
    ei, el = rel(img, out["image"]), rel(ill, out["illum"])
    say("image rel=%.3e illum rel=%.3e bitwise=%s/%s" % (ei, el, np.array_equal(img, out["image"]),
                                                         np.array_equal(ill, out["illum"])))
    assert ei <= TOL and el <= TOL


def test_rtm_resident_traces():
    """sfd_set_traces: RTM and LSRTM reading the observed data from HBM give
    the images of the host-traces calls bitwise; a resident copy whose length
    does not match the shots is rejected."""
    say("case: rtm resident traces")
    nz, nx, nt = 48, 72, 320
    vel, w, rx, rz, data = _rtm_case(nz=nz, nx=nx, nt=nt)
    gs, _ = both_shots(18.0, 355.0, w, rz, rx)
    e = engine(nz, nx, 10.0, 10.0, 1e-3, 12, vel=vel)
    img_h, ill_h = e.rtm([gs], [data])
    ls_h, _ = e.lsrtm([gs], [data])
    e.set_traces([data], [gs])
    img_r, ill_r = e.rtm([gs], None)
    ls_r, _ = e.lsrtm([gs], None)
    assert np.array_equal(img_h, img_r) and np.array_equal(ill_h, ill_r) and np.array_equal(ls_h, ls_r)
    short = sa.Shot(gs.sz, gs.sx, w[: nt // 2], rz, rx)
    with pytest.raises(sa.SeisFDError):
        e.rtm([short], None)
    e.set_traces(None, None)
    with pytest.raises(sa.SeisFDError):
        e.rtm([gs], None)


def test_rtm_dense_receivers():
    """Receivers every half cell: a tile holds more injection contributions
    (about 3000) than its LDS stage (2048), so the reverse-time injection
    stages them in two chunks; the image must not change."""
    say("case: rtm dense receivers")
    nz, nx, nt = 40, 72, 200
    vel, w, _, _, _ = _rtm_case(nz=nz, nx=nx, nt=nt, seed=5)
    rx = np.arange(20.0, 10.0 * (nx - 3), 5.0).astype(np.float32)
    rz = np.full_like(rx, 14.0)
    data = (np.random.default_rng(5).standard_normal((len(rx), nt)) * 1e-2).astype(np.float32)
    gs, os_ = both_shots(18.0, 300.0, w, rz, rx)
    e = engine(nz, nx, 10.0, 10.0, 1e-3, 12, vel=vel)
    img, ill = e.rtm([gs], [data])
    out = O.rtm(O.make_grid(nz, nx, 10.0, 10.0, 1e-3, ntaper=12), vel, os_, data)
    ei, el = rel(img, out["image"]), rel(ill, out["illum"])
    say("%d receivers: image rel=%.3e illum rel=%.3e bitwise=%s" % (len(rx), ei, el,
                                                                    np.array_equal(img, out["image"])))
    assert ei <= TOL and el <= TOL


def test_rtm_vel_bwd():
    """Optional second velocity for the backward pass (3DRTM.f90 vel2)."""
    say("case: rtm with vel2")
    nz, nx, nt = 40, 64, 240
    vel, w, rx, rz, data = _rtm_case(nz=nz, nx=nx, nt=nt, seed=8)
    vel2 = (vel * 1.05).astype(np.float32)
    gs, os_ = both_shots(18.0, 300.0, w, rz, rx)
    e = engine(nz, nx, 10.0, 10.0, 1e-3, 12, vel=vel, vel_bwd=vel2)
    img, ill = e.rtm([gs], [data])
    out = O.rtm(O.make_grid(nz, nx, 10.0, 10.0, 1e-3, ntaper=12), vel, os_, data, vel_bwd=vel2)
    assert rel(img, out["image"]) <= TOL and rel(ill, out["illum"]) <= TOL


def test_rtm_fp16_snapshots_vs_fp64():
    """FP16 max-scaled snapshots: image error against the FP64 oracle, next to
    the FP32-snapshot error."""
    say("case: rtm fp16 snapshots")
    nz, nx, nt = 48, 72, 320
    vel, w, rx, rz, data = _rtm_case(nz=nz, nx=nx, nt=nt, seed=9)
    gs, os_ = both_shots(18.0, 355.0, w, rz, rx)
    e = engine(nz, nx, 10.0, 10.0, 1e-3, 12, vel=vel)
    i32, _ = e.rtm([gs], [data], snap_mode=sa.SFD_SNAP_FP32)
    i16, _ = e.rtm([gs], [data], snap_mode=sa.SFD_SNAP_FP16S)
    og = O.make_grid(nz, nx, 10.0, 10.0, 1e-3, ntaper=12)
    o64 = O.rtm(og, vel, os_, data, prec="d")
    o32 = O.rtm(og, vel, os_, data, prec="f")
    e32, e16, eo = rel(i32, o64["image"]), rel(i16, o64["image"]), rel(o32["image"], o64["image"])
    say("image vs FP64: fp32-snap %.3e  fp16-snap %.3e  (fp32 oracle itself %.3e)" % (e32, e16, eo))
    # FP32 storage reproduces the FP32 oracle; its distance to FP64 is FP32's own
    assert rel(i32, o32["image"]) <= TOL
    assert e32 <= 2 * eo + 1e-7
    # FP16 max-scaled storage: within 1e-3 of the FP64 image
    assert e16 <= 1e-3


def test_fp16_snapshots_with_denormal_amplitudes_stay_finite():
    """A wavefield whose tiles hold only tiny (down to denormal) values: the
    per-tile FP16 scale must not overflow (1/max) into Inf/NaN images."""
    say("case: fp16 snapshots, tiny amplitudes")
    nz, nx, nt = 48, 72, 320
    vel, w, rx, rz, data = _rtm_case(nz=nz, nx=nx, nt=nt, seed=9)
    w = (w * np.float32(1e-36)).astype(np.float32)
    gs = sa.Shot(18.0, 355.0, w, rz, rx)
    e = engine(nz, nx, 10.0, 10.0, 1e-3, 12, vel=vel)
    img, ill = e.rtm([gs], [data], snap_mode=sa.SFD_SNAP_FP16S)
    assert np.isfinite(img).all() and np.isfinite(ill).all()


def test_born_matches_oracle():
    say("case: born small")
    nz, nx, nt = 44, 70, 300
    vel = smooth_random(nz, nx, seed=13)
    refl = np.zeros((nz, nx), np.float32)
    refl[25, 10:60] = 0.1
    refl[33, 20:40] = -0.05
    w = ricker(20.0, nt, 1e-3)
    rx = np.linspace(10.0, 10.0 * (nx - 2), 25).astype(np.float32)
    rz = np.full_like(rx, 12.0)
    gs, os_ = both_shots(18.0, 345.0, w, rz, rx)
    e = engine(nz, nx, 10.0, 10.0, 1e-3, 12, vel=vel)
    gpu = e.born([gs], np.asfortranarray(refl))[0]
    ref = O.born(O.make_grid(nz, nx, 10.0, 10.0, 1e-3, ntaper=12), vel, np.asfortranarray(refl), os_)
    say("born rel=%.3e bitwise=%s" % (rel(gpu, ref), np.array_equal(gpu, ref)))
    assert rel(gpu, ref) <= TOL


def test_error_codes():
    say("case: error codes")
    g = sa.make_grid(40, 40, 10.0, 10.0, 1e-3, ntaper=5)
    with pytest.raises(sa.SeisFDError) as ei:
        sa.Engine(g)
    assert ei.value.code == -1
    e = engine(40, 40, 10.0, 10.0, 1e-3, 12)
    w = ricker(20.0, 50, 1e-3)
    s = sa.Shot(20.0, 200.0, w, [20.0], [100.0])
    with pytest.raises(sa.SeisFDError) as ei:
        e.forward([s])
    assert ei.value.code == -5   # no model yet
    vel = np.full((40, 40), 5000.0, np.float32)   # Courant 5.0 > 0.45
    with pytest.raises(sa.SeisFDError) as ei:
        e.set_model(vel)
    assert ei.value.code == -6


@pytest.mark.parametrize("snap,illupow", [("fp32", 1.0), ("fp32", 0.5), ("fp16", 1.0)])
def test_fwi_gradient_matches_oracle(snap, illupow):
    """compute_fct_gdt: AMOD, residual/misfit, AGRAD, scatter, illumination
    normalisation, mask and triangle smoothing, over a 3-shot batch."""
    say("case: fwi gradient snap=%s illupow=%g" % (snap, illupow))
    nz, nx, nt = 40, 96, 300
    true = smooth_random(nz, nx, seed=21, vmin=1600, vmax=2600)
    init = smooth_random(nz, nx, seed=22, vmin=1600, vmax=2500)
    og = O.make_grid(nz, nx, 10.0, 10.0, 1e-3, ntaper=12)
    w = ricker(18.0, nt, 1e-3)
    gsh, osh, dobs, mute = [], [], [], []
    for k, sx in enumerate([200.0, 470.0, 760.0]):
        rx = np.linspace(max(0.0, sx - 300), min(10.0 * (nx - 1), sx + 300), 41).astype(np.float32)
        rz = np.full_like(rx, 15.0)
        a, b = both_shots(12.0, sx, w, rz, rx)
        gsh.append(a)
        osh.append(b)
        dobs.append(O.modeling(og, true, b, extract_kind=1))
        m = np.ones((len(rx), nt), np.float32)
        m[:, :20] = np.linspace(0, 1, 20, dtype=np.float32)
        mute.append(m)
    mask = np.ones((nz, nx), np.float32)
    mask[:3, :] = 0.0
    mask = np.asfortranarray(mask)
    kw = dict(mute=mute, vpmask=mask, illupow=illupow, rect_z=3, rect_x=4, do_z=1, do_x=1)
    ref = O.fwi_gradient(og, init, osh, dobs, **kw)
    e = engine(nz, nx, 10.0, 10.0, 1e-3, 12, vel=init)
    f, grad, illu, res = e.fwi_gradient(gsh, dobs, snap_mode=(sa.SFD_SNAP_FP16S if snap == "fp16" else
                                                              sa.SFD_SNAP_FP32), **kw)
    ef = abs(f - ref["f"]) / abs(ref["f"])
    eg, ei = rel(grad, ref["grad"]), rel(illu, ref["illum"])
    er = rel(np.concatenate([r.ravel() for r in res]), ref["resid"])
    say("misfit rel=%.3e grad rel=%.3e illum rel=%.3e resid rel=%.3e grad bitwise=%s" %
        (ef, eg, ei, er, np.array_equal(grad, ref["grad"])))
    assert ef <= 1e-12 and er <= TOL
    if snap == "fp32":
        assert eg <= TOL and ei <= TOL
    else:
        assert eg <= 1e-3 and ei <= 1e-3


def test_fwi_lbfgs_matches_oracle_driver():
    """FWI by l_bfgs (Inversion.f90:29-189) with compute_fct_gdt as the
    evaluation: sfd_fwi_lbfgs against the oracle driver calling the oracle
    gradient — same misfit at every evaluation, misfit decreasing."""
    say("case: fwi l_bfgs")
    nz, nx, nt = 40, 96, 300
    true = smooth_random(nz, nx, seed=21, vmin=1600, vmax=2600)
    init = smooth_random(nz, nx, seed=22, vmin=1600, vmax=2500)
    og = O.make_grid(nz, nx, 10.0, 10.0, 1e-3, ntaper=12)
    w = ricker(18.0, nt, 1e-3)
    gsh, osh, dobs = [], [], []
    for sx in [200.0, 470.0, 760.0]:
        rx = np.linspace(max(0.0, sx - 300), min(10.0 * (nx - 1), sx + 300), 41).astype(np.float32)
        rz = np.full_like(rx, 15.0)
        a, b = both_shots(12.0, sx, w, rz, rx)
        gsh.append(a)
        osh.append(b)
        dobs.append(O.modeling(og, true, b, extract_kind=1))
    kw = dict(rect_z=2, rect_x=3, do_z=1, do_x=1)

    def fct(xv):
        r = O.fwi_gradient(og, xv.reshape(nx, nz).T, osh, dobs, **kw)
        return r["f"], np.asfortranarray(r["grad"]).ravel(order="F")

    lk = dict(niter=3, neval=12, bound=2, vpmin=1500.0, vpmax=3000.0)
    x0 = np.asfortranarray(init).ravel(order="F")
    xo, ro = O.l_bfgs(fct, x0, niter=3, neval=12, bound=2, xmin=1500.0, xmax=3000.0)
    e = engine(nz, nx, 10.0, 10.0, 1e-3, 12, vel=init)
    xg, rg = e.fwi_lbfgs(gsh, dobs, init, **lk, **kw)
    say("oracle f %s\ndevice f %s" % (np.array2string(ro["f_eval"], precision=6),
                                       np.array2string(rg["f_eval"], precision=6)))
    assert rg["evaluations"] == ro["evaluations"] and rg["iterations"] == ro["iterations"] == 3
    assert np.allclose(rg["f_eval"], ro["f_eval"], rtol=1e-5)
    assert rel(xg, xo.reshape(nx, nz).T) <= 1e-5
    fa = rg["f_eval"][rg["accepted"]]
    assert fa[-1] < rg["f_eval"][0]


def test_rtm_repeated_calls_reuse_the_plan():
    """Repeated RTMs of one shot reuse the device tables (the plan is keyed by
    the shots, the model and the tile settings): the images stay bitwise equal
    to a fresh engine's, also after a new model, another kind of call or a
    tile change in between, each of which plans anew."""
    say("case: rtm plan reuse")
    nz, nx, nt = 48, 72, 320
    vel, w, rx, rz, data = _rtm_case(nz=nz, nx=nx, nt=nt, seed=6)
    gs, os_ = both_shots(18.0, 355.0, w, rz, rx)
    e = engine(nz, nx, 10.0, 10.0, 1e-3, 12, vel=vel)
    first = e.rtm([gs], [data])
    again = e.rtm([gs], [data])
    assert all(np.array_equal(a, b) for a, b in zip(first, again))
    vel2 = (vel * 1.03).astype(np.float32)
    e.set_model(vel2)
    want = engine(nz, nx, 10.0, 10.0, 1e-3, 12, vel=vel2).rtm([gs], [data])
    same = lambda got: all(np.array_equal(a, b) for a, b in zip(got, want))
    assert same(e.rtm([gs], [data]))
    e.forward([gs])
    assert same(e.rtm([gs], [data]))
    e.lsrtm([gs], [data])
    assert same(e.rtm([gs], [data]))
    e.set_tile(64, 64)
    assert same(e.rtm([gs], [data]))
    e.set_option(sa.abi.SFD_OPT_PLAN_REUSE, 0)
    assert same(e.rtm([gs], [data])) and same(e.rtm([gs], [data]))
    out = O.rtm(O.make_grid(nz, nx, 10.0, 10.0, 1e-3, ntaper=12), vel2, os_, data)
    assert rel(want[0], out["image"]) <= TOL and rel(want[1], out["illum"]) <= TOL


@pytest.mark.parametrize("lsrtm", [False, True])
def test_rtm_batch_of_ragged_shots(lsrtm):
    """A batch of three shots with different record lengths and receiver
    spreads (batched launches: lean step kernel plus the separate imaging
    kernel): the images are the oracle's per-shot images summed in shot
    order."""
    say("case: %s batch of 3 ragged shots" % ("lsrtm" if lsrtm else "rtm"))
    nz, nx = 44, 90
    vel = smooth_random(nz, nx, seed=12)
    rng = np.random.default_rng(12)
    og = O.make_grid(nz, nx, 10.0, 10.0, 1e-3, ntaper=12)
    e = engine(nz, nx, 10.0, 10.0, 1e-3, 12, vel=vel)
    gsh, datas, img_ref, ill_ref = [], [], 0.0, 0.0
    for sx, nt, nr in ((150.0, 300, 21), (455.0, 260, 33), (700.0, 330, 17)):
        w = ricker(18.0, nt, 1e-3)
        rx = np.linspace(max(10.0, sx - 200), min(10.0 * (nx - 2), sx + 200), nr).astype(np.float32)
        rz = np.full_like(rx, 14.0)
        a, b = both_shots(18.0, sx, w, rz, rx)
        data = (rng.standard_normal((nr, nt)) * 1e-2).astype(np.float32)
        out = O.rtm(og, vel, b, data, lsrtm=lsrtm)
        img_ref = img_ref + out["image"]
        ill_ref = ill_ref + out["illum"]
        gsh.append(a)
        datas.append(data)
    img, ill = (e.lsrtm if lsrtm else e.rtm)(gsh, datas)
    ei, el = rel(img, img_ref), rel(ill, ill_ref)
    say("image rel=%.3e illum rel=%.3e bitwise=%s/%s" % (ei, el, np.array_equal(img, img_ref),
                                                         np.array_equal(ill, ill_ref)))
    assert ei <= TOL and el <= TOL

"""Shot waves, ragged-batch launches and engine options, on the GPU through
the C-ABI.

A call whose shots do not fit in device memory runs them in consecutive
waves (shot order).  Images, gradients and misfits are summed over the waves
in the same order as over one batch, so a call split into waves (forced here
with a small SFD_OPT_BATCH_MIB) must give results bitwise equal to the same
call in one batch.  This replaces the reference's OpenMP shot loop
(progs/afwi/compute_function_gradient.f90:246-256), which has no memory
limit of this kind."""
import numpy as np
import pytest

import oracle_lib as O
from models import ricker, smooth_random

sa = pytest.importorskip("seismic_amd")
import seismic_amd.abi as abi  # noqa: E402

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)


NZ, NX, DX, DT, NTAPER, SNAPI = 40, 96, 10.0, 1e-3, 12, 4


def ragged_shots(n=5, seed=0):
    """n shots along the surface with uneven spreads and record lengths."""
    rng = np.random.default_rng(seed)
    shots = []
    for k, sx in enumerate(np.linspace(100.0, DX * (NX - 1) - 100.0, n)):
        nt = int(rng.integers(180, 320))
        nr = int(rng.integers(8, 40))
        rx = np.clip(sx + DX * (np.arange(nr) - nr // 2), 0.0, DX * (NX - 1)).astype(np.float32)
        rx = np.unique(rx)
        shots.append(sa.Shot(20.0, float(sx), ricker(18.0, nt, DT), np.full_like(rx, 20.0), rx))
    return shots


def engine(vel, cap_mib=0):
    e = sa.Engine(sa.make_grid(NZ, NX, DX, DX, DT, ntaper=NTAPER, snapi=SNAPI))
    e.set_model(vel)
    e.set_option(abi.SFD_OPT_BATCH_MIB, cap_mib)
    return e


def test_fwi_gradient_in_waves_is_bitwise_one_batch():
    say("case: fwi gradient, 5 ragged shots, one batch vs waves")
    true = smooth_random(NZ, NX, seed=1)
    mig = smooth_random(NZ, NX, seed=2)
    shots = ragged_shots()
    e = engine(true)
    dobs = e.forward(shots, extract_kind=sa.SFD_EXTRACT_NEAREST)
    e.set_model(mig)
    f1, g1, i1, r1 = e.fwi_gradient(shots, dobs, illupow=1.0, rect_z=2, rect_x=3, do_z=1, do_x=1)
    assert e.last_waves() == 1
    e.set_option(abi.SFD_OPT_BATCH_MIB, 2)
    f2, g2, i2, r2 = e.fwi_gradient(shots, dobs, illupow=1.0, rect_z=2, rect_x=3, do_z=1, do_x=1)
    nw = e.last_waves()
    say("waves=%d misfit %.9e / %.9e" % (nw, f1, f2))
    assert nw > 1
    assert f1 == f2
    assert np.array_equal(g1, g2) and np.array_equal(i1, i2)
    for a, b in zip(r1, r2):
        assert np.array_equal(a, b)


def test_fwi_gradient_resident_dobs_is_bitwise_host_dobs():
    say("case: fwi gradient with observed data resident in HBM")
    true = smooth_random(NZ, NX, seed=3)
    mig = smooth_random(NZ, NX, seed=4)
    shots = ragged_shots(seed=5)
    e = engine(true)
    dobs = e.forward(shots, extract_kind=sa.SFD_EXTRACT_NEAREST)
    e.set_model(mig)
    f1, g1, _, _ = e.fwi_gradient(shots, dobs, snap_mode=sa.SFD_SNAP_FP16S)
    e.set_traces(dobs, shots)
    f2, g2, _, _ = e.fwi_gradient(shots, None, snap_mode=sa.SFD_SNAP_FP16S)
    assert f1 == f2 and np.array_equal(g1, g2)
    e.set_option(abi.SFD_OPT_BATCH_MIB, 2)
    f3, g3, _, _ = e.fwi_gradient(shots, None, snap_mode=sa.SFD_SNAP_FP16S)
    assert e.last_waves() > 1
    assert f1 == f3 and np.array_equal(g1, g3)


@pytest.mark.parametrize("lsrtm", [False, True])
def test_rtm_in_waves_is_bitwise_one_batch(lsrtm):
    say("case: rtm lsrtm=%s, one batch vs waves" % lsrtm)
    true = smooth_random(NZ, NX, seed=6)
    mig = smooth_random(NZ, NX, seed=7)
    shots = ragged_shots(seed=8)
    e = engine(true)
    d = e.forward(shots, extract_kind=sa.SFD_EXTRACT_SINC_V2)
    e.set_model(mig)
    call = e.lsrtm if lsrtm else e.rtm
    im1, il1 = call(shots, d)
    e.set_option(abi.SFD_OPT_BATCH_MIB, 2)
    im2, il2 = call(shots, d)
    say("waves=%d" % e.last_waves())
    assert e.last_waves() > 1
    assert np.array_equal(im1, im2) and np.array_equal(il1, il2)
    # and against the oracle, shot by shot summed in order (one shot each)
    g = O.make_grid(NZ, NX, DX, DX, DT, ntaper=NTAPER, snapi=SNAPI)
    ref_im = np.zeros((NZ, NX))
    for s, ds in zip(shots, d):
        os_ = O.Shot(s.sz, s.sx, s.wavelet, s.rz, s.rx)
        ref_im += O.rtm(g, mig, os_, ds, lsrtm=lsrtm)["image"]
    err = float(np.linalg.norm(im2 - ref_im) / np.linalg.norm(ref_im))
    say("rtm in waves vs oracle rel L2 = %.3e" % err)
    assert err <= 1e-5


def test_born_and_forward_in_waves_are_bitwise_one_batch():
    say("case: born and forward, one batch vs waves")
    vel = smooth_random(NZ, NX, seed=9)
    refl = np.asfortranarray(np.random.default_rng(1).standard_normal((NZ, NX)).astype(np.float32) * 10.0)
    shots = ragged_shots(seed=10)
    e = engine(vel)
    b1 = e.born(shots, refl)
    f1 = e.forward(shots, extract_kind=sa.SFD_EXTRACT_SINC_V2)
    e.set_option(abi.SFD_OPT_BATCH_MIB, 1)
    b2 = e.born(shots, refl)
    assert e.last_waves() > 1
    f2 = e.forward(shots, extract_kind=sa.SFD_EXTRACT_SINC_V2)
    for a, b in zip(b1 + f1, b2 + f2):
        assert np.array_equal(a, b)


def test_ragged_batch_times_every_step():
    """Past the shortest record only the running shots are launched, and the
    kernel totals cover every step of both passes with their point-updates."""
    say("case: ragged batch kernel totals")
    vel = smooth_random(NZ, NX, seed=11)
    shots = ragged_shots(seed=12)
    e = engine(vel)
    d = e.forward(shots, extract_kind=sa.SFD_EXTRACT_NEAREST)
    f, g, _, _ = e.fwi_gradient(shots, d)
    t = e.kernel_totals()
    ntmax = max(s.nt for s in shots)
    pts = e.last_timing()["pts"]
    say("steps %s, pts %s (call %.0f), us %s" % (t["steps"], t["pts"], pts, t["us"]))
    assert t["steps"] == [ntmax, ntmax]
    assert abs(sum(t["pts"]) - pts) <= 1e-9 * pts
    assert all(u > 0 for u in t["us"]) and all(b > 24 * p for b, p in zip(t["bytes"], t["pts"]))


def test_strip_width_must_match_tile():
    """A strip width the tile shape does not compile is refused, and the
    handle keeps working: a later forward still matches the oracle (an
    accepted bad pair used to fail inside a graph capture and leave the
    stream capturing)."""
    say("case: strip width 16 on 64x32 tiles")
    nz, nx, nt = 47, 83, 260
    vel = smooth_random(nz, nx, seed=13)
    w = ricker(20.0, nt, 1e-3)
    rx = np.linspace(30.0, 10.0 * (nx - 1) - 30.0, 21).astype(np.float32)
    rz = np.full_like(rx, 20.0)
    e = sa.Engine(sa.make_grid(nz, nx, 10.0, 10.0, 1e-3, ntaper=12))
    e.set_model(vel)
    lib = abi.load()
    assert lib.sfd_set_option(e.h, abi.SFD_OPT_STRIP_PX, 16) == abi.SFD_EINVAL
    assert b"not compiled" in lib.sfd_last_error(e.h)
    e.set_tile(64, 64)
    e.set_option(abi.SFD_OPT_STRIP_PX, 16)
    e.set_tile(64, 32)            # falls back to the 64x32 default strip
    gpu = e.forward([sa.Shot(20.0, 400.0, w, rz, rx)])[0]
    ref = O.modeling(O.make_grid(nz, nx, 10.0, 10.0, 1e-3, ntaper=12), vel, O.Shot(20.0, 400.0, w, rz, rx),
                     extract_kind=0)
    assert np.array_equal(gpu, ref)


def test_bad_shot_in_a_planned_batch_is_an_error_and_the_engine_recovers():
    """Shots are planned on several host threads: a shot whose source
    footprint leaves the FD region (here the eighth of twelve) makes the call
    fail with SFD_EINVAL and the planner's message, and the engine then runs
    the good shots exactly as a fresh engine does."""
    say("case: bad shot in a 12-shot batch")
    vel = smooth_random(40, 96, seed=6, vmin=1700, vmax=2500)
    w = ricker(15.0, 250, 1e-3)
    good = []
    for k in range(12):
        rx = (np.arange(2 + k % 3, 94, 3) * 10.0).astype(np.float32)
        good.append(sa.Shot(20.0, 60.0 + 70.0 * k, w, np.full(len(rx), 20.0, np.float32), rx))
    bad = list(good)
    bad[7] = sa.Shot(-2000.0, good[7].sx, w, good[7].rz, good[7].rx)
    e = engine(vel)
    with pytest.raises(sa.SeisFDError) as ex:
        e.forward(bad)
    assert ex.value.code == abi.SFD_EINVAL and "footprint" in str(ex.value)
    got = e.forward(good)
    want = engine(vel).forward(good)
    for a, b in zip(got, want):
        assert np.array_equal(a, b)

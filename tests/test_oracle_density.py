"""Variable density (withRho): the oracle against the reference's own
kernels — FD_2D_derivatives_acoustic_forward_grid with
allocate_sxx2d_szz2d_delp2d, Interpolate, FD_acoustic_rho_init_coefs with
coefpower 1, Injection_sinc's v^2 rho gain — composed in Propagator.f90 order
by oracle/ref_harness.f90; readsou's half-step source shift and the operator
itself by what the mathematics fixes."""
import numpy as np
import pytest

import oracle_lib as O
from models import ricker, smooth_random

needs_ref = pytest.mark.skipif(not O.ref_available(), reason="oracle/_ref not built (no reference tree)")


def rel(a, b):
    return float(np.linalg.norm((np.asarray(a, np.float64) - b).ravel()) /
                 max(np.linalg.norm(np.asarray(b, np.float64).ravel()), 1e-300))


def _case(seed=1, nz=36, nx=48, nt=240, types=(0, 0, 0)):
    g = O.make_grid(nz, nx, 10.0, 10.0, 1e-3, ntaper=12, snapi=4)
    g.rec_type, g.shot_type, g.surf_type = types
    vel = smooth_random(nz, nx, seed=seed)
    rho = smooth_random(nz, nx, seed=seed + 100, vmin=1.0, vmax=2.6)
    w = ricker(25.0, nt, 1e-3)
    rx = np.linspace(13.0, 10.0 * (nx - 1) - 17.0, 21).astype(np.float32)
    rz = np.full_like(rx, 23.0)
    shot = O.Shot(37.0, 241.0, w, rz, rx)
    return g, vel, rho, shot


@needs_ref
def test_mksinc9_half_shift_bitwise():
    assert np.array_equal(O.mksinc(9, 0.5), O.ref_mksinc(9, 0.5))


@needs_ref
@pytest.mark.parametrize("ext", [0, 1, 2, 3])
@pytest.mark.parametrize("types", [(0, 0, 0), (1, 1, 1)])
def test_forward_with_density_matches_reference(ext, types):
    g, vel, rho, shot = _case(types=types)
    if types[2] > 0:
        shot = O.Shot(20.0, shot.sx, shot.wavelet, shot.rz, shot.rx)
    O.set_density(rho)
    try:
        vpad, oxw, nxw, b = O.padded_vel(g, vel, shot)
        rpad, _, _, _ = O.padded_vel(g, rho, shot)
        src = O.half_step(shot.wavelet)          # readsou's shift, applied before injection
        ref_tr, ref_w = O.ref_propagate(g, vpad, oxw, nxw, 0, 1, [shot.sz], [shot.sx], src[None, :], ext,
                                        shot.rz, shot.rx, shot.nt, 1, shot.nt // g.snapi, rho_pad=rpad)
        tr, snaps = O.modeling(g, vel, shot, extract_kind=ext, snaps=True)
    finally:
        O.set_density(None)
    assert np.array_equal(tr, ref_tr), rel(tr, ref_tr)
    assert np.array_equal(snaps, ref_w), rel(snaps, ref_w)


@needs_ref
def test_rtm_backward_fields_with_density_match_reference():
    g, vel, rho, shot = _case(seed=2)
    rng = np.random.default_rng(5)
    data = rng.standard_normal((shot.ntr, shot.nt)).astype(np.float32)
    O.set_density(rho)
    try:
        vpad, oxw, nxw, b = O.padded_vel(g, vel, shot)
        rpad, _, _, _ = O.padded_vel(g, rho, shot)
        _, ref_w = O.ref_propagate(g, vpad, oxw, nxw, 1, 1, shot.rz, shot.rx, data, -1, [], [],
                                   shot.nt, -1, shot.nt // g.snapi, rho_pad=rpad)
        cap = O.rtm(g, vel, shot, data, capture=True)["capture"]
    finally:
        O.set_density(None)
    assert np.array_equal(cap, ref_w), rel(cap, ref_w)


def test_half_step_shift_is_half_a_sample():
    """A band-limited wavelet shifted by readsou's interpolator equals the
    wavelet evaluated half a sample later (interior), and a constant trace
    stays constant up to the ends (the renormalisation)."""
    nt, dt, f = 400, 1e-3, 20.0
    w = ricker(f, nt, dt).astype(np.float64)
    s = O.half_step(w, prec="d")
    t = np.arange(nt) * dt + 0.5 * dt - 1.2 / f
    exact = (1 - 2 * (np.pi * f * t) ** 2) * np.exp(-(np.pi * f * t) ** 2)
    assert rel(s[10:-10], exact[10:-10]) < 2e-3
    # a shift of -0.5 instead of +0.5 (wrong direction) fails by far
    t2 = np.arange(nt) * dt - 0.5 * dt - 1.2 / f
    wrong = (1 - 2 * (np.pi * f * t2) ** 2) * np.exp(-(np.pi * f * t2) ** 2)
    assert rel(s[10:-10], wrong[10:-10]) > 0.1
    c = O.half_step(np.full(50, 3.0), prec="d")
    assert np.allclose(c, 3.0, rtol=1e-12)


def test_density_scaling_scales_the_gathers():
    """Scaling rho by k leaves the propagation operator unchanged (v^2 rho
    times 1/rho2) and scales the injected source by k (gain v^2 rho): the
    gathers scale by k."""
    g, vel, rho, shot = _case(seed=3)
    outs = []
    for k in (1.0, 2.5):
        O.set_density(rho * k)
        try:
            outs.append(O.modeling(g, vel, shot, extract_kind=0, prec="d"))
        finally:
            O.set_density(None)
    assert rel(outs[1], 2.5 * outs[0]) < 1e-6   # rho k is rounded to FP32 before the FP64 oracle


def _zero_offset_reflection(rho2):
    nz, nx, nt = 70, 80, 420
    g = O.make_grid(nz, nx, 10.0, 10.0, 1e-3, ntaper=12, snapi=4)
    vel = np.full((nz, nx), 2000.0, np.float32)
    rho = np.ones((nz, nx), np.float32)
    w = ricker(20.0, nt, 1e-3)
    rx = np.array([400.0, 620.0], np.float32)
    shot = O.Shot(100.0, 400.0, w, np.full_like(rx, 100.0), rx)
    O.set_density(rho)
    try:
        ref = O.modeling(g, vel, shot, extract_kind=1, prec="d")
        rho_c = rho.copy()
        rho_c[21:, :] = rho2          # interface at z = 210 m
        O.set_density(rho_c)
        got = O.modeling(g, vel, shot, extract_kind=1, prec="d")
    finally:
        O.set_density(None)
    a, b = (got - ref)[0][110:280], ref[1][110:280]
    return float(np.dot(a, b) / np.dot(b, b))


def test_density_contrast_reflects_with_the_impedance_coefficient():
    """Constant velocity, a flat density step rho 1 -> rho2 110 m below the
    source: at zero offset the reflection is R = (rho2 - 1)/(rho2 + 1) times
    the direct wave at the image distance (220 m), angle-independent for
    equal velocities.  Measured 0.376 against 1/3: the 9-point sinc that puts
    rho at the half cells rings at the step (its overshoot raises the
    contrast).  A density-free operator gives 0, an inverted 1/rho the
    opposite sign, and R(rho2) = -R(1/rho2) must hold."""
    up, down = _zero_offset_reflection(2.0), _zero_offset_reflection(0.5)
    assert abs(up - 1.0 / 3.0) < 0.25 / 3.0, up
    assert abs(down + 1.0 / 3.0) < 0.25 / 3.0, down
    assert abs(up + down) < 0.05 * abs(up)

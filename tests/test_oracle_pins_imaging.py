"""Pins of the oracle's Born, LSRTM, illumination-normalisation and
window/aperture functions by what the mathematics or the geometry fixes
(no reference run covers them: Imaging.f90, ExtractPadModel.f90 and the
compute_fct_gdt driver need SEPlib).  Each check is chosen so that a
plausible slip fails it; where a slip can be emulated exactly through the
oracle's own inputs, the test shows the same bound rejecting it.
"""
import numpy as np
import pytest

import oracle_lib as O
from models import ricker, smooth_random
from test_oracle_gradient import BUMPS, bump, problem


def rel(a, b):
    return float(np.linalg.norm(np.ravel(a) - np.ravel(b)) / np.linalg.norm(np.ravel(b)))


# ---------------------------------------------------------------------------
# Born (Injection_Born, Imaging.f90:747-788; readsou's second difference,
# ReadSouVelRho.f90:318-326)

NZ, NX, NT, DX, DT = 50, 80, 700, 10.0, 1e-3


def born_case(snapi):
    g = O.make_grid(NZ, NX, DX, DX, DT, ntaper=15, snapi=snapi)
    v = smooth_random(NZ, NX, seed=2, vmin=1800, vmax=2400).astype(np.float64)
    w = ricker(15.0, NT, DT)
    rx = (np.arange(3, NX - 3, 2) * DX).astype(np.float32)
    shot = O.Shot(20.0, 400.0, w, np.full(len(rx), 20.0, np.float32), rx)
    zz, xx = np.mgrid[0:NZ, 0:NX]
    # a smooth perturbation in the interior: away from the source and the
    # receivers (whose injection / extraction gains read v), the Higdon edges
    # and the window taper's first and last ten columns
    dv = np.exp(-((zz - 30) ** 2 + (xx - 45) ** 2) / (2 * 3.0 ** 2))
    return g, v, shot, dv


def modeling_derivative(g, v, shot, dv, eps=0.5):
    """Central difference of forward modeling along dv, with the Born data's
    extraction (sinc, gain 1/(dz dx)), in FP64."""
    fp = O.modeling(g, v + eps * dv, shot, extract_kind=2, prec="d")
    fm = O.modeling(g, v - eps * dv, shot, extract_kind=2, prec="d")
    return (fp - fm) / (2 * eps)


def test_born_is_the_derivative_of_modeling():
    """The scheme's velocity enters as v^2 lap(u), so the exact linearisation
    of one step adds 2 v dv lap(u) = (2 dv / v) (v^2 lap u), and v^2 lap u is
    d2u/dt2 away from the source.  Born adds 2 m S / v with S the background
    field of the second-differenced wavelet (= d2u/dt2 of the same linear,
    time-invariant scheme).  So born(m = dv) must equal the derivative of
    modeling along dv up to the difference between the centred second
    difference and the scheme's own time operator, O((w dt)^2) ~ 1e-3 at 15 Hz
    and 1 ms.  A factor 2 slip, 1/v^2 in place of 1/v, or a missing
    second difference of the wavelet breaks the 1e-2 bound by orders of
    magnitude."""
    g, v, shot, dv = born_case(snapi=1)
    D = modeling_derivative(g, v, shot, dv)
    B = O.born(g, v, dv, shot, prec="d")
    err = rel(B, D)
    print("born vs d(modeling)/dv: rel L2 %.3e, norm ratio %.6f" % (err, np.linalg.norm(B) / np.linalg.norm(D)))
    assert err <= 1e-2
    # the same bound rejects the slips (emulated exactly through the inputs:
    # the Born term is linear in m, so 2 m S / v^2 is born(m / v))
    assert rel(2 * B, D) > 0.5
    assert rel(O.born(g, v, dv / v, shot, prec="d"), D) > 0.5


def test_born_injects_every_snapi_steps_without_rescaling():
    """With snapi > 1 the Born source is added on snapshot steps only and is
    not scaled by snapi (Imaging.f90:779-783), so the scattered data are
    1/snapi of the snapi = 1 data to first order."""
    g, v, shot, dv = born_case(snapi=2)
    D = modeling_derivative(g, v, shot, dv)
    B2 = O.born(g, v, dv, shot, prec="d")
    err = rel(2 * B2, D)
    print("snapi 2: rel L2 of 2 * born vs d(modeling)/dv %.3e" % err)
    assert err <= 2e-2
    assert rel(B2, D) > 0.4


# ---------------------------------------------------------------------------
# LSRTM imaging (Imaging.f90:698-745) against the exact transpose of Born

def test_lsrtm_image_is_the_born_adjoint():
    """The reverse-time pass runs the 5-level scheme backward with the
    transposed stencil lap(v^2 .), which is the transpose of the forward
    recursion (the coefficients 20, -6, -4, 1 land on the same lags), and the
    LSRTM condition 2 t u S / v is the transpose of the Born term.  So the
    LSRTM image of data d must equal B^T d (the oracle's exact discrete
    adjoint, pinned by its FP64 dot test) up to the Higdon band, which is not
    transposed: ~1e-3 here.  A missing factor 2 or 1/v^2 for 1/v fails the
    5e-3 bound; so does a sign slip."""
    rng = np.random.default_rng(0)
    for snapi in (1, 4):
        g, v, shot, _ = born_case(snapi)
        d = rng.standard_normal((shot.ntr, shot.nt))
        mb = O.born_adjoint(g, v, shot, d, prec="d")
        im = O.rtm(g, v, shot, d, prec="d", lsrtm=True)["image"]
        err = rel(im, mb)
        print("snapi %d: LSRTM image vs B^T d: rel L2 %.3e, norm ratio %.6f" %
              (snapi, err, np.linalg.norm(im) / np.linalg.norm(mb)))
        assert err <= 5e-3
        assert rel(0.5 * im, mb) > 0.4 and rel(im / v, mb) > 0.5 and rel(-im, mb) > 1.5


# ---------------------------------------------------------------------------
# Illumination normalisation (compute_function_gradient.f90:332-343):
# illu = illu^illupow; illu = (illu + max(illu) / 1e4) / sqrt(sum(illu^2) / N);
# grad = grad / illu.

def test_constant_illumination_normalises_to_one_plus_1e4():
    """illupow = 0 makes the illumination 1 everywhere, so it normalises to
    exactly (1 + 1 / 1e4) / sqrt(1) (FP64; the model cells outside every
    window, whose illumination is 0, give 0^0 = 1 too)."""
    g, v0, shots, dobs = problem(1)
    out = O.fwi_gradient(g, v0, shots, dobs, illupow=0.0, prec="d")
    assert np.all(out["illum"] == 1.0 + 1e-4)


@pytest.mark.parametrize("illupow", [1.0, 0.5])
def test_normalisation_is_invariant_to_the_illumination_scale(illupow):
    """Scaling the wavelet and the observed data by 2 scales the residual,
    both wavefields and the raw gradient by 2, 2 and 4, and the illumination
    by 4 (by 4^illupow after the power).  The normalised illumination is
    homogeneous of degree 0 in the illumination, so it is unchanged, and the
    final gradient and misfit scale by 4: bit for bit at illupow = 1 (powers
    of 2 scale exactly), to the last ulps of pow() otherwise.  Dividing by
    the squared norm, dropping the norm, or adding a fixed 1e-4 in place of
    max / 1e4 breaks the equalities."""
    g, v0, shots, dobs = problem(1)
    a = O.fwi_gradient(g, v0, shots, dobs, illupow=illupow, prec="d")
    shots2 = [O.Shot(s.sz, s.sx, 2 * s.wavelet, s.rz, s.rx) for s in shots]
    b = O.fwi_gradient(g, v0, shots2, [2 * d for d in dobs], illupow=illupow, prec="d")
    assert b["f"] == 4 * a["f"]
    if illupow == 1.0:
        assert np.array_equal(b["illum"], a["illum"])
        assert np.array_equal(b["grad"], 4 * a["grad"])
    else:
        np.testing.assert_allclose(b["illum"], a["illum"], rtol=1e-13, atol=0)
        np.testing.assert_allclose(b["grad"], 4 * a["grad"], rtol=1e-12, atol=1e-12 * np.abs(a["grad"]).max())


@pytest.mark.parametrize("illupow", [1.0, 0.5])
def test_normalised_illumination_from_the_snapshots(illupow):
    """The illumination is, per shot, sum over the snapshots of taper * S^2
    (Imaging.f90:366), added into the model by the window columns; computed
    here from the oracle's forward snapshots and normalised as the
    reference's formula says, it must match the oracle's returned
    illumination."""
    g, v0, shots, dobs = problem(1)
    out = O.fwi_gradient(g, v0, shots, dobs, illupow=illupow, prec="d")
    raw = np.zeros((g.nz, g.nx))
    for s in shots:
        _, wave = O.modeling(g, v0, s, extract_kind=1, prec="d", snaps=True)   # (ntsnap, nxw, nz)
        oxw, nxw = O.shot_window(g, s)
        win = (wave ** 2).sum(axis=0).T * O.taper(nxw, prec="d")[None, :]
        raw = O.copy_image(g, oxw, nxw, win, raw)
    ip = raw ** illupow
    want = (ip + ip.max() / 1e4) / np.sqrt(np.mean(ip ** 2))
    err = float(np.abs(out["illum"] - want).max() / want.max())
    print("illupow %.1f: normalised illumination max rel diff %.2e" % (illupow, err))
    assert err <= 1e-10


def test_gradient_at_illupow_one_matches_finite_differences():
    """At the reference default illupow = 1 the gradient divided by the
    normalised illumination, multiplied back by it, is -dJ/dv / snapi: the
    same central-difference check as at illupow = 0, now through the
    illumination-weighted path."""
    g, v0, shots, dobs = problem(1)
    J = lambda v: O.fwi_gradient(g, v, shots, dobs, illupow=1.0, prec="d")["f"]
    out = O.fwi_gradient(g, v0, shots, dobs, illupow=1.0, prec="d")
    raw = out["grad"] * out["illum"]
    for cz, cx in BUMPS:
        dv = bump(g.nz, g.nx, cz, cx)
        D = (J(v0 + dv) - J(v0 - dv)) / 2.0
        P = float(np.sum(raw * dv))
        print("bump (%d,%d): dJ = %.6e, <grad * illum, dv> = %.6e, ratio %.6f" % (cz, cx, D, P, D / P))
        assert abs(D / P + 1.0) <= 2e-3
        # without multiplying the illumination back the ratio is far off
        assert abs(D / float(np.sum(out["grad"] * dv)) + 1.0) > 0.05


# ---------------------------------------------------------------------------
# Shot window, padding and image scatter (ExtractPadModel.f90:19-85 window
# and aperture, :177 nxw, :509-530 mod_window_pad, :662-695 mod_copy_image).
# Expected values derived by hand from the geometry, dx = 10 m, model x from
# 0 to 290 m (30 columns); centre = midpoint of the first and last
# source-receiver midpoints, window = [min(spread, centre - aperture/2),
# max(spread, centre + aperture/2)], nxw = nint(width / dx) + 1.

W = ricker(20.0, 10, 1e-3)


def window(aperture, sx, rx):
    g = O.make_grid(12, 30, 10.0, 10.0, 1e-3, ntaper=10, aperture=aperture)
    rx = np.asarray(rx, np.float32)
    return g, O.Shot(10.0, sx, W, np.full(len(rx), 10.0, np.float32), rx)


@pytest.mark.parametrize("aperture,sx,rx,oxw,nxw", [
    # split spread 50..250, source 100: midpoints 75, 175 -> centre 125;
    # no aperture: the spread itself, 200 m wide
    (0.0, 100.0, np.arange(50, 251, 10), 50.0, 21),
    # 400 m aperture around 125: [-75, 325], wider than the spread
    (400.0, 100.0, np.arange(50, 251, 10), -75.0, 41),
    # receivers all right of the source at 20: [20, 200]
    (0.0, 20.0, np.arange(100, 201, 10), 20.0, 19),
    # receivers on both sides, 30 and 280, source 150: midpoints 90, 215 ->
    # centre 152.5; 100 m aperture [102.5, 202.5] lies inside the spread
    (100.0, 150.0, [30.0, 280.0], 30.0, 26),
    # spread beyond both model edges [-50, 350]
    (0.0, 150.0, np.arange(-50, 351, 10), -50.0, 41),
])
def test_shot_window_table(aperture, sx, rx, oxw, nxw):
    g, s = window(aperture, sx, rx)
    assert O.shot_window(g, s) == (oxw, nxw)


def test_shot_window_off_grid():
    """Off-grid spread 51.2 .. 248.9: the window starts at the first receiver
    and is nint(197.7 / 10) + 1 = 21 columns wide."""
    g, s = window(0.0, 103.7, np.linspace(51.2, 248.9, 20))
    oxw, nxw = O.shot_window(g, s)
    assert oxw == np.float32(51.2) and nxw == 21


def test_window_columns_off_the_model():
    """Window [-75, 325] on the 0..290 model: column j sits at x = 10 (j - 1)
    - 75, model column nint(x / 10) + 1, with nint rounding halves away from
    zero: j = 8 (x = -5) -> 0, j = 9 (x = 5) -> 2, j = 37 (x = 285) -> 30,
    j = 38 (x = 295) -> 31.  So the image scatter fills model columns 2..30
    from window columns 9..37 (model column 1, between two window columns,
    gets nothing) and skips the rest; the padding clamps: window columns 1..8
    copy model column 1 and 38..41 copy column 30."""
    g, s = window(400.0, 100.0, np.arange(50, 251, 10))
    oxw, nxw = O.shot_window(g, s)
    small = np.tile(np.arange(1, nxw + 1, dtype=np.float64), (g.nz, 1))   # value = window column j
    big = O.copy_image(g, oxw, nxw, small)
    assert np.all(big[:, 0] == 0)
    np.testing.assert_array_equal(big[0, 1:], np.arange(9, 38))
    assert np.all(big == big[0])
    # padding: the model's column c holds 100 + c (1-based)
    vel = np.asfortranarray(np.tile(100.0 + np.arange(1, g.nx + 1), (g.nz, 1)).astype(np.float32))
    vp, _, _, _ = O.padded_vel(g, vel, s)
    i0 = 4 + g.ntaper
    got = vp[i0, i0:i0 + nxw] - 100.0
    want = np.concatenate([np.ones(8), np.arange(2, 31), np.full(4, 30.0)])
    np.testing.assert_array_equal(got, want)

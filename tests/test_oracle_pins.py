"""Pins of the oracle that do not need the reference build: closed forms,
invariants and brute-force sums (each chosen so a plausible slip fails)."""
import numpy as np
import pytest

import oracle_lib as O
from models import ricker, smooth_random


def test_fd_coefficients_polynomial_exactness():
    """c0..c4 (FDcoefs.f90:25-36): the 9-point second derivative is exact for
    polynomials up to degree 9 and wrong at degree 10 (8th order)."""
    c = np.array([-205 / 72, 8 / 5, -1 / 5, 8 / 315, -1 / 560])
    k = np.arange(1, 5)
    for p in range(0, 10):
        # second derivative of x^p at 0 is 2 if p == 2 else 0
        s = (c[0] * (0.0 ** p if p else 1.0)) + sum(c[m] * (m ** p + (-m) ** p) for m in k)
        assert abs(s - (2.0 if p == 2 else 0.0)) < 1e-9, p
    s10 = sum(c[m] * (2 * m ** 10) for m in k)
    assert abs(s10) > 1.0


def test_mksinc_delta_and_unit_sum():
    s0 = O.mksinc(7, 0.0, prec="d")
    assert abs(s0[3] - 1.0) < 1e-8 and np.all(np.abs(np.delete(s0, 3)) < 1e-8)
    for sh in [0.1, 0.3, -0.45, 0.5]:
        assert abs(O.mksinc(7, sh, prec="d").sum() - 1.0) < 1e-12
    # symmetric: shift -> reversed weights at -shift
    a, b = O.mksinc(7, 0.3, prec="d"), O.mksinc(7, -0.3, prec="d")
    assert np.allclose(a, b[::-1], atol=1e-12)


def test_higdon_coeff_closed_form():
    """Boundary.f90:66-87: the 8 coefficients are the products of the two
    one-way Higdon factors; check against an independent expansion."""
    for nu in [0.1, 0.3, 0.45]:
        b1, b2, w = 1.0, 0.7, 0.5

        def fac(b):
            d = (b + nu) * (1 - w)
            return (w * (b + nu) - nu) / d, (w * (b + nu) - b) / d, w / (w - 1)
        qx, qt, qxt = fac(b1)
        rx, rt, rxt = fac(b2)
        # coefficient vector of (1 + qx X + qt T + qxt X T)(1 + rx X + rt T + rxt X T)
        # in the order X, X^2, T, XT, X^2T, T^2, XT^2, X^2T^2
        ref = [qx + rx, qx * rx, qt + rt, qx * rt + qt * rx + qxt + rxt, qx * rxt + rx * qxt,
               qt * rt, qt * rxt + rt * qxt, qxt * rxt]
        assert np.allclose(O.higdon_coeff(nu, prec="d"), ref, rtol=1e-12)


def test_triangle_impulse_is_triangle_and_mass_preserving():
    n, nr = 41, 5
    u = np.zeros((n, 1))
    u[20, 0] = 1.0
    out = O.triangle2(u, nr, 1, doz=1, dox=0, prec="d")[:, 0]
    k = np.arange(-(nr - 1), nr)
    expect = (nr - np.abs(k)) / nr ** 2
    assert np.allclose(out[20 - (nr - 1):20 + nr], expect, atol=1e-14)
    assert abs(out.sum() - 1.0) < 1e-14
    rng = np.random.default_rng(0)
    v = rng.standard_normal((17, 23))
    sm = O.triangle2(v, 4, 6, prec="d")
    assert abs(sm.sum() - v.sum()) < 1e-10   # end folds preserve the total


def test_padding_and_window():
    """Window + edge-replication padding: interior equals the model window,
    pads replicate the edge rows/columns (ReadSouVelRho.f90:545-596)."""
    nz, nx = 12, 30
    g = O.make_grid(nz, nx, 10.0, 10.0, 1e-3, ntaper=10)
    vel = np.asfortranarray(np.arange(nz * nx, dtype=np.float32).reshape(nz, nx) + 1000)
    w = ricker(20.0, 10, 1e-3)
    shot = O.Shot(10.0, 100.0, w, [10.0, 10.0], [60.0, 200.0])
    vp, oxw, nxw, b = O.padded_vel(g, vel, shot)
    assert oxw == 60.0 and nxw == 15
    i0 = 4 + 10
    np.testing.assert_array_equal(vp[i0:i0 + nz, i0:i0 + nxw], vel[:, 6:21])
    np.testing.assert_array_equal(vp[:i0, i0:i0 + nxw], np.repeat(vel[:1, 6:21], i0, axis=0))
    np.testing.assert_array_equal(vp[:, :i0], np.repeat(vp[:, i0:i0 + 1], i0, axis=1))


def test_rtm_image_is_zero_lag_crosscorrelation():
    """Brute force: image = sum over imaging steps of taper * u_bwd * S_fwd,
    illum = sum taper * S^2, from independently captured fields."""
    nz, nx, nt = 24, 40, 120
    g = O.make_grid(nz, nx, 10.0, 10.0, 1e-3, ntaper=10, snapi=4)
    vel = smooth_random(nz, nx, seed=3)
    w = ricker(25.0, nt, 1e-3)
    rx = np.linspace(20.0, 350.0, 12).astype(np.float32)
    rz = np.full_like(rx, 15.0)
    shot = O.Shot(15.0, 190.0, w, rz, rx)
    data = (np.random.default_rng(1).standard_normal((len(rx), nt)) * 1e-2)
    out = O.rtm(g, vel, shot, data, prec="d", capture=True)
    _, snaps = O.modeling(g, vel, shot, extract_kind=0, prec="d", snaps=True)
    cap = out["capture"]               # backward u at imaging steps, first = it=nt
    ntsnap = nt // 4
    tap = O.taper(out["nxw"], prec="d")
    img = np.zeros((out["nxw"], nz))
    ill = np.zeros_like(img)
    for c in range(ntsnap):
        idx = ntsnap - c - 1           # forward snapshot of the same time
        img += cap[c] * snaps[idx] * tap[:, None]
        ill += snaps[idx] ** 2 * tap[:, None]
    assert np.allclose(out["win_image"], img.T, rtol=1e-12, atol=1e-300)
    assert np.allclose(out["win_illum"], ill.T, rtol=1e-12, atol=1e-300)


def test_forward_linear_in_source():
    nz, nx, nt = 20, 30, 100
    g = O.make_grid(nz, nx, 10.0, 10.0, 1e-3, ntaper=10)
    vel = smooth_random(nz, nx, seed=4)
    # quantised so that 2*w1 - 3*w2 is exact in FP32 (the Shot stores FP32)
    w1 = np.round(ricker(20.0, nt, 1e-3) * 4096) / 4096
    w2 = np.round(ricker(30.0, nt, 1e-3) * 4096) / 4096
    rx, rz = np.array([50.0, 250.0], np.float32), np.array([30.0, 30.0], np.float32)
    t = [O.modeling(g, vel, O.Shot(40.0, 150.0, w, rz, rx), prec="d") for w in (w1, w2, 2 * w1 - 3 * w2)]
    assert np.allclose(t[2], 2 * t[0] - 3 * t[1], rtol=1e-10, atol=1e-12 * np.abs(t[0]).max())


def test_reciprocity_nearest():
    """Source-receiver reciprocity of the forward operator on grid points
    (Injection_simple-like sinc at shift 0 is a near delta)."""
    nz, nx, nt = 30, 40, 300
    g = O.make_grid(nz, nx, 10.0, 10.0, 1e-3, ntaper=12)
    vel = np.asfortranarray(np.full((nz, nx), 2000.0, np.float32))
    w = ricker(20.0, nt, 1e-3)
    a = O.modeling(g, vel, O.Shot(50.0, 100.0, w, [50.0, 80.0], [300.0, 300.0]), extract_kind=3, prec="d")
    b = O.modeling(g, vel, O.Shot(50.0, 300.0, w, [50.0, 80.0], [100.0, 100.0]), extract_kind=3, prec="d")
    # receiver 0 of shot A (at x=300) vs receiver 0 of shot B (at x=100): same pair swapped
    assert np.allclose(a[0], b[0], rtol=1e-6, atol=1e-6 * np.abs(a[0]).max())


def test_lbfgs_quadratic_finite_termination():
    """On an SPD quadratic, L-BFGS with the More-Thuente search converges to
    the minimiser; f decreases monotonically over accepted iterations."""
    rng = np.random.default_rng(0)
    n = 12
    Q = rng.standard_normal((n, n))
    A = Q @ Q.T + n * np.eye(n)
    bvec = rng.standard_normal(n)
    xs = np.linalg.solve(A, bvec)
    lb = O.LBFGS(n, m=5, boundt=0)
    x = np.zeros(n)
    fs = []
    for _ in range(200):
        f = 0.5 * x @ A @ x - bvec @ x
        gr = A @ x - bvec
        iflag, info = lb.step(x, f, gr, eps=1e-12)
        if info == 1:
            fs.append(f)
        if iflag <= 0:
            break
    assert np.linalg.norm(x - xs) / np.linalg.norm(xs) < 1e-6
    assert all(b <= a + 1e-12 for a, b in zip(fs, fs[1:]))


def test_higdon_absorbs_outgoing_energy():
    """Energy decay with the absorbing boundary: in a homogeneous model the
    wavefield energy over the interior, after the wavelet has left the
    60 x 80 grid, falls below 1e-5 of its peak (a reflecting edge would keep
    it near 1)."""
    nz, nx, nt = 60, 80, 900
    g = O.make_grid(nz, nx, 10.0, 10.0, 1e-3, ntaper=20, snapi=10)
    vel = np.full((nz, nx), 2000.0, np.float32)
    w = ricker(15.0, nt, 1e-3)
    rx = np.linspace(0.0, 790.0, 80).astype(np.float32)
    rz = np.full_like(rx, 100.0)
    _, wave = O.modeling(g, vel, O.Shot(300.0, 400.0, w, rz, rx), snaps=True)
    assert wave.shape[1:] == (nx, nz)
    E = (wave.astype(np.float64) ** 2).sum(axis=(1, 2))
    k = int(E.argmax())
    assert E[-1] < 1e-5 * E[k]
    # 400 ms after the peak the wavelet has crossed the 800 m model at 2000 m/s
    assert np.all(E[k + 40:] < 1e-4 * E[k])


def _frame(nz=16, nxw=16, ntaper=4, dx=10.0, dt=2.0 ** -9, v=2000.0):
    """Padded frame of the oracle's single step: centred coordinates (X, Z)
    of every padded cell, (NX, NZ) arrays, and the mask of FD cells."""
    g = O.make_grid(nz, nxw, dx, dx, dt, ntaper=ntaper)
    NZ, NX = nz + 2 * ntaper + 8, nxw + 2 * ntaper + 8
    z = (np.arange(NZ) - (ntaper + 3) - 1) * dx
    x = (np.arange(NX) - (ntaper + 3) - 1) * dx
    X, Z = np.meshgrid(x - x.mean(), z - z.mean(), indexing="ij")
    fd = np.zeros((NX, NZ), bool)
    fd[4:NX - 4, 4:NZ - 4] = True
    return g, nxw, X, Z, np.full((NX, NZ), v), fd


def _one_step_error(frame, f, dt):
    g, nxw, X, Z, vel, fd = frame
    u3 = O.test_step(g, nxw, vel, [f(X, Z, t) for t in (-3 * dt, -2 * dt, -dt, 0.0)])
    ex = f(X, Z, dt)
    return float(np.abs(u3 - ex)[fd].max() / np.abs(ex[fd]).max())


def test_time_scheme_exact_for_polynomials_of_degree_4():
    """u = (d.x - v t)^k solves the wave equation; the 8th-order stencil is
    exact for it and the 5-level update (20u2 - 6u1 - 4u0 + u_-1 + 12dt^2
    v^2 lap u2) / 11 reproduces polynomials of degree <= 4 in t, so one step
    of the oracle gives u(t + dt) to FP64 rounding for k <= 4, along x, z and
    an oblique direction.  k = 5 is not reproduced (the scheme's order), which
    shows the check can fail.  dt = 2^-9 s is exact in the grid's FP32 field."""
    dt, v = 2.0 ** -9, 2000.0
    fr = _frame(dt=dt, v=v)
    for cx, cz in ((1.0, 0.0), (0.0, -1.0), (0.6, 0.8)):
        for k in range(1, 6):
            e = _one_step_error(fr, lambda X, Z, t: (cx * X + cz * Z - v * t) ** k, dt)
            print("direction (%.1f, %.1f) k=%d: one-step relative error %.2e" % (cx, cz, k, e))
            if k <= 4:
                assert e <= 1e-13
            else:
                assert e >= 1e-9


def test_plane_wave_dispersion():
    """A plane wave cos(k.x - w t) with the continuous relation w = v |k|,
    advanced one step from its exact history: the one-step error over w dt
    estimates the scheme's relative phase-velocity error.  It must stay below
    1e-3 at 5 points per wavelength and 1e-4 at 10 (Courant 0.2), and fall as
    (w dt)^4 (the time scheme's order) as the sampling doubles.  A 1% error in
    the phase velocity, or a stencil coefficient off by a few percent, breaks
    the bounds."""
    dt, v, dx = 2.0 ** -10, 2000.0, 10.0
    fr = _frame(nz=40, nxw=40, dt=dt, v=v, dx=dx)
    err = {}
    for ppw in (5, 10, 20):
        kx = 2 * np.pi / (ppw * dx)
        kz = 0.5 * kx
        w = v * np.hypot(kx, kz)
        e = _one_step_error(fr, lambda X, Z, t: np.cos(kx * X + kz * Z - w * t), dt)
        err[ppw] = e / (w * dt)
        print("%d points per wavelength: phase-velocity error ~ %.2e" % (ppw, err[ppw]))
    assert err[5] < 1e-3 and err[10] < 1e-4
    assert 12.0 < err[10] / err[20] < 20.0

"""3D forward modeling: the oracle against the reference's own 3D kernels
(FD_2nd_3D_derivatives_scalar_forward_grid, Higdon with its y tables,
Boundary_3d with nbound 4, Injection_sinc -> Injection_source_sinc_xyz,
Extraction_* -> Extraction_1trace_sinc_xyz) composed in Propagator.f90 order
by oracle/ref_harness.f90, plus what the mathematics fixes."""
import numpy as np
import pytest

import oracle_lib as O
from models import ricker

needs_ref = pytest.mark.skipif(not O.ref_available(), reason="oracle/_ref not built (no reference tree)")


def rel(a, b):
    return float(np.linalg.norm((np.asarray(a, np.float64) - b).ravel()) /
                 max(np.linalg.norm(np.asarray(b, np.float64).ravel()), 1e-300))


def model3d(nz, nx, ny, seed=1):
    rng = np.random.default_rng(seed)
    z = np.arange(nz)[:, None, None] / nz
    x = np.arange(nx)[None, :, None] / nx
    y = np.arange(ny)[None, None, :] / ny
    v = 1600.0 + 700.0 * z + 150.0 * np.sin(2 * np.pi * (x + rng.uniform())) * np.cos(2 * np.pi * (y + rng.uniform()))
    return v.astype(np.float32)


def _case(nt=150, seed=1):
    nz, nx, ny = 20, 26, 22
    g = O.make_grid3d(nz, nx, ny, 10.0, 10.0, 10.0, 1e-3, ntaper=10, apx=60.0, apy=40.0)
    vel = model3d(nz, nx, ny, seed)
    w = ricker(30.0, nt, 1e-3)
    rx = np.array([33.0, 97.5, 141.0, 200.2, 77.7], np.float32)
    ry = np.array([51.0, 88.8, 120.1, 60.0, 150.3], np.float32)
    rz = np.array([23.0, 31.0, 47.3, 28.0, 55.5], np.float32)
    shot = O.Shot3D(44.4, 121.2, 99.9, w, rz, rx, ry)
    return g, vel, shot


@needs_ref
@pytest.mark.parametrize("ext", [0, 1, 2, 3])
def test_forward3d_matches_reference(ext):
    g, vel, shot = _case()
    vpad, win, b = O.padded_vel3d(g, vel, shot)
    ref = O.ref_propagate3d(g, vpad, win, shot, ext)
    got = O.modeling3d(g, vel, shot, extract_kind=ext)
    assert np.array_equal(got, ref), rel(got, ref)


def test_window3d_by_hand():
    """Source (x 121.2, y 99.9), receivers spanning x 33..200.2, y 51..150.3;
    apertures 60 (x) and 40 (y) about the midpoint centre."""
    g, vel, shot = _case()
    oxw, nxw, oyw, nyw = O.window3d(g, shot)
    cx = (33.0 + 121.2) / 2 + abs((200.2 + 121.2) / 2 - (33.0 + 121.2) / 2) / 2
    cy = (51.0 + 99.9) / 2 + abs((150.3 + 99.9) / 2 - (51.0 + 99.9) / 2) / 2
    assert abs(oxw - min(33.0, cx - 30.0)) < 1e-4 and abs(oyw - min(51.0, cy - 20.0)) < 1e-4
    assert nxw == int(np.floor((max(200.2, cx + 30.0) - oxw) / 10.0 + 0.5)) + 1
    assert nyw == int(np.floor((max(150.3, cy + 20.0) - oyw) / 10.0 + 0.5)) + 1


def test_forward3d_homogeneous_is_symmetric_and_travels_at_v():
    """Homogeneous medium, source at a grid point, receivers at equal distance
    along +x, -x, +y, -y and +z, -z: equal traces (the stencil is the same
    along every axis; a 30-cell pad keeps the edges' residuals out of the
    record), first peak at
    distance / v after the wavelet's delay.  Swapping two axes' stencil
    coefficients or dropping the y term breaks the equality."""
    nz = nx = ny = 31
    g = O.make_grid3d(nz, nx, ny, 10.0, 10.0, 10.0, 1e-3, ntaper=30)   # edges far: no residuals in nt
    vel = np.full((nz, nx, ny), 2000.0, np.float32)
    nt = 130
    w = ricker(30.0, nt, 1e-3)
    c = 150.0
    d = 80.0
    rz = np.array([c, c, c, c, c + d, c - d], np.float32)
    rx = np.array([c + d, c - d, c, c, c, c], np.float32)
    ry = np.array([c, c, c + d, c - d, c, c], np.float32)
    shot = O.Shot3D(c, c, c, w, rz, rx, ry)
    tr = O.modeling3d(g, vel, shot, extract_kind=1)
    for k in range(1, 6):
        assert rel(tr[k], tr[0]) < 1e-5, k
    t_peak = np.argmax(tr[0]) * 1e-3
    assert abs(t_peak - (1.2 / 30.0 + d / 2000.0)) < 3e-3


# ---- variable density (withRho) in 3D ----

def rho3d(nz, nx, ny, seed=11):
    rng = np.random.default_rng(seed)
    z = np.arange(nz)[:, None, None] / nz
    x = np.arange(nx)[None, :, None] / nx
    y = np.arange(ny)[None, None, :] / ny
    r = 1.8 + 0.5 * z + 0.3 * np.sin(2 * np.pi * (y + rng.uniform())) + 0.2 * np.cos(2 * np.pi * (x + rng.uniform()))
    return r.astype(np.float32)


@needs_ref
@pytest.mark.parametrize("ext", [0, 1, 2, 3])
def test_forward3d_density_matches_reference(ext):
    """The density stencil (FD_3D_derivatives_acoustic_forward_grid with
    allocate_sxx3d_syy3d_szz3d_delp3d, the 3D Interpolate of rho, coefpower 1)
    and the v^2 rho injection gain, against the compiled kernels; readsou's
    half-step shift (SEPlib, not compiled) is applied to the harness's
    wavelet by the oracle's pinned half_step."""
    g, vel, shot = _case(nt=140)
    rho = rho3d(g.nz, g.nx, g.ny)
    vpad, win, b = O.padded_vel3d(g, vel, shot)
    rpad, _, _ = O.padded_vel3d(g, rho, shot)
    w = O.half_step(shot.wavelet, g.lsinc)
    ref = O.ref_propagate3d(g, vpad, win, shot, ext, rho_pad=rpad, wavelet=w)
    got = O.modeling3d(g, vel, shot, extract_kind=ext, rho=rho)
    assert np.abs(ref).max() > 0
    assert np.array_equal(got, ref), rel(got, ref)


def test_rho_half3d_constant_and_separable():
    """A constant density interpolates to itself (the 9-point sinc sums to
    1); a density varying along one axis only is interpolated along that axis
    alone, and equals the 1D sinc at shift 0.5 with end extension applied to
    that profile: catches a pass over the wrong axis or a dropped y pass."""
    g, vel, shot = _case()
    vpad, win, b = O.padded_vel3d(g, vel, shot)
    shape = (b.NY, b.NX, b.NZ)   # C order of the column-major padded box
    d = O.rho_half3d(g, np.full(vpad.size, 2.0, np.float32), b)
    assert np.allclose(d, 0.5, rtol=1e-6, atol=0)   # three FP32 9-term sums
    w = np.zeros(9, np.float32)
    O.oracle().orc_f_mksinc.argtypes = [O.C.c_void_p, O.C.c_int, O.C.c_float]
    O.oracle().orc_f_mksinc(O.ptr(w), 9, 0.5)
    for ax, n in ((0, b.NY), (1, b.NX), (2, b.NZ)):
        prof = (1.5 + np.linspace(0, 1, n) ** 2).astype(np.float32)
        sh = [1, 1, 1]
        sh[ax] = n
        r = np.broadcast_to(prof.reshape(sh), shape).astype(np.float32)
        got = 1.0 / O.rho_half3d(g, np.ascontiguousarray(r).ravel(), b).reshape(shape).astype(np.float64)
        ext = np.concatenate([np.full(4, prof[0]), prof, np.full(4, prof[-1])]).astype(np.float64)
        want = np.array([np.dot(w, ext[i:i + 9]) for i in range(n)])
        want_b = np.broadcast_to(want.reshape(sh), shape)
        assert np.allclose(got, want_b, rtol=1e-6), ax


def test_forward3d_density_scaling_is_exact():
    """Doubling rho doubles the gathers bit for bit: the operator
    v^2 rho div(1/rho grad) is invariant and the v^2 rho source gain doubles
    (every rounding commutes with a power-of-two scale)."""
    g, vel, shot = _case(nt=100)
    rho = rho3d(g.nz, g.nx, g.ny)
    a = O.modeling3d(g, vel, shot, extract_kind=0, rho=rho)
    b = O.modeling3d(g, vel, shot, extract_kind=0, rho=2 * rho)
    assert np.array_equal(b, 2 * a)


def test_forward3d_density_x_and_y_axes_agree():
    """A model whose velocity and density vary along x, and the same model
    turned to vary along y with the geometry turned with it (dx = dy, every
    position on the grid so the x-sinc y weights are the delta): the traces
    agree to FP32 rounding.  A y term with the wrong coefficients, gradient
    or padding rule fails this."""
    nz, n = 18, 24
    g = O.make_grid3d(nz, n, n, 10.0, 10.0, 10.0, 1e-3, ntaper=10)
    x = np.arange(n) / n
    v1 = (1800.0 + 400.0 * np.sin(2 * np.pi * x)).astype(np.float32)
    r1 = (2.0 + 0.6 * np.cos(2 * np.pi * x)).astype(np.float32)
    vel_x = np.broadcast_to(v1[None, :, None], (nz, n, n)).astype(np.float32)
    rho_x = np.broadcast_to(r1[None, :, None], (nz, n, n)).astype(np.float32)
    vel_y = np.ascontiguousarray(vel_x.transpose(0, 2, 1))
    rho_y = np.ascontiguousarray(rho_x.transpose(0, 2, 1))
    nt = 120
    w = ricker(30.0, nt, 1e-3)
    rz = np.array([40.0, 60.0, 90.0], np.float32)
    ra = np.array([50.0, 120.0, 200.0], np.float32)
    rb = np.array([110.0, 70.0, 150.0], np.float32)
    tx = O.modeling3d(g, vel_x, O.Shot3D(50.0, 130.0, 100.0, w, rz, ra, rb), extract_kind=1, rho=rho_x)
    ty = O.modeling3d(g, vel_y, O.Shot3D(50.0, 100.0, 130.0, w, rz, rb, ra), extract_kind=1, rho=rho_y)
    assert np.abs(tx).max() > 0
    assert rel(ty, tx) < 2e-5, rel(ty, tx)


# ---- 3D RTM (3DRTM.x with genpar%twoD false) ----

def rtm_case(density=False, vel_bwd=False, snapi=2):
    nz, nx, ny = 20, 26, 22
    g = O.make_grid3d(nz, nx, ny, 10.0, 10.0, 10.0, 1e-3, ntaper=10, apx=60.0, apy=40.0, snapi=snapi)
    vel = model3d(nz, nx, ny, 1)
    rho = rho3d(nz, nx, ny) if density else None
    vb = (vel * 1.03).astype(np.float32) if vel_bwd else None
    w = ricker(30.0, 120, 1e-3)
    rx = np.array([33.0, 97.5, 141.0, 200.2, 77.7, 120.0], np.float32)
    ry = np.array([51.0, 88.8, 120.1, 60.0, 150.3, 99.0], np.float32)
    rz = np.array([23.0, 31.0, 47.3, 28.0, 55.5, 40.0], np.float32)
    shot = O.Shot3D(44.4, 121.2, 99.9, w, rz, rx, ry)
    tr = O.modeling3d(g, (vel * 0.97).astype(np.float32), shot, extract_kind=0, rho=rho)
    return g, vel, vb, rho, shot, tr


@needs_ref
@pytest.mark.parametrize("density,vel_bwd", [(False, False), (False, True), (True, False)])
def test_rtm3d_fields_match_reference(density, vel_bwd):
    """The forward snapshots and the backward u2 at every imaging step,
    bitwise against the compiled kernels run forward (source, Extraction_
    wavefield) and in reverse time (the receivers' traces by Injection_sinc,
    the same forward stencil, the backward model's own Higdon tables)."""
    g, vel, vb, rho, shot, tr = rtm_case(density, vel_bwd)
    r = O.rtm3d(g, vel, shot, tr, vel_bwd=vb, rho=rho)
    ns = shot.nt // g.snapi
    vpad, win, b = O.padded_vel3d(g, vel, shot)
    rpad = None if rho is None else O.padded_vel3d(g, rho, shot)[0]
    w = shot.wavelet if rho is None else O.half_step(shot.wavelet, g.lsinc)
    _, snaps = O.ref_propagate3d(g, vpad, win, shot, 0, rho_pad=rpad, wavelet=w, snapi=g.snapi, nsnap=ns)
    assert np.array_equal(snaps, r["snaps"])
    vbpad = vpad if vb is None else O.padded_vel3d(g, vb, shot)[0]
    _, back = O.ref_propagate3d(g, vbpad, win, shot, 0, rho_pad=rpad, sources=(shot.rz, shot.rx, shot.ry, tr),
                                tdir=-1, snapi=g.snapi, nsnap=ns)
    assert np.abs(back).max() > 0
    assert np.array_equal(back, r["capture"])


def test_rtm3d_image_is_the_zero_lag_sum():
    """The window image and illumination are the sums over imaging steps of
    u_b S_index taperx(j) tapery(k) and S^2 taperx tapery (index counting back
    from the last snapshot), and the model image is the window added at the
    nearest model columns and lines."""
    g, vel, vb, rho, shot, tr = rtm_case(snapi=3)
    r = O.rtm3d(g, vel, shot, tr)
    ns = len(r["snaps"])
    oxw, nxw, oyw, nyw = O.window3d(g, shot)
    tx = O.taper(nxw).astype(np.float32)
    ty = O.taper(nyw).astype(np.float32)
    t = (tx[None, :] * ty[:, None])[:, :, None].astype(np.float64)
    S = r["snaps"].astype(np.float64)
    img = sum(r["capture"][c].astype(np.float64) * S[ns - 1 - c] * t for c in range(ns))
    ill = sum(S[k] ** 2 * t for k in range(ns))
    assert rel(r["win_image"], img) < 1e-5 and rel(r["win_illum"], ill) < 1e-5
    full = np.zeros((g.nz, g.nx, g.ny))
    for k in range(nyw):
        iy = int(np.floor((k * g.dy + oyw - g.oy) / g.dy + 0.5))
        for j in range(nxw):
            ix = int(np.floor((j * g.dx + oxw - g.ox) / g.dx + 0.5))
            if 0 <= ix < g.nx and 0 <= iy < g.ny:
                full[:, ix, iy] += r["win_image"][k, j]
    assert rel(r["image"], full) < 1e-6
    # deterministic: a second call gives the same image
    r2 = O.rtm3d(g, vel, shot, tr)
    assert np.array_equal(r2["image"], r["image"])


# ---- 3DAFWI's gradient in 3D ----

def bump3(shape, c, s):
    zz, xx, yy = np.mgrid[0:shape[0], 0:shape[1], 0:shape[2]]
    return np.exp(-((zz - c[0]) ** 2 + (xx - c[1]) ** 2 + (yy - c[2]) ** 2) / (2 * s ** 2))


def afwi3d_case(snapi=1, density=False, fpeak=25.0):
    nz, nx, ny = 24, 30, 26
    g = O.make_grid3d(nz, nx, ny, 10.0, 10.0, 10.0, 1e-3, ntaper=10, apx=300.0, apy=300.0, snapi=snapi)
    w = ricker(fpeak, 220, 1e-3)
    rx, ry = np.meshgrid(np.arange(3, nx - 3, 3) * 10.0, np.arange(3, ny - 3, 3) * 10.0)
    rx, ry = rx.ravel().astype(np.float32), ry.ravel().astype(np.float32)
    shots = [O.Shot3D(20.0, 150.0, 130.0, w, np.full_like(rx, 20.0), rx, ry)]
    v0 = np.full((nz, nx, ny), 2000.0, np.float32)
    vt = v0.copy()
    vt[12:18, 8:22, 8:18] += 150.0
    rho0 = rt = None
    if density:
        rho0 = np.full((nz, nx, ny), 2.0, np.float32)
        rt = rho0.copy()
        rt[15:20, 6:20, 8:20] += 0.3
    dobs = [O.modeling3d(g, vt, s, extract_kind=3, rho=rt) for s in shots]
    return g, v0, rho0, shots, dobs


@pytest.mark.parametrize("snapi", [1, 2])
def test_afwi3d_gradient_constant_density(snapi):
    """dJ / <grad, dv> = -2 snapi in a homogeneous model (the 2D pin's
    relation; FP32 here, so 2%)."""
    g, v0, _, shots, dobs = afwi3d_case(snapi)
    r = O.afwi3d(g, v0, shots, dobs)
    J = lambda v: O.afwi3d(g, v, shots, dobs)["f"]
    for c in ((14, 12, 12), (15, 18, 14), (18, 15, 10)):
        dv = bump3(v0.shape, c, 2.5).astype(np.float32)
        D = (J(v0 + dv) - J(v0 - dv)) / 2
        q = D / float(np.sum(r["grad"][0].astype(np.float64) * dv))
        assert abs(q + 2 * snapi) < 0.02 * 2 * snapi, (c, q)


def test_afwi3d_rhovp_vp_gradient():
    """With a density the vp gradient carries d(v^2)/dv's 2: ratio -snapi."""
    g, v0, rho0, shots, dobs = afwi3d_case(1, density=True, fpeak=15.0)
    r = O.afwi3d(g, v0, shots, dobs, rho=rho0)
    J = lambda v: O.afwi3d(g, v, shots, dobs, rho=rho0)["f"]
    for c in ((14, 12, 12), (18, 15, 10)):
        dv = bump3(v0.shape, c, 2.5).astype(np.float32)
        D = (J(v0 + dv) - J(v0 - dv)) / 2
        q = D / float(np.sum(r["grad"][0].astype(np.float64) * dv))
        assert abs(q + 1.0) < 0.06, (c, q)


def test_afwi3d_rhovp_y_term_mirrors_x_term():
    """A model varying along x, and the same model turned to vary along y with
    the geometry turned with it (dx = dy, positions on the grid so the x-sinc
    y weights are the delta): both gradients come out turned, to FP32
    rounding.  The 2D pin fixes the x and z terms of the v-rho condition; this
    fixes its y term (dyi, kp/km) and the 3D bookkeeping to match them."""
    nz, n = 18, 22
    g = O.make_grid3d(nz, n, n, 10.0, 10.0, 10.0, 1e-3, ntaper=10, apx=300.0, apy=300.0, snapi=2)
    x = np.arange(n) / n
    v1 = (1900.0 + 300.0 * np.sin(2 * np.pi * x)).astype(np.float32)
    r1 = (2.0 + 0.5 * np.cos(2 * np.pi * x)).astype(np.float32)
    vx = np.ascontiguousarray(np.broadcast_to(v1[None, :, None], (nz, n, n))).astype(np.float32)
    rx_ = np.ascontiguousarray(np.broadcast_to(r1[None, :, None], (nz, n, n))).astype(np.float32)
    vy = np.ascontiguousarray(vx.transpose(0, 2, 1))
    ry_ = np.ascontiguousarray(rx_.transpose(0, 2, 1))
    w = ricker(20.0, 160, 1e-3)
    ra = np.array([40.0, 90.0, 160.0, 120.0], np.float32)
    rb = np.array([60.0, 140.0, 80.0, 110.0], np.float32)   # window centres on the grid in x and y
    rz = np.full(4, 30.0, np.float32)
    sx = O.Shot3D(20.0, 100.0, 120.0, w, rz, ra, rb)
    sy = O.Shot3D(20.0, 120.0, 100.0, w, rz, rb, ra)
    dx = [O.modeling3d(g, (vx * 0.95).astype(np.float32), sx, extract_kind=3, rho=rx_)]
    dy = [O.modeling3d(g, (vy * 0.95).astype(np.float32), sy, extract_kind=3, rho=ry_)]
    gx = O.afwi3d(g, vx, [sx], dx, rho=rx_)
    gy = O.afwi3d(g, vy, [sy], dy, rho=ry_)
    assert abs(gy["f"] - gx["f"]) < 1e-5 * abs(gx["f"])
    for k in range(2):
        a, b = gx["grad"][k], gy["grad"][k].transpose(0, 2, 1)
        assert np.abs(a).max() > 0
        assert rel(b, a) < 1e-4, (k, rel(b, a))


def test_afwi3d_vprho_change_of_variables():
    g, v0, rho0, shots, dobs = afwi3d_case(1, density=True, fpeak=15.0)
    rho0 = (rho0 + 0.1 * np.sin(np.arange(v0.size).reshape(v0.shape) * 0.01)).astype(np.float32)
    a = O.afwi3d(g, v0, shots, dobs, rho=rho0, vprho=0)
    b = O.afwi3d(g, v0, shots, dobs, rho=rho0, vprho=1)
    assert np.array_equal(b["grad"][1], ((1.0 / v0).astype(np.float32) * a["grad"][1]).astype(np.float32))
    assert np.array_equal(b["grad"][0], (a["grad"][0] - (rho0 / v0).astype(np.float32) * a["grad"][1]).astype(np.float32))


# ---- 3D LSRTM and Born ----

def second_diff32(w, dt):
    o = np.asarray(w, np.float32)
    out = np.zeros_like(o)
    out[1:-1] = ((np.float32(-2.0) * o[1:-1] + o[:-2]) + o[2:]) / np.float32(np.float32(dt) * np.float32(dt))
    return out


@needs_ref
@pytest.mark.parametrize("vel_bwd", [False, True])
def test_lsrtm3d_fields_match_reference(vel_bwd):
    """LSRTM's reverse pass in 3D: the adjoint-style stencil
    (FD_2nd_3D_derivatives_scalar_adjoint_grid) with Injection_LSRTM_sinc of
    the receivers' traces, on the backward model when given: the u2 at every
    imaging step bitwise against the compiled kernels; the image the
    zero-lag sum 2 t u S / v with S the snapshots of the second-differenced
    source, also from the compiled kernels."""
    g, vel, vb, _, shot, tr = rtm_case(False, vel_bwd)
    r = O.lsrtm3d(g, vel, shot, tr, vel_bwd=vb)
    ns = shot.nt // g.snapi
    vpad, win, b = O.padded_vel3d(g, vel, shot)
    vbpad = vpad if vb is None else O.padded_vel3d(g, vb, shot)[0]
    _, back = O.ref_propagate3d(g, vbpad, win, shot, 0, sources=(shot.rz, shot.rx, shot.ry, tr), tdir=-1,
                                snapi=g.snapi, nsnap=ns, stencil=1, inj=3)
    assert np.abs(back).max() > 0
    assert np.array_equal(back, r["capture"])
    _, snaps = O.ref_propagate3d(g, vpad, win, shot, 0, wavelet=second_diff32(shot.wavelet, g.dt), snapi=g.snapi,
                                 nsnap=ns)
    oxw, nxw, oyw, nyw = win
    tx = O.taper(nxw).astype(np.float32)
    ty = O.taper(nyw).astype(np.float32)
    t = (tx[None, :] * ty[:, None])[:, :, None].astype(np.float64)
    vw = (vb if vb is not None else vel).astype(np.float64)
    # the window's velocity (nearest model column / line) as [nyw, nxw, nz]
    cols = [min(max(int(np.floor((j * g.dx + oxw - g.ox) / g.dx + 0.5)), 0), g.nx - 1) for j in range(nxw)]
    lins = [min(max(int(np.floor((k * g.dy + oyw - g.oy) / g.dy + 0.5)), 0), g.ny - 1) for k in range(nyw)]
    vwin = np.stack([np.stack([vw[:, c, l] for c in cols]) for l in lins])
    img = sum(2 * t * back[c].astype(np.float64) * snaps[ns - 1 - c] / vwin for c in range(ns))
    full = np.zeros((g.nz, g.nx, g.ny))
    for k in range(nyw):
        iy = int(np.floor((k * g.dy + oyw - g.oy) / g.dy + 0.5))
        for j in range(nxw):
            ix = int(np.floor((j * g.dx + oxw - g.ox) / g.dx + 0.5))
            if 0 <= ix < g.nx and 0 <= iy < g.ny:
                full[:, ix, iy] += img[k, j]
    assert rel(r["image"], full) < 1e-5


def test_born3d_is_the_derivative_of_modeling():
    """As in 2D: born(m = dv) equals the derivative of modeling (LSRTM sinc
    extraction) along dv up to the scheme's own time operator, O((w dt)^2);
    a factor 2 or 1/v^2 slip fails; with snapi 2 the Born source is added on
    snapshot steps only, not rescaled (half the data to first order)."""
    nz, nx, ny = 20, 26, 22
    vel = model3d(nz, nx, ny, 1)
    zz, xx, yy = np.mgrid[0:nz, 0:nx, 0:ny]
    dv = (40.0 * np.exp(-((zz - 13) ** 2 + (xx - 14) ** 2 + (yy - 11) ** 2) / 8.0)).astype(np.float32)
    w = ricker(25.0, 200, 1e-3)
    rx = np.array([60.0, 120.0, 180.0, 100.0], np.float32)
    ry = np.array([80.0, 110.0, 90.0, 140.0], np.float32)
    rz = np.full(4, 25.0, np.float32)
    shot = O.Shot3D(30.0, 121.0, 99.0, w, rz, rx, ry)
    # apertures wide enough that the perturbation lies inside the imaging
    # taper's flat part (Injection_Born multiplies by taperx tapery)
    out = {}
    for snapi in (1, 2):
        g = O.make_grid3d(nz, nx, ny, 10.0, 10.0, 10.0, 1e-3, ntaper=10, apx=500.0, apy=500.0, snapi=snapi)
        out[snapi] = O.born3d(g, vel, dv, shot)
    g = O.make_grid3d(nz, nx, ny, 10.0, 10.0, 10.0, 1e-3, ntaper=10, apx=500.0, apy=500.0)
    eps = 0.5
    D = (O.modeling3d(g, (vel + eps * dv).astype(np.float32), shot, extract_kind=2).astype(np.float64) -
         O.modeling3d(g, (vel - eps * dv).astype(np.float32), shot, extract_kind=2)) / (2 * eps)
    B = out[1].astype(np.float64)
    assert rel(B, D) < 1e-2, rel(B, D)
    assert rel(2 * B, D) > 0.5
    assert rel(O.born3d(g, vel, (dv / vel).astype(np.float32), shot), D) > 0.5
    assert rel(2 * out[2].astype(np.float64), D) < 3e-2

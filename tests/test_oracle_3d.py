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

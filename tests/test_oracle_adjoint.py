"""Pin of the oracle's exact discrete adjoint of forward modeling by the
dot-product test in FP64: <F s, d> = <s, F^T d> for random s, d, where F is
the oracle's forward modeling (itself bitwise equal to the reference's
kernels).  A dropped Higdon term, a wrong band order, a missing time-level
coefficient or a transposed stencil would break the identity."""
import numpy as np
import pytest

import oracle_lib as O
from models import smooth_random


@pytest.mark.parametrize("ext", [0, 1, 2, 3])
def test_modeling_adjoint_dot_product_fp64(ext):
    nz, nx, nt = 30, 44, 260
    vel = smooth_random(nz, nx, seed=3 + ext, vmin=1500, vmax=2600)
    g = O.make_grid(nz, nx, 10.0, 10.0, 1e-3, ntaper=12)
    rng = np.random.default_rng(ext)
    rx = np.linspace(12.3, 10.0 * (nx - 2), 9).astype(np.float32)
    rz = np.full_like(rx, 33.7)
    s = rng.standard_normal(nt)
    shot = O.Shot(26.4, 201.7, s.astype(np.float32), rz, rx)
    d = O.modeling(g, vel, shot, extract_kind=ext, prec="d")
    db = rng.standard_normal(d.shape)
    sb = O.modeling_adjoint(g, vel, shot, db, extract_kind=ext, prec="d")
    # the forward reads the FP32 wavelet; use exactly that s in the dot product
    lhs = float(np.sum(d * db))
    rhs = float(np.sum(np.asarray(shot.wavelet, np.float64) * sb))
    print("ext=%d <Fs,d>=%.15e <s,F^T d>=%.15e rel=%.2e" % (ext, lhs, rhs, abs(lhs - rhs) / abs(lhs)))
    assert abs(lhs - rhs) <= 1e-12 * abs(lhs)


def test_modeling_adjoint_sees_the_boundary():
    """Long record: energy reaches the Higdon bands, so the identity above
    also exercises their transpose (a mistake there shows at this length)."""
    nz, nx, nt = 24, 30, 900
    vel = np.full((nz, nx), 2000.0, np.float32)
    g = O.make_grid(nz, nx, 10.0, 10.0, 1e-3, ntaper=10)
    rng = np.random.default_rng(9)
    rx = np.array([45.0, 150.0, 233.0], np.float32)
    rz = np.array([20.0, 180.0, 100.0], np.float32)
    shot = O.Shot(110.0, 140.0, rng.standard_normal(nt).astype(np.float32), rz, rx)
    d = O.modeling(g, vel, shot, extract_kind=0, prec="d")
    db = rng.standard_normal(d.shape)
    sb = O.modeling_adjoint(g, vel, shot, db, extract_kind=0, prec="d")
    lhs, rhs = float(np.sum(d * db)), float(np.sum(np.asarray(shot.wavelet, np.float64) * sb))
    print("<Fs,d>=%.15e <s,F^T d>=%.15e rel=%.2e" % (lhs, rhs, abs(lhs - rhs) / abs(lhs)))
    assert abs(lhs - rhs) <= 1e-10 * abs(lhs)   # 900 FP64 steps: rounding ~1e-12


@pytest.mark.parametrize("snapi,aperture", [(1, 0.0), (4, 0.0), (2, 900.0)])
def test_born_adjoint_dot_product_fp64(snapi, aperture):
    """<B m, d> = <m, B^T d> for the oracle's Born modeling B (background
    pass with the second-differenced wavelet, Injection_Born every snapi
    steps from the stored snapshots, LSRTM sinc extraction).  A wide
    aperture puts window columns outside the model, so the transpose of the
    clamped nearest-column window gather is exercised too."""
    nz, nx, nt = 28, 40, 240
    vel = smooth_random(nz, nx, seed=7, vmin=1500, vmax=2600)
    g = O.make_grid(nz, nx, 10.0, 10.0, 1e-3, ntaper=12, snapi=snapi, aperture=aperture)
    rng = np.random.default_rng(snapi)
    rx = np.linspace(31.3, 10.0 * (nx - 4), 11).astype(np.float32)
    rz = np.full_like(rx, 27.2)
    w = np.zeros(nt, np.float32)
    w[5:60] = np.sin(np.linspace(0, 3 * np.pi, 55)) * np.hanning(55)
    shot = O.Shot(36.4, 181.7, w, rz, rx)
    m = rng.standard_normal((nz, nx))
    d = O.born(g, vel, np.asfortranarray(m), shot, prec="d")
    db = rng.standard_normal(d.shape)
    mb = O.born_adjoint(g, vel, shot, db, prec="d")
    lhs, rhs = float(np.sum(d * db)), float(np.sum(m * mb))
    print("snapi=%d aperture=%g <Bm,d>=%.15e <m,B^T d>=%.15e rel=%.2e" % (snapi, aperture, lhs, rhs,
                                                                        abs(lhs - rhs) / abs(lhs)))
    assert abs(lhs) > 0 and abs(lhs - rhs) <= 1e-12 * abs(lhs)

"""The device FWI gradient (sfd_fwi_gradient) against central finite
differences of the device misfit: the same check the oracle passes in FP64
(test_oracle_gradient.py), here in FP32 with FP32 and FP16 snapshots.  With
illupow = 0 and no smoothing, grad = -dJ/dv / (snapi (1 + 1e-4))."""
import numpy as np
import pytest

from models import smooth_random, ricker

sa = pytest.importorskip("seismic_amd")
pytestmark = pytest.mark.gpu

BUMPS = ((23, 22), (22, 42), (30, 32), (15, 32))


@pytest.mark.parametrize("snap_mode,tol", [(0, 5e-3), (1, 5e-3)], ids=["fp32", "fp16s"])
def test_device_gradient_matches_finite_differences(snap_mode, tol):
    print("case: device gradient vs finite differences, snap_mode %d" % snap_mode, flush=True)
    nz, nx, nt, snapi = 40, 64, 700, 1
    e = sa.Engine(sa.make_grid(nz, nx, 10.0, 10.0, 1e-3, ntaper=12, snapi=snapi))
    v0 = smooth_random(nz, nx, seed=2, vmin=1800, vmax=2400).astype(np.float32)
    vt = v0.copy()
    vt[18:26, 20:44] += 150.0
    w = ricker(15.0, nt, 1e-3)
    rx = (np.arange(2, nx - 2, 2) * 10.0).astype(np.float32)
    shots = [sa.Shot(20.0, sx, w, np.full(len(rx), 20.0, np.float32), rx) for sx in (150.0, 470.0)]
    e.set_model(vt)
    dobs = e.forward(shots, extract_kind=1)

    def fg(v):
        e.set_model(v)
        f, grad, _, _ = e.fwi_gradient(shots, dobs, illupow=0.0, snap_mode=snap_mode)
        return f, grad

    _, grad = fg(v0)
    want = -snapi * (1.0 + 1e-4)
    zz, xx = np.mgrid[0:nz, 0:nx]
    eps = 5.0
    for cz, cx in BUMPS:
        dv = (eps * np.exp(-((zz - cz) ** 2 + (xx - cx) ** 2) / (2 * 3.0 ** 2))).astype(np.float32)
        D = (fg(v0 + dv)[0] - fg(v0 - dv)[0]) / 2.0
        P = float(np.sum(grad.astype(np.float64) * dv))
        print("bump (%d,%d): dJ = %.6e, <grad, dv> = %.6e, ratio %.6f" % (cz, cx, D, P, D / P), flush=True)
        assert abs(D / P - want) <= tol * abs(want)


def test_gradient_without_residual_download_is_identical():
    """residuals=False leaves the residual gathers on the device (as the
    evaluations of sfd_fwi_lbfgs do): misfit, gradient and illumination are
    the same to the bit (20 ragged shots, planned on several host threads)."""
    print("case: gradient without residual download", flush=True)
    nz, nx, nt = 40, 96, 300
    e = sa.Engine(sa.make_grid(nz, nx, 10.0, 10.0, 1e-3, ntaper=12, snapi=2))
    v0 = smooth_random(nz, nx, seed=4, vmin=1800, vmax=2400).astype(np.float32)
    w = ricker(15.0, nt, 1e-3)
    shots = []
    for k, sx in enumerate(np.linspace(100.0, 850.0, 20)):
        rx = (np.arange(2 + k % 3, nx - 2, 3) * 10.0).astype(np.float32)
        shots.append(sa.Shot(20.0, float(sx), w[: nt - 7 * k], np.full(len(rx), 20.0, np.float32), rx))
    e.set_model((v0 * 1.03).astype(np.float32))
    dobs = e.forward(shots, extract_kind=1)
    e.set_model(v0)
    f1, g1, i1, r1 = e.fwi_gradient(shots, dobs, snap_mode=1)
    f2, g2, i2, r2 = e.fwi_gradient(shots, dobs, snap_mode=1, residuals=False)
    assert r2 is None and len(r1) == len(shots)
    assert f1 == f2 and np.array_equal(g1, g2) and np.array_equal(i1, i2)
    assert np.abs(g1).max() > 0 and all(np.abs(r).max() > 0 for r in r1)

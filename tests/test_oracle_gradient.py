"""Gradient check of the oracle's FWI gradient (compute_fct_gdt,
progs/afwi/compute_function_gradient.f90:208-350) against central finite
differences of its own misfit, in FP64.

With illupow = 0 the illumination is 1 everywhere, so the normalisation
(:332-343) divides by 1 + 1e-4, and with rect 1 nothing is smoothed.  The
imaging condition sums u_b (S(k-1) - 2 S(k) + S(k+1)) / (snapi dt)^2 / v^3
over every snapi-th step, which makes the reference's grad equal to
-dJ/dv / (snapi (1 + 1e-4)): negative (l_bfgs passes -grad to LBFGS,
Inversion.f90:128) and scaled by the imaging's step.  Smooth bumps in the
illuminated interior, away from the source and the window's taper, must
show that ratio; a wrong sign, power of v, factor 2 or snapshot step would
move it by far more than the tolerance."""
import numpy as np
import pytest

import oracle_lib as O
from models import smooth_random, ricker

BUMPS = ((23, 22), (22, 42), (30, 32), (15, 32))


def problem(snapi):
    nz, nx, nt = 40, 64, 700
    g = O.make_grid(nz, nx, 10.0, 10.0, 1e-3, ntaper=12, snapi=snapi)
    v0 = smooth_random(nz, nx, seed=2, vmin=1800, vmax=2400).astype(np.float64)
    vt = v0.copy()
    vt[18:26, 20:44] += 150.0
    w = ricker(15.0, nt, 1e-3)
    rx = (np.arange(2, nx - 2, 2) * 10.0).astype(np.float32)
    shots = [O.Shot(20.0, sx, w, np.full(len(rx), 20.0, np.float32), rx) for sx in (150.0, 470.0)]
    dobs = [O.modeling(g, vt, s, extract_kind=1, prec="d") for s in shots]
    return g, v0, shots, dobs


def bump(nz, nx, cz, cx):
    zz, xx = np.mgrid[0:nz, 0:nx]
    return np.exp(-((zz - cz) ** 2 + (xx - cx) ** 2) / (2 * 3.0 ** 2))


@pytest.mark.parametrize("snapi,tol", [(1, 2e-3), (2, 4e-3)])
def test_fwi_gradient_matches_finite_differences(snapi, tol):
    g, v0, shots, dobs = problem(snapi)
    J = lambda v: O.fwi_gradient(g, v, shots, dobs, illupow=0.0, prec="d")["f"]
    grad = O.fwi_gradient(g, v0, shots, dobs, illupow=0.0, prec="d")["grad"]
    want = -snapi * (1.0 + 1e-4)
    for cz, cx in BUMPS:
        dv = bump(g.nz, g.nx, cz, cx)
        D = (J(v0 + dv) - J(v0 - dv)) / 2.0
        P = float(np.sum(grad * dv))
        print("snapi %d bump (%d,%d): dJ = %.6e, <grad, dv> = %.6e, ratio %.6f" % (snapi, cz, cx, D, P, D / P))
        assert abs(D / P - want) <= tol * abs(want)

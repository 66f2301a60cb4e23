"""Pins of the oracle's L-BFGS restatements against the reference's own
compiled routines, and closed-form checks of what the out-of-core variant
adds.

- orc_lbfgsoc (nlinv_lbfgs.f LBFGSOC: per-parameter clip FREEZE_XOC, STP1 =
  |x|/|g|/100 or 1/|g| or step_init) against the reference LBFGSOC built in
  place from libs/libnlinv/nlinv_lbfgs.f (its state written to and read back
  from its text files every call), driven as progs/inversion/L_BFGS.f90 does:
  one call per evaluation, x, f and g passing through FP32 files between calls.
- orc_lbfgs (libs/libinv/lbfgs.f LBFGS) against the reference LBFGS.
Both sum their dot products serially (the reference's OpenMP reductions run
on one thread here), so the iterates must agree bitwise.  These pins found
two constants of the reference that a restatement easily misses: MCSTEP's
safeguard factor 0.66 is a default-REAL literal (lbfgs.f:1203), and libinv's
MCSRCH extrapolates up to XTRAPF = 4.0D1 (lbfgs.f:792) where MCSRCHOC uses
4.0D0 (nlinv_lbfgs.f:752).
"""
import numpy as np
import pytest

import oracle_lib as O


def rosen(x):
    f = np.sum(100.0 * (x[1::2] - x[::2] ** 2) ** 2 + (1 - x[::2]) ** 2)
    g = np.zeros_like(x)
    g[::2] = -400.0 * x[::2] * (x[1::2] - x[::2] ** 2) - 2 * (1 - x[::2])
    g[1::2] = 200.0 * (x[1::2] - x[::2] ** 2)
    return f, g


def bounded_quad(n=40, seed=5):
    rng = np.random.default_rng(seed)
    Q = rng.standard_normal((n, n))
    A = Q @ Q.T / n + np.eye(n)
    xs = np.concatenate([2.0 + rng.standard_normal(n // 2), -1.0 + rng.standard_normal(n - n // 2)])

    def fg(x):
        r = x - xs
        return 0.5 * r @ A @ r, A @ r
    return fg


def drive(opt, fg, x0, fp32_files, nmax=300):
    """Calls until IFLAG <= 0; returns the per-call (iflag, info, x) record."""
    x = x0.copy()
    rec = []
    for _ in range(nmax):
        f, g = fg(x)
        if fp32_files:   # L_BFGS.f90: fct, gdt and mod are FP32 SEP files
            f, g = float(np.float32(f)), g.astype(np.float32).astype(np.float64)
        iflag, info = opt.step(x, f, g)
        if fp32_files:
            x[:] = x.astype(np.float32)
        rec.append((iflag, info, x.copy()))
        if iflag <= 0:
            break
    return rec


OC_CASES = [
    # (problem, nparam bounds, stp1_opt, step_init, boundt, fp32 files)
    ("quad", ((0.5, -2.0), (3.0, 0.3)), 1, 999999.0, 2, True),
    ("quad", ((0.5, -2.0), (3.0, 0.3)), 0, 999999.0, 1, False),
    ("quad", ((-1e30,), (1e30,)), 1, 0.05, 0, False),
    ("rosen", ((-1.5,), (1.5,)), 1, 999999.0, 2, False),
    ("rosen", ((-1e30,), (1e30,)), 0, 999999.0, 0, True),
]


@pytest.mark.skipif(not O.nlref_available(), reason="reference LBFGSOC not built (oracle/_ref)")
@pytest.mark.parametrize("case", OC_CASES, ids=lambda c: "%s-stp1opt%d-boundt%d-fp32%d" % (c[0], c[2], c[4], c[5]))
def test_lbfgsoc_oracle_matches_reference(case, tmp_path):
    prob, (xmin, xmax), stp1_opt, step_init, boundt, fp32 = case
    if prob == "quad":
        fg, n = bounded_quad(), 40
        x0 = np.full(n, 1.0)
    else:
        fg, n = rosen, 20
        x0 = np.tile([-1.2, 1.0], n // 2)
    kw = dict(m=5, xmin=xmin, xmax=xmax, stp1_opt=stp1_opt, step_init=step_init, boundt=boundt)
    ro = drive(O.RefLBFGSOC(n, str(tmp_path), **kw), fg, x0, fp32)
    oo = drive(O.LBFGSOC(n, **kw), fg, x0, fp32)
    assert len(ro) == len(oo) and len(ro) >= 8
    for k, ((f1, i1, x1), (f2, i2, x2)) in enumerate(zip(ro, oo)):
        assert (f1, i1) == (f2, i2), "call %d" % k
        assert np.array_equal(x1, x2), "call %d" % k
    print("%s: %d calls, final iflag %d, accepted %d" % (case, len(ro), ro[-1][0], sum(r[1] == 1 for r in ro)))


@pytest.mark.skipif(not O.ref_available(), reason="reference library not built (oracle/_ref)")
@pytest.mark.parametrize("boundt", [0, 1, 2])
@pytest.mark.parametrize("freeze_soft", [1, 0])
def test_lbfgs_oracle_matches_reference(boundt, freeze_soft):
    n = 40
    fg = bounded_quad()
    x0 = np.full(n, 1.0)
    mask = np.ones(n, np.float32)
    mask[::7] = 0.0
    mask[3::11] = 2.0
    xinit = np.linspace(0.0, 1.0, n)
    kw = dict(m=5, xmin=-0.5, xmax=2.5, mask=mask, boundt=boundt, freeze_soft=freeze_soft, xinit=xinit)
    ro = drive(O.RefLBFGS(n, **kw), fg, x0, False)
    oo = drive(O.LBFGS(n, **kw), fg, x0, False)
    assert len(ro) == len(oo) and len(ro) >= 5
    for k, ((f1, i1, x1), (f2, i2, x2)) in enumerate(zip(ro, oo)):
        assert (f1, i1) == (f2, i2), "call %d" % k
        assert np.array_equal(x1, x2), "call %d" % k


def test_freeze_xoc_blocks():
    """FREEZE_XOC: block j of n // nparam values clipped to [xmin_j, xmax_j];
    the remainder of an uneven split is left alone."""
    import ctypes as C
    L = O.oracle()
    L.orc_freeze_xoc.argtypes = [C.c_void_p, C.c_void_p, C.c_int, C.c_void_p, C.c_int]
    x = np.linspace(-5.0, 5.0, 11)
    lo, hi = np.array([-1.0, 0.0, 2.0]), np.array([1.0, 0.5, 3.0])
    y = x.copy()
    L.orc_freeze_xoc(O.ptr(lo), O.ptr(hi), 3, O.ptr(y), len(y))
    want = x.copy()
    for j in range(3):
        want[3 * j:3 * j + 3] = np.clip(x[3 * j:3 * j + 3], lo[j], hi[j])
    assert np.array_equal(y, want)
    assert y[9] == x[9] and y[10] == x[10]


@pytest.mark.parametrize("stp1_opt,step_init", [(1, 999999.0), (0, 999999.0), (0, 0.125), (1, 0.125)])
def test_lbfgsoc_first_trial_step(stp1_opt, step_init):
    """The first trial point is x0 - STP1 g with STP1 = |x0|/|g|/100
    (stp1_opt 1) or 1/|g| (otherwise), replaced by step_init when that is
    not 999999 (nlinv_lbfgs.f:336-344)."""
    n = 30
    fg = bounded_quad(n)
    x0 = np.linspace(0.5, 2.0, n)
    f, g = fg(x0)
    lb = O.LBFGSOC(n, stp1_opt=stp1_opt, step_init=step_init, boundt=0)
    x = x0.copy()
    iflag, _ = lb.step(x, f, g)
    assert iflag == 1
    if step_init != 999999.0:
        stp1 = step_init
    elif stp1_opt == 1:
        stp1 = np.linalg.norm(x0) / np.linalg.norm(g) / 100
    else:
        stp1 = 1.0 / np.linalg.norm(g)
    np.testing.assert_allclose(x, x0 - stp1 * g, rtol=1e-14, atol=1e-14)

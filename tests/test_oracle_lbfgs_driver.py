"""Pins of the oracle's l_bfgs driver (Inversion.f90:29-189) by what its
control flow fixes in closed form: the first trial model, the stop rules and
the model returned at each stop.  LBFGS itself is pinned against the
reference's compiled lbfgs.f in test_oracle_lbfgsoc.py."""
import numpy as np

import oracle_lib as O


def quad_problem(n=40, seed=3):
    rng = np.random.default_rng(seed)
    Q = rng.standard_normal((n, n))
    A = Q @ Q.T / n + np.eye(n)
    xs = (2500.0 + 300.0 * rng.standard_normal(n)).astype(np.float64)

    def fct(x):
        r = x.astype(np.float64) - xs
        # reference sign convention: the evaluation returns the descent direction
        return 0.5 * r @ A @ r, -(A @ r)
    return fct, xs


def recorder(fct):
    seen = []

    def f(x):
        seen.append(x.copy())
        return fct(x)
    return f, seen


def test_first_trial_model_closed_form():
    """LBFGS gets gd = -g0*g; its first trial is x0 - stp1*gd with stp1 = 1/|gd|
    (bound 0) or |x0|/|gd|/100 (bound 2): a step along +g of length 1, or of
    1% of |x0| (lbfgs.f:336, Inversion.f90:117-129)."""
    fct, _ = quad_problem()
    x0 = np.full(40, 2000.0, np.float32)
    _, g = fct(x0)
    g0 = np.float32(x0.max() / (100 * np.abs(g.astype(np.float32)).max()))
    gs = (g0 * g.astype(np.float32)).astype(np.float64)
    for bound, length in [(0, 1.0), (2, np.linalg.norm(x0.astype(np.float64)) / 100)]:
        f, seen = recorder(fct)
        O.l_bfgs(f, x0, niter=5, neval=2, bound=bound, xmin=0.0, xmax=1e6)
        want = x0.astype(np.float64) + length * gs / np.linalg.norm(gs)
        assert len(seen) == 2
        assert np.allclose(seen[1], want.astype(np.float32), rtol=1e-6, atol=0)


def test_stop_rules_and_returned_model():
    fct, xs = quad_problem()
    x0 = np.full(40, 2000.0, np.float32)
    # neval reached before any accepted iteration: the starting model comes back
    x, rep = O.l_bfgs(fct, x0, niter=5, neval=1, bound=0, xmin=0.0, xmax=1e6)
    assert rep["evaluations"] == 1 and rep["iterations"] == 0
    assert np.array_equal(x, x0)
    # niter reached: the next trial model is returned (cont stays true)
    f2, seen = recorder(fct)
    _, rep2 = O.l_bfgs(f2, x0, niter=2, neval=100, bound=0, xmin=0.0, xmax=1e6)
    first_acc = int(np.argmax(rep2["accepted"]))
    x1, rep1 = O.l_bfgs(fct, x0, niter=1, neval=100, bound=0, xmin=0.0, xmax=1e6)
    assert rep1["iterations"] == 1 and rep1["evaluations"] == first_acc + 1
    assert np.array_equal(x1, seen[first_acc + 1])
    # misfit decreases over accepted iterations and converges on an SPD quadratic
    x3, rep3 = O.l_bfgs(fct, x0, niter=30, neval=200, bound=0, xmin=0.0, xmax=1e6)
    fa = rep3["f_eval"][rep3["accepted"]]
    assert np.all(np.diff(fa) < 0)
    assert fa[-1] < 1e-6 * rep3["f_eval"][0]


def test_bounds_clip_accepted_models():
    """bound 2: each accepted model passes through FREEZE_X (clip to
    [vpmin, vpmax]); the returned model respects the bounds."""
    fct, xs = quad_problem(seed=4)
    x0 = np.full(40, 2000.0, np.float32)
    x, rep = O.l_bfgs(fct, x0, niter=8, neval=60, bound=2, xmin=1900.0, xmax=2600.0)
    assert rep["iterations"] >= 1
    assert x.min() >= 1900.0 - 1e-3 and x.max() <= 2600.0 + 1e-3


def test_nan_misfit_ends_the_run_with_iflag_minus_one():
    """Fault injection (Inversion.f90:158-169): a misfit that turns NaN makes
    the line search fail (INFO != 1), LBFGS returns IFLAG = -1 and l_bfgs
    stops, returning the last accepted model (finite)."""
    fct, _ = quad_problem()
    calls = []

    def bad(x):
        calls.append(1)
        f, g = fct(x)
        return (float("nan") if len(calls) > 4 else f), g
    x0 = np.full(40, 2000.0, np.float32)
    x, rep = O.l_bfgs(bad, x0, niter=20, neval=200, bound=0)
    assert rep["iflag"] == -1
    assert rep["iterations"] >= 1 and rep["evaluations"] < 200
    assert np.all(np.isfinite(x))

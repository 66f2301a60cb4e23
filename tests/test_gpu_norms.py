"""The device FWI gradient under every data misfit norm (sfd_fwi_opts.norm,
norm_thresh: data_nrm_type / data_threshold, libs/libinv/Norm.f90:12-141 via
Objective_Function_Residual_Adjsource.f90:17-40) against the FP32 oracle,
whose norms equal the reference's compiled routines bit for bit
(test_oracle_norms.py).

- Residuals and per-shot thresholds are bitwise; the misfit is bitwise for
  every norm but Cauchy (its log is the device's logf, which may differ from
  the host's by an ulp per sample: relative 1e-6); gradients within 1e-5.
- A 3-shot batch with ragged spreads and a 0/1 mute that also mutes one whole
  trace inside the second shot (a zero residual under the shot's THRESH).
- A shot whose first trace is all zero with data_threshold 0 gives the
  reference a NaN misfit (Huber, Cauchy) or adjoint source (hyperbolic):
  the device reports SFD_ENOTFINITE.
- A split into shot waves is bitwise the one-batch result under a norm with
  per-shot thresholds."""
import numpy as np
import pytest

import oracle_lib as O
from models import smooth_random, ricker

sa = pytest.importorskip("seismic_amd")
import seismic_amd.abi as abi  # noqa: E402

pytestmark = pytest.mark.gpu

NORMS = {"L1": 1, "L2": 2, "Huber": 12, "Cauchy": 3, "hyperbolic": 4}


def problem():
    nz, nx = 40, 72
    v0 = smooth_random(nz, nx, seed=2, vmin=1800, vmax=2400).astype(np.float32)
    vt = v0.copy()
    vt[18:26, 20:50] += 150.0
    shots = []
    w = ricker(15.0, 600, 1e-3)   # one record length: the reference's misfit scaling has one n1
    for sx, n0, n1 in ((150.0, 2, 60), (400.0, 10, 70), (620.0, 4, 48)):
        rx = (np.arange(n0, n1, 2) * 10.0).astype(np.float32)
        shots.append((20.0, sx, w, np.full(len(rx), 20.0, np.float32), rx))
    return nz, nx, v0, vt, shots


def mutes(shots, zero_trace=None):
    """A 0/1 mute of the first 40 samples; zero_trace = (shot, trace) mutes
    that whole trace."""
    out = []
    for k, s in enumerate(shots):
        m = np.ones((len(s[4]), len(s[2])), np.float32)
        m[:, :40] = 0.0
        if zero_trace is not None and k == zero_trace[0]:
            m[zero_trace[1]] = 0.0
        out.append(m)
    return out


def run_both(norm, thresh, mute, waves_mib=None):
    nz, nx, v0, vt, sh = problem()
    e = sa.Engine(sa.make_grid(nz, nx, 10.0, 10.0, 1e-3, ntaper=12, snapi=2))
    shots = [sa.Shot(*s) for s in sh]
    e.set_model(vt)
    dobs = e.forward(shots, extract_kind=1)
    e.set_model(v0)
    if waves_mib:
        e.set_option(abi.SFD_OPT_BATCH_MIB, waves_mib)
    dev = e.fwi_gradient(shots, dobs, mute=mute, illupow=1.0, norm=norm, norm_thresh=thresh)
    og = O.make_grid(nz, nx, 10.0, 10.0, 1e-3, ntaper=12, snapi=2)
    oshots = [O.Shot(*s) for s in sh]
    orc = O.fwi_gradient(og, v0, oshots, dobs, mute=mute, illupow=1.0, prec="f", norm=norm, norm_thresh=thresh)
    return e, dev, orc


def rel(a, b):
    a, b = np.asarray(a, np.float64), np.asarray(b, np.float64)
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


@pytest.mark.parametrize("name", list(NORMS))
@pytest.mark.parametrize("thresh", [0.0, 2e-3])
def test_norm_matches_oracle(name, thresh):
    print("case: norm %s, data_threshold %g, 3 ragged shots with a 0/1 mute" % (name, thresh), flush=True)
    _, nx, _, _, sh = problem()
    mute = mutes(sh, zero_trace=(1, 3))
    _, (f, g, il, res), orc = run_both(NORMS[name], thresh, mute)
    rd = np.concatenate([r.ravel() for r in res])
    assert np.array_equal(rd, orc["resid"])
    ef, eg = abs(f - orc["f"]) / abs(orc["f"]), rel(g, orc["grad"])
    print("misfit %.9e vs %.9e (rel %.2e, bitwise %s); gradient rel L2 %.3e (bitwise %s)" %
          (f, orc["f"], ef, f == orc["f"], eg, np.array_equal(g, orc["grad"])), flush=True)
    if name == "Cauchy":
        assert ef <= 1e-6
    else:
        assert f == orc["f"]
    assert eg <= 1e-5


@pytest.mark.parametrize("name", ["Huber", "Cauchy", "hyperbolic"])
def test_zero_first_trace_is_not_finite(name):
    print("case: norm %s, first trace of shot 0 muted, data_threshold 0" % name, flush=True)
    nz, nx, v0, vt, sh = problem()
    e = sa.Engine(sa.make_grid(nz, nx, 10.0, 10.0, 1e-3, ntaper=12, snapi=2))
    shots = [sa.Shot(*s) for s in sh]
    e.set_model(vt)
    dobs = e.forward(shots, extract_kind=1)
    e.set_model(v0)
    with pytest.raises(sa.SeisFDError) as err:
        e.fwi_gradient(shots, dobs, mute=mutes(sh, zero_trace=(0, 0)), illupow=1.0, norm=NORMS[name])
    assert err.value.code == abi.SFD_ENOTFINITE


def test_norm_waves_equal_one_batch():
    print("case: Huber, data_threshold 0, one batch vs shot waves", flush=True)
    _, _, _, _, sh = problem()
    mute = mutes(sh, zero_trace=(1, 3))
    e1, (f1, g1, _, _), _ = run_both(12, 0.0, mute)
    e2, (f2, g2, _, _), _ = run_both(12, 0.0, mute, waves_mib=1)
    print("waves %d vs %d: misfit equal %s, gradient bitwise %s" % (e1.last_waves(), e2.last_waves(), f1 == f2,
                                                                    np.array_equal(g1, g2)), flush=True)
    assert e1.last_waves() == 1 and e2.last_waves() > 1
    assert f1 == f2 and np.array_equal(g1, g2)

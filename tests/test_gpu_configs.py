"""The benchmark workloads at their full sizes on the GPU, through the C-ABI.

Single-shot RTM 751 x 2301 (nt 7500), the 96-shot ragged batch on the same
model, the 201 x 601 FWI (a 4-shot subset at full size) against golden
fixtures the CPU oracle wrote (tests/golden/make_config_fixtures.py:
SHA-256 of each full FP32 array for the bitwise check, decimated copies for
the 1e-5 relative-L2 bar and the FP64 image for the snapshot-precision
report); the 1911 x 5395 salt FWI (4 of its 64 shots, nt 8000) through
size-independent properties: gradient against central differences of the
misfit, FP16 snapshot forms against FP32, shot waves bitwise equal to one
batch.  Reference drivers: progs/rtm/to_disk_to_memory_RTM.f90:358-492,
progs/afwi/compute_function_gradient.f90:208-350."""
import os

import numpy as np
import pytest

from golden.make_config_fixtures import (BATCHED_SAMPLE, DEC_GATHER, DEC_GRAD, DEC_IMG, FWI_SUBSET, digest)

sa = pytest.importorskip("seismic_amd")
import seismic_amd.abi as abi  # noqa: E402
from seismic_amd import synthetic as syn  # noqa: E402

pytestmark = pytest.mark.gpu
GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def say(*a):
    print(*a, flush=True)


def rel(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    return float(np.linalg.norm((a - b).ravel()) / np.linalg.norm(b.ravel()))


def fixture(name):
    return np.load(os.path.join(GOLDEN, name + ".npz"))


def engine(w):
    nz, nx, dz, dx, dt = w["grid"]
    return sa.Engine(sa.make_grid(nz, nx, dz, dx, dt, ntaper=w["ntaper"], snapi=w["snapi"]))


def shots_of(w, idx=None):
    idx = range(len(w["shots"])) if idx is None else idx
    return [sa.Shot(*w["shots"][k]) for k in idx]


def scaling(shots):
    return sum(2.0 * s.nt * s.ntr for s in shots)


def test_rtm751x2301_fp32_matches_oracle_and_fp16_forms_vs_fp64():
    """Observed gather and FP32-snapshot image against the FP32 oracle (bar:
    relative L2 <= 1e-5; bitwise reported), then the image error of each
    snapshot storage form against the FP64 oracle."""
    say("case: rtm 751x2301 nt 7500 (observed data, FP32 / FP16 per 64x4 block / FP16 per snapshot)")
    fx = fixture("rtm751x2301")
    w = syn.rtm_751x2301_workload()
    e = engine(w)
    s = shots_of(w)
    e.set_model(w["true"])
    d = e.forward(s)[0]
    say("gather: bitwise %s, decimated rel L2 %.3e" % (digest(d) == str(fx["data_sha"]),
                                                       rel(d[DEC_GATHER], fx["data_dec"])))
    assert rel(d[DEC_GATHER], fx["data_dec"]) <= 1e-5
    assert digest(d) == str(fx["data_sha"])
    e.set_model(w["mig"])
    e.set_traces([d], s)
    im32, il32 = e.rtm(s, None, snap_mode=sa.SFD_SNAP_FP32)
    err = rel(im32[DEC_IMG], fx["image32_dec"])
    bit = digest(im32) == str(fx["image32_sha"])
    say("FP32 image vs FP32 oracle: decimated rel L2 %.3e, bitwise %s; illumination bitwise %s" %
        (err, bit, digest(il32) == str(fx["illum32_sha"])))
    assert err <= 1e-5
    assert rel(il32[DEC_IMG], fx["illum32_dec"]) <= 1e-5
    ref64 = fx["image64_dec"]
    errs = {"fp32": rel(im32[DEC_IMG], ref64)}
    for name, mode in (("fp16_block", sa.SFD_SNAP_FP16S), ("fp16_snapshot", sa.SFD_SNAP_FP16G)):
        im, _ = e.rtm(s, None, snap_mode=mode)
        errs[name] = rel(im[DEC_IMG], ref64)
    say("image rel L2 vs the FP64 oracle (every 8th depth and column): " +
        ", ".join("%s %.3e" % kv for kv in errs.items()))
    # FP32 arithmetic over 7500 steps already departs from FP64 (the
    # fixture's image norms differ by 1.6e-4); the FP16 forms must not add
    # more than a few times that
    assert errs["fp32"] <= 1e-2
    assert errs["fp16_block"] <= 3e-2 and errs["fp16_snapshot"] <= 5e-2


def test_batched96_gathers_and_gradient_sums():
    """All 96 ragged shots modelled in one call: the two sampled gathers equal
    the oracle's; the batch's misfit and raw gradient (illupow 0, so the
    normalisation is the constant 1 + 1e-4) are the sums of two half
    batches'; the batch split into waves is bitwise the one-batch result."""
    say("case: 96-shot ragged batch, 751x2301 (gathers, gradient sums, waves)")
    fx = fixture("batched96")
    w = syn.batched_96shot_workload()
    e = engine(w)
    shots = shots_of(w)
    e.set_model(w["true"])
    d = e.forward(shots, extract_kind=sa.SFD_EXTRACT_NEAREST)
    for k in BATCHED_SAMPLE:
        err = rel(d[k][DEC_GATHER], fx["gather%d_dec" % k])
        say("shot %d gather: decimated rel L2 %.3e, bitwise %s" % (k, err, digest(d[k]) == str(fx["gather%d_sha" % k])))
        assert err <= 1e-5
        assert digest(d[k]) == str(fx["gather%d_sha" % k])
    e.set_model(w["mig"])
    e.set_traces(d, shots)
    fs = sa.SFD_SNAP_FP16S
    f, g, _, _ = e.fwi_gradient(shots, None, illupow=0.0, snap_mode=fs)
    assert e.last_waves() == 1
    h = len(shots) // 2
    f1, g1, _, _ = e.fwi_gradient(shots[:h], d[:h], illupow=0.0, snap_mode=fs)
    f2, g2, _, _ = e.fwi_gradient(shots[h:], d[h:], illupow=0.0, snap_mode=fs)
    S, S1, S2 = scaling(shots), scaling(shots[:h]), scaling(shots[h:])
    ef = abs(f * S - (f1 * S1 + f2 * S2)) / (f * S)
    eg = rel(g1.astype(np.float64) * S1 + g2.astype(np.float64) * S2, g.astype(np.float64) * S)
    say("misfit sum rel diff %.3e, gradient sum rel L2 %.3e" % (ef, eg))
    assert ef <= 1e-12 and eg <= 1e-5
    e.set_option(abi.SFD_OPT_BATCH_MIB, 50000)
    fw, gw, _, _ = e.fwi_gradient(shots, None, illupow=0.0, snap_mode=fs)
    say("waves %d: misfit equal %s, gradient bitwise %s" % (e.last_waves(), fw == f, np.array_equal(gw, g)))
    assert e.last_waves() > 1 and fw == f and np.array_equal(gw, g)


def test_fwi201x601_subset_gradient_matches_oracle():
    """4 of the 32 shots on the full 201 x 601 grid, nt 3000: observed
    gathers, misfit and illupow-1 gradient against the FP32 oracle."""
    say("case: fwi 201x601, shots %s, nt 3000" % (FWI_SUBSET,))
    fx = fixture("fwi201x601")
    w = syn.fwi_201x601_workload()
    e = engine(w)
    shots = shots_of(w, FWI_SUBSET)
    e.set_model(w["true"])
    d = e.forward(shots, extract_kind=sa.SFD_EXTRACT_NEAREST)
    for k, dk in zip(FWI_SUBSET, d):
        assert digest(dk) == str(fx["dobs%d_sha" % k])
    e.set_model(w["mig"])
    f, g, il, _ = e.fwi_gradient(shots, d, illupow=1.0)
    dec = DEC_GRAD
    ef = abs(f - float(fx["misfit"])) / float(fx["misfit"])
    eg = rel(g[dec], fx["grad_dec"])
    say("misfit rel diff %.3e; gradient decimated rel L2 %.3e, bitwise %s; illumination bitwise %s" %
        (ef, eg, digest(g) == str(fx["grad_sha"]), digest(il) == str(fx["illum_sha"])))
    assert ef <= 1e-10 and eg <= 1e-5
    assert rel(il[dec], fx["illum_dec"]) <= 1e-5


def test_salt_fwi_gradient_properties():
    """4 of the 64 shots of the 1911 x 5395 salt model at full size (1200
    receivers, nt 8000, snapi 10): the gradient times the normalised
    illumination against central differences of the misfit along compact
    bumps (ratio -snapi within 5%; measured -10.1 to -10.2: the second
    difference of snapshots 10 ms apart underestimates a 7 Hz wavelet's
    second derivative by ~1-2%), both FP16 snapshot forms against FP32
    snapshots, and a split into shot waves bitwise equal to one batch."""
    say("case: salt FWI 1911x5395, 4 shots, nt 8000")
    w = syn.salt_fwi_workload()
    idx = (8, 24, 40, 56)
    shots = shots_of(w, idx)
    e = engine(w)
    e.set_model(w["true"])
    d = e.forward(shots, extract_kind=sa.SFD_EXTRACT_NEAREST)
    v0 = w["mig"]

    def fg(v, mode=sa.SFD_SNAP_FP16G):
        e.set_model(v)
        f, g, il, _ = e.fwi_gradient(shots, d, illupow=1.0, snap_mode=mode)
        return f, g, il

    f, g, il = fg(v0)
    raw = g.astype(np.float64) * il
    nz, nx = v0.shape
    # Candidate bumps: 20 m/s raised cosines of radius 40 cells (500 m) at
    # 1 km depth, 2.5 km either side of each source.  They vanish on the top
    # rows: the model's first row is replicated into the 40-cell pad above
    # it, where the surface source and receivers make the misfit very
    # sensitive, and the reference's gradient has no pad cells, so a bump
    # reaching the surface changes the misfit by more than its gradient says.
    # The check uses the three bumps the gradient predicts the largest misfit
    # change for (deep bumps change this 4-shot misfit by less than its FP32
    # rounding, ~1e-8 of 4e-2).
    zc = np.arange(nz)[:, None]
    xc = np.arange(nx)[None, :]

    def bump(cz, cx, R=40.0):
        r = np.sqrt((zc - cz) ** 2 + (xc - cx) ** 2)
        return np.where(r < R, 20.0 * np.cos(np.pi * r / (2 * R)) ** 2, 0.0).astype(np.float32)

    cands = [(80, int(round(s.sx / w["grid"][3])) + dx) for s in shots for dx in (-200, 200)]
    pred = [abs(float(np.sum(raw * bump(cz, cx)))) for cz, cx in cands]
    for k in np.argsort(pred)[::-1][:3]:
        cz, cx = cands[k]
        dv = bump(cz, cx)
        D = (fg(v0 + dv)[0] - fg(v0 - dv)[0]) / 2.0
        P = float(np.sum(raw * dv))
        say("bump (%d,%d): dJ %.6e, <grad * illum, dv> %.6e, ratio %.4f" % (cz, cx, D, P, D / P))
        assert abs(D / P + w["snapi"]) <= 0.05 * w["snapi"]
    _, g32, _ = fg(v0, sa.SFD_SNAP_FP32)
    _, g16s, _ = fg(v0, sa.SFD_SNAP_FP16S)
    e16g, e16s = rel(g, g32), rel(g16s, g32)
    say("gradient rel L2 vs FP32 snapshots: FP16 per snapshot %.3e, FP16 per block %.3e" % (e16g, e16s))
    assert e16g <= 5e-2 and e16s <= 5e-2
    e.set_option(abi.SFD_OPT_BATCH_MIB, 12000)
    fw, gw, _ = fg(v0)
    say("waves %d: misfit equal %s, gradient bitwise %s" % (e.last_waves(), fw == f, np.array_equal(gw, g)))
    assert e.last_waves() > 1 and fw == f and np.array_equal(gw, g)

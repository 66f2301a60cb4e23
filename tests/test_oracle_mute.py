"""Pins of the oracle's mute mask (MuteParam_compute_mask, Mute.f90:50-98 with
compute_top_mute :171-195 and compute_bottom_mute :197-219).  Mute.f90 needs
SEPlib (`use sep`) and is not compiled here, so the restatement is pinned by
what the mute's definition fixes:

- the top mute reaches 1 at the sample of time sqrt(t0^2 + (h/v0)^2)
  (1-based index (t - ot)/dt + 1.5 truncated, i.e. round half up) and is 0
  more than nramp = int(tramp/dt + 1.5) samples before it;
- in between it is sin^2 of k pi / (2 nramp), k = 0..nramp: monotone, 1/2 at
  the ramp's middle for even nramp, and 1 - (the cos^2 bottom ramp) sample by
  sample;
- the bottom mute is 1 up to its onset sample, falls as cos^2 over nramp
  samples and is 0 after;
- the mask depends on the offset only through (h/v)^2: mirror-symmetric in
  receiver position about the source;
- zero offset and t0 = 0 put the top onset at the first sample.
"""
import numpy as np

import oracle_lib as O


def one(nt, dt, sx, rx, **kw):
    w = np.zeros(nt, np.float32)
    s = O.Shot(0.0, sx, w, np.zeros(len(rx), np.float32), np.asarray(rx, np.float32))
    return O.mute_mask(dt, [s], **kw)[0]


def test_top_mute_onset_and_ramp():
    dt, nt, v0, t0, tramp = 0.004, 600, 1500.0, 0.1, 0.2
    offs = np.array([-1500.0, -300.0, 0.0, 450.0, 3000.0], np.float32)
    m = one(nt, dt, 1000.0, 1000.0 + offs, v0=v0, t0=t0, tramp=tramp)
    nramp = int(np.float32(tramp) / np.float32(dt) + np.float32(1.5))
    assert nramp == 51
    for h, mk in zip(offs, m):
        t = np.sqrt(t0 ** 2 + (h / v0) ** 2)
        it = int(np.floor(t / dt + 1.5))          # 1-based sample of the mute time
        assert mk[it - 1] == 1.0 and np.all(mk[it - 1:] == 1.0)
        assert np.all(mk[:max(0, it - nramp - 1)] == 0.0)
        i = np.arange(max(1, it - nramp), it + 1)            # 1-based samples of the ramp
        ramp = mk[i - 1]
        assert np.allclose(ramp, np.sin((i - it + nramp) * np.pi / (2 * nramp)) ** 2, atol=2e-7)
        assert np.all(np.diff(ramp) >= 0)
    # symmetric in offset sign
    m2 = one(nt, dt, 1000.0, 1000.0 - offs, v0=v0, t0=t0, tramp=tramp)
    assert np.array_equal(m, m2)


def test_bottom_mute_and_product():
    dt, nt = 0.002, 900
    kw = dict(v0=1500.0, t0=0.0, v1=3000.0, t1=0.8, tramp=0.1)
    offs = np.array([0.0, 300.0, 900.0], np.float32)   # top onsets before the bottom ones
    m = one(nt, dt, 0.0, offs, **kw)
    nramp = int(np.float32(0.1) / np.float32(dt) + np.float32(1.5))
    for h, mk in zip(offs, m):
        itop = int(np.floor(np.sqrt((h / 1500.0) ** 2) / dt + 1.5))
        ibot = int(np.floor(np.sqrt(0.8 ** 2 + (h / 3000.0) ** 2) / dt + 1.5))
        assert np.all(mk[itop - 1:ibot] == 1.0)                 # between the mutes
        fall = mk[ibot:ibot + nramp]
        k = np.arange(nramp)
        assert np.allclose(fall, np.cos(k * np.pi / (2 * nramp)) ** 2, atol=2e-7)
        assert np.all(np.diff(fall) <= 0)
        assert np.all(mk[ibot + nramp:] == 0.0)
    # zero offset, t0 = 0: the top onset is the first sample
    assert m[0][0] == 1.0


def test_ramps_are_complementary():
    """sin^2 and cos^2 of the same angle sum to 1: a top ramp read forwards
    and a bottom ramp of the same length match sample by sample."""
    dt, nt = 0.004, 400
    m = one(nt, dt, 0.0, [0.0], v0=1500.0, t0=0.4, v1=1e9, t1=10.0, tramp=0.1)[0]
    nramp = int(np.float32(0.1) / np.float32(dt) + np.float32(1.5))
    itop = int(np.floor(0.4 / dt + 1.5))
    top = m[itop - nramp - 1:itop - 1]                       # k = 0..nramp-1
    k = np.arange(nramp)
    assert np.allclose(top + np.cos(k * np.pi / (2 * nramp)) ** 2, 1.0, atol=3e-7)

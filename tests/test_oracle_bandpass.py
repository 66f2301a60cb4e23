"""Pins of the oracle's Butterworth band-pass (libs/libwave/bandpass.c:72-266
highcut_/lowcut_, wrapped by Bandpass.f90:89-147 bandpass).

- The filters equal the reference's bandpass.c, compiled from its source into
  oracle/_ref, bit for bit: zero and minimum phase, odd and even pole counts,
  cut-offs across the band.
- The wrapper (Bandpass.f90 needs SEPlib and is not compiled) is pinned by the
  filter's definition: a Butterworth section cascade built by the bilinear
  transform with the cut-off pre-warped (a = 2 tan(pi f)) passes half the
  amplitude at the cut-off when run forward and backward (|H(fc)|^2 = 1/2 per
  pass), passes DC and low frequencies unchanged and removes frequencies an
  octave past the cut-off; zero phase leaves a symmetric pulse symmetric.
"""
import ctypes as C
import os

import numpy as np
import pytest

import oracle_lib as O

REF_BP = os.path.join(O.ROOT, "oracle", "_ref", "libbandpassref.so")


def orc():
    L = O.oracle()
    for f in (L.orc_highcut, L.orc_lowcut):
        f.argtypes = [C.c_float, C.c_int, C.c_int, C.c_int, C.c_void_p, C.c_void_p, C.c_void_p]
    L.orc_f_bandpass.argtypes = [C.c_float, C.c_float, C.c_int, C.c_int, C.c_int, C.c_float, C.c_int, C.c_int,
                                 C.c_void_p]
    return L


def run_filter(fn, f, npole, phase, x, ref):
    d = np.ascontiguousarray(x, np.float32).copy()
    nw = np.zeros_like(d)
    tmp = np.zeros_like(d)
    if ref:
        fn(C.byref(C.c_float(f)), C.byref(C.c_int(npole)), C.byref(C.c_int(phase)), C.byref(C.c_int(d.size)),
           O.ptr(d), O.ptr(nw), O.ptr(tmp))
    else:
        fn(f, npole, phase, d.size, O.ptr(d), O.ptr(nw), O.ptr(tmp))
    return d


@pytest.mark.skipif(not os.path.exists(REF_BP), reason="reference bandpass.c not built (oracle/_ref)")
@pytest.mark.parametrize("kind", ["highcut", "lowcut"])
@pytest.mark.parametrize("npole,phase", [(14, 0), (7, 0), (6, 1), (5, 1), (2, 0), (1, 1)])
def test_filters_equal_reference(kind, npole, phase):
    R = C.CDLL(REF_BP, mode=os.RTLD_LOCAL)
    L = orc()
    rng = np.random.default_rng(npole * 10 + phase)
    for f in (0.01, 0.05, 0.12, 0.3, 0.45):
        x = rng.standard_normal(517).astype(np.float32)
        a = run_filter(getattr(R, kind + "_"), f, npole, phase, x, True)
        b = run_filter(getattr(L, "orc_" + kind), f, npole, phase, x, False)
        assert np.array_equal(a, b), (kind, npole, phase, f)


def bandpass(x, flo, fhi, dt, nplo=14, nphi=14, phase=0):
    tr = np.ascontiguousarray(np.atleast_2d(x), np.float32).copy()
    rc = orc().orc_f_bandpass(flo, fhi, nplo, nphi, phase, dt, tr.shape[1], tr.shape[0], O.ptr(tr))
    assert rc == 0
    return tr


def gain(f, flo, fhi, dt=0.002, nt=4000):
    t = np.arange(nt) * dt
    x = np.sin(2 * np.pi * f * t).astype(np.float32)
    y = bandpass(x, flo, fhi, dt)[0]
    mid = slice(nt // 4, 3 * nt // 4)
    return float(np.sqrt(np.mean(y[mid] ** 2) / np.mean(x[mid] ** 2)))


def test_highcut_response():
    assert abs(gain(20.0, 0.0, 20.0) - 0.5) < 0.01          # cut-off: 1/sqrt(2) per pass, two passes
    assert abs(gain(4.0, 0.0, 20.0) - 1.0) < 1e-3
    assert gain(40.0, 0.0, 20.0) < 1e-3
    y = bandpass(np.full(4000, 3.0, np.float32), 0.0, 20.0, 0.002)[0]
    assert np.allclose(y[1000:3000], 3.0, rtol=1e-5)         # DC passes (away from the ends' ringing)


def test_lowcut_response():
    assert abs(gain(8.0, 8.0, 250.0) - 0.5) < 0.01
    assert abs(gain(40.0, 8.0, 250.0) - 1.0) < 2e-3
    assert gain(3.0, 8.0, 250.0) < 2e-3


def test_zero_phase_keeps_a_symmetric_pulse_symmetric():
    nt, dt = 801, 0.002
    t = (np.arange(nt) - 400) * dt
    x = (1 - 2 * (np.pi * 10 * t) ** 2) * np.exp(-(np.pi * 10 * t) ** 2)   # Ricker at the middle
    y = bandpass(x.astype(np.float32), 3.0, 15.0, dt)[0]
    core = y[200:601]
    assert np.argmax(np.abs(core)) == 200                   # peak stays at the centre
    assert np.max(np.abs(core - core[::-1])) < 2e-3 * np.max(np.abs(core))


def test_no_cut_requested_is_refused():
    tr = np.zeros((1, 100), np.float32)
    assert orc().orc_f_bandpass(0.0, 300.0, 14, 14, 0, 0.002, 100, 1, O.ptr(tr)) == -1

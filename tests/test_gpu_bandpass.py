"""sfd_bandpass (Bandpass.f90:89-147 bandpass with the filters of
libs/libwave/bandpass.c) against the oracle's restatement, which equals the
reference's compiled bandpass.c bit for bit (test_oracle_bandpass.py): every
sample of every trace must be equal, for the reference's defaults, a band
with both cuts, minimum phase, odd and single-pole filters, a high cut above
Nyquist (the reference still applies it below fhi dt = 4.999999), trace
counts around a wavefront and record lengths around the 64-sample tiles."""
import ctypes as C

import numpy as np
import pytest

import oracle_lib as O

sa = pytest.importorskip("seismic_amd")
import seismic_amd.abi as abi  # noqa: E402

pytestmark = pytest.mark.gpu

CASES = [
    ("defaults", dict()),
    ("band 3-15 Hz", dict(flo=3.0, fhi=15.0)),
    ("min phase", dict(flo=2.0, fhi=25.0, nplo=6, nphi=8, phase=1)),
    ("odd poles", dict(flo=5.0, fhi=40.0, nplo=5, nphi=9)),
    ("one pole", dict(flo=0.0, fhi=12.0, nphi=1, phase=1)),
    ("high cut near Nyquist", dict(flo=4.0, fhi=240.0)),
    ("high cut past Nyquist", dict(flo=4.0, fhi=300.0)),   # tan wraps: NaN, as in the reference
]


def oracle_bandpass(tr, dt, flo=0.0, fhi=20.0, nplo=14, nphi=14, phase=0):
    L = O.oracle()
    L.orc_f_bandpass.argtypes = [C.c_float, C.c_float, C.c_int, C.c_int, C.c_int, C.c_float, C.c_int, C.c_int,
                                 C.c_void_p]
    out = np.ascontiguousarray(tr, np.float32).copy()
    assert L.orc_f_bandpass(flo, fhi, nplo, nphi, phase, dt, out.shape[1], out.shape[0], O.ptr(out)) == 0
    return out


@pytest.mark.parametrize("name,kw", CASES, ids=[c[0] for c in CASES])
@pytest.mark.parametrize("ntr,nt", [(1, 37), (63, 700), (130, 1029)])
def test_bandpass_matches_oracle(name, kw, ntr, nt):
    print("case: bandpass %s, %d traces of %d samples" % (name, ntr, nt), flush=True)
    dt = 2e-3
    e = sa.Engine(sa.make_grid(40, 40, 10.0, 10.0, dt, ntaper=12))
    rng = np.random.default_rng(ntr + nt)
    tr = rng.standard_normal((ntr, nt)).astype(np.float32) * rng.uniform(0.1, 5.0, (ntr, 1)).astype(np.float32)
    dev = e.bandpass(tr, **kw)
    ref = oracle_bandpass(tr, dt, **kw)
    same = np.array_equal(dev, ref, equal_nan=True)
    print("finite %s, bitwise %s" % (bool(np.isfinite(ref).all()), same), flush=True)
    assert same


def test_bandpass_refusals():
    print("case: bandpass refusals", flush=True)
    e = sa.Engine(sa.make_grid(40, 40, 10.0, 10.0, 2e-3, ntaper=12))
    tr = np.zeros((2, 100), np.float32)
    with pytest.raises(sa.SeisFDError) as err:
        e.bandpass(tr, flo=0.0, fhi=300.0)        # neither cut: the reference's erexit
    assert err.value.code == abi.SFD_EINVAL
    with pytest.raises(sa.SeisFDError) as err:
        e.bandpass(tr, fhi=20.0, nphi=40)         # 10 sections at zero phase
    assert err.value.code == abi.SFD_EUNSUPPORTED

"""The allocation-free entry point sfd_forward_noalloc (no reference
counterpart; for callers that cannot allocate device memory) against
sfd_forward, through the C-ABI: bitwise equal gathers on two calls in a row,
the second with a smaller input (fewer shots and receivers, shorter records),
and a clear SFD_ENOMEM for a batch beyond the 64 MiB scratch array."""
import ctypes as C

import numpy as np
import pytest

from models import ricker, layered_anomaly

sa = pytest.importorskip("seismic_amd")
import seismic_amd.abi as abi  # noqa: E402

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)


def noalloc(e, shots, kind):
    lib = abi.load()
    out = np.zeros(sum(s.nt * s.ntr for s in shots), np.float32)
    rc = lib.sfd_forward_noalloc(e.h, e._shots(shots), len(shots), out.ctypes.data_as(C.c_void_p), kind)
    return rc, e._split(out, shots) if rc == 0 else None


def shot(sx, nt, nr, dx=10.0):
    rx = (sx + dx * (np.arange(nr) - nr // 2)).astype(np.float32)
    return sa.Shot(20.0, float(sx), ricker(15.0, nt, 1e-3), np.full_like(rx, 20.0), rx)


def test_forward_noalloc_matches_forward_twice():
    nz, nx = 121, 201
    vel = layered_anomaly(nz, nx, 10.0, 10.0, seed=5)
    e = sa.Engine(sa.make_grid(nz, nx, 10.0, 10.0, 1e-3, ntaper=20))
    e.set_model(vel)
    calls = [([shot(700.0, 600, 61), shot(1300.0, 520, 45)], sa.SFD_EXTRACT_SINC_V2),
             ([shot(1000.0, 300, 17)], sa.SFD_EXTRACT_NEAREST)]
    for k, (shots, kind) in enumerate(calls):
        say("case: noalloc call %d, %d shots, nt %s" % (k, len(shots), [s.nt for s in shots]))
        ref = e.forward(shots, extract_kind=kind)
        rc, got = noalloc(e, shots, kind)
        assert rc == 0, abi.load().sfd_last_error(e.h)
        for a, b in zip(got, ref):
            assert np.array_equal(a, b)
        say("noalloc call %d bitwise equal to sfd_forward" % k)


def test_forward_noalloc_rejects_batch_beyond_scratch():
    nz, nx = 2000, 2400
    vel = np.full((nz, nx), 2000.0, np.float32, order="F")
    e = sa.Engine(sa.make_grid(nz, nx, 10.0, 10.0, 1e-3, ntaper=20))
    e.set_model(vel)
    say("case: noalloc 2000x2400 grid, 2001 receivers (needs more than 64 MiB)")
    rc, _ = noalloc(e, [shot(12000.0, 50, 2001)], sa.SFD_EXTRACT_SINC_V2)
    msg = abi.load().sfd_last_error(e.h).decode()
    say("rc=%d: %s" % (rc, msg))
    assert rc == abi.SFD_ENOMEM and "scratch array" in msg
    # the handle keeps working: a small shot on the same grid fits
    say("case: noalloc small shot after the rejection")
    small = [shot(12000.0, 60, 11)]
    ref = e.forward(small, extract_kind=sa.SFD_EXTRACT_SINC_V2)
    rc, got = noalloc(e, small, sa.SFD_EXTRACT_SINC_V2)
    assert rc == 0, abi.load().sfd_last_error(e.h)
    assert np.array_equal(got[0], ref[0])

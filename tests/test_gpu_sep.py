"""An FWI gradient driven from SEP files, as 2DAFWI.x reads its inputs
(progs/afwi/2DAFWI.f90:53-73: readsou from the wavelet file, readgathercoords
from `coordfile` with keygx/keygy/keysx/keysy/keyselev/keygelev,
readgathertraces from `traces`, read_vel from `vel`): the gradient from the
files equals the one from the same arrays bit for bit, and the gradient
written back as a SEP file reads back unchanged."""
import os

import numpy as np
import pytest

from models import ricker, smooth_random

sa = pytest.importorskip("seismic_amd")
from seismic_amd import sep  # noqa: E402

pytestmark = pytest.mark.gpu


def test_fwi_gradient_from_sep_files(tmp_path):
    print("case: FWI gradient from SEP files (3 shots)", flush=True)
    nz, nx, dx, dt, nt = 40, 72, 10.0, 1e-3, 600
    v0 = smooth_random(nz, nx, seed=2, vmin=1800, vmax=2400).astype(np.float32)
    vt = v0.copy()
    vt[18:26, 20:50] += 150.0
    w = ricker(15.0, nt, dt)
    geo = [(150.0, np.arange(2, 60, 2)), (400.0, np.arange(10, 70, 2)), (620.0, np.arange(4, 48, 2))]
    shots = [sa.Shot(20.0, sx, w, np.full(len(ix), 20.0, np.float32), (ix * dx).astype(np.float32))
             for sx, ix in geo]
    e = sa.Engine(sa.make_grid(nz, nx, dx, dx, dt, ntaper=12, snapi=2))
    e.set_model(vt)
    dobs = e.forward(shots, extract_kind=1)
    e.set_model(v0)
    f0, g0, _, _ = e.fwi_gradient(shots, dobs, illupow=1.0)

    d = str(tmp_path)
    sep.write(os.path.join(d, "wavelet.H"), w, d=(dt,))
    sep.write(os.path.join(d, "vel.H"), v0.T, d=(dx, dx))           # (nz, nx) column-major: n1 = nz
    rows = [[s.sx, 0.0, s.sz, rx, 0.0, rz] for s in shots for rz, rx in zip(s.rz, s.rx)]
    sep.write(os.path.join(d, "coord.H"), np.array(rows, np.float32))
    # the traces as SEPlib writes them by default: big-endian, data_format absent
    np.concatenate([np.asarray(t, np.float32) for t in dobs]).astype(">f4").tofile(os.path.join(d, "traces.H@"))
    with open(os.path.join(d, "traces.H"), "w") as fo:
        fo.write("\tn1=%d n2=%d d1=%g o1=0\n\tin=\"traces.H@\" esize=4\n" % (nt, sum(len(s.rx) for s in shots), dt))

    wav, hw = sep.read(os.path.join(d, "wavelet.H"))
    vel, hv = sep.read(os.path.join(d, "vel.H"))
    gath = sep.gathers(os.path.join(d, "coord.H"), keygx=4, keygy=5, keysx=1, keysy=2, keyselev=3, keygelev=6)
    tr, ht = sep.read(os.path.join(d, "traces.H"))
    assert ht["n"][0] == hw["n"][0] == nt            # readgathertraces: n1 of traces equals the source's nt
    fshots = [sa.Shot(sz, sx, wav, rz, rx) for sz, sx, rz, rx in gath]
    fobs, o = [], 0
    for s in fshots:
        fobs.append(tr[o:o + s.ntr])
        o += s.ntr
    e2 = sa.Engine(sa.make_grid(hv["n"][0], hv["n"][1], hv["d"][0], hv["d"][1], hw["d"][0], ntaper=12, snapi=2))
    e2.set_model(vel.T)
    f1, g1, _, _ = e2.fwi_gradient(fshots, fobs, illupow=1.0)
    print("misfit equal %s, gradient bitwise %s" % (f1 == f0, np.array_equal(g1, g0)), flush=True)
    assert f1 == f0 and np.array_equal(g1, g0)
    sep.write(os.path.join(d, "grad.H"), g1.T, o=(0.0, 0.0), d=(dx, dx))
    back, _ = sep.read(os.path.join(d, "grad.H"))
    assert np.array_equal(back.T, g1)

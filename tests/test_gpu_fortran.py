"""The Fortran host side (seismic_amd/fortran): a Fortran program calling the
engine through the iso_c_binding module gets the same gather as the Python
path on the same inputs, and drives l_bfgs with a Fortran fctgdt callback."""
import os
import subprocess

import numpy as np
import pytest

sa = pytest.importorskip("seismic_amd")
pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "seismic_amd", "lib", "seisfd_example")


def test_fortran_example_matches_python(tmp_path):
    if not os.path.exists(EXE):
        pytest.fail("seisfd_example not built (run __graft_entry__.build())")
    print("case: fortran example", flush=True)
    r = subprocess.run([EXE, str(tmp_path)], capture_output=True, text=True, timeout=120)
    print(r.stdout, r.stderr, flush=True)
    assert r.returncode == 0
    nz, nx, nt, ntr = 60, 80, 300, 40
    vel = np.fromfile(tmp_path / "vel.f32", np.float32).reshape(nx, nz).T
    w = np.fromfile(tmp_path / "wavelet.f32", np.float32)
    gat = np.fromfile(tmp_path / "gather.f32", np.float32).reshape(ntr, nt)
    rx = (20.0 * np.arange(ntr)).astype(np.float32)
    rz = np.full_like(rx, 20.0)
    e = sa.Engine(sa.make_grid(nz, nx, 10.0, 10.0, 1e-3, ntaper=20))
    e.set_model(vel)
    py = e.forward([sa.Shot(20.0, 400.0, w, rz, rx)])[0]
    assert np.abs(py).max() > 0
    assert np.array_equal(gat, py)
    it, ev, iflag, err = open(tmp_path / "lbfgs.txt").read().split()
    assert int(it) >= 5 and float(err) < 1.0

"""CPU-side checks of the C-ABI library: it loads without a GPU and exports
every symbol include/seisfd.h declares (no compute calls)."""
import ctypes
import os
import re

import pytest

from seismic_amd import abi

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def header_symbols():
    txt = open(os.path.join(ROOT, "include", "seisfd.h")).read()
    return sorted(set(re.findall(r"\b(sfd_[a-z0-9_]+)\s*\(", txt)))


def test_header_matches_abi_list():
    assert header_symbols() == sorted(abi.EXPORTS)


@pytest.mark.skipif(not os.path.exists(abi.LIB_PATH), reason="libseisfd.so not built")
def test_library_exports_all_symbols():
    lib = ctypes.CDLL(abi.LIB_PATH)
    for name in header_symbols():
        assert hasattr(lib, name), name


def test_struct_sizes():
    assert ctypes.sizeof(abi.sfd_grid) == 15 * 4
    assert ctypes.sizeof(abi.sfd_shot) == 8 + 8 + 3 * 8
    assert ctypes.sizeof(abi.sfd_fwi_opts) == 8 * 4
    # int m, nparam; two pointers; int stp1_opt (+4 pad); double; int clip_type (+4 pad)
    assert ctypes.sizeof(abi.sfd_lbfgsoc_opts) == 48
    assert ctypes.sizeof(abi.sfd_grid3d) == 16 * 4
    assert ctypes.sizeof(abi.sfd_precond_opts) == 11 * 4
    assert ctypes.sizeof(abi.sfd_shot3d) == 12 + 8 + 4 + 4 * 8   # 3 floats, 2 ints, pad, 4 pointers


def test_fortran_binding_struct_sizes():
    """The bind(C) types of seismic_amd/fortran/seisfd_mod.f90 have the C
    structs' sizes (a Fortran caller passes them by reference)."""
    import subprocess
    exe = os.path.join(ROOT, "seismic_amd", "lib", "seisfd_example")
    if not os.path.exists(exe):
        pytest.skip("Fortran example not built")
    out = subprocess.run([exe, "--sizes"], capture_output=True, text=True, timeout=60).stdout.split()
    want = [ctypes.sizeof(t) for t in (abi.sfd_grid, abi.sfd_shot, abi.sfd_lbfgs_opts, abi.sfd_lbfgs_report,
                                  abi.sfd_fwi_opts, abi.sfd_lbfgsoc_opts, abi.sfd_grid3d, abi.sfd_shot3d)]
    assert [int(v) for v in out] == want

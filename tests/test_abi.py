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
    assert ctypes.sizeof(abi.sfd_grid) == 14 * 4
    assert ctypes.sizeof(abi.sfd_shot) == 8 + 8 + 3 * 8
    assert ctypes.sizeof(abi.sfd_fwi_opts) == 6 * 4

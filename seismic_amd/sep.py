"""SEP files through the C-ABI (sfd_sep_*): the history-file + binary format
the reference reads and writes every array in (SEPlib's from_aux / sreed /
srite, libs/libwave/IO.f90:14-51), and the shot geometry of a coordinate
file (readgathercoords, ReadSouVelRho.f90:161-241)."""
import ctypes as C
import os

import numpy as np

from . import abi


class SepError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__("seisfd SEP error %d: %s" % (code, msg))
        self.code = code


def _chk(rc, what):
    if rc != 0:
        raise SepError(rc, what)


def header(path):
    """The header of a history file: dict of n, o, d (6 axes each), ndim,
    xdr, in."""
    lib = abi.load()
    h = abi.sfd_sep_header()
    _chk(lib.sfd_sep_read_header(os.fsencode(path), C.byref(h)), "cannot read the header of %s" % path)
    return {"ndim": h.ndim, "n": list(h.n), "o": list(h.o), "d": list(h.d), "xdr": h.xdr,
            "in": h.in_.decode()}


def read(path):
    """The whole file as float32, shaped (n_ndim, ..., n2, n1) (n1 fastest,
    as the reference's column-major arrays); returns (array, header)."""
    lib = abi.load()
    hd = header(path)
    n = hd["n"][:max(1, hd["ndim"])]
    a = np.zeros(int(np.prod(n)), np.float32)
    _chk(lib.sfd_sep_read(os.fsencode(path), 0, a.size, a.ctypes.data), "cannot read %s" % path)
    return a.reshape(list(reversed(n))), hd


def write(path, array, o=None, d=None):
    """Writes array (shaped (..., n2, n1), n1 fastest) as native floats to
    path@ with its history file at path."""
    lib = abi.load()
    a = np.ascontiguousarray(array, np.float32)
    shape = list(reversed(a.shape)) or [1]
    h = abi.sfd_sep_header()
    h.ndim = len(shape)
    for k in range(6):
        h.n[k] = shape[k] if k < len(shape) else 1
        h.o[k] = (o[k] if o is not None and k < len(o) else 0.0)
        h.d[k] = (d[k] if d is not None and k < len(d) else 1.0)
    h.esize = 4
    _chk(lib.sfd_sep_write(os.fsencode(path), C.byref(h), a.ctypes.data, a.size), "cannot write %s" % path)


def gathers(coordfile, keygx, keygy, keysx, keysy, keyselev, keygelev):
    """Shots of a coordinate file as the reference groups them: a list of
    (sz, sx, rz, rx) with rz, rx float32 arrays of the shot's traces."""
    lib = abi.load()
    keys = (C.c_int * 6)(keygx, keygy, keysx, keysy, keyselev, keygelev)
    ns = C.c_int()
    cf = os.fsencode(coordfile)
    _chk(lib.sfd_sep_gathers(cf, keys, C.byref(ns), None, None, None, None, None, None, 0, 0), "bad coordinate file")
    ntr = int(np.prod(header(coordfile)["n"][1:]))
    first = np.zeros(ns.value, np.int32)
    cnt = np.zeros(ns.value, np.int32)
    sz = np.zeros(ns.value, np.float32)
    sx = np.zeros(ns.value, np.float32)
    rz = np.zeros(ntr, np.float32)
    rx = np.zeros(ntr, np.float32)
    _chk(lib.sfd_sep_gathers(cf, keys, C.byref(ns), first.ctypes.data, cnt.ctypes.data, sz.ctypes.data,
                             sx.ctypes.data, rz.ctypes.data, rx.ctypes.data, ns.value, ntr), "bad coordinate file")
    return [(float(sz[k]), float(sx[k]), rz[first[k]:first[k] + cnt[k]].copy(), rx[first[k]:first[k] + cnt[k]].copy())
            for k in range(ns.value)]

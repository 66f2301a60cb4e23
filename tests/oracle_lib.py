"""ctypes bindings to the CPU oracle (oracle/liboracle.so) and to the
reference's own Fortran kernels (oracle/_ref/libseisref.so).

Test infrastructure: used by tests/, __graft_entry__.smoke() and bench.py's
cpu_baseline leg as the checker only.
"""
import ctypes as C
import os
import subprocess

import numpy as np

from seismic_amd.abi import sfd_grid, sfd_shot, sfd_fwi_opts

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_SO = os.path.join(ROOT, "oracle", "liboracle.so")
REF_SO = os.path.join(ROOT, "oracle", "_ref", "libseisref.so")
NLREF_SO = os.path.join(ROOT, "oracle", "_ref", "libnlinvref.so")

_FP = C.POINTER(C.c_float)
_DP = C.POINTER(C.c_double)


class ORC_BOUNDS(C.Structure):
    _fields_ = [("nmin1", C.c_int), ("nmax1", C.c_int), ("nmin2", C.c_int),
                ("nmax2", C.c_int), ("NZ", C.c_int), ("NX", C.c_int)]


def _ensure_built():
    if not os.path.exists(ORACLE_SO):
        subprocess.check_call(["make", "-s", "-C", os.path.join(ROOT, "oracle")])


_orc = None
_ref = None


def oracle():
    global _orc
    if _orc is None:
        _ensure_built()
        _orc = C.CDLL(ORACLE_SO)
    return _orc


def ref_available():
    return os.path.exists(REF_SO)


def ref():
    global _ref
    if _ref is None:
        # Lazy binding: Extraction.f90's disk-spill routine references SEPlib's
        # srite_, which is never called on the in-memory path exercised here.
        _ref = C.CDLL(REF_SO, mode=os.RTLD_LAZY | os.RTLD_LOCAL)
    return _ref


def ptr(a):
    return a.ctypes.data_as(C.c_void_p)


def colmajor(a, dtype):
    """(nz, nx) array -> contiguous column-major (z fastest) buffer."""
    return np.ascontiguousarray(np.asarray(a, dtype=dtype).T)


def make_grid(nz, nx, dz, dx, dt, ntaper=20, lsinc=7, snapi=4, oz=0.0, ox=0.0, aperture=0.0):
    g = sfd_grid()
    g.nz, g.nx = nz, nx
    g.oz, g.ox, g.dz, g.dx, g.dt = oz, ox, dz, dx, dt
    g.ntaper, g.lsinc, g.snapi = ntaper, lsinc, snapi
    g.aperture_x = aperture
    g.rec_type = g.shot_type = g.surf_type = 0
    return g


class Shot:
    """Keeps the numpy arrays behind an sfd_shot alive."""

    def __init__(self, sz, sx, wavelet, rz, rx):
        self.wavelet = np.ascontiguousarray(wavelet, dtype=np.float32)
        self.rz = np.ascontiguousarray(rz, dtype=np.float32)
        self.rx = np.ascontiguousarray(rx, dtype=np.float32)
        self.sz, self.sx = float(sz), float(sx)
        self.nt = len(self.wavelet)
        self.ntr = len(self.rz)

    def c(self):
        s = sfd_shot()
        s.sz, s.sx = self.sz, self.sx
        s.nt, s.ntr = self.nt, self.ntr
        s.wavelet = self.wavelet.ctypes.data_as(_FP)
        s.rz = self.rz.ctypes.data_as(_FP)
        s.rx = self.rx.ctypes.data_as(_FP)
        return s


def _dt(prec):
    return np.float32 if prec == "f" else np.float64


def shot_window(g, shot):
    L = oracle()
    oxw = C.c_float()
    nxw = C.c_int()
    s = shot.c()
    L.orc_f_shot_window(C.byref(g), C.byref(s), C.byref(oxw), C.byref(nxw))
    return oxw.value, nxw.value


def bounds(nz, nxw, ntaper, lsinc=7, shot_type=0, surf_type=0):
    b = ORC_BOUNDS()
    oracle().orc_bounds_t(C.byref(b), nz, nxw, ntaper, lsinc, shot_type, surf_type)
    return b


def set_elevation(elev):
    """The oracle's free-surface elevation per model column (None: 0)."""
    L = oracle()
    if elev is None:
        L.orc_set_elevation(None, 0)
        return
    e = np.ascontiguousarray(elev, dtype=np.float32)
    L.orc_set_elevation(ptr(e), len(e))


def padded_vel(g, vel, shot, prec="f"):
    """Window + edge-replication pad of the full model for one shot."""
    L = oracle()
    oxw, nxw = shot_window(g, shot)
    b = bounds(g.nz, nxw, g.ntaper, g.lsinc, g.shot_type, g.surf_type)
    v = colmajor(vel, _dt(prec))
    out = np.zeros(b.NZ * b.NX, dtype=_dt(prec))
    fn = getattr(L, "orc_%s_window_pad" % prec)
    fn.argtypes = [C.c_void_p, C.c_void_p, (C.c_float if prec == "f" else C.c_double), C.c_int,
                   C.c_void_p, C.c_void_p]
    fn(C.byref(g), ptr(v), oxw, nxw, C.byref(b), ptr(out))
    return out.reshape(b.NX, b.NZ).T, oxw, nxw, b


def copy_image(g, oxw, nxw, small, big=None, prec="d"):
    """mod_copy_image: add the window image small (nz, nxw) into the model
    image big (nz, nx), nearest index, skipping columns off the model."""
    L = oracle()
    T = _dt(prec)
    sm = colmajor(small, T)
    bg = colmajor(np.zeros((g.nz, g.nx)) if big is None else big, T)
    fn = getattr(L, "orc_%s_copy_image" % prec)
    fn.argtypes = [C.c_void_p, (C.c_float if prec == "f" else C.c_double), C.c_int, C.c_void_p, C.c_void_p]
    fn(C.byref(g), oxw, nxw, ptr(sm), ptr(bg))
    return bg.reshape(g.nx, g.nz).T


def modeling(g, vel, shot, extract_kind=0, prec="f", snaps=False):
    L = oracle()
    T = _dt(prec)
    v = colmajor(vel, T)
    w = np.ascontiguousarray(shot.wavelet, dtype=T)
    tr = np.zeros((shot.ntr, shot.nt), dtype=T)
    snap = None
    sp = None
    if snaps:
        oxw, nxw = shot_window(g, shot)
        snap = np.zeros((shot.nt // g.snapi, nxw, g.nz), dtype=T)
        sp = ptr(snap)
    s = shot.c()
    fn = getattr(L, "orc_%s_modeling" % prec)
    fn.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_int, C.c_void_p]
    rc = fn(C.byref(g), ptr(v), C.byref(s), ptr(w), ptr(tr), extract_kind, sp)
    assert rc == 0
    return (tr, snap) if snaps else tr


def rtm(g, vel, shot, traces_in, prec="f", lsrtm=False, vel_bwd=None, capture=False):
    """Returns (image, illum) full-model (nz,nx) as column-major -> array [nx, nz]^T."""
    L = oracle()
    T = _dt(prec)
    v = colmajor(vel, T)
    vb = None if vel_bwd is None else colmajor(vel_bwd, T)
    w = np.ascontiguousarray(shot.wavelet, dtype=T)
    d = np.ascontiguousarray(traces_in, dtype=T)
    img = np.zeros(g.nz * g.nx, dtype=T)
    ill = np.zeros(g.nz * g.nx, dtype=T)
    oxw, nxw = shot_window(g, shot)
    wimg = np.zeros(g.nz * nxw, dtype=T)
    will = np.zeros(g.nz * nxw, dtype=T)
    cap = None
    if capture:
        cap = np.zeros((shot.nt // g.snapi, nxw, g.nz), dtype=T)
    s = shot.c()
    fn = getattr(L, "orc_%s_rtm" % prec)
    fn.argtypes = [C.c_void_p] * 8 + [C.c_int] + [C.c_void_p] * 3
    rc = fn(C.byref(g), ptr(v), None if vb is None else ptr(vb), C.byref(s), ptr(w), ptr(d),
            ptr(img), ptr(ill), int(lsrtm), None if cap is None else ptr(cap), ptr(wimg), ptr(will))
    assert rc == 0
    out = dict(image=img.reshape(g.nx, g.nz).T, illum=ill.reshape(g.nx, g.nz).T,
               win_image=wimg.reshape(nxw, g.nz).T, win_illum=will.reshape(nxw, g.nz).T,
               oxw=oxw, nxw=nxw)
    if capture:
        out["capture"] = cap
    return out


def born(g, vel, refl, shot, prec="f"):
    L = oracle()
    T = _dt(prec)
    v = colmajor(vel, T)
    m = colmajor(refl, T)
    w = np.ascontiguousarray(shot.wavelet, dtype=T)
    tr = np.zeros((shot.ntr, shot.nt), dtype=T)
    s = shot.c()
    fn = getattr(L, "orc_%s_born" % prec)
    fn.argtypes = [C.c_void_p] * 6
    rc = fn(C.byref(g), ptr(v), ptr(m), C.byref(s), ptr(w), ptr(tr))
    assert rc == 0
    return tr


def born_adjoint(g, vel, shot, dbar, prec="d"):
    """Exact adjoint of born(): mbar (nz, nx) = B^T dbar, dbar (ntr, nt)."""
    L = oracle()
    T = _dt(prec)
    v = colmajor(vel, T)
    w = np.ascontiguousarray(shot.wavelet, dtype=T)
    db = np.ascontiguousarray(np.asarray(dbar, dtype=T).reshape(shot.ntr, shot.nt))
    mb = np.zeros(g.nz * g.nx, dtype=T)
    s = shot.c()
    fn = getattr(L, "orc_%s_born_adjoint" % prec)
    fn.argtypes = [C.c_void_p] * 6
    rc = fn(C.byref(g), ptr(v), C.byref(s), ptr(w), ptr(db), ptr(mb))
    assert rc == 0
    return mb.reshape(g.nx, g.nz).T


def fwi_gradient(g, vel, shots, dobs, mute=None, vpmask=None, illupow=1.0, rect_z=1, rect_x=1,
                 do_z=0, do_x=0, prec="f", norm=2, norm_thresh=0.0):
    """compute_fct_gdt restatement.  dobs: list of (ntr, nt) arrays per shot;
    norm / norm_thresh: data_nrm_type / data_threshold (Norm.f90 codes)."""
    L = oracle()
    T = _dt(prec)
    v = colmajor(vel, T)
    cs = (sfd_shot * len(shots))(*[s.c() for s in shots])
    ob = np.ascontiguousarray(np.concatenate([np.asarray(d, dtype=T).ravel() for d in dobs]))
    mu = None if mute is None else np.ascontiguousarray(
        np.concatenate([np.asarray(m, dtype=T).ravel() for m in mute]))
    vm = None if vpmask is None else colmajor(vpmask, T)
    o = sfd_fwi_opts()
    o.illupow, o.rect_z, o.rect_x, o.do_z, o.do_x, o.snap_mode = illupow, rect_z, rect_x, do_z, do_x, 0
    o.norm, o.norm_thresh = norm, norm_thresh
    f = C.c_double()
    grad = np.zeros(g.nz * g.nx, dtype=T)
    illu = np.zeros(g.nz * g.nx, dtype=T)
    res = np.zeros_like(ob)
    w = np.ascontiguousarray(shots[0].wavelet, dtype=T)
    fn = getattr(L, "orc_%s_fwi_gradient" % prec)
    fn.argtypes = [C.c_void_p, C.c_void_p, C.c_int, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p,
                   C.c_void_p, C.c_void_p, C.POINTER(C.c_double), C.c_void_p, C.c_void_p, C.c_void_p]
    rc = fn(C.byref(g), ptr(v), len(shots), cs, ptr(w), ptr(ob), None if mu is None else ptr(mu),
            None if vm is None else ptr(vm), C.byref(o), C.byref(f), ptr(grad), ptr(illu), ptr(res))
    assert rc == 0
    return dict(f=f.value, grad=grad.reshape(g.nx, g.nz).T, illum=illu.reshape(g.nx, g.nz).T,
                resid=res)


def afwi(g, vel, shots, dobs, mute=None, norm=2, norm_thresh=0.0, vprho=0, prec="f"):
    """3DAFWI.x's gradient (AFWI_to_memory + mod_copy_gradient_to_disk) in 2D,
    summed over the shots; with the oracle's density (set_density) two
    gradients (vp, then rho or Ip).  Returns f, grad [nparam, nz, nx], illum."""
    L = oracle()
    T = _dt(prec)
    v = colmajor(vel, T)
    cs = (sfd_shot * len(shots))(*[s.c() for s in shots])
    ob = np.ascontiguousarray(np.concatenate([np.asarray(d, dtype=T).ravel() for d in dobs]))
    mu = None if mute is None else np.ascontiguousarray(
        np.concatenate([np.asarray(m, dtype=T).ravel() for m in mute]))
    f = C.c_double()
    grad = np.zeros(2 * g.nz * g.nx, dtype=T)
    illu = np.zeros(g.nz * g.nx, dtype=T)
    w = np.ascontiguousarray(shots[0].wavelet, dtype=T)
    fn = getattr(L, "orc_%s_afwi" % prec)
    fn.argtypes = [C.c_void_p, C.c_void_p, C.c_int, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_int,
                   C.c_double if prec == "d" else C.c_float, C.c_int, C.POINTER(C.c_double), C.c_void_p, C.c_void_p]
    rc = fn(C.byref(g), ptr(v), len(shots), cs, ptr(w), ptr(ob), None if mu is None else ptr(mu), norm,
            norm_thresh, vprho, C.byref(f), ptr(grad), ptr(illu))
    assert rc == 0, rc
    gr = grad.reshape(2, g.nx, g.nz).transpose(0, 2, 1)
    return dict(f=f.value, grad=gr, illum=illu.reshape(g.nx, g.nz).T)


def mute_mask(dt, shots, ot=0.0, v0=1500.0, t0=0.0, v1=10000.0, t1=40.0, tramp=0.5):
    """Mute.f90 MuteParam_compute_mask restatement: one (ntr, nt) mask per shot."""
    L = oracle()
    cs = (sfd_shot * len(shots))(*[s.c() for s in shots])
    opt = np.array([v0, t0, v1, t1, tramp], np.float32)
    n = [len(s.rx) * s.nt for s in shots]
    out = np.zeros(sum(n), np.float32)
    L.orc_f_mute_mask.argtypes = [C.c_int, C.c_void_p, C.c_float, C.c_float, C.c_void_p, C.c_void_p]
    L.orc_f_mute_mask(len(shots), cs, dt, ot, ptr(opt), ptr(out))
    res, o = [], 0
    for s, k in zip(shots, n):
        res.append(out[o:o + k].reshape(len(s.rx), s.nt))
        o += k
    return res


def norm_fct(ntype, rd, thresh, lib="oracle"):
    """fct_compute (Norm.f90:12-79) on one FP32 trace: the oracle's or, with
    lib="ref", the reference's own; returns (f, thresh after the call)."""
    r = np.ascontiguousarray(rd, np.float32)
    th = C.c_float(thresh)
    if lib == "ref":
        fn = ref().ref_fct_compute
    else:
        fn = oracle().orc_f_fct_compute
    fn.restype = C.c_double
    fn.argtypes = [C.c_int, C.c_void_p, C.c_int, C.POINTER(C.c_float)]
    f = fn(ntype, ptr(r), r.size, C.byref(th))
    return f, th.value


def norm_gdt(ntype, rd, thresh, lib="oracle"):
    """gdt_compute (Norm.f90:81-141) on a copy of one FP32 trace."""
    r = np.ascontiguousarray(rd, np.float32).copy()
    if lib == "ref":
        th = C.c_float(thresh)
        fn = ref().ref_gdt_compute
        fn.argtypes = [C.c_int, C.c_void_p, C.c_int, C.POINTER(C.c_float)]
        fn(ntype, ptr(r), r.size, C.byref(th))
    else:
        fn = oracle().orc_f_gdt_compute
        fn.argtypes = [C.c_int, C.c_void_p, C.c_int, C.c_float]
        fn(ntype, ptr(r), r.size, thresh)
    return r


def mksinc(lsinc, shift, prec="f"):
    T = _dt(prec)
    out = np.zeros(lsinc, dtype=T)
    fn = getattr(oracle(), "orc_%s_mksinc" % prec)
    fn.argtypes = [C.c_void_p, C.c_int, (C.c_float if prec == "f" else C.c_double)]
    fn(ptr(out), lsinc, shift)
    return out


def higdon_coeff(nu, prec="f"):
    T = _dt(prec)
    out = np.zeros(8, dtype=T)
    fn = getattr(oracle(), "orc_%s_higdon_coeff" % prec)
    R = C.c_float if prec == "f" else C.c_double
    fn.argtypes = [R, R, R, R, C.c_void_p]
    fn(nu, 1.0, 0.7, 0.5, ptr(out))
    return out


def taper(n, prec="f"):
    T = _dt(prec)
    out = np.zeros(n, dtype=T)
    fn = getattr(oracle(), "orc_%s_compute_taper" % prec)
    fn.argtypes = [C.c_void_p, C.c_int]
    fn(ptr(out), n)
    return out


def triangle2(u, rect_z, rect_x, doz=1, dox=1, prec="f"):
    T = _dt(prec)
    a = np.asfortranarray(np.array(u, dtype=T))
    fn = getattr(oracle(), "orc_%s_triangle2" % prec)
    fn.argtypes = [C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.c_void_p]
    fn(a.shape[0], a.shape[1], rect_z, rect_x, doz, dox, a.ctypes.data_as(C.c_void_p))
    return np.array(a)


# ---- reference (flang build) ----
def ref_mksinc(lsinc, shift):
    out = np.zeros(lsinc, dtype=np.float32)
    fn = ref().ref_mksinc
    fn.argtypes = [C.c_void_p, C.c_int, C.c_float]
    fn(ptr(out), lsinc, shift)
    return out


def ref_higdon_coeff(nu):
    out = np.zeros(8, dtype=np.float32)
    fn = ref().ref_higdon_coeff
    fn.argtypes = [C.c_float, C.c_void_p]
    fn(nu, ptr(out))
    return out


def ref_taper(n):
    out = np.zeros(n, dtype=np.float32)
    fn = ref().ref_taper
    fn.argtypes = [C.c_int, C.c_void_p]
    fn(n, ptr(out))
    return out


def ref_propagate(g, vel_pad, oxw, nxw, stencil, inj, sou_z, sou_x, sou_tr, ext, dat_z, dat_x,
                  nt, tdir, nsnap, elev_fd=None, rho_pad=None):
    """Runs the reference kernels in Propagator.f90 order.  vel_pad: [NZ, NX].
    g's rec/shot/surf types are passed on; elev_fd: elevation per FD column
    nmin2..nmax2 (None: 0); rho_pad: padded density [NZ, NX] (withRho).

    With a density the reference's FD_2D_derivatives_acoustic_forward_grid
    reads a row and a column of its gradient arrays that nothing writes (they
    only reach cells the Higdon bands overwrite, and the halo beyond them);
    glibc's M_PERTURB makes those fresh allocations zero for the call, the
    value the oracle takes."""
    fn = ref().ref_propagate
    fn.argtypes = ([C.c_int] * 3 + [C.c_float] * 5 + [C.c_int] * 2 + [C.c_void_p] + [C.c_int] * 3
                   + [C.c_void_p] * 3 + [C.c_int] * 2 + [C.c_void_p] * 3 + [C.c_int] * 3 + [C.c_void_p]
                   + [C.c_int] * 3 + [C.c_void_p] + [C.c_int] + [C.c_void_p])
    ev = np.zeros(nxw + 2 * g.ntaper, np.float32) if elev_fd is None else \
        np.ascontiguousarray(elev_fd, dtype=np.float32)
    vp = np.asfortranarray(vel_pad, dtype=np.float32)
    sz = np.ascontiguousarray(sou_z, dtype=np.float32)
    sx = np.ascontiguousarray(sou_x, dtype=np.float32)
    st = np.ascontiguousarray(sou_tr, dtype=np.float32).reshape(len(sz), nt) if len(sz) else \
        np.zeros((1, nt), np.float32)
    dz_ = np.ascontiguousarray(dat_z, dtype=np.float32)
    dx_ = np.ascontiguousarray(dat_x, dtype=np.float32)
    dtr = np.zeros((max(len(dz_), 1), nt), dtype=np.float32)
    wave = np.zeros((max(nsnap, 1), nxw, g.nz), dtype=np.float32)
    rp = np.zeros(1, np.float32) if rho_pad is None else np.asfortranarray(rho_pad, dtype=np.float32)
    libc = C.CDLL(None)
    if rho_pad is not None:
        libc.mallopt(-6, 255)   # M_PERTURB: fresh allocations filled with zero bytes
    try:
        fn(g.nz, nxw, g.ntaper, g.dz, g.dx, g.dt, g.oz, oxw, g.lsinc, g.snapi, vp.ctypes.data_as(C.c_void_p),
           stencil, inj, len(sz), ptr(sz), ptr(sx), ptr(st), ext, len(dz_), ptr(dz_), ptr(dx_), ptr(dtr),
           nt, tdir, nsnap, ptr(wave), g.rec_type, g.shot_type, g.surf_type, ptr(ev),
           int(rho_pad is not None), rp.ctypes.data_as(C.c_void_p))
    finally:
        if rho_pad is not None:
            libc.mallopt(-6, 0)
    return dtr[:len(dz_)], wave[:nsnap]


def ref_interpolate_rho(g, rho_pad, nxw):
    """The reference's Interpolate on a padded density: rho2 [NZ, NX]."""
    fn = ref().ref_interpolate_rho
    fn.argtypes = [C.c_int] * 3 + [C.c_void_p] * 2
    rp = np.asfortranarray(rho_pad, dtype=np.float32)
    out = np.zeros_like(rp, order="F")
    fn(g.nz, nxw, g.ntaper, rp.ctypes.data_as(C.c_void_p), out.ctypes.data_as(C.c_void_p))
    return out


def set_density(rho):
    """The oracle's density (nz, nx) (None: constant density)."""
    L = oracle()
    L.orc_set_density.argtypes = [C.c_void_p, C.c_size_t]
    if rho is None:
        L.orc_set_density(None, 0)
        return
    r = colmajor(rho, np.float32)
    L.orc_set_density(ptr(r), r.size)


def half_step(trace, lsinc=7, prec="f"):
    """readsou's half-time-step source shift (withRho)."""
    L = oracle()
    T = _dt(prec)
    t = np.ascontiguousarray(trace, dtype=T).copy()
    fn = getattr(L, "orc_%s_half_step" % prec)
    fn.argtypes = [C.c_void_p, C.c_int, C.c_int]
    fn(ptr(t), len(t), lsinc)
    return t


# ---- L-BFGS oracle ----
class LBFGS:
    """orc_lbfgs state machine (restatement of lbfgs.f LBFGS)."""

    def __init__(self, n, m=5, xmin=-1e30, xmax=1e30, mask=None, boundt=0, freeze_soft=1, xinit=None):
        L = oracle()
        L.orc_lbfgs_new.restype = C.c_void_p
        L.orc_lbfgs.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_void_p, C.c_double, C.c_void_p,
                                C.c_void_p, C.c_double, C.c_double, C.c_void_p, C.POINTER(C.c_int),
                                C.POINTER(C.c_int), C.c_double, C.c_double, C.c_void_p, C.c_int,
                                C.c_int, C.c_void_p]
        L.orc_lbfgs_free.argtypes = [C.c_void_p]
        self.L = L
        self.S = L.orc_lbfgs_new()
        self.n, self.m = n, m
        self.diag = np.zeros(n)
        self.w = np.zeros(n * (2 * m + 1) + 2 * m)
        self.iflag = C.c_int(0)
        self.myinfo = C.c_int(0)
        self.xmin, self.xmax = xmin, xmax
        self.mask = np.ones(n, np.float32) if mask is None else np.asarray(mask, np.float32)
        self.boundt, self.type = boundt, freeze_soft
        self.xinit = np.zeros(n) if xinit is None else np.asarray(xinit, np.float64)

    def step(self, x, f, g, eps=1e-20, xtol=np.finfo(np.float64).eps):
        gg = np.ascontiguousarray(g, np.float64)
        self.L.orc_lbfgs(self.S, self.n, self.m, ptr(x), f, ptr(gg), ptr(self.diag), eps, xtol,
                         ptr(self.w), C.byref(self.iflag), C.byref(self.myinfo), self.xmin, self.xmax,
                         ptr(self.mask), self.boundt, self.type, ptr(self.xinit))
        return self.iflag.value, self.myinfo.value

    def __del__(self):
        try:
            self.L.orc_lbfgs_free(self.S)
        except Exception:
            pass


class LBFGSOC:
    """orc_lbfgsoc state machine (restatement of nlinv_lbfgs.f LBFGSOC; the
    SAVE state stays in memory between calls)."""

    def __init__(self, n, m=5, xmin=(-1e30,), xmax=(1e30,), stp1_opt=1, step_init=999999.0, boundt=2):
        L = oracle()
        L.orc_lbfgs_new.restype = C.c_void_p
        L.orc_lbfgsoc.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_void_p, C.c_double, C.c_void_p,
                                  C.c_void_p, C.c_double, C.c_double, C.c_double, C.c_void_p,
                                  C.POINTER(C.c_int), C.POINTER(C.c_int), C.c_void_p, C.c_void_p,
                                  C.c_int, C.c_int, C.c_int]
        L.orc_lbfgs_free.argtypes = [C.c_void_p]
        self.L = L
        self.S = L.orc_lbfgs_new()
        self.n, self.m = n, m
        self.diag = np.zeros(n)
        self.w = np.zeros(n * (2 * m + 1) + 2 * m)
        self.iflag = C.c_int(0)
        self.myinfo = C.c_int(0)
        self.xmin = np.ascontiguousarray(xmin, np.float64)
        self.xmax = np.ascontiguousarray(xmax, np.float64)
        self.stp1_opt, self.step_init, self.boundt = stp1_opt, step_init, boundt

    def step(self, x, f, g, eps=1e-20, xtol=np.finfo(np.float64).eps):
        gg = np.ascontiguousarray(g, np.float64)
        self.L.orc_lbfgsoc(self.S, self.n, self.m, ptr(x), f, ptr(gg), ptr(self.diag), eps, xtol,
                           self.step_init, ptr(self.w), C.byref(self.iflag), C.byref(self.myinfo),
                           ptr(self.xmin), ptr(self.xmax), len(self.xmin), self.stp1_opt, self.boundt)
        return self.iflag.value, self.myinfo.value

    def __del__(self):
        try:
            self.L.orc_lbfgs_free(self.S)
        except Exception:
            pass


def nlref_available():
    return os.path.exists(NLREF_SO)


def _fstr(s):
    return C.create_string_buffer(s.encode().ljust(1024), 1024)


class RefLBFGSOC:
    """The reference's own LBFGSOC (oracle/_ref/libnlinvref.so), its scalar
    state in the text files LbfgsDat.dat / McsrchDat.dat of `state_dir`
    (read and rewritten in place every call); W and DIAG are the caller's
    arrays, as L_BFGS.f90 keeps them in WorkArray.dat / Diag.dat."""

    def __init__(self, n, state_dir, m=5, xmin=(-1e30,), xmax=(1e30,), stp1_opt=1, step_init=999999.0,
                 boundt=2):
        self.L = C.CDLL(NLREF_SO, mode=os.RTLD_LOCAL)
        self.n, self.m = n, m
        self.diag = np.zeros(n)
        self.w = np.zeros(n * (2 * m + 1) + 2 * m)
        self.iflag = C.c_int(0)
        self.myinfo = C.c_int(0)
        self.xmin = np.ascontiguousarray(xmin, np.float64)
        self.xmax = np.ascontiguousarray(xmax, np.float64)
        self.stp1_opt, self.step_init, self.boundt = stp1_opt, step_init, boundt
        self.lb = _fstr(os.path.join(state_dir, "LbfgsDat.dat"))
        self.mc = _fstr(os.path.join(state_dir, "McsrchDat.dat"))
        self.iprint = (C.c_int * 2)(-1, 0)

    def step(self, x, f, g, eps=1e-20, xtol=np.finfo(np.float64).eps):
        gg = np.ascontiguousarray(g, np.float64)
        I, D = lambda v: C.byref(C.c_int(v)), lambda v: C.byref(C.c_double(v))
        L1024 = C.c_int64(1024)
        self.L.lbfgsoc_(I(self.n), I(self.m), ptr(x), D(f), ptr(gg), I(0), ptr(self.diag), self.iprint, D(eps),
                        D(xtol), D(self.step_init), ptr(self.w), C.byref(self.iflag), C.byref(self.myinfo),
                        self.lb, self.lb, self.mc, self.mc, ptr(self.xmin), ptr(self.xmax), I(len(self.xmin)),
                        I(self.stp1_opt), I(self.boundt), L1024, L1024, L1024, L1024)
        return self.iflag.value, self.myinfo.value


def ref_lbfgs_x(state_dir, x, f, g, m=5, xmin=(-1e30,), xmax=(1e30,), stp1_opt=1, step_init=999999.0, boundt=2):
    """One run of progs/inversion/L_BFGS.f90 (lbfgs_type 1) on the state files
    of `state_dir`, by the reference's LBFGSOC and the program's own Fortran
    file handling (oracle/ref_lbfgs_x.f90); x is updated in place; returns
    (iflag, info)."""
    L = C.CDLL(NLREF_SO, mode=os.RTLD_LOCAL)
    d = os.fsencode(state_dir)
    xmin = np.ascontiguousarray(xmin, np.float64)
    xmax = np.ascontiguousarray(xmax, np.float64)
    gg = np.ascontiguousarray(g, np.float64)
    iflag, info = C.c_int(0), C.c_int(0)
    L.ref_lbfgs_x(C.c_char_p(d), C.c_int(len(d)), C.c_int(x.size), C.c_int(m), ptr(x), C.c_double(f), ptr(gg),
                  C.c_int(xmin.size), ptr(xmin), ptr(xmax), C.c_int(stp1_opt), C.c_double(step_init),
                  C.c_int(boundt), C.byref(iflag), C.byref(info))
    return iflag.value, info.value


class RefLBFGS:
    """The reference's own LBFGS of libs/libinv/lbfgs.f (oracle/_ref/libseisref.so)."""

    def __init__(self, n, m=5, xmin=-1e30, xmax=1e30, mask=None, boundt=0, freeze_soft=1, xinit=None):
        self.L = ref()
        self.n, self.m = n, m
        self.diag = np.zeros(n)
        self.w = np.zeros(n * (2 * m + 1) + 2 * m)
        self.iflag = C.c_int(0)
        self.myinfo = C.c_int(0)
        self.xmin, self.xmax = xmin, xmax
        self.mask = np.ones(n, np.float32) if mask is None else np.asarray(mask, np.float32)
        self.boundt, self.type = boundt, 1 if freeze_soft else 0
        self.xinit = np.zeros(n) if xinit is None else np.asarray(xinit, np.float64)
        self.iprint = (C.c_int * 2)(-1, 0)

    def step(self, x, f, g, eps=1e-20, xtol=np.finfo(np.float64).eps):
        gg = np.ascontiguousarray(g, np.float64)
        I, D = lambda v: C.byref(C.c_int(v)), lambda v: C.byref(C.c_double(v))
        self.L.lbfgs_(I(self.n), I(self.m), ptr(x), D(f), ptr(gg), I(0), ptr(self.diag), self.iprint, D(eps),
                      D(xtol), ptr(self.w), C.byref(self.iflag), C.byref(self.myinfo), D(self.xmin),
                      D(self.xmax), ptr(self.mask), I(self.boundt), I(self.type), ptr(self.xinit))
        return self.iflag.value, self.myinfo.value


FCTGDT = C.CFUNCTYPE(C.c_int, C.c_void_p, C.POINTER(C.c_float), C.POINTER(C.c_float), C.POINTER(C.c_double))


def l_bfgs(fct, x0, niter=10, neval=None, m=5, bound=2, freeze_soft=True, xmin=1500.0, xmax=4500.0,
           vpmask=None, vpinit=None, eval0=0):
    """orc_l_bfgs (restatement of Inversion.f90 l_bfgs); fct(x) -> (f, g).
    Returns (x, dict(f_eval, accepted, evaluations, iterations, iflag))."""
    L = oracle()
    L.orc_l_bfgs.restype = C.c_int
    L.orc_l_bfgs.argtypes = [C.c_int, C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_int, C.c_double, C.c_double,
                             C.c_void_p, C.c_int, C.c_int, C.c_void_p, FCTGDT, C.c_void_p, C.c_void_p,
                             C.c_void_p, C.c_int, C.POINTER(C.c_int), C.POINTER(C.c_int)]
    x = np.ascontiguousarray(x0, np.float32).copy()
    n = x.size
    neval = 3 * niter if neval is None else neval
    mask = np.ones(n, np.float32) if vpmask is None else np.ascontiguousarray(vpmask, np.float32)
    vi = x.astype(np.float64) if vpinit is None else np.ascontiguousarray(vpinit, np.float64)
    fe = np.zeros(neval)
    acc = np.zeros(neval, np.int32)
    ne, ni = C.c_int(), C.c_int()

    def cb(_ctx, xp, gp, fp):
        f, g = fct(np.ctypeslib.as_array(xp, shape=(n,)).copy())
        np.ctypeslib.as_array(gp, shape=(n,))[:] = np.asarray(g, np.float32)
        fp[0] = float(f)
        return 0

    cf = FCTGDT(cb)
    iflag = L.orc_l_bfgs(n, ptr(x), niter, neval, eval0, m, xmin, xmax, ptr(mask), bound, int(bool(freeze_soft)),
                         ptr(vi), cf, None, ptr(fe), ptr(acc), neval, C.byref(ne), C.byref(ni))
    k = ne.value
    return x, dict(f_eval=fe[:k], accepted=acc[:k].astype(bool), evaluations=k, iterations=ni.value, iflag=iflag)


def modeling_adjoint(g, vel, shot, dbar, extract_kind=0, prec="d"):
    """Exact adjoint of modeling(): sbar = F^T dbar (nt samples)."""
    L = oracle()
    T = _dt(prec)
    v = colmajor(vel, T)
    fn = getattr(L, "orc_%s_modeling_adjoint" % prec)
    fn.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_int]
    db = np.ascontiguousarray(np.asarray(dbar, dtype=T).reshape(shot.ntr, shot.nt))
    sb = np.zeros(shot.nt, dtype=T)
    cs = shot.c()
    rc = fn(C.byref(g), ptr(v), C.byref(cs), ptr(db), ptr(sb), extract_kind)
    assert rc == 0
    return sb


def test_step(g, nxw, vel, levels, prec="d"):
    """One interior step of the oracle (stencil + 5-level update, no boundary):
    levels = (u_-1, u0, u1, u2) as padded (NX, NZ) arrays; returns u3 (NX, NZ),
    written on FD cells (zeros elsewhere)."""
    L = oracle()
    T = _dt(prec)
    arrs = [np.ascontiguousarray(a, dtype=T) for a in (vel,) + tuple(levels)]
    u3 = np.zeros_like(arrs[1])
    fn = getattr(L, "orc_%s_test_step" % prec)
    fn.argtypes = [C.c_void_p, C.c_int] + [C.c_void_p] * 6
    rc = fn(C.byref(g), nxw, *[ptr(a) for a in arrs], ptr(u3))
    assert rc == 0
    return u3


# ---- 3D forward modeling ----
from seismic_amd.abi import sfd_grid3d, sfd_shot3d  # noqa: E402


class ORC_B3(C.Structure):
    _fields_ = [("nmin1", C.c_int), ("nmax1", C.c_int), ("nmin2", C.c_int), ("nmax2", C.c_int),
                ("nmin3", C.c_int), ("nmax3", C.c_int), ("NZ", C.c_int), ("NX", C.c_int), ("NY", C.c_int)]


def make_grid3d(nz, nx, ny, dz, dx, dy, dt, ntaper=12, lsinc=7, oz=0.0, ox=0.0, oy=0.0, apx=0.0, apy=0.0, snapi=1):
    g = sfd_grid3d()
    g.snapi = snapi
    g.nz, g.nx, g.ny = nz, nx, ny
    g.oz, g.ox, g.oy, g.dz, g.dx, g.dy, g.dt = oz, ox, oy, dz, dx, dy, dt
    g.ntaper, g.lsinc = ntaper, lsinc
    g.aperture_x, g.aperture_y = apx, apy
    g.fmax = 0.0
    return g


class Shot3D:
    def __init__(self, sz, sx, sy, wavelet, rz, rx, ry):
        self.wavelet = np.ascontiguousarray(wavelet, dtype=np.float32)
        self.rz = np.ascontiguousarray(rz, dtype=np.float32)
        self.rx = np.ascontiguousarray(rx, dtype=np.float32)
        self.ry = np.ascontiguousarray(ry, dtype=np.float32)
        self.sz, self.sx, self.sy = float(sz), float(sx), float(sy)
        self.nt, self.ntr = int(self.wavelet.size), int(self.rz.size)

    def c(self):
        s = sfd_shot3d()
        s.sz, s.sx, s.sy, s.nt, s.ntr = self.sz, self.sx, self.sy, self.nt, self.ntr
        s.wavelet = self.wavelet.ctypes.data_as(_FP)
        s.rz, s.rx, s.ry = (a.ctypes.data_as(_FP) for a in (self.rz, self.rx, self.ry))
        return s


def model3d_buf(vel):
    """(nz, nx, ny) array -> column-major float32 buffer (z fastest)."""
    return np.ascontiguousarray(np.asarray(vel, np.float32).transpose(2, 1, 0))


def window3d(g, shot):
    L = oracle()
    oxw, oyw, nxw, nyw = C.c_float(), C.c_float(), C.c_int(), C.c_int()
    s = shot.c()
    L.orc_window3d(C.byref(g), C.byref(s), C.byref(oxw), C.byref(nxw), C.byref(oyw), C.byref(nyw))
    return oxw.value, nxw.value, oyw.value, nyw.value


def padded_vel3d(g, vel, shot):
    L = oracle()
    oxw, nxw, oyw, nyw = window3d(g, shot)
    b = ORC_B3()
    L.orc_bounds3d(C.byref(b), g.nz, nxw, nyw, g.ntaper)
    out = np.zeros(b.NZ * b.NX * b.NY, np.float32)
    L.orc_window_pad3d.argtypes = [C.c_void_p, C.c_void_p, C.c_float, C.c_int, C.c_float, C.c_int, C.c_void_p,
                                   C.c_void_p]
    L.orc_window_pad3d(C.byref(g), ptr(model3d_buf(vel)), oxw, nxw, oyw, nyw, C.byref(b), ptr(out))
    return out, (oxw, nxw, oyw, nyw), b


def modeling3d(g, vel, shot, extract_kind=0, rho=None):
    """rho (nz, nx, ny) or None: variable density (withRho)."""
    L = oracle()
    tr = np.zeros((shot.ntr, shot.nt), np.float32)
    s = shot.c()
    L.orc_f_modeling3d_rho.argtypes = [C.c_void_p] * 5 + [C.c_int]
    rb = None if rho is None else model3d_buf(rho)
    rc = L.orc_f_modeling3d_rho(C.byref(g), ptr(model3d_buf(vel)), None if rb is None else ptr(rb), C.byref(s),
                                ptr(tr), extract_kind)
    assert rc == 0
    return tr


def rho_half3d(g, rho_pad, b):
    """orc_rho_half3d: 1 / (rho sinc-interpolated half a cell along z, x, y)."""
    L = oracle()
    out = np.zeros_like(rho_pad)
    L.orc_rho_half3d.argtypes = [C.c_void_p] * 3
    L.orc_rho_half3d(C.byref(b), ptr(rho_pad), ptr(out))
    return out


def ref_propagate3d(g, vel_pad, win, shot, ext, rho_pad=None, wavelet=None, sources=None, tdir=1, snapi=1,
                    nsnap=0, stencil=0, inj=1):
    """rho_pad: padded density (withRho); the wavelet is passed as given
    (readsou's half-step shift is the caller's, as in the drivers).
    sources: (z, x, y, traces (n, nt)) injected instead of the shot's source
    (RTM's receivers); tdir -1 runs reverse time; nsnap > 0 also returns u2
    on the window at every snapi-th step (Extraction_wavefield), in the order
    taken: (traces, waves [nsnap, nyw, nxw, nz])."""
    oxw, nxw, oyw, nyw = win
    fn = ref().ref_propagate3d
    fn.argtypes = ([C.c_int] * 4 + [C.c_float] * 7 + [C.c_int] + [C.c_void_p] + [C.c_int] + [C.c_void_p] * 4
                   + [C.c_int] * 2 + [C.c_void_p] * 4 + [C.c_int] * 2 + [C.c_void_p] + [C.c_int] * 3 + [C.c_void_p]
                   + [C.c_int] * 2)
    if sources is None:
        sz = np.array([shot.sz], np.float32)
        sx = np.array([shot.sx], np.float32)
        sy = np.array([shot.sy], np.float32)
        w = shot.wavelet if wavelet is None else wavelet
        st = np.ascontiguousarray(np.asarray(w, np.float32)[None, :])
    else:
        sz, sx, sy = (np.ascontiguousarray(a, np.float32) for a in sources[:3])
        st = np.ascontiguousarray(sources[3], np.float32)
    nsou = len(sz)
    dtr = np.zeros((shot.ntr, shot.nt), np.float32)
    wave = np.zeros((max(nsnap, 1), nyw, nxw, g.nz), np.float32)
    rp = np.zeros(1, np.float32) if rho_pad is None else np.ascontiguousarray(rho_pad, np.float32)
    libc = C.CDLL(None)
    if rho_pad is not None:
        libc.mallopt(-6, 255)   # M_PERTURB: the never-written gradient cells start at zero, as in the oracle
    try:
        fn(g.nz, nxw, nyw, g.ntaper, g.dz, g.dx, g.dy, g.dt, g.oz, oxw, oyw, g.lsinc, ptr(vel_pad),
           nsou, ptr(sz), ptr(sx), ptr(sy), ptr(st), ext, shot.ntr, ptr(shot.rz), ptr(shot.rx), ptr(shot.ry), ptr(dtr),
           shot.nt, 0 if rho_pad is None else 1, ptr(rp), tdir, snapi, nsnap, ptr(wave), stencil, inj)
    finally:
        if rho_pad is not None:
            libc.mallopt(-6, 0)
    return (dtr, wave[:nsnap]) if nsnap > 0 else dtr


def afwi3d(g, vel, shots, dobs, rho=None, mute=None, norm=2, norm_thresh=0.0, vprho=0):
    """orc_f_afwi3d: dict(f, grad [nparam, nz, nx, ny], illum [nz, nx, ny])."""
    L = oracle()
    cs = (sfd_shot3d * len(shots))(*[s.c() for s in shots])
    ob = np.ascontiguousarray(np.concatenate([np.asarray(d, np.float32).ravel() for d in dobs]))
    mu = None if mute is None else np.ascontiguousarray(np.concatenate([np.asarray(m, np.float32).ravel()
                                                                        for m in mute]))
    N = g.nz * g.nx * g.ny
    grad = np.zeros(2 * N, np.float32)
    ill = np.zeros(N, np.float32)
    f = C.c_double()
    rb = None if rho is None else model3d_buf(rho)
    L.orc_f_afwi3d.argtypes = [C.c_void_p] * 3 + [C.c_int] + [C.c_void_p] * 3 + [C.c_int, C.c_float, C.c_int,
                                                                                   C.POINTER(C.c_double)] + [C.c_void_p] * 2
    rc = L.orc_f_afwi3d(C.byref(g), ptr(model3d_buf(vel)), None if rb is None else ptr(rb), len(shots), cs, ptr(ob),
                        None if mu is None else ptr(mu), norm, norm_thresh, vprho, C.byref(f), ptr(grad), ptr(ill))
    assert rc == 0, rc
    t = lambda a: a.reshape(g.ny, g.nx, g.nz).transpose(2, 1, 0)
    return dict(f=f.value, grad=np.stack([t(grad[:N]), t(grad[N:])]), illum=t(ill))


def lsrtm3d(g, vel, shot, traces_in, vel_bwd=None):
    """orc_f_lsrtm3d: dict(image, illum (nz, nx, ny), capture [ntsnap, nyw, nxw, nz])."""
    L = oracle()
    oxw, nxw, oyw, nyw = window3d(g, shot)
    ns = shot.nt // g.snapi
    img = np.zeros((g.ny, g.nx, g.nz), np.float32)
    ill = np.zeros_like(img)
    cp = np.zeros((max(ns, 1), nyw, nxw, g.nz), np.float32)
    s = shot.c()
    L.orc_f_lsrtm3d.argtypes = [C.c_void_p] * 8
    vb = None if vel_bwd is None else model3d_buf(vel_bwd)
    rc = L.orc_f_lsrtm3d(C.byref(g), ptr(model3d_buf(vel)), None if vb is None else ptr(vb), C.byref(s),
                         ptr(np.ascontiguousarray(traces_in, np.float32)), ptr(img), ptr(ill), ptr(cp))
    assert rc == 0, rc
    t = lambda a: a.transpose(2, 1, 0)
    return dict(image=t(img), illum=t(ill), capture=cp[:ns])


def born3d(g, vel, refl, shot, vel_bwd=None):
    """orc_f_born3d: traces (ntr, nt)."""
    L = oracle()
    tr = np.zeros((shot.ntr, shot.nt), np.float32)
    s = shot.c()
    L.orc_f_born3d.argtypes = [C.c_void_p] * 6
    vb = None if vel_bwd is None else model3d_buf(vel_bwd)
    rc = L.orc_f_born3d(C.byref(g), ptr(model3d_buf(vel)), None if vb is None else ptr(vb), ptr(model3d_buf(refl)),
                        C.byref(s), ptr(tr))
    assert rc == 0, rc
    return tr


def rtm3d(g, vel, shot, traces_in, vel_bwd=None, rho=None):
    """orc_f_rtm3d: dict(image, illum (nz, nx, ny), win_image, win_illum
    [nyw, nxw, nz], snaps, capture [ntsnap, nyw, nxw, nz])."""
    L = oracle()
    oxw, nxw, oyw, nyw = window3d(g, shot)
    ns = shot.nt // g.snapi
    img = np.zeros((g.ny, g.nx, g.nz), np.float32)
    ill = np.zeros_like(img)
    wi = np.zeros((nyw, nxw, g.nz), np.float32)
    wl = np.zeros_like(wi)
    sn = np.zeros((max(ns, 1), nyw, nxw, g.nz), np.float32)
    cp = np.zeros_like(sn)
    s = shot.c()
    tr = np.ascontiguousarray(traces_in, np.float32)
    L.orc_f_rtm3d.argtypes = [C.c_void_p] * 12
    vb = None if vel_bwd is None else model3d_buf(vel_bwd)
    rb = None if rho is None else model3d_buf(rho)
    rc = L.orc_f_rtm3d(C.byref(g), ptr(model3d_buf(vel)), None if vb is None else ptr(vb),
                       None if rb is None else ptr(rb), C.byref(s), ptr(tr), ptr(img), ptr(ill), ptr(wi), ptr(wl),
                       ptr(sn), ptr(cp))
    assert rc == 0, rc
    t = lambda a: a.transpose(2, 1, 0)
    return dict(image=t(img), illum=t(ill), win_image=wi, win_illum=wl, snaps=sn[:ns], capture=cp[:ns])

"""Host-side mirror of the reference drivers over the seisfd C-ABI.

Each function keeps the name and argument meaning of the reference routine it
replaces (Fortran drivers under progs/ and libs/libwave), on numpy arrays in
the reference's column-major layout (vel[nz, nx] with z fastest on disk is
``np.asfortranarray`` here; traces are (ntr, nt) row-major = (nt, ntr)
column-major).  All compute runs in libseisfd.so on the GPU; there is no CPU
fallback: a missing library or device raises.
"""
import ctypes as C
import os

import numpy as np

from . import abi
from .abi import sfd_grid, sfd_shot, sfd_fwi_opts

_FP = C.POINTER(C.c_float)


class SeisFDError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__("seisfd error %d: %s" % (code, msg))
        self.code = code


def make_grid(nz, nx, dz, dx, dt, ntaper=20, lsinc=7, snapi=4, oz=0.0, ox=0.0, aperture_x=0.0, fmax=30.0,
              rec_type=0, shot_type=0, surf_type=0):
    """GeneralParam subset (ReadParam.f90:18-36 defaults: ntaper=20, snapi=4,
    lsinc=7, fmax=30 Hz, rec_type = shot_type = surf_type = 0)."""
    g = sfd_grid()
    g.nz, g.nx = int(nz), int(nx)
    g.oz, g.ox, g.dz, g.dx, g.dt = oz, ox, dz, dx, dt
    g.ntaper, g.lsinc, g.snapi = ntaper, lsinc, snapi
    g.aperture_x = aperture_x
    g.rec_type, g.shot_type, g.surf_type = int(rec_type), int(shot_type), int(surf_type)
    g.fmax = fmax
    return g


class Shot:
    """One source with its receiver spread (sourcevec/datavec TraceSpace)."""

    def __init__(self, sz, sx, wavelet, rz, rx):
        self.wavelet = np.ascontiguousarray(wavelet, dtype=np.float32)
        self.rz = np.ascontiguousarray(rz, dtype=np.float32)
        self.rx = np.ascontiguousarray(rx, dtype=np.float32)
        self.sz, self.sx = float(sz), float(sx)
        self.nt = int(self.wavelet.size)
        self.ntr = int(self.rz.size)
        if self.rx.size != self.ntr:
            raise ValueError("rz and rx differ in length")

    def c(self):
        s = sfd_shot()
        s.sz, s.sx, s.nt, s.ntr = self.sz, self.sx, self.nt, self.ntr
        s.wavelet = self.wavelet.ctypes.data_as(_FP)
        s.rz = self.rz.ctypes.data_as(_FP)
        s.rx = self.rx.ctypes.data_as(_FP)
        return s


def _ptr(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


def _f32c(a):
    return np.ascontiguousarray(a, dtype=np.float32)


def _model_buf(v, nz, nx):
    """(nz, nx) array -> contiguous column-major float32 buffer."""
    v = np.asarray(v, dtype=np.float32)
    if v.shape != (nz, nx):
        raise ValueError("model shape %s != (%d, %d)" % (v.shape, nz, nx))
    return np.ascontiguousarray(v.T)


class Engine:
    """One seisfd handle (one GPU stream, its device buffers and graphs)."""

    def __init__(self, grid):
        self.lib = abi.load()
        self.g = grid
        h = C.c_void_p()
        rc = self.lib.sfd_create(C.byref(h), C.byref(grid))
        if rc != 0:
            raise SeisFDError(rc, self.lib.sfd_last_error(None).decode())
        self.h = h

    def close(self):
        if getattr(self, "h", None):
            self.lib.sfd_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _chk(self, rc):
        if rc != 0:
            raise SeisFDError(rc, self.lib.sfd_last_error(self.h).decode())

    def _shots(self, shots):
        arr = (sfd_shot * len(shots))(*[s.c() for s in shots])
        return arr

    def set_model(self, vel, vel_bwd=None):
        self._vel = _model_buf(vel, self.g.nz, self.g.nx)
        self._velb = None if vel_bwd is None else _model_buf(vel_bwd, self.g.nz, self.g.nx)
        self._chk(self.lib.sfd_set_model(self.h, _ptr(self._vel), _ptr(self._velb)))

    def set_density(self, rho=None):
        """Variable density rho(nz, nx) (genpar%withRho); None: constant density."""
        self._rho = None if rho is None else _model_buf(rho, self.g.nz, self.g.nx)
        self._chk(self.lib.sfd_set_density(self.h, _ptr(self._rho)))

    def set_elevation(self, elev=None):
        """Free-surface elevation per model column (ModelSpace_elevation%elev);
        None: 0 everywhere, as the reference's drivers set it."""
        self._elev = None if elev is None else _f32c(elev)
        if self._elev is not None and self._elev.size != self.g.nx:
            raise ValueError("elevation needs nx = %d values" % self.g.nx)
        self._chk(self.lib.sfd_set_elevation(self.h, _ptr(self._elev)))

    def set_tile(self, tz, tx):
        self._chk(self.lib.sfd_set_tile(self.h, tz, tx))

    # Acoustic_modeling_sep.f90 / AMOD_to_memory
    def get_tile(self):
        tz, tx, px, tb = C.c_int(), C.c_int(), C.c_int(), C.c_int()
        buf = C.create_string_buffer(512)
        self._chk(self.lib.sfd_get_tile(self.h, C.byref(tz), C.byref(tx), C.byref(px), C.byref(tb), buf, 512))
        return dict(tile=(tz.value, tx.value), strip=px.value, time_block=tb.value, log=buf.value.decode())

    def forward(self, shots, extract_kind=abi.SFD_EXTRACT_SINC_V2):
        tot = sum(s.nt * s.ntr for s in shots)
        out = np.zeros(tot, dtype=np.float32)
        self._chk(self.lib.sfd_forward(self.h, self._shots(shots), len(shots), _ptr(out), extract_kind))
        return self._split(out, shots)

    @staticmethod
    def _split(flat, shots):
        res, o = [], 0
        for s in shots:
            n = s.nt * s.ntr
            res.append(flat[o:o + n].reshape(s.ntr, s.nt))
            o += n
        return res

    @staticmethod
    def _cat(traces, shots):
        return np.ascontiguousarray(np.concatenate(
            [np.asarray(t, dtype=np.float32).reshape(s.ntr, s.nt).ravel() for t, s in zip(traces, shots)]))

    # RTM_to_memory
    def rtm(self, shots, traces, snap_mode=abi.SFD_SNAP_FP32, image=None, illum=None):
        """traces None: the copy kept in HBM by set_traces."""
        d = None if traces is None else self._cat(traces, shots)
        nz, nx = self.g.nz, self.g.nx
        img = np.zeros(nz * nx, np.float32) if image is None else np.ascontiguousarray(np.asarray(image, np.float32).T).ravel()
        ill = np.zeros(nz * nx, np.float32) if illum is None else np.ascontiguousarray(np.asarray(illum, np.float32).T).ravel()
        self._chk(self.lib.sfd_rtm(self.h, self._shots(shots), len(shots), _ptr(d), _ptr(img), _ptr(ill), snap_mode))
        return img.reshape(nx, nz).T, ill.reshape(nx, nz).T

    def set_traces(self, traces, shots):
        """Keep the shots' traces in HBM for rtm/lsrtm(shots, None); traces None releases them."""
        if traces is None:
            self._chk(self.lib.sfd_set_traces(self.h, None, 0))
            return
        self._resident = self._cat(traces, shots)
        self._chk(self.lib.sfd_set_traces(self.h, _ptr(self._resident), self._resident.size))

    # LSRTM_to_memory
    def lsrtm(self, shots, traces):
        d = None if traces is None else self._cat(traces, shots)
        nz, nx = self.g.nz, self.g.nx
        img = np.zeros(nz * nx, np.float32)
        ill = np.zeros(nz * nx, np.float32)
        self._chk(self.lib.sfd_lsrtm(self.h, self._shots(shots), len(shots), _ptr(d), _ptr(img), _ptr(ill)))
        return img.reshape(nx, nz).T, ill.reshape(nx, nz).T

    # BORN_to_memory
    def born(self, shots, refl):
        m = _model_buf(refl, self.g.nz, self.g.nx)
        tot = sum(s.nt * s.ntr for s in shots)
        out = np.zeros(tot, dtype=np.float32)
        self._chk(self.lib.sfd_born(self.h, self._shots(shots), len(shots), _ptr(m), _ptr(out)))
        return self._split(out, shots)

    def born_adjoint_exact(self, shots, dbar):
        """B^T dbar summed over the shots (B: refl -> traces of born()); returns (nz, nx)."""
        db = self._cat(dbar, shots)
        mb = np.zeros(self.g.nz * self.g.nx, np.float32)
        self._chk(self.lib.sfd_born_adjoint_exact(self.h, self._shots(shots), len(shots), _ptr(db), _ptr(mb)))
        return mb.reshape(self.g.nx, self.g.nz).T.copy()

    def bandpass(self, traces, flo=0.0, fhi=20.0, nplo=14, nphi=14, phase=0):
        """Bandpass.f90 bandpass (defaults of Init_BandPassParam) on a (ntr, nt)
        array of traces at the grid's dt; returns the filtered copy."""
        tr = np.ascontiguousarray(np.atleast_2d(traces), np.float32).copy()
        o = abi.sfd_bandpass_opts(flo, fhi, nplo, nphi, phase)
        self._chk(self.lib.sfd_bandpass(self.h, C.byref(o), tr.shape[1], tr.shape[0], _ptr(tr)))
        return tr

    def mute_mask(self, shots, ot=0.0, v0=1500.0, t0=0.0, v1=10000.0, t1=40.0, tramp=0.5):
        """MuteParam_compute_mask (Mute.f90:50-98; defaults of Init_MuteParam):
        one (ntr, nt) mask per shot."""
        o = abi.sfd_mute_opts(v0, t0, v1, t1, tramp)
        out = np.zeros(sum(s.nt * s.ntr for s in shots), np.float32)
        self._chk(self.lib.sfd_mute_mask(self.h, self._shots(shots), len(shots), float(ot), C.byref(o), _ptr(out)))
        return self._split(out, shots)

    # compute_fct_gdt
    def fwi_gradient(self, shots, dobs, mute=None, vpmask=None, illupow=1.0, rect_z=1, rect_x=1, do_z=0,
                     do_x=0, snap_mode=abi.SFD_SNAP_FP32, norm=abi.SFD_NORM_L2, norm_thresh=0.0, residuals=True):
        """dobs None: the copy kept in HBM by set_traces.  norm / norm_thresh:
        data_nrm_type / data_threshold (SFD_NORM_*).  residuals False: the
        residual gathers stay on the device (None is returned for them), as
        in the evaluations of sfd_fwi_lbfgs."""
        nz, nx = self.g.nz, self.g.nx
        ob = None if dobs is None else self._cat(dobs, shots)
        mu = None if mute is None else self._cat(mute, shots)
        vm = None if vpmask is None else _model_buf(vpmask, nz, nx)
        o = sfd_fwi_opts()
        o.illupow, o.rect_z, o.rect_x, o.do_z, o.do_x, o.snap_mode = illupow, rect_z, rect_x, do_z, do_x, snap_mode
        o.norm, o.norm_thresh = norm, norm_thresh
        f = C.c_double()
        grad = np.zeros(nz * nx, np.float32)
        illu = np.zeros(nz * nx, np.float32)
        res = np.zeros(sum(s.nt * s.ntr for s in shots), np.float32) if residuals else None
        self._chk(self.lib.sfd_fwi_gradient(self.h, self._shots(shots), len(shots), _ptr(ob), _ptr(mu), _ptr(vm),
                                            C.byref(o), C.byref(f), _ptr(grad), _ptr(illu), _ptr(res)))
        return (f.value, grad.reshape(nx, nz).T, illu.reshape(nx, nz).T,
                self._split(res, shots) if residuals else None)

    def fwi_lbfgs(self, shots, dobs, x0, niter=10, neval=None, vpmin=1500.0, vpmax=4500.0, bound=2,
                  freeze_soft=True, m=5, mute=None, vpmask=None, illupow=1.0, rect_z=1, rect_x=1, do_z=0,
                  do_x=0, snap_mode=abi.SFD_SNAP_FP32, precond=None, norm=abi.SFD_NORM_L2, norm_thresh=0.0):
        """FWI by Inversion.f90 l_bfgs with compute_fct_gdt as the evaluation
        (defaults of Inversion_types.f90:48-57).  precond: None for plain
        L-BFGS, or a dict of sfd_precond_opts fields (enable is implied) for
        the learned CNN preconditioner.  Returns (model, report)."""
        nz, nx = self.g.nz, self.g.nx
        ob = self._cat(dobs, shots)
        mu = None if mute is None else self._cat(mute, shots)
        vm = None if vpmask is None else _model_buf(vpmask, nz, nx)
        fo = sfd_fwi_opts()
        fo.illupow, fo.rect_z, fo.rect_x, fo.do_z, fo.do_x, fo.snap_mode = illupow, rect_z, rect_x, do_z, do_x, snap_mode
        fo.norm, fo.norm_thresh = norm, norm_thresh
        lo = lbfgs_opts(niter, neval, m, bound, freeze_soft, vpmin, vpmax)
        pc = precond_opts(**(precond or {})) if precond is not None else None
        x = _model_buf(x0, nz, nx).copy()
        rep, keep = _report(lo.neval)
        self._chk(self.lib.sfd_fwi_lbfgs(self.h, self._shots(shots), len(shots), _ptr(ob), _ptr(mu), _ptr(vm),
                                         C.byref(fo), C.byref(lo), None if pc is None else C.byref(pc), _ptr(x),
                                         C.byref(rep)))
        return x.reshape(nx, nz).T.copy(), _report_dict(rep, keep)

    def afwi_gradient(self, shots, dobs=None, mute=None, snap_mode=abi.SFD_SNAP_FP32, norm=0, norm_thresh=0.0,
                      vprho_param=0):
        """3DAFWI.x's gradient summed over the shots: dict(f, grad [nparam, nz, nx]
        (vp; with a density also rho or Ip), illum [nz, nx]).  dobs None: the
        traces kept by set_traces."""
        o = abi.sfd_afwi_opts()
        o.snap_mode, o.norm, o.norm_thresh, o.vprho_param = snap_mode, norm, norm_thresh, vprho_param
        ob = None if dobs is None else self._cat(dobs, shots)
        mu = None if mute is None else self._cat(mute, shots)
        npar = 2 if getattr(self, "_rho", None) is not None else 1
        f = C.c_double()
        grad = np.zeros(npar * self.g.nz * self.g.nx, np.float32)
        illum = np.zeros(self.g.nz * self.g.nx, np.float32)
        self._chk(self.lib.sfd_afwi_gradient(self.h, self._shots(shots), len(shots), _ptr(ob), _ptr(mu), C.byref(o),
                                             C.byref(f), _ptr(grad), _ptr(illum)))
        return dict(f=f.value, grad=grad.reshape(npar, self.g.nx, self.g.nz).transpose(0, 2, 1),
                    illum=illum.reshape(self.g.nx, self.g.nz).T)

    def adjoint_exact(self, shots, dbar, extract_kind=abi.SFD_EXTRACT_SINC_V2):
        """F^T dbar for each shot (F: wavelet -> traces of forward()); returns a list of nt arrays."""
        db = self._cat(dbar, shots)
        sb = np.zeros(sum(s.nt for s in shots), np.float32)
        self._chk(self.lib.sfd_adjoint_exact(self.h, self._shots(shots), len(shots), _ptr(db), _ptr(sb),
                                             extract_kind))
        out, o = [], 0
        for s in shots:
            out.append(sb[o:o + s.nt].copy())
            o += s.nt
        return out

    def set_option(self, opt, value):
        self._chk(self.lib.sfd_set_option(self.h, opt, int(value)))

    def kernel_stats(self):
        us = (C.c_double * 2)()
        n = (C.c_int * 2)()
        self._chk(self.lib.sfd_kernel_stats(self.h, us, n))
        return dict(fwd_us=us[0], bwd_us=us[1], fwd_n=n[0], bwd_n=n[1])

    def last_waves(self):
        """Shot waves the last call ran in (1: one batch)."""
        w = C.c_int()
        self._chk(self.lib.sfd_last_waves(self.h, C.byref(w)))
        return w.value

    def kernel_totals(self):
        """Per pass (forward time, reverse time) of the last call: event-timed
        kernel microseconds, algorithmic HBM bytes, FD point-updates, steps."""
        us, by, pts = (C.c_double * 2)(), (C.c_double * 2)(), (C.c_double * 2)()
        n = (C.c_int * 2)()
        self._chk(self.lib.sfd_kernel_totals(self.h, us, by, pts, n))
        return dict(us=[us[0], us[1]], bytes=[by[0], by[1]], pts=[pts[0], pts[1]], steps=[n[0], n[1]])

    def last_timing(self):
        ms = (C.c_double * 4)()
        pts = (C.c_double * 1)()
        self._chk(self.lib.sfd_last_timing(self.h, ms, pts))
        return dict(total_ms=ms[0], fwd_ms=ms[1], bwd_ms=ms[2], setup_ms=ms[3], pts=pts[0])


def make_grid3d(nz, nx, ny, dz, dx, dy, dt, ntaper=20, lsinc=7, oz=0.0, ox=0.0, oy=0.0, aperture_x=0.0,
                aperture_y=0.0, fmax=30.0, snapi=1):
    """sfd_grid3d: model (nz, nx, ny) at (oz, ox, oy) with spacing (dz, dx, dy)."""
    g = abi.sfd_grid3d()
    g.nz, g.nx, g.ny = nz, nx, ny
    g.oz, g.ox, g.oy, g.dz, g.dx, g.dy, g.dt = oz, ox, oy, dz, dx, dy, dt
    g.ntaper, g.lsinc = ntaper, lsinc
    g.aperture_x, g.aperture_y, g.fmax = aperture_x, aperture_y, fmax
    g.snapi = snapi
    return g


class Shot3D:
    """One 3D shot: source (sz, sx, sy), wavelet[nt], receivers (rz, rx, ry)[ntr]."""

    def __init__(self, sz, sx, sy, wavelet, rz, rx, ry):
        self.sz, self.sx, self.sy = float(sz), float(sx), float(sy)
        self.wavelet = _f32c(wavelet)
        self.rz, self.rx, self.ry = _f32c(rz), _f32c(rx), _f32c(ry)
        if not (self.rz.size == self.rx.size == self.ry.size):
            raise ValueError("rz, rx and ry differ in length")
        self.nt, self.ntr = int(self.wavelet.size), int(self.rz.size)

    def c(self):
        s = abi.sfd_shot3d()
        s.sz, s.sx, s.sy, s.nt, s.ntr = self.sz, self.sx, self.sy, self.nt, self.ntr
        fp = C.POINTER(C.c_float)
        s.wavelet = self.wavelet.ctypes.data_as(fp)
        s.rz, s.rx, s.ry = (a.ctypes.data_as(fp) for a in (self.rz, self.rx, self.ry))
        return s


class Engine3D:
    """One 3D seisfd handle: constant-density forward modeling (genpar%twoD false)."""

    def __init__(self, grid):
        self.lib = abi.load()
        self.g = grid
        h = C.c_void_p()
        rc = self.lib.sfd_create3d(C.byref(h), C.byref(grid))
        if rc != 0:
            raise SeisFDError(rc, self.lib.sfd_last_error3d(None).decode())
        self.h = h

    def close(self):
        if getattr(self, "h", None):
            self.lib.sfd_destroy3d(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _chk(self, rc):
        if rc != 0:
            raise SeisFDError(rc, self.lib.sfd_last_error3d(self.h).decode())

    def _model(self, m, what):
        v = np.asarray(m, dtype=np.float32)
        if v.shape != (self.g.nz, self.g.nx, self.g.ny):
            raise ValueError("%s shape %s != (%d, %d, %d)" % (what, v.shape, self.g.nz, self.g.nx, self.g.ny))
        return np.ascontiguousarray(v.transpose(2, 1, 0))

    def set_model(self, vel, vel_bwd=None):
        """vel (nz, nx, ny); vel_bwd: RTM's backward-pass model (mod%vel2)."""
        self._vel = self._model(vel, "model")
        self._velb = None if vel_bwd is None else self._model(vel_bwd, "backward model")
        self._chk(self.lib.sfd_set_model3d(self.h, _ptr(self._vel), _ptr(self._velb)))

    def rtm(self, shots, traces, image=None, illum=None, lsrtm=False):
        """3D RTM (LSRTM with lsrtm=True); traces per shot (ntr, nt); returns
        (image, illum) (nz, nx, ny): new arrays holding image / illum (when
        given; they are not modified) plus this call's images, as the library
        accumulates."""
        n = (self.g.nz, self.g.nx, self.g.ny)
        img = np.zeros(n[::-1], np.float32) if image is None else self._model(image, "image").copy()
        ill = np.zeros(n[::-1], np.float32) if illum is None else self._model(illum, "illum").copy()
        arr = (abi.sfd_shot3d * len(shots))(*[s.c() for s in shots])
        tr = Engine._cat(traces, shots)
        fn = self.lib.sfd_lsrtm3d if lsrtm else self.lib.sfd_rtm3d
        self._chk(fn(self.h, arr, len(shots), _ptr(tr), _ptr(img), _ptr(ill)))
        return img.transpose(2, 1, 0), ill.transpose(2, 1, 0)

    def born(self, shots, refl):
        """3D Born modeling of the reflectivity refl (nz, nx, ny): traces per shot (ntr, nt)."""
        arr = (abi.sfd_shot3d * len(shots))(*[s.c() for s in shots])
        m = self._model(refl, "reflectivity")
        out = np.zeros(sum(s.nt * s.ntr for s in shots), np.float32)
        self._chk(self.lib.sfd_born3d(self.h, arr, len(shots), _ptr(m), _ptr(out)))
        return Engine._split(out, shots)

    def set_density(self, rho=None):
        """Variable density rho (nz, nx, ny) (genpar%withRho); None: constant density."""
        if rho is None:
            self._rho = None
        else:
            r = np.asarray(rho, dtype=np.float32)
            if r.shape != (self.g.nz, self.g.nx, self.g.ny):
                raise ValueError("density shape %s != (%d, %d, %d)" % (r.shape, self.g.nz, self.g.nx, self.g.ny))
            self._rho = np.ascontiguousarray(r.transpose(2, 1, 0))
        self._chk(self.lib.sfd_set_density3d(self.h, _ptr(self._rho)))

    def afwi_gradient(self, shots, dobs, mute=None, norm=0, norm_thresh=0.0, vprho_param=0):
        """3DAFWI.x's gradient summed over the shots: dict(f, grad [nparam, nz,
        nx, ny] (vp; with a density also rho or Ip), illum [nz, nx, ny])."""
        o = abi.sfd_afwi_opts()
        o.snap_mode, o.norm, o.norm_thresh, o.vprho_param = abi.SFD_SNAP_FP32, norm, norm_thresh, vprho_param
        arr = (abi.sfd_shot3d * len(shots))(*[s.c() for s in shots])
        ob = Engine._cat(dobs, shots)
        mu = None if mute is None else Engine._cat(mute, shots)
        npar = 2 if getattr(self, "_rho", None) is not None else 1
        N = self.g.nz * self.g.nx * self.g.ny
        grad = np.zeros(npar * N, np.float32)
        illum = np.zeros(N, np.float32)
        f = C.c_double()
        self._chk(self.lib.sfd_afwi_gradient3d(self.h, arr, len(shots), _ptr(ob), _ptr(mu), C.byref(o), C.byref(f),
                                               _ptr(grad), _ptr(illum)))
        t = lambda a: a.reshape(self.g.ny, self.g.nx, self.g.nz).transpose(2, 1, 0)
        return dict(f=f.value, grad=np.stack([t(grad[k * N:(k + 1) * N]) for k in range(npar)]), illum=t(illum))

    def forward(self, shots, extract_kind=abi.SFD_EXTRACT_SINC_V2):
        """Traces per shot, (ntr, nt) each."""
        arr = (abi.sfd_shot3d * len(shots))(*[s.c() for s in shots])
        out = np.zeros(sum(s.nt * s.ntr for s in shots), dtype=np.float32)
        self._chk(self.lib.sfd_forward3d(self.h, arr, len(shots), _ptr(out), extract_kind))
        return Engine._split(out, shots)


def device_info():
    lib = abi.load()
    name = C.create_string_buffer(256)
    cu = C.c_int()
    mem = C.c_size_t()
    rc = lib.sfd_device_info(name, 256, C.byref(cu), C.byref(mem))
    if rc != 0:
        raise SeisFDError(rc, "no HIP device")
    return dict(name=name.value.decode(), cus=cu.value, hbm_bytes=mem.value)


def lbfgs_opts(niter=10, neval=None, m=5, bound=2, freeze_soft=True, vpmin=1500.0, vpmax=4500.0, eval0=0):
    o = abi.sfd_lbfgs_opts()
    o.niter, o.neval, o.m, o.const_type = int(niter), int(3 * niter if neval is None else neval), int(m), int(bound)
    o.freeze_soft, o.eval0, o.xmin, o.xmax = int(bool(freeze_soft)), int(eval0), float(vpmin), float(vpmax)
    return o


def precond_opts(train_after=3, train_steps=300, lr=2e-3, seed=1, max_evals=4, max_fallbacks=3, enable=1,
                 target=0, use_net=1, illum_h0=0, coarsen=0):
    o = abi.sfd_precond_opts()
    o.target, o.use_net, o.illum_h0, o.coarsen = int(target), int(use_net), int(illum_h0), int(coarsen)
    o.enable, o.train_after, o.train_steps, o.lr = int(enable), int(train_after), int(train_steps), float(lr)
    o.seed, o.max_evals, o.max_fallbacks = int(seed), int(max_evals), int(max_fallbacks)
    return o


def _report(cap):
    keep = dict(f=np.zeros(cap), t=np.zeros(cap), a=np.zeros(cap, np.int32))
    r = abi.sfd_lbfgs_report()
    r.hist_cap = cap
    r.f_eval = keep["f"].ctypes.data_as(C.POINTER(C.c_double))
    r.t_eval = keep["t"].ctypes.data_as(C.POINTER(C.c_double))
    r.accepted = keep["a"].ctypes.data_as(C.POINTER(C.c_int))
    return r, keep


def _report_dict(r, keep):
    k = r.nhist
    return dict(iterations=r.iterations, evaluations=r.evaluations, iflag=r.iflag, info=r.info, seconds=r.seconds,
                pc_iterations=r.pc_iterations, pc_fallbacks=r.pc_fallbacks, pc_rejects=r.pc_rejects,
                pc_train_seconds=r.pc_train_seconds, pc_loss=(r.pc_loss_first, r.pc_loss_last),
                f_eval=keep["f"][:k].copy(), t_eval=keep["t"][:k].copy(), accepted=keep["a"][:k].astype(bool))


def lbfgs_run(fct, x0, niter=10, neval=None, m=5, bound=2, freeze_soft=True, xmin=1500.0, xmax=4500.0,
              vpmask=None, vpinit=None, eval0=0):
    """Inversion.f90 l_bfgs on the device L-BFGS: fct(x) -> (f, g) with x, g
    FP32 vectors (g in the reference's sign convention: LBFGS is given -g).
    Returns (x, report)."""
    lib = abi.load()
    x = np.ascontiguousarray(x0, np.float32).copy()
    n = x.size
    lo = lbfgs_opts(niter, neval, m, bound, freeze_soft, xmin, xmax, eval0)
    rep, keep = _report(lo.neval)
    err = []

    def cb(_ctx, xp, gp, fp):
        try:
            xv = np.ctypeslib.as_array(xp, shape=(n,)).copy()
            f, g = fct(xv)
            np.ctypeslib.as_array(gp, shape=(n,))[:] = np.asarray(g, np.float32)
            fp[0] = float(f)
            return 0
        except Exception as ex:   # reported after the run
            err.append(ex)
            return -5

    cfun = abi.SFD_FCTGDT(cb)
    vm = None if vpmask is None else np.ascontiguousarray(vpmask, np.float32)
    vi = None if vpinit is None else np.ascontiguousarray(vpinit, np.float64)
    rc = lib.sfd_lbfgs_run(C.byref(lo), n, _ptr(x), _ptr(vm), _ptr(vi), cfun, None, C.byref(rep))
    if err:
        raise err[0]
    if rc != 0:
        raise SeisFDError(rc, "sfd_lbfgs_run failed")
    return x, _report_dict(rep, keep)


class CNN:
    """The preconditioner's network through sfd_cnn_* (parameters: FP32 vector
    of SFD_CNN_NPARAM, layer l weights W[t][co][ci] then 16 biases)."""

    def __init__(self, nz, nx, nimg_max=4, seed=1):
        self.lib = abi.load()
        self.nz, self.nx, self.nimg_max = int(nz), int(nx), int(nimg_max)
        h = C.c_void_p()
        rc = self.lib.sfd_cnn_create(C.byref(h), self.nz, self.nx, self.nimg_max, int(seed))
        if rc != 0:
            raise SeisFDError(rc, "sfd_cnn_create failed")
        self.h = h

    def _chk(self, rc, what):
        if rc != 0:
            raise SeisFDError(rc, what)

    def _imgs(self, x):
        x = np.asarray(x, np.float32)
        n = x.size // (self.nz * self.nx)
        # (nimg, nz, nx) -> column-major images stacked
        return np.ascontiguousarray(np.stack([np.asfortranarray(im).ravel(order="F") for im in
                                              x.reshape(n, self.nz, self.nx)])), n

    def params(self):
        th = np.zeros(abi.SFD_CNN_NPARAM, np.float32)
        self._chk(self.lib.sfd_cnn_get_params(self.h, _ptr(th)), "sfd_cnn_get_params")
        return th

    def set_params(self, th):
        th = np.ascontiguousarray(th, np.float32)
        self._chk(self.lib.sfd_cnn_set_params(self.h, _ptr(th)), "sfd_cnn_set_params")

    def forward(self, x):
        xs, n = self._imgs(x)
        y = np.zeros_like(xs)
        self._chk(self.lib.sfd_cnn_forward(self.h, n, _ptr(xs), _ptr(y)), "sfd_cnn_forward")
        return np.stack([im.reshape(self.nx, self.nz).T for im in y])

    def loss_grad(self, x, t):
        xs, n = self._imgs(x)
        ts, _ = self._imgs(t)
        loss = C.c_double()
        g = np.zeros(abi.SFD_CNN_NPARAM, np.float32)
        self._chk(self.lib.sfd_cnn_loss_grad(self.h, n, _ptr(xs), _ptr(ts), C.byref(loss), _ptr(g)), "loss_grad")
        return loss.value, g

    def train(self, x, t, steps, lr):
        xs, n = self._imgs(x)
        ts, _ = self._imgs(t)
        l0, l1 = C.c_double(), C.c_double()
        self._chk(self.lib.sfd_cnn_train(self.h, n, _ptr(xs), _ptr(ts), int(steps), float(lr), C.byref(l0),
                                         C.byref(l1)), "sfd_cnn_train")
        return l0.value, l1.value

    def __del__(self):
        try:
            if getattr(self, "h", None):
                self.lib.sfd_cnn_destroy(self.h)
                self.h = None
        except Exception:
            pass


class LBFGSOC:
    """Out-of-core L-BFGS (LBFGSOC of libs/libnlinv/nlinv_lbfgs.f): every
    step(x, f, g) is one sfd_lbfgsoc_step call that reads the optimiser state
    from `state_path` (except the first) and writes it back, so the object
    holds nothing but the options, IFLAG and MYINFO and a run can be resumed
    from the file by a new object (resume=True) in another process."""

    def __init__(self, n, state_path, m=5, xmin=(-1e30,), xmax=(1e30,), stp1_opt=1, step_init=999999.0,
                 clip_type=2, resume=False, info=0):
        self.lib = abi.load()
        self.n = int(n)
        self.path = os.fsencode(state_path)
        self._xmin = np.ascontiguousarray(xmin, np.float64)
        self._xmax = np.ascontiguousarray(xmax, np.float64)
        if self._xmin.shape != self._xmax.shape or self._xmin.ndim != 1:
            raise ValueError("xmin and xmax need one value per parameter block")
        self.opts = abi.sfd_lbfgsoc_opts(int(m), self._xmin.size, self._xmin.ctypes.data, self._xmax.ctypes.data,
                                         int(stp1_opt), float(step_init), int(clip_type))
        self.iflag = C.c_int(1 if resume else 0)
        self.info = C.c_int(int(info))

    def step(self, x, f, g):
        assert x.dtype == np.float64 and x.flags.c_contiguous and x.size == self.n
        gg = np.ascontiguousarray(g, np.float64)
        rc = self.lib.sfd_lbfgsoc_step(self.path, self.path, C.byref(self.opts), self.n, _ptr(x), float(f),
                                       _ptr(gg), C.byref(self.iflag), C.byref(self.info))
        if rc != 0:
            raise SeisFDError(rc, "sfd_lbfgsoc_step failed")
        return self.iflag.value, self.info.value


def lbfgsoc_step_files(state_dir, x, f, g, m=5, xmin=(-1e30,), xmax=(1e30,), stp1_opt=1, step_init=999999.0,
                       clip_type=2):
    """One run of progs/inversion/L_BFGS.f90 (lbfgs_type 1) on the reference's
    state files in `state_dir` (Iflag.dat, Info.dat, Diag.dat, WorkArray.dat,
    LbfgsDat.dat, McsrchDat.dat in; the *Out.dat files out): sfd_lbfgsoc_step_files.
    x (float64, n) is updated in place; returns (iflag, info)."""
    lib = abi.load()
    assert x.dtype == np.float64 and x.flags.c_contiguous
    xmin = np.ascontiguousarray(xmin, np.float64)
    xmax = np.ascontiguousarray(xmax, np.float64)
    if xmin.shape != xmax.shape or xmin.ndim != 1:
        raise ValueError("xmin and xmax need one value per parameter block")
    opts = abi.sfd_lbfgsoc_opts(int(m), xmin.size, xmin.ctypes.data, xmax.ctypes.data, int(stp1_opt),
                                float(step_init), int(clip_type))
    gg = np.ascontiguousarray(g, np.float64)
    iflag, info = C.c_int(0), C.c_int(0)
    rc = lib.sfd_lbfgsoc_step_files(os.fsencode(state_dir), C.byref(opts), x.size, _ptr(x), float(f), _ptr(gg),
                                    C.byref(iflag), C.byref(info))
    if rc != 0:
        raise SeisFDError(rc, "sfd_lbfgsoc_step_files failed")
    return iflag.value, info.value


def lbfgsoc_rotate_files(state_dir):
    """Moves the *Out.dat files of a run onto the names the next run reads
    (the step between two runs of the reference program)."""
    for name in ("Iflag", "Info", "Diag", "WorkArray", "LbfgsDat", "McsrchDat"):
        src = os.path.join(state_dir, name + "Out.dat")
        if os.path.exists(src):
            os.replace(src, os.path.join(state_dir, name + ".dat"))


class LBFGS:
    """Device L-BFGS with the reference's reverse-communication protocol
    (LBFGS(N,M,X,F,G,...) of libs/libinv/lbfgs.f): call step(x, f, g) with
    f, g evaluated at x; x is updated in place; returns (iflag, info)."""

    def __init__(self, n, m=5, xmin=-1e30, xmax=1e30, mask=None, boundt=0, freeze_soft=True, xinit=None):
        self.lib = abi.load()
        self.n = int(n)
        self._mask = None if mask is None else np.ascontiguousarray(mask, np.float32)
        self._xinit = None if xinit is None else np.ascontiguousarray(xinit, np.float64)
        h = C.c_void_p()
        rc = self.lib.sfd_lbfgs_create(C.byref(h), self.n, int(m), float(xmin), float(xmax), _ptr(self._mask),
                                       int(boundt), int(bool(freeze_soft)), _ptr(self._xinit))
        if rc != 0:
            raise SeisFDError(rc, "sfd_lbfgs_create failed")
        self.h = h
        self.iflag = C.c_int(0)
        self.info = C.c_int(0)

    def step(self, x, f, g):
        assert x.dtype == np.float64 and x.flags.c_contiguous and x.size == self.n
        gg = np.ascontiguousarray(g, np.float64)
        rc = self.lib.sfd_lbfgs_step(self.h, _ptr(x), float(f), _ptr(gg), C.byref(self.iflag), C.byref(self.info))
        if rc != 0:
            raise SeisFDError(rc, "sfd_lbfgs_step failed")
        return self.iflag.value, self.info.value

    def __del__(self):
        try:
            if getattr(self, "h", None):
                self.lib.sfd_lbfgs_destroy(self.h)
                self.h = None
        except Exception:
            pass

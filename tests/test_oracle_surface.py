"""Free surface and ghost sources/receivers: the oracle against the
reference's own kernels (Boundary_set_free_surface_grid, the ghost branches
of Injection_source_sinc_xz / _simple_xyz and Extraction_1trace_sinc_xy /
_array_simple_afwi_noomp, Boundary_3d's surf_type rules, Extraction_wavefield's
mask), composed in Propagator.f90 order by oracle/ref_harness.f90.  Skipped
where the reference tree is absent."""
import numpy as np
import pytest

import oracle_lib as O
from models import ricker, smooth_random

pytestmark = pytest.mark.skipif(not O.ref_available(), reason="oracle/_ref not built (no reference tree)")


def rel(a, b):
    return float(np.linalg.norm((a - b).ravel()) / max(np.linalg.norm(b.ravel()), 1e-30))


def nint(x):
    x = np.float32(x)
    return int(np.sign(x) * np.floor(np.abs(x) + np.float32(0.5)))


def elev_fd(g, elev, oxw, nxw):
    """Elevation of FD columns nmin2..nmax2: the nearest window column's model
    column, as the engine and the oracle map them."""
    out = []
    for j in range(1 - g.ntaper, nxw + g.ntaper + 1):
        jw = min(max(j, 1), nxw)
        cxx = np.float32(jw - 1) * np.float32(g.dx) + np.float32(oxw)
        ix = nint((cxx - np.float32(g.ox)) / np.float32(g.dx)) + 1
        ix = min(max(ix, 1), g.nx)
        out.append(0.0 if elev is None else elev[ix - 1])
    return np.array(out, np.float32)


def _case(types, sz=37.0, rz=23.0, seed=1, nz=36, nx=48, nt=240, oz=0.0):
    g = O.make_grid(nz, nx, 10.0, 10.0, 1e-3, ntaper=12, snapi=4, oz=oz)
    g.rec_type, g.shot_type, g.surf_type = types
    vel = smooth_random(nz, nx, seed=seed)
    w = ricker(25.0, nt, 1e-3)
    rx = np.linspace(13.0, 10.0 * (nx - 1) - 17.0, 21).astype(np.float32)
    shot = O.Shot(sz, 241.0, w, np.full_like(rx, rz), rx)
    return g, vel, shot


# (rec_type, shot_type, surf_type)
TYPES = [(0, 0, 1), (0, 0, -1), (0, 1, 0), (1, 0, 0), (1, 1, 0), (1, 0, 1), (0, 1, 1), (1, 1, 1)]


@pytest.mark.parametrize("types", TYPES)
@pytest.mark.parametrize("ext", [0, 1])
def test_forward_with_surface_and_ghosts_matches_reference(types, ext):
    # a surface with shot_type > 0 keeps lsinc-1 rows above the model: the
    # source's mirror footprint must fit there (source row 3)
    sz = 20.0 if types[1] and types[2] > 0 else 37.0
    g, vel, shot = _case(types, sz=sz)
    O.set_elevation(None)
    vpad, oxw, nxw, b = O.padded_vel(g, vel, shot)
    assert b.nmin1 == {1: -6}.get(types[2] > 0 and 1, -(12 + 3 + 2) + 1 if types[1] > 0 else -11)
    ref_tr, ref_w = O.ref_propagate(g, vpad, oxw, nxw, 0, 1, [shot.sz], [shot.sx], shot.wavelet[None, :],
                                    ext, shot.rz, shot.rx, shot.nt, 1, shot.nt // g.snapi)
    tr, snaps = O.modeling(g, vel, shot, extract_kind=ext, snaps=True)
    assert np.array_equal(tr, ref_tr), rel(tr, ref_tr)
    assert np.array_equal(snaps, ref_w), rel(snaps, ref_w)


@pytest.mark.parametrize("elev", ["offgrid", "below_top", "varying"])
def test_free_surface_elevations_match_reference(elev):
    """Off-grid surface (sinc mirror), a surface below model row 1 (zeroed
    snapshot rows), and a per-column elevation mixing both."""
    g, vel, shot = _case((0, 0, 1), sz=57.0, rz=53.0)
    if elev == "offgrid":
        e = np.full(g.nx, 3.0, np.float32)
    elif elev == "below_top":
        e = np.full(g.nx, 20.0, np.float32)
    else:
        x = np.arange(g.nx, dtype=np.float32)
        e = (12.0 + 9.0 * np.sin(x / 5.0)).astype(np.float32)
    O.set_elevation(e)
    try:
        vpad, oxw, nxw, b = O.padded_vel(g, vel, shot)
        ref_tr, ref_w = O.ref_propagate(g, vpad, oxw, nxw, 0, 1, [shot.sz], [shot.sx], shot.wavelet[None, :],
                                        0, shot.rz, shot.rx, shot.nt, 1, shot.nt // g.snapi,
                                        elev_fd=elev_fd(g, e, oxw, nxw))
        tr, snaps = O.modeling(g, vel, shot, extract_kind=0, snaps=True)
    finally:
        O.set_elevation(None)
    assert np.array_equal(tr, ref_tr), rel(tr, ref_tr)
    assert np.array_equal(snaps, ref_w), rel(snaps, ref_w)
    if elev == "below_top":
        assert not np.any(snaps[:, :, :2])   # rows 1..2 lie above the surface at row 3


@pytest.mark.parametrize("types", [(0, 1, 0), (0, 1, 1), (1, 0, 1)])
@pytest.mark.parametrize("inj,stencil", [(1, 1), (3, 1), (2, 0)])
def test_backward_fields_with_surface_and_ghosts_match_reference(types, inj, stencil):
    """Receivers injected in reverse time: sinc ghosts follow shot_type, the
    simple (AGRAD) ghost rec_type; the surface is set every step."""
    g, vel, shot = _case(types, sz=37.0, rz=23.0, seed=2)
    O.set_elevation(None)
    rng = np.random.default_rng(5)
    data = rng.standard_normal((shot.ntr, shot.nt)).astype(np.float32)
    vpad, oxw, nxw, b = O.padded_vel(g, vel, shot)
    _, ref_w = O.ref_propagate(g, vpad, oxw, nxw, stencil, inj, shot.rz, shot.rx, data, -1, [], [],
                               shot.nt, -1, shot.nt // g.snapi)
    if inj in (1, 3) and stencil == 1:
        cap = O.rtm(g, vel, shot, data, lsrtm=(inj == 3), capture=True)["capture"]
    else:
        from test_oracle_vs_ref import _agrad_capture
        cap = _agrad_capture(g, vel, shot, data)
    assert np.array_equal(cap, ref_w), rel(cap, ref_w)


def test_free_surface_is_an_odd_mirror():
    """What the physics fixes: a flat surface at ielev = 1 odd-mirrors the
    wavefield about row 0, so in a homogeneous medium the surface-reflected
    wave at zero offset equals minus the direct wave at the same distance
    (the source's depth + its mirror's: 2 x 11 rows) with the top absorbing."""
    nz, nx, nt = 60, 80, 500
    g = O.make_grid(nz, nx, 10.0, 10.0, 1e-3, ntaper=12, snapi=4)
    vel = np.full((nz, nx), 2000.0, np.float32)
    w = ricker(25.0, nt, 1e-3)
    rx = np.array([400.0, 620.0], np.float32)
    shot = O.Shot(100.0, 400.0, w, np.full_like(rx, 100.0), rx)
    O.set_elevation(None)
    ab = O.modeling(g, vel, shot, extract_kind=1)
    g.surf_type = 1
    fs = O.modeling(g, vel, shot, extract_kind=1)
    refl = (fs - ab)[0]            # zero offset: the surface reflection alone
    direct = ab[1]                 # 220 m away, absorbing top
    win = slice(100, 260)          # around 48 ms + 110 ms
    a, b = refl[win], direct[win]
    corr = float(np.dot(a, b) / np.linalg.norm(a) / np.linalg.norm(b))
    assert corr < -0.99, corr
    assert 0.9 < np.linalg.norm(a) / np.linalg.norm(b) < 1.1

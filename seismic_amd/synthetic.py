"""Seeded synthetic models, wavelets and acquisitions for the benchmark
configurations (the reference ships no data; its recipes build inputs with
SEP tools, e.g. `Wavelet wavelet=ricker2`, test/Makefile:14-31).

Velocities are (nz, nx) float32 arrays in Fortran order (z fastest).
"""
import numpy as np


def ricker(f, nt, dt):
    """r(t) = (1 - 2 pi^2 f^2 tau^2) exp(-pi^2 f^2 tau^2), tau = t - 1.2/f."""
    t = np.arange(nt, dtype=np.float64) * dt
    tau = t - 1.2 / f
    a = (np.pi * f * tau) ** 2
    return ((1 - 2 * a) * np.exp(-a)).astype(np.float32)


def _grid(nz, nx, dz, dx):
    z = np.arange(nz) * dz
    x = np.arange(nx) * dx
    return np.meshgrid(z, x, indexing="ij")


def layered_anomaly(nz, nx, dz, dx, seed=0):
    """Config 1: flat layers 1500/2500/3500 m/s (interfaces at 600 m and
    1300 m) plus a +300 m/s Gaussian anomaly (sigma 150 m) at a seeded
    position."""
    rng = np.random.default_rng(seed)
    Z, X = _grid(nz, nx, dz, dx)
    v = np.where(Z < 600, 1500.0, np.where(Z < 1300, 2500.0, 3500.0))
    xc = rng.uniform(0.25, 0.75) * X.max()
    zc = rng.uniform(0.15, 0.85) * Z.max()
    v = v + 300.0 * np.exp(-((Z - zc) ** 2 + (X - xc) ** 2) / (2 * 150.0 ** 2))
    return np.asfortranarray(v.astype(np.float32))


def smooth_random(nz, nx, seed=0, vmin=1500.0, vmax=2500.0):
    """Smooth, seeded, laterally varying model for parity tests."""
    rng = np.random.default_rng(seed)
    z = np.linspace(0, 1, nz)[:, None]
    x = np.linspace(0, 1, nx)[None, :]
    v = vmin + (vmax - vmin) * (0.5 * z + 0.25 * (1 + np.sin(6 * x + rng.uniform(0, 6))) * z
                                + 0.1 * rng.uniform(0, 1) * np.cos(9 * z * x))
    return np.asfortranarray(v.astype(np.float32))


def faulted_layers(nz, nx, dz, dx, seed=0, nlayers=12, nfaults=3, vmin=1500.0, vmax=4500.0):
    """Configs 2/3: ~12 layers, velocity increasing with depth from vmin to
    vmax, each interface with a random dip in [-15, 15] degrees, cut by
    `nfaults` normal faults (60-degree planes, seeded position and throw).
    The top velocity is capped just below vmax so that vmax*dt/dx stays at
    or below the reference's 0.45 Courant limit in FP32 arithmetic."""
    rng = np.random.default_rng(seed)
    Z, X = _grid(nz, nx, dz, dx)
    zmax, xmax = Z.max(), X.max()
    base = np.sort(rng.uniform(0.05, 0.95, nlayers - 1)) * zmax
    dips = np.deg2rad(rng.uniform(-15, 15, nlayers - 1))
    xc = 0.5 * xmax
    # fault planes: x_f(z) = x0 + z / tan(60deg); hanging wall (x > x_f) drops by throw
    fx0 = np.sort(rng.uniform(0.15, 0.85, nfaults)) * xmax
    throw = rng.uniform(0.02, 0.06, nfaults) * zmax
    shift = np.zeros_like(Z)
    for k in range(nfaults):
        xf = fx0[k] + Z / np.tan(np.deg2rad(60.0))
        shift = shift + np.where(X > xf, throw[k], 0.0)
    zeff = Z - shift
    layer = np.zeros(Z.shape, dtype=np.int32)
    for k in range(nlayers - 1):
        interface = base[k] + np.tan(dips[k]) * (X - xc)
        layer += (zeff > interface).astype(np.int32)
    vel_layer = vmin + (vmax - vmin) * np.arange(nlayers) / (nlayers - 1)
    v = vel_layer[layer]
    v = np.minimum(v, vmax * (1 - 1e-6))
    return np.asfortranarray(v.astype(np.float32))


def gaussian_smooth(v, sigma_cells):
    """Gaussian smoothing (migration / initial models), edges replicated."""
    from scipy.ndimage import gaussian_filter
    return np.asfortranarray(gaussian_filter(np.asarray(v, np.float64), sigma_cells, mode="nearest")
                             .astype(np.float32))


def gradient_anomalies(nz, nx, dz, dx, seed=0, v0=1500.0, v1=3500.0, namp=250.0, nanom=4, sigma=300.0):
    """Config 4: 1500->3500 m/s depth gradient plus `nanom` seeded Gaussian
    anomalies of +/- namp m/s (sigma 300 m)."""
    rng = np.random.default_rng(seed)
    Z, X = _grid(nz, nx, dz, dx)
    v = v0 + (v1 - v0) * Z / Z.max()
    for k in range(nanom):
        zc = rng.uniform(0.25, 0.8) * Z.max()
        xc = rng.uniform(0.15, 0.85) * X.max()
        sgn = 1.0 if k % 2 == 0 else -1.0
        v = v + sgn * namp * np.exp(-((Z - zc) ** 2 + (X - xc) ** 2) / (2 * sigma ** 2))
    return np.asfortranarray(v.astype(np.float32))


# ---------------------------------------------------------------------------
# The benchmark workloads.  Each returns a dict with the grid (nz, nx, dz, dx,
# dt), ntaper, snapi, the true and the starting/migration models, and the
# shots as (sz, sx, wavelet, rz, rx) tuples (positions in m).

def faulted_workload_models(seed=2, nz=751, nx=2301, dx=4.0):
    """The 751 x 2301 faulted-layer model at 4 m and its migration model
    (Gaussian sigma 60 m)."""
    true = faulted_layers(nz, nx, dx, dx, seed=seed)
    return true, gaussian_smooth(true, 60.0 / dx)


def rtm_751x2301_workload(nt=7500, seed=2):
    """Single-shot RTM: 751 x 2301 at 4 m, 20 Hz Ricker at the surface mid-x,
    2301 surface receivers, dt 0.4 ms, snapshots every 4 steps, 40-cell pad."""
    nz, nx, dx, dt = 751, 2301, 4.0, 4e-4
    true, mig = faulted_workload_models(seed, nz, nx, dx)
    rx = (np.arange(nx) * dx).astype(np.float32)
    shot = (0.0, dx * (nx // 2), ricker(20.0, nt, dt), np.zeros_like(rx), rx)
    return dict(grid=(nz, nx, dx, dx, dt), ntaper=40, snapi=4, true=true, mig=mig, shots=[shot])


def batched_96shot_workload(seed=3, nshots=96):
    """96 shots uniformly over the surface of the 751 x 2301 faulted model,
    each with a centred spread of U{200..1200} traces (clipped to the model)
    and nt in U{4000..7500} at dt 0.4 ms, 20 Hz Ricker, snapshots every 4."""
    nz, nx, dx, dt = 751, 2301, 4.0, 4e-4
    true, mig = faulted_workload_models(2, nz, nx, dx)
    rng = np.random.default_rng(seed)
    shots = []
    for sx in np.linspace(0.0, dx * (nx - 1), nshots):
        nr = int(rng.integers(200, 1201))
        nt = int(rng.integers(4000, 7501))
        j0 = int(round(sx / dx)) - nr // 2
        rx = np.unique((np.clip(np.arange(j0, j0 + nr), 0, nx - 1) * dx).astype(np.float32))
        shots.append((0.0, float(sx), ricker(20.0, nt, dt), np.zeros_like(rx), rx))
    return dict(grid=(nz, nx, dx, dx, dt), ntaper=40, snapi=4, true=true, mig=mig, shots=shots)


def fwi_201x601_workload(seed=4, nshots=32, nt=3000):
    """2D FWI: 201 x 601 at 10 m, 1500->3500 m/s gradient + 4 Gaussians of
    +/-250 m/s, initial model Gaussian sigma 200 m; 32 shots of a 5 Hz Ricker
    over the surface, 601 surface receivers, dt 1 ms, nt 3000."""
    nz, nx, dx, dt = 201, 601, 10.0, 1e-3
    true = gradient_anomalies(nz, nx, dx, dx, seed=seed)
    init = gaussian_smooth(true, 200.0 / dx)
    w = ricker(5.0, nt, dt)
    rx = (np.arange(nx) * dx).astype(np.float32)
    shots = [(0.0, float(sx), w, np.zeros_like(rx), rx) for sx in np.linspace(0.0, dx * (nx - 1), nshots)]
    return dict(grid=(nz, nx, dx, dx, dt), ntaper=40, snapi=4, true=true, mig=init, shots=shots)


def salt_body(nz, nx, dz, dx, seed=5, nellipses=6):
    """Mask of an irregular salt body: the union of `nellipses` seeded,
    rotated ellipses centred at 15-45% of the depth and 25-75% of the width,
    semi-axes 0.6-2.0 km vertically and 2-7 km laterally."""
    rng = np.random.default_rng(seed)
    Z, X = _grid(nz, nx, dz, dx)
    zmax, xmax = Z.max(), X.max()
    mask = np.zeros(Z.shape, bool)
    for _ in range(nellipses):
        zc = rng.uniform(0.15, 0.45) * zmax
        xc = rng.uniform(0.25, 0.75) * xmax
        az = rng.uniform(600.0, 2000.0)
        ax = rng.uniform(2000.0, 7000.0)
        th = np.deg2rad(rng.uniform(-25.0, 25.0))
        u = (X - xc) * np.cos(th) + (Z - zc) * np.sin(th)
        w = -(X - xc) * np.sin(th) + (Z - zc) * np.cos(th)
        mask |= (u / ax) ** 2 + (w / az) ** 2 <= 1.0
    return mask


def salt_fwi_workload(seed=5, nshots=64, nt=8000, nrec=1200, nz=1911, nx=5395, dx=12.5):
    """Large 2D FWI: 1911 x 5395 at 12.5 m, a 1500->4000 m/s depth gradient
    plus a seeded irregular 4500 m/s salt body (union of 6 ellipses); the
    starting model is the salt-free model smoothed with sigma 500 m.  64
    shots of a 7 Hz Ricker spread over the surface, 1200 receivers per shot
    (12.5 m apart, centred on the source, slid inside the model near its
    edges), dt 1 ms, nt 8000.  Courant number 4500 * 1e-3 / 12.5 = 0.36."""
    dt = 1e-3
    Z, X = _grid(nz, nx, dx, dx)
    back = 1500.0 + 2500.0 * Z / Z.max()
    true = np.where(salt_body(nz, nx, dx, dx, seed=seed), 4500.0, back)
    true = np.asfortranarray(true.astype(np.float32))
    init = gaussian_smooth(np.asfortranarray(back.astype(np.float32)), 500.0 / dx)
    w = ricker(7.0, nt, dt)
    shots = []
    for sx in np.linspace(0.0, dx * (nx - 1), nshots):
        j0 = min(max(int(round(sx / dx)) - nrec // 2, 0), nx - nrec)
        rx = (np.arange(j0, j0 + nrec) * dx).astype(np.float32)
        shots.append((0.0, float(sx), w, np.zeros_like(rx), rx))
    return dict(grid=(nz, nx, dx, dx, dt), ntaper=40, snapi=10, true=true, mig=init, shots=shots)

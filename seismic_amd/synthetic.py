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

"""Seeded synthetic inputs shared by tests, smoke() and bench.py."""
import numpy as np


def ricker(f, nt, dt):
    """r(t) = (1 - 2 pi^2 f^2 tau^2) exp(-pi^2 f^2 tau^2), tau = t - 1.2/f."""
    t = np.arange(nt, dtype=np.float64) * dt
    tau = t - 1.2 / f
    a = (np.pi * f * tau) ** 2
    return ((1 - 2 * a) * np.exp(-a)).astype(np.float32)


def layered_anomaly(nz, nx, dz, dx, seed=0):
    """Three flat layers 1500/2500/3500 m/s (interfaces 600 m, 1300 m) plus a
    +300 m/s Gaussian anomaly (sigma 150 m) at a seeded position."""
    rng = np.random.default_rng(seed)
    z = np.arange(nz) * dz
    x = np.arange(nx) * dx
    Z, X = np.meshgrid(z, x, indexing="ij")
    v = np.where(Z < 600, 1500.0, np.where(Z < 1300, 2500.0, 3500.0))
    xc = rng.uniform(0.25, 0.75) * x[-1]
    zc = rng.uniform(0.15, 0.85) * z[-1]
    v = v + 300.0 * np.exp(-((Z - zc) ** 2 + (X - xc) ** 2) / (2 * 150.0 ** 2))
    return np.asfortranarray(v.astype(np.float32))


def smooth_random(nz, nx, seed=0, vmin=1500.0, vmax=2500.0):
    rng = np.random.default_rng(seed)
    z = np.linspace(0, 1, nz)[:, None]
    x = np.linspace(0, 1, nx)[None, :]
    v = vmin + (vmax - vmin) * (0.5 * z + 0.25 * (1 + np.sin(6 * x + rng.uniform(0, 6))) * z
                                + 0.1 * rng.uniform(0, 1) * np.cos(9 * z * x))
    return np.asfortranarray(v.astype(np.float32))

"""Shared test inputs: small seeded states on the analytic sounding."""
import numpy as np

from nimbus_amd import lib as L
from oracle import oracle_py as O


def small_case(nx=14, ny=12, nz=10, dx=1000.0, dy=1000.0, dz=500.0, seed=7, noise=0.1, cloudy=False,
               wind_noise=0.0, dtype=np.float32):
    """Initial state with a bubble centred in the small domain, optional random
    wind perturbations and a cloudy/rainy layer for microphysics coverage."""
    g = L.make_grid(nx, ny, nz, dx, dy, dz)
    base = L.base_state(g)
    xc = (nx - 2) * dx / 2.0
    yc = (ny - 2) * dy / 2.0
    prm = L.ic_params(g, bubbles=[(xc, yc, 1400.0, 4.0 * dx, 1400.0, 2.0)], noise_amp=noise, seed=seed)
    st = L.initial_state(g, prm, np.float64)
    r = np.random.default_rng(seed)
    if wind_noise:
        for f in ("u", "v", "w"):
            st[f] = st[f] + wind_noise * r.standard_normal(st[f].shape)
        st["w"][:, :, -1] = 0.0
    if cloudy:
        k1 = max(1, nz // 4)
        k2 = max(k1 + 1, nz // 2)
        st["qv"][:, :, k1:k2] *= 1.3
        st["qc"][:, :, k1:k2] = 1.5e-3 * r.random((nx, ny, k2 - k1))
        st["qr"][:, :, 1:k2] = 2.0e-3 * r.random((nx, ny, k2 - 1))
    out = {k: (v.astype(dtype) if k in L.FIELDS else v) for k, v in st.items()}
    return g, base, out


def oracle_grid(g):
    return O.grid(g.nx, g.ny, g.nz, g.dx, g.dy, g.dz)


def rel_err(a, b, floor):
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    return float(np.max(np.abs(a - b) / np.maximum(np.abs(b), floor)))


def max_abs(a, b):
    return float(np.max(np.abs(np.asarray(a, dtype=np.float64) - np.asarray(b, dtype=np.float64))))

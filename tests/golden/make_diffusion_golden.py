#!/usr/bin/env python3
"""Writes tests/golden/diffusion.npz by running the reference's own explicit
diffusion scheme (src/numerics/diffusion/schemes/explicit/explicit.cpp, built
by `make -C oracle ref` into oracle/_ref/ref_diffusion) on seeded FP32 states:
for each case the inputs (u, v, w, theta, qv, rho, optional per-cell vertical
diffusivity, layer thicknesses), the settings and the reference's u, v, w,
theta and qv tendencies. Run in the development container:

    make -C oracle ref && python tests/golden/make_diffusion_golden.py
"""
import os
import subprocess
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
HARNESS = os.path.join(ROOT, "oracle", "_ref", "ref_diffusion")
NX, NY, NZ, DX, DY = 10, 9, 8, 1000.0, 800.0
FIELDS = ("u", "v", "w", "theta", "qv")
OUT = ("du", "dv", "dw", "dtheta", "dqv")
# name: K_h, K_v, c_grid, periodic x, periodic y, variable K, apply_to (0 scalars, 1 momentum, 2 all), stretched dz
CASES = {
    "cgrid_closed_all": (50.0, 10.0, 1, 0, 0, 0, 2, 0),
    "cgrid_periodic_all_stretched": (80.0, 5.0, 1, 1, 1, 0, 2, 1),
    "cgrid_periodic_x_variable_k": (30.0, 0.0, 1, 1, 0, 1, 2, 1),
    "agrid_closed_all": (40.0, 20.0, 0, 0, 0, 0, 2, 0),
    "cgrid_horizontal_scalars": (120.0, 0.0, 1, 0, 1, 0, 0, 0),
    "cgrid_vertical_momentum_stretched": (0.0, 25.0, 1, 0, 0, 0, 1, 1),
}


def state(seed, stretched):
    rng = np.random.default_rng(seed)
    dz = np.full(NZ, 500.0) if not stretched else 250.0 * 1.3 ** np.arange(NZ)
    z = np.cumsum(dz) - 0.5 * dz
    x = np.arange(NX)[:, None, None] * DX
    y = np.arange(NY)[None, :, None] * DY
    zz = z[None, None, :]
    shape = (NX, NY, NZ)
    u = 8.0 * np.sin(2e-4 * x + 3e-4 * y) + 0.004 * zz + 3.0 * rng.standard_normal(shape)
    v = -4.0 * np.cos(3e-4 * x - 1e-4 * y) + 2.0 * rng.standard_normal(shape)
    w = 2.0 * rng.standard_normal(shape)
    theta = 300.0 + 0.004 * zz + 0.8 * rng.standard_normal(shape)
    qv = np.clip(0.012 * np.exp(-zz / 3000.0) * (1.0 + 0.2 * rng.standard_normal(shape)), 0.0, None)
    rho = 1.2 * np.exp(-zz / 9000.0) * (1.0 + 2e-3 * rng.standard_normal(shape))
    kv = 15.0 * (1.0 + 0.5 * rng.random(shape))
    f = {k: np.ascontiguousarray(np.broadcast_to(a, shape), dtype=np.float32) for k, a in
         (("u", u), ("v", v), ("w", w), ("theta", theta), ("qv", qv), ("rho", rho), ("kv", kv))}
    f["dz"] = np.asarray(dz, np.float32).astype(np.float64)  # the device contract carries dz in FP32
    return f


def main():
    out = {}
    for q, (name, c) in enumerate(CASES.items()):
        kh, kv, cg, px, py, var, apply, stretched = c
        st = state(200 + q, stretched)
        with tempfile.TemporaryDirectory() as d:
            for k in FIELDS + ("rho", "kv"):
                st[k].tofile(os.path.join(d, k + ".f32"))
            st["dz"].tofile(os.path.join(d, "dz.f64"))
            with open(os.path.join(d, "params.txt"), "w") as f:
                f.write(" ".join(repr(x) for x in (NX, NY, NZ, DX, DY, kh, kv, cg, px, py, var, apply)))
            subprocess.run([HARNESS, d], check=True)
            res = {k: np.fromfile(os.path.join(d, k + ".f32"), dtype=np.float32).reshape(NX, NY, NZ) for k in OUT}
        for k, a in st.items():
            out[f"{name}/in/{k}"] = a
        for k, a in res.items():
            out[f"{name}/out/{k}"] = a
        out[f"{name}/settings"] = np.array(c, np.float64)
    out["grid"] = np.array([NX, NY, NZ, DX, DY])
    np.savez_compressed(os.path.join(HERE, "diffusion.npz"), **out)
    print("wrote", os.path.join(HERE, "diffusion.npz"))


if __name__ == "__main__":
    main()
